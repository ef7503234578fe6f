#!/usr/bin/env python3
"""Summarise rocprofv3 output directories into per-kernel averages.

usage: summarize_rocprof.py <dir> [<dir> ...]
       summarize_rocprof.py --pmc-json OUT --copy <copy dir> [...] -- <dir> [...]

Reads *_kernel_stats.csv (kernel trace) and *_counter_collection.csv (--pmc)
files and prints, per kernel name, the average duration and the average value
of every counter per dispatch.

With --pmc-json, also writes the per-launch counters of the step kernels that
bench.py reads: HBM bytes (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, each
scaled by known bytes / counter value of the runtime's copy kernel found in
the --copy directories: perf/copy_kernel.py copies 256 MiB) and
SQ_INSTS_VALU per dispatch, keyed by the step's kernel roles (stage1,
stage1_fused, stage2, diffuse_half1, diffuse_half2)."""
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict

COPY_BYTES = 256 << 20


def short(name):
    name = name.replace("fvm::", "")
    name = name.split("(")[0]
    return name[5:] if name.startswith("void ") else name


def role(k):
    m = re.match(r"k_stage<\w+, (\d+), (\d+), (\d), (\w+)>", k)
    if m:
        return {"0": "stage1", "1": "stage2", "2": "stage1_fused"}[m.group(3)], f"{m.group(1)}x{m.group(2)}"
    m = re.match(r"k_diffuse<\w+, (\d+), (\d+), (\w+)>", k)
    if m:
        return ("diffuse_half2" if m.group(3) == "true" else "diffuse_half1"), f"{m.group(1)}x{m.group(2)}"
    return None, None


def read(dirs):
    stats = {}
    pmc = defaultdict(lambda: defaultdict(list))
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
            for row in csv.DictReader(open(f)):
                stats[short(row["Name"])] = (int(row["Calls"]), float(row["AverageNs"]) / 1e3)
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for row in csv.DictReader(open(f)):
                k = short(row.get("Kernel_Name", row.get("Kernel-Name", "")))
                pmc[k][row["Counter_Name"]].append(float(row["Counter_Value"]))
    return stats, pmc


def avg(v):
    return sum(v) / len(v) if v else float("nan")


def print_tables(stats, pmc):
    if stats:
        print("| kernel | calls | avg us |")
        print("|---|---|---|")
        for k, (n, us) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
            print(f"| `{k}` | {n} | {us:.2f} |")
    if pmc:
        print()
        names = sorted({c for k in pmc for c in pmc[k]})
        print("| kernel | dispatches | " + " | ".join(names) + " |")
        print("|---|---|" + "---|" * len(names))
        for k in sorted(pmc):
            n = max(len(v) for v in pmc[k].values())
            print(f"| `{k}` | {n} | " + " | ".join(f"{avg(pmc[k][c]):.4g}" for c in names) + " |")


def pmc_json(out, pmc, copy_pmc, source):
    copy = [k for k in copy_pmc if "copyBuffer" in k]
    if not copy:
        sys.exit("no copy kernel in the --copy directories")
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        v = avg(copy_pmc[copy[0]].get(c, []))
        scale[c] = COPY_BYTES / 1024.0 / v if v == v and v > 0 else None
    res = {"hbm_bytes_per_launch": {}, "fetch_kib_per_dispatch": {}, "write_kib_per_dispatch": {},
           "valu_insts_per_dispatch": {}, "scale": scale, "tile": None, "source": source,
           "copy_kernel": f"{copy[0]} of 256 MiB: " + ", ".join(
               f"{c} {avg(copy_pmc[copy[0]].get(c, [])):.1f} KiB" for c in ("FETCH_SIZE", "WRITE_SIZE")) + " per dispatch"}
    for k, cs in pmc.items():
        r, tile = role(k)
        if not r:
            continue
        res["tile"] = tile
        f, w = avg(cs.get("FETCH_SIZE", [])), avg(cs.get("WRITE_SIZE", []))
        if f == f:
            res["fetch_kib_per_dispatch"][r] = f
        if w == w:
            res["write_kib_per_dispatch"][r] = w
        if f == f and w == w and scale["FETCH_SIZE"] and scale["WRITE_SIZE"]:
            res["hbm_bytes_per_launch"][r] = int(round((f * scale["FETCH_SIZE"] + w * scale["WRITE_SIZE"]) * 1024))
        v = avg(cs.get("SQ_INSTS_VALU", []))
        if v == v:
            res["valu_insts_per_dispatch"][r] = v
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")


def main(argv):
    if argv and argv[0] == "--pmc-json":
        out = argv[1]
        i = argv.index("--")
        copy_dirs = [a for a in argv[2:i] if a != "--copy"]
        dirs = argv[i + 1:]
        stats, pmc = read(dirs)
        _, copy_pmc = read(copy_dirs)
        print_tables(stats, pmc)
        pmc_json(out, pmc, copy_pmc, "profiles/step_kernels_2048x1024_pmc.md")
        return
    print_tables(*read(argv))


if __name__ == "__main__":
    main(sys.argv[1:])
