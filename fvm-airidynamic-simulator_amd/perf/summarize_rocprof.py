#!/usr/bin/env python3
"""Summarise rocprofv3 output directories into per-kernel averages.

usage: summarize_rocprof.py <dir> [<dir> ...]
Reads *_kernel_stats.csv (kernel trace) and *_counter_collection.csv (--pmc)
files and prints, per kernel name, the average duration and the average value
of every counter per dispatch."""
import csv
import glob
import os
import sys
from collections import defaultdict


def short(name):
    name = name.replace("fvm::", "")
    return name.split("(")[0]


def main(dirs):
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
            vals = [(sum(pmc[k][c]) / len(pmc[k][c])) if pmc[k][c] else float("nan") for c in names]
            print(f"| `{k}` | {n} | " + " | ".join(f"{v:.4g}" for v in vals) + " |")


if __name__ == "__main__":
    main(sys.argv[1:])
