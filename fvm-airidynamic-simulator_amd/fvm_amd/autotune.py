"""Tile-shape autotuner for the stage / viscous kernels.

Every stage kernel is compiled for a set of tile shapes (cells per workgroup
tile, 256 threads).  The shapes trade halo recomputation (small tiles) against
register pressure and workgroup-level imbalance (large tiles) differently for
each grid size and for inviscid vs viscous cases.  A viscous step also has two
schedules: the first half step as a launch of its own, or fused into stage 1
(recomputed on a wider halo: fewer launches, more arithmetic).  This module
times a step on each shape (and schedule) for a set of workloads, keeps the
fastest per bucket, and writes

* tiles.json        -- the measured table (ms per step for every shape), and
* tile_table.inc    -- the default choice per bucket, compiled into
                       libfvm_amd.so (fvm_set_tile(0, 0) = "auto" uses it).

usage: python -m fvm_amd.autotune [--out DIR] [--quick]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

from . import params as P
from .batch import BatchEngine
from .solver import CFDSolver

TILES = [(64, 4), (64, 6), (64, 7), (64, 8), (64, 12), (64, 16), (32, 2), (32, 3), (32, 4), (32, 7), (32, 8), (32, 16), (128, 4), (128, 8),
         (96, 8)]

# (name, builder, steps): one workload per bucket of the default table
WORKLOADS = [
    ("inviscid_32k", lambda: P.euler_bow_shock(), 400),
    ("inviscid_512k", lambda: P.euler_bow_shock(nx=1024, ny=512, radius=64.0, center=(256.0, 256.0)), 200),
    ("inviscid_2m", lambda: P.euler_bow_shock(nx=2048, ny=1024, radius=64.0, center=(512.0, 512.0)), 200),
    ("inviscid_8m", lambda: P.euler_bow_shock(nx=4096, ny=2048, radius=128.0, center=(1024.0, 1024.0)), 50),
    ("viscous_32k", lambda: P.viscous_cylinder(nx=256, ny=128, diameter=16.0, reynolds=100.0), 400),
    ("viscous_512k", lambda: P.viscous_cylinder(nx=1024, ny=512, diameter=32.0, reynolds=200.0), 200),
    ("viscous_2m", lambda: P.viscous_cylinder(nx=2048, ny=1024, diameter=64.0, reynolds=200.0,
                                               center=(512.0, 512.0)), 200),
    ("viscous_8m", lambda: P.viscous_cylinder(nx=4096, ny=2048, diameter=128.0, reynolds=200.0), 50),
]


def time_plain(p: P.FvmParams, tile, steps: int, warmup: int = 50, schedule: int = P.SCHEDULE_SPLIT) -> float:
    s = CFDSolver(0)
    s.set_tile(*tile)
    s.set_viscous_schedule(schedule)
    q = p.copy()
    q.dt_mode = P.DT_DEVICE_CFL
    s.initialize(q)
    s.prepare_graphs(q)   # no graph capture inside the timed steps
    s.advance(q, warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.advance(q, steps)
    s.synchronize()
    dt = (time.perf_counter() - t0) / steps
    s.close()
    return dt * 1e3


def time_batch(cases, tile, steps: int, schedule: int = P.SCHEDULE_SPLIT) -> float:
    b = BatchEngine(0, mem_budget_bytes=16 << 30, max_cases=len(cases))
    b.set_tile(*tile)
    b.set_viscous_schedule(schedule)
    for p in cases:
        b.add(p)
    b.prepare_graphs()
    b.advance(10)
    b.synchronize()
    t0 = time.perf_counter()
    b.advance(steps)
    b.synchronize()
    dt = (time.perf_counter() - t0) / steps
    b.close()
    return dt * 1e3


def settle_clocks(seconds: float = 1.0):
    """Steps of the 2048x1024 viscous case for `seconds` before any timing:
    the GPU's clocks ramp up over the first few hundred steps after start-up,
    which would otherwise slow the first measurements."""
    p = P.viscous_cylinder(nx=2048, ny=1024, diameter=64.0, reynolds=200.0, center=(512.0, 512.0))
    p.dt_mode = P.DT_DEVICE_CFL
    s = CFDSolver(0)
    s.initialize(p)
    q = p.copy()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.advance(q, 100)
        s.synchronize()
    s.close()


def parse_choice(key: str):
    """'64x7' or '32x4+fused' -> (tile_x, tile_y, fused)."""
    shape, _, sched = key.partition("+")
    tx, ty = (int(x) for x in shape.split("x"))
    return tx, ty, int(sched == "fused")


def bucket_of(cells: int) -> int:
    """log4 bucket of the cell count, clamped to 0..4 (32K, 128K, 512K, 2M, 8M+);
    halves round away from zero, as the library's lookup (std::lround) does."""
    x = math.log(max(cells, 1) / 32768.0) / math.log(4.0)
    r = math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)
    return max(0, min(4, int(r)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc"))
    ap.add_argument("--quick", action="store_true", help="only the 2M-cell workloads and the batch")
    a = ap.parse_args(argv)
    settle_clocks()
    result = {"plain": {}, "batch": {}}
    table = {}
    for name, build, steps in WORKLOADS:
        if a.quick and not name.endswith("2m"):
            continue
        p = build()
        times = {}
        for t in TILES:
            times["%dx%d" % t] = time_plain(p, t, steps)
            if p.enable_viscosity and CFDSolver.fused_schedule_supported(*t):
                times["%dx%d+fused" % t] = time_plain(p, t, steps, schedule=P.SCHEDULE_FUSED)
        best = min(times, key=times.get)
        result["plain"][name] = {"cells": p.nx * p.ny, "viscous": bool(p.enable_viscosity), "ms_per_step": times,
                                 "best": best}
        table[(int(p.enable_viscosity), bucket_of(p.nx * p.ny))] = parse_choice(best)
        print(f"{name}: best {best} {times[best]:.4f} ms/step (64x8: {times['64x8']:.4f})", flush=True)
    cases = P.mixed_batch(512)
    times = {}
    for t in TILES:
        times["%dx%d" % t] = time_batch(cases, t, 40)
        if CFDSolver.fused_schedule_supported(*t):
            times["%dx%d+fused" % t] = time_batch(cases, t, 40, schedule=P.SCHEDULE_FUSED)
    best = min(times, key=times.get)
    result["batch"]["mixed_512"] = {"cells": sum(c.nx * c.ny for c in cases), "ms_per_step": times, "best": best}
    print(f"mixed_512 batch: best {best} {times[best]:.4f} ms/step", flush=True)
    with open(os.path.join(a.out, "tiles.json"), "w") as f:
        json.dump(result, f, indent=1)
    bx, by, bf = parse_choice(best)
    with open(os.path.join(a.out, "tile_table.inc"), "w") as f:
        f.write("// Generated by fvm_amd/autotune.py from measured step times on one MI355X.\n")
        f.write("// entries {viscous, log4(cells / 32768) bucket, tile_x, tile_y, "
                "first viscous half step fused into stage 1}\n")
        f.write("#define FVM_TILE_ENTRIES " + ", ".join(
            f"{{{v}, {bk}, {tx}, {ty}, {fu}}}" for (v, bk), (tx, ty, fu) in sorted(table.items())) + "\n")
        f.write(f"// batched engine default\n#define FVM_BATCH_TILE_X {bx}\n#define FVM_BATCH_TILE_Y {by}\n"
                f"#define FVM_BATCH_FUSED {bf}\n")
    return result


if __name__ == "__main__":
    main()
