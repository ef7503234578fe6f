#!/usr/bin/env python3
"""Long run of the full-size viscous cylinder on the GPU alone.

Case: 2048x1024 cells, cylinder of diameter 64 at (512, 512), Re 200,
Mach 0.3, HLLC + van Leer, SSP-RK2 with Strang viscous half steps, CFL 0.4
dt on the device, +-1e-3 transverse-velocity noise.  Advances `--steps`
steps, records the pressure force on the cylinder every `--every` steps and
prints one JSON line with the drag / lift statistics and the Strouhal number
over the last half of the run (fvm_amd.diagnostics.shedding), with the wall
time.  The FP64 oracle cannot follow this case for tens of thousands of
steps; the long-run comparison against it is made at 256x128 (GPU tests).
--precision fp64 runs the GPU's FP64 path (which matches the FP64 oracle to
round-off) as the full-size reference for the FP32 run.

usage: python perf/shedding_full.py [--steps 60000] [--every 50] [--precision fp32|fp64] [--out FILE.npz]
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import fvm_amd  # noqa: E402
from fvm_amd import diagnostics as D  # noqa: E402
from fvm_amd import params as P  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--out", default=None, help="optional .npz with the force history (t, Fx, Fy)")
    a = ap.parse_args()
    p = P.viscous_cylinder(nx=2048, ny=1024, diameter=64.0, reynolds=200.0, mach=0.3, center=(512.0, 512.0))
    p.dt_mode = P.DT_DEVICE_CFL
    if a.precision == "fp64":
        p.precision = P.PRECISION_FP64
    s = fvm_amd.CFDSolver(0)
    s.initialize(p)
    t0 = time.perf_counter()
    h = D.force_history(s, p, a.steps, a.every)
    wall = time.perf_counter() - t0
    st = D.shedding(p, h[:, 0], h[:, 1], h[:, 2], window=0.5)
    if a.out:
        np.savez(a.out, t=h[:, 0], fx=h[:, 1], fy=h[:, 2])
    out = {"case": "viscous cylinder 2048x1024, D 64, Re 200, Mach 0.3", "steps": a.steps, "every": a.every,
           "precision": a.precision, "sim_time": float(h[-1, 0]), "wall_s": wall, "window": "last half"}
    out.update(dataclasses.asdict(st))
    print(json.dumps(out))
    s.close()


if __name__ == "__main__":
    main()
