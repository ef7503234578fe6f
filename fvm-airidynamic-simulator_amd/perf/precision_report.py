#!/usr/bin/env python3
"""Precision variants of the FP32 path against the FP64 oracle.

Case: inviscid Mach-3 flow past a circle (256x128, r = 16 at (64, 64),
HLLC + minmod, +-1e-3 density noise, fixed dt = 0.1).  For FP32, FP32 with a
compensated SSP blend (Kahan) and FP64 on the GPU, and the FP32 oracle, prints
the per-field relative L1 (sum|a-b| / sum|b|) and Linf (max|a-b| / max|b|)
against the FP64 oracle after 200 and 2000 steps, and the drift of the total
mass and energy relative to the FP64 GPU run, (M - M_fp64) / M_fp64.  The open
domain has no exact conservation identity (edge cells are overwritten by the
boundary conditions every stage), so the drift of a variant is measured
against the FP64 run of the same scheme, which isolates rounding.

usage: python perf/precision_report.py [--steps 200 2000] [--out FILE.md]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
REPO = os.path.dirname(PKG)
sys.path[:0] = [PKG, os.path.join(REPO, "oracle")]

import fvm_amd  # noqa: E402
import oracle as O  # noqa: E402
from fvm_amd import params as P  # noqa: E402

FIELDS = ["rho", "rho_u", "rho_v", "E", "rho_e"]


def case():
    p = P.euler_bow_shock()
    p.dt = 0.1
    return p


def gpu_run(p, precision, steps):
    s = fvm_amd.CFDSolver(0)
    q = p.copy()
    q.precision = precision
    s.initialize(q)
    s.advance(q, steps)
    st = s.state64()
    cons = s.conservation_sums()
    s.close()
    return st, cons


def errors(a, b):
    return [(np.abs(a[c] - b[c]).sum() / np.abs(b[c]).sum(), np.abs(a[c] - b[c]).max() / np.abs(b[c]).max())
            for c in range(5)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, nargs="+", default=[200, 2000])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    p = case()
    lines = ["# Precision variants vs the FP64 oracle", "",
             "Inviscid Mach 3, 256x128, circle r=16, HLLC + minmod, dt = 0.1, density noise seed 1.", ""]
    for n in a.steps:
        ref = O.Oracle(p, fp64=True)
        ref.step(n, p.copy())
        r = ref.state()
        ref.close()
        o32 = O.Oracle(p, fp64=False)
        o32.step(n, p.copy())
        s_o32 = o32.state()
        c_o32 = o32.conservation_sums()
        o32.close()
        runs = {"FP32 (HIP)": gpu_run(p, P.PRECISION_FP32, n),
                "FP32 + Kahan blend (HIP)": gpu_run(p, P.PRECISION_FP32_KAHAN, n),
                "FP64 (HIP)": gpu_run(p, P.PRECISION_FP64, n),
                "FP32 oracle (CPU)": (s_o32, c_o32)}
        c64 = runs["FP64 (HIP)"][1]
        lines += [f"## {n} steps", "",
                  "| variant | " + " | ".join(f"{f} L1" for f in FIELDS) + " | " + " | ".join(f"{f} Linf" for f in FIELDS)
                  + " | mass drift | energy drift |",
                  "|---|" + "---|" * 12]
        for name, (st, cons) in runs.items():
            e = errors(st, r)
            md = (cons[0] - c64[0]) / c64[0]
            ed = (cons[3] - c64[3]) / c64[3]
            lines.append(f"| {name} | " + " | ".join(f"{x[0]:.2e}" for x in e) + " | "
                         + " | ".join(f"{x[1]:.2e}" for x in e) + f" | {md:.2e} | {ed:.2e} |")
        lines.append("")
    text = "\n".join(lines)
    print(text, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
