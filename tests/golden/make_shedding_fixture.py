#!/usr/bin/env python3
"""Writes tests/golden/shedding_256x128_re200.npz: the FP64 oracle's pressure
force history on the obstacle for a viscous cylinder that sheds vortices.

Case: 256x128 cells, dx = dy = 1, cylinder D = 16 at (64, 64), Mach 0.3,
Re 200, Sutherland viscosity in nondimensional units, no-slip adiabatic wall,
HLLC + van Leer, CFL 0.4 / 0.4 with the CFL dt recomputed every step, seeded
+-1e-3 transverse-velocity noise (seed 2).  30000 steps, (t, Fx, Fy) recorded
every 20 steps.  The GPU test compares mean drag, lift amplitude and
shedding frequency over the last half of the run against this history.

usage: python tests/golden/make_shedding_fixture.py [--steps N] [--threads T]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "fvm-airidynamic-simulator_amd")]

import oracle as O  # noqa: E402
from fvm_amd import params as P  # noqa: E402
from fvm_amd import diagnostics as D  # noqa: E402

NX, NY, DIAM, RE, MACH = 256, 128, 16.0, 200.0, 0.3
EVERY = 20


def case():
    p = P.viscous_cylinder(nx=NX, ny=NY, diameter=DIAM, reynolds=RE, mach=MACH)
    p.dt_mode = P.DT_DEVICE_CFL
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--out", default=os.path.join(HERE, "shedding_256x128_re200.npz"))
    a = ap.parse_args()
    O.set_threads(a.threads)
    p = case()
    o = O.Oracle(p, fp64=True)
    q = p.copy()
    rows = []
    t0 = time.time()
    for k in range(a.steps // EVERY):
        o.step(EVERY, q)
        f = o.forces()
        rows.append((o.time()[0], f[0], f[1]))
        if k % 250 == 0:
            print(f"step {(k + 1) * EVERY}  {time.time() - t0:.0f} s  F = {f[0]:.5f}, {f[1]:.5f}", flush=True)
    h = np.array(rows)
    st = D.shedding(p, h[:, 0], h[:, 1], h[:, 2], window=0.5)
    print(st, flush=True)
    np.savez_compressed(a.out, t=h[:, 0], fx=h[:, 1], fy=h[:, 2], every=EVERY, steps=a.steps,
                        nx=NX, ny=NY, diameter=DIAM, reynolds=RE, mach=MACH,
                        cd_mean=st.cd_mean, cl_mean=st.cl_mean, cl_amplitude=st.cl_amplitude,
                        period=st.period, strouhal=st.strouhal)
    o.close()


if __name__ == "__main__":
    main()
