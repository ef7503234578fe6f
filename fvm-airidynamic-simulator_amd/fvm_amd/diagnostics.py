"""Force and shedding diagnostics over a force history.

The solver reports the pressure force on the obstacle per unit depth
(`CFDSolver.forces()`, `BatchEngine.forces()`): the normal-momentum wall flux
of every face between an updated fluid cell and a solid/ghost cell, as the
step applies it.  That flux always uses the pressure of the face's left/lower
cell (src/solver.cu:1589-1598, 1673-1682), the ghost cell on the body's right
and top sides, the fluid cell on its left and bottom sides, so a symmetric
body at rest in a symmetric flow already carries a small steady lift.  The
helpers below turn a history of (t, Fx, Fy) into drag/lift coefficients and a
shedding frequency.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import params as P


def reference_length(p: P.FvmParams) -> float:
    """Frontal length used for the coefficients: the circle's diameter 2r."""
    return 2.0 * float(p.obstacle_r)


def coefficients(p: P.FvmParams, fx, fy):
    """Cd, Cl = F / (1/2 rho_inf |u_inf|^2 L) per unit depth."""
    q = 0.5 * float(p.rho_inf) * (float(p.u_inf) ** 2 + float(p.v_inf) ** 2) * reference_length(p)
    return np.asarray(fx) / q, np.asarray(fy) / q


@dataclasses.dataclass
class SheddingStats:
    cd_mean: float
    cl_mean: float
    cl_amplitude: float      # half the peak-to-peak of the lift coefficient over whole periods
    period: float            # shedding period in solver time units
    strouhal: float          # L / (period |u_inf|)
    periods: int             # whole lift periods in the window


def shedding(p: P.FvmParams, t, fx, fy, window: float = 0.5) -> SheddingStats:
    """Statistics over the last `window` fraction of a force history sampled
    at times t.  The period is the mean spacing of the upward zero crossings
    of Cl - mean(Cl) (linear interpolation between samples); the means and
    the amplitude are taken over the whole periods between the first and the
    last crossing, so a partial period does not bias them."""
    t = np.asarray(t, dtype=np.float64)
    cd, cl = coefficients(p, fx, fy)
    n0 = int(len(t) * (1.0 - window))
    t, cd, cl = t[n0:], cd[n0:], cl[n0:]
    x = cl - cl.mean()
    up = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    if len(up) < 2:
        return SheddingStats(float(cd.mean()), float(cl.mean()), float(0.5 * (cl.max() - cl.min())),
                             float("nan"), float("nan"), 0)
    tc = t[up] - x[up] * (t[up + 1] - t[up]) / (x[up + 1] - x[up])
    period = float(np.mean(np.diff(tc)))
    a, b = up[0], up[-1] + 1
    u = float(np.hypot(p.u_inf, p.v_inf))
    return SheddingStats(cd_mean=float(cd[a:b].mean()), cl_mean=float(cl[a:b].mean()),
                         cl_amplitude=float(0.5 * (cl[a:b].max() - cl[a:b].min())),
                         period=period, strouhal=reference_length(p) / (period * u),
                         periods=len(up) - 1)


def force_history(solver, p: P.FvmParams, nsteps: int, every: int) -> np.ndarray:
    """Advance `solver` nsteps in chunks of `every`, recording (t, Fx, Fy)
    after each chunk.  p.dt_mode decides host or device dt."""
    q = p.copy()
    rows = []
    for _ in range(nsteps // every):
        solver.advance(q, every)
        f = solver.forces()
        rows.append((solver.time()[0], f[0], f[1]))
    return np.array(rows)
