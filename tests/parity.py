"""Per-field parity bars shared by the GPU tests.

FP32 HIP states are compared, field by field, with the FP64 oracle; the IEEE
FP32 oracle run on the same input gives the rounding floor.  Bars: relative L1
<= max(1e-4, 1.5 x floor) and relative Linf <= max(1e-3, 2 x floor).
"""
import numpy as np

import oracle as O

L1_BAR = 1e-4
LINF_BAR = 1e-3
L1_FLOOR_FACTOR = 1.5     # HIP takes v_rcp / v_sqrt (1 ulp), the oracle IEEE division
LINF_FLOOR_FACTOR = 2.0
FIELDS = ("rho", "rho_u", "rho_v", "E", "rho_e")


def rel_errors(g, r):
    out = []
    for c in range(5):
        l1 = np.abs(g[c] - r[c]).sum() / np.abs(r[c]).sum()
        li = np.abs(g[c] - r[c]).max() / np.abs(r[c]).max()
        out.append((l1, li))
    return out


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


RESTART_LINF_BAR = 1e-5   # the FP64 oracle restarted from the HIP state, a few steps
MAX_FLIPPED_CELLS = 16    # cells allowed beyond 1e-4 of the field's scale after a flip


def field_parity(g, r64, r32, label="", restart=None):
    """Per-field bars of the FP32 HIP state g against the FP64 oracle r64,
    with the FP32 oracle r32 on the same input as the rounding floor:
    rel L1 <= max(L1_BAR, 1.5 floor) and rel Linf <= max(LINF_BAR, 2 floor).
    Prints the three numbers of every field; returns the worst L1.

    The reference step has branches that are discontinuous in the state: the
    Harten-type push-out of the HLLC wave speeds (|S| < 0.1 max(cL, cR) sets
    S = -+0.1 max(cL, cR), solver.cu:1146-1150, 1344-1348), the upwind choice of the
    internal-energy flux and the first-order fallback.  Where a rough
    near-wall state sits within rounding of such a threshold, two correct FP32
    runs can take different branches, and the cells next to it then differ by
    ~1e-3 while the rest of the field stays at the floor.  With
    restart = (p, g_prev, n), a field whose Linf exceeds its bar passes when
    the FP64 oracle restarted from the HIP state g_prev, n steps before g,
    reproduces g to RESTART_LINF_BAR (the HIP step is faithful on that very
    input) and at most MAX_FLIPPED_CELLS cells deviate beyond 1e-4 of the
    field's scale."""
    he, fe = rel_errors(g, r64), rel_errors(r32, r64)
    worst = 0.0
    restarted = None
    for c, name in enumerate(FIELDS):
        (l1, li), (fl1, fli) = he[c], fe[c]
        b1, bi = max(L1_BAR, L1_FLOOR_FACTOR * fl1), max(LINF_BAR, LINF_FLOOR_FACTOR * fli)
        print(f"  {label}{name:6s} L1 {l1:.2e} (FP32 floor {fl1:.2e}, bar {b1:.1e})"
              f"  Linf {li:.2e} (floor {fli:.2e}, bar {bi:.1e})", flush=True)
        assert l1 <= b1, (name, l1, fl1)
        if li > bi:
            assert restart is not None, (name, li, fli)
            if restarted is None:
                p, g_prev, n = restart
                restarted = oracle_state_from(p, g_prev, n)
            scale = np.abs(r64[c]).max()
            rli = np.abs(g[c] - restarted[c]).max() / scale
            flipped = int((np.abs(g[c] - r64[c]) > 1e-4 * scale).sum())
            print(f"  {label}{name:6s} Linf over the bar: FP64 oracle restarted from the HIP state "
                  f"{restart[2]} steps earlier gives Linf {rli:.2e}; {flipped} cells beyond 1e-4", flush=True)
            assert rli <= RESTART_LINF_BAR, (name, rli)
            assert flipped <= MAX_FLIPPED_CELLS, (name, flipped)
        worst = max(worst, l1)
    return worst


def oracle_state_from(p, s, nsteps, fp64=True):
    """FP64 oracle state after nsteps from state s (5, ny, nx)."""
    o = O.Oracle(p, fp64=fp64)
    o.set_state(np.asarray(s, dtype=np.float64))
    o.step(nsteps, p.copy())
    out = o.state()
    o.close()
    return out


def oracle_state(p, nsteps, fp64=False):
    """State of the oracle (FP32 by default: the rounding floor) after nsteps."""
    o = O.Oracle(p, fp64=fp64)
    o.step(nsteps, p.copy())
    s = o.state()
    o.close()
    return s
