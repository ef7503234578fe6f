"""Precision variants on the GPU (the 256x128 Mach-3 bow-shock configuration:
AUSM+ and minmod, CFL 0.5 with the CFL dt recomputed on the device every step,
density noise), against the FP64 CPU oracle at 200 and 2000 steps.

Variants: FP32 (the reference's arithmetic); FP32 with the compensated (Kahan)
SSP blend and the CFL maxima and dt in FP64; FP64 throughout.  Per field:
relative L1 = sum|a - b| / sum|b| and Linf = max|a - b| / max|b| against the
FP64 oracle; conservation drift = (total - FP64 GPU total) / FP64 GPU total of
mass and energy (FP64 sums on the device).  The open domain has no exact
conservation identity (the boundary pass overwrites edge cells every stage),
so a variant's drift is measured against the FP64 run of the same scheme.

Bars at 200 steps: FP32 and compensated per field as parity.field_parity
(1e-4 L1 or the FP32 floor); the compensated variant no worse than FP32
(1.5x + 1e-7 in L1); mass and energy drift below 1e-5; FP64 within 1e-10 of
the FP64 oracle (the FP64 path takes its dt from FP64 maxima as the FP64
oracle does, so device CFL runs agree to round-off) and its clock within
1e-12.

At 2000 steps the unsteady wake behind the cylinder has amplified round-off
of any size: the FP64 GPU run, 1e-12 from the FP64 oracle at step 200, is
~5e-5 (L1) from it at step 2000, as far as the FP32 runs are.  So the FP64
path is judged per step there: the FP64 oracle restarted from the FP64 GPU
state two steps before the end reproduces it to 1e-11.  The FP32 variants
keep the floor bars of field_parity (the FP32 oracle decorrelates the same
way), and the drift bar is 1e-4.
"""
import numpy as np
import pytest

import fvm_amd
import oracle as O
from fvm_amd import params as P
from parity import FIELDS, announce, field_parity, rel_errors

pytestmark = pytest.mark.gpu

VARIANTS = (("FP32", P.PRECISION_FP32), ("FP32 + Kahan, FP64 dt", P.PRECISION_FP32_KAHAN), ("FP64", P.PRECISION_FP64))


def config():
    p = P.euler_bow_shock(flux=P.FLUX_AUSM_PLUS, limiter=P.LIMITER_MINMOD)
    p.dt_mode = P.DT_DEVICE_CFL
    return p


@pytest.mark.timeout(600)
@pytest.mark.parametrize("steps", [200, 2000])
def test_precision_variants(steps):
    p = config()
    chaotic = steps > 200
    announce(f"precision variants, 256x128 Mach 3 AUSM+ minmod, device CFL, {steps} steps")
    ref = O.Oracle(p, fp64=True)
    ref.step(steps, p.copy())
    r64, t64 = ref.state(), ref.time()[0]
    ref.close()
    o32 = O.Oracle(p, fp64=False)
    o32.step(steps, p.copy())
    r32 = o32.state()
    o32.close()
    runs = {}
    s = fvm_amd.CFDSolver(0)
    for name, prec in VARIANTS:
        q = p.copy()
        q.precision = prec
        s.initialize(q)
        s.advance(q, steps - 2)
        prev = s.state64()
        s.advance(q, 2)
        runs[name] = (s.state64(), s.conservation_sums(), s.time()[0], prev)
    s.close()
    c64 = runs["FP64"][1]
    print("  variant                  " + " ".join(f"{f + ' L1':>10s}" for f in FIELDS)
          + " " + " ".join(f"{f + ' Linf':>11s}" for f in FIELDS) + "  mass drift  energy drift", flush=True)
    errs = {}
    for name, (st, cons, t, _) in list(runs.items()) + [("FP32 oracle (CPU)", (r32, None, None, None))]:
        e = rel_errors(st, r64)
        errs[name] = e
        drift = "" if cons is None else f"  {(cons[0] - c64[0]) / c64[0]:+.2e}   {(cons[3] - c64[3]) / c64[3]:+.2e}"
        print(f"  {name:24s} " + " ".join(f"{x[0]:10.2e}" for x in e) + " "
              + " ".join(f"{x[1]:11.2e}" for x in e) + drift, flush=True)
    drift_bar = 1e-4 if chaotic else 1e-5
    for name in ("FP32", "FP32 + Kahan, FP64 dt"):
        field_parity(runs[name][0], r64, r32, f"{name}: ", restart=(p, runs[name][3], 2))
        cons = runs[name][1]
        assert abs(cons[0] - c64[0]) / c64[0] < drift_bar
        assert abs(cons[3] - c64[3]) / c64[3] < drift_bar
        assert runs[name][2] == pytest.approx(t64, rel=1e-5 if not chaotic else 1e-3)
    g64, _, tg64, prev64 = runs["FP64"]
    if not chaotic:
        for (a, _), (b, _) in zip(errs["FP32"], errs["FP32 + Kahan, FP64 dt"]):
            assert b <= 1.5 * a + 1e-7
        for l1, _ in errs["FP64"]:
            assert l1 < 1e-10
        assert tg64 == pytest.approx(t64, rel=1e-12)
    else:
        q = p.copy()
        q.precision = P.PRECISION_FP64
        o = O.Oracle(q, fp64=True)
        o.set_state(prev64)
        o.step(2, q.copy())
        step_err = max(np.abs(g64[c] - o.state()[c]).max() / np.abs(r64[c]).max() for c in range(5))
        o.close()
        print(f"  FP64: oracle restarted from the FP64 GPU state at step {steps - 2}: Linf {step_err:.2e} "
              f"after 2 steps", flush=True)
        assert step_err < 1e-11
        for l1, _ in errs["FP64"]:
            assert l1 < 1e-3


def test_fp64_device_cfl_matches_oracle_dt():
    """The FP64 path's CFL dt comes from FP64 maxima and the FP64 formula, as
    in the FP64 oracle: the first computeStableTimeStep and the clock after 50
    device-CFL steps agree to round-off (the FP32 dt formula would leave
    ~1e-7)."""
    p = P.viscous_cylinder(nx=128, ny=64, diameter=12.0, reynolds=100.0, mach=0.5)
    p.dt_mode = P.DT_DEVICE_CFL
    p.precision = P.PRECISION_FP64
    announce("FP64 device CFL dt vs FP64 oracle, viscous 128x64, 50 steps")
    o = O.Oracle(p, fp64=True)
    dt0 = o.stable_dt(p)
    with fvm_amd.CFDSolver(0) as s:
        s.initialize(p)
        assert s.computeStableTimeStep(p) == pytest.approx(dt0, rel=1e-7)   # float return value
        s.advance(p.copy(), 50)
        t, n = s.time()
        o.step(50, p.copy())
        assert n == 50
        assert t == pytest.approx(o.time()[0], rel=1e-13)
        for l1, _ in rel_errors(s.state64(), o.state()):
            assert l1 < 1e-10
    o.close()



def test_kahan_rejects_viscous_cases():
    """The compensated variant is the precision variant of the inviscid
    configuration: a viscous case is refused at initialize, and a step call
    that switches viscosity on is refused before any work is queued.  The
    state is then unchanged, and the run continues bit for bit as one that
    never made the call."""
    p = P.euler_bow_shock(nx=128, ny=64, radius=8.0, center=(32.0, 32.0))
    p.dt = 0.1
    p.precision = P.PRECISION_FP32_KAHAN
    announce("compensated variant: viscous cases rejected, state unchanged")
    with fvm_amd.CFDSolver(0) as s:
        q = p.copy()
        q.enable_viscosity = 1
        with pytest.raises(fvm_amd.FvmError, match="inviscid"):
            s.initialize(q)
        s.initialize(p)
        a = p.copy()
        s.advance(a, 10)
        before = s.state()
        with pytest.raises(fvm_amd.FvmError, match="inviscid"):
            s.advance(q, 5)
        assert np.array_equal(s.state(), before)
        assert s.time()[1] == 10
        s.advance(a, 10)
        after = s.state()
        s.initialize(p)
        b = p.copy()
        s.advance(b, 20)
        assert np.array_equal(s.state(), after)
