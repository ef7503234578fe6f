"""The HIP path, through the C ABI, against the closed forms of closed_forms.py:
conservation, characteristic boundaries, flat-wall reflection, Sutherland's
law and the non-finite fallback.  The same checks pin the oracle
(test_oracle_closed_forms.py); here they hold for the GPU's FP32, Kahan and
FP64 arithmetic, with bars set by each one's rounding."""
import numpy as np
import pytest

import closed_forms as C
import fvm_amd
from fvm_amd import params as P

pytestmark = pytest.mark.gpu

PRECISIONS = {"fp32": P.PRECISION_FP32, "kahan": P.PRECISION_FP32_KAHAN, "fp64": P.PRECISION_FP64}


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def solver():
    s = fvm_amd.CFDSolver(0)
    yield s
    s.close()


def start(solver, p, s0, precision):
    p.precision = PRECISIONS[precision]
    solver.initialize(p)
    solver.set_state(s0)


@pytest.mark.parametrize("viscous,precision", [(False, "fp32"), (False, "kahan"), (False, "fp64"),
                                               (True, "fp32"), (True, "fp64")])   # Kahan is inviscid only
def test_conservation_telescopes(solver, viscous, precision):
    # The state is rounded to float32, so the freestream it holds differs from
    # the one the boundary pass derives from the parameters by ~1e-8: the edge
    # cells move by that much, which bounds the FP64 path's drift.
    p, s0 = C.conservation_case(viscous)
    announce(f"conservation viscous={viscous} {precision} 96x64 {C.CONSERVATION_STEPS} steps")
    start(solver, p, s0, precision)
    solver.advance(p.copy(), C.CONSERVATION_STEPS)
    s1 = solver.state64()
    drift = C.conservation_drift(s0, s1)
    print(f"  drift {drift:.1e}", flush=True)
    assert C.bump_moved(s0, s1) > 1e-3
    assert drift < (1e-8 if precision == "fp64" else 1e-6)
    sums = solver.conservation_sums()
    assert np.allclose(sums, s1[:4].sum(axis=(1, 2)), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("side", C.OUTFLOW_SIDES)
def test_supersonic_outflow_edge_unchanged(solver, side, precision):
    for supersonic in (True, False):
        p, s0 = C.outflow_case(side, supersonic)
        announce(f"outflow {side} supersonic={supersonic} {precision} 64x48 1 step")
        start(solver, p, s0, precision)
        solver.advance(p.copy(), 1)
        change = C.outflow_change(p, side, s0, solver.state64())
        print(f"  edge change {change:.1e}", flush=True)
        if supersonic:
            # FP32 takes p / rho^gamma through exp and log
            assert change < (1e-12 if precision == "fp64" else 2e-6), change
        else:
            assert change > 1e-3, change


@pytest.mark.parametrize("precision", ["fp32", "kahan", "fp64"])
@pytest.mark.parametrize("kind", list(C.CHARACTERISTIC_KINDS))
@pytest.mark.parametrize("side", C.OUTFLOW_SIDES)
def test_characteristic_edge_closed_form(solver, side, kind, precision):
    """After one step the edge cells equal B(U0/2 + B(U0)/2), B the textbook
    characteristic boundary (outflow at Mach ~0.5 and ~3, inflow at ~0.5).
    FP32 bar: the boundary takes rho from (c^2 / (gamma s))^(1/(gamma-1)), which
    amplifies the FP32 rounding of c and s about fivefold, twice over."""
    supersonic, inward = C.CHARACTERISTIC_KINDS[kind]
    p, s0 = C.outflow_case(side, supersonic, inward=inward)
    announce(f"characteristic edge {side} {kind} {precision} 64x48 1 step")
    start(solver, p, s0, precision)
    solver.advance(p.copy(), 1)
    err = C.outflow_error(p, side, s0, solver.state64())
    print(f"  edge error {err:.1e}", flush=True)
    assert err < (1e-12 if precision == "fp64" else 2e-5), err


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("mode", C.FLAT_WALL_MODES)
def test_flat_wall_ghost_reflection(solver, mode, precision):
    p, s0 = C.flat_wall_case(mode)
    announce(f"flat wall {mode} {precision} 96x64 one boundary pass")
    start(solver, p, s0, precision)
    solver.updateObstacleGeometry(p)     # same geometry: one boundary pass
    err, fluid = C.flat_wall_error(p, mode, s0, solver.getCellTypes(), solver.state64())
    print(f"  ghost error {err:.1e}", flush=True)
    assert err < (1e-9 if precision == "fp64" else 1e-6), err
    assert fluid == 0.0


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("mode", C.FLAT_WALL_MODES)
def test_oblique_wall_ghost_reflection(solver, mode, precision):
    p, s0 = C.oblique_wall_case(mode)
    announce(f"45-degree wall {mode} {precision} 96x64 one boundary pass")
    start(solver, p, s0, precision)
    solver.updateObstacleGeometry(p)
    err, count = C.oblique_wall_error(p, mode, s0, solver.getCellTypes(), solver.getSDF(), solver.state64())
    print(f"  ghost error {err:.1e} over {count} cells", flush=True)
    assert count > 100
    assert err < 2e-6, err


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("t_mult", C.SUTHERLAND_T)
def test_sutherland_through_cfl_dt(solver, t_mult, precision):
    p = C.sutherland_case(t_mult)
    p.precision = PRECISIONS[precision]
    announce(f"Sutherland T={t_mult} T_ref {precision} 32x32 CFL dt")
    solver.initialize(p)
    dt, _ = C.sutherland_dt(p)
    assert solver.computeStableTimeStep(p) == pytest.approx(dt, rel=2e-6)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_nonfinite_cell_stays_confined(solver, precision):
    p, s0 = C.nonfinite_case()
    announce(f"NaN in one cell {precision} 64x48 {C.NAN_STEPS} steps")
    start(solver, p, s0, precision)
    solver.advance(p.copy(), C.NAN_STEPS)
    s1 = solver.state64()
    spread, dev = C.nonfinite_spread(s0, s1, C.NAN_STEPS)
    print(f"  other non-finite cells {spread}, far change {dev:.1e}", flush=True)
    assert spread == 0
    assert dev < (1e-7 if precision == "fp64" else 1e-6), dev
    assert np.abs(s1[0] - s0[0]).max() > 1e-2


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("viscous", [False, True])
def test_freestream_preserved(solver, viscous, precision):
    """Uniform flow with no obstacle stays uniform (computeFluxesKernel and
    updateKernel, solver.cu:1450-1876; the viscous terms vanish).  Away from
    the edges every face carries the same flux, so rho, rho_u, rho_v do not
    change by a single bit; E and rho_e are re-derived from each other every
    stage (dual-energy sync) and may move by an ulp.  The edges take the
    characteristic boundary state through p / rho^gamma, which FP32 rounds
    (exp and log) to ~1e-6."""
    if viscous:
        p = P.viscous_cylinder(nx=96, ny=48, diameter=8.0, reynolds=50.0, mach=0.5)
    else:
        p = P.euler_bow_shock(nx=96, ny=48)
    C.no_obstacle(p)
    p.dt = 0.1
    p.precision = PRECISIONS[precision]
    announce(f"freestream viscous={viscous} {precision} 96x48 50 steps")
    solver.initialize(p)
    s0 = solver.state64()
    solver.advance(p.copy(), 50)
    s1 = solver.state64()
    dev = [float(np.abs(s1[c] - s0[c]).max() / max(np.abs(s0[c]).max(), 1.0)) for c in range(5)]
    inner = (slice(12, -12), slice(12, -12))
    dev_in = [float(np.abs(s1[c][inner] - s0[c][inner]).max() / max(np.abs(s0[c]).max(), 1.0)) for c in range(5)]
    print(f"  largest change {max(dev):.1e}, interior {max(dev_in):.1e}", flush=True)
    if precision == "fp64":
        assert max(dev) < 1e-13, dev
    else:
        assert max(dev) < 2e-5, dev
        for c in range(3):
            assert np.array_equal(s1[c][inner], s0[c][inner]), c
        assert max(dev_in[3:]) < 3e-7, dev_in


def viscous_and_inviscid(solver, case, precision):
    """One step of a viscous closed-form case from the same float32 state,
    with and without viscosity, through the C ABI."""
    p, s0 = case()
    out = []
    for visc in (1, 0):
        q = p.copy()
        q.enable_viscosity = visc
        start(solver, q, s0, precision)
        solver.advance(q.copy(), 1)
        out.append(solver.state64())
    return p, out[0], out[1]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_heat_conduction_closed_form(solver, precision):
    """Heat conduction k = mu cp / Pr, q = -k grad T (solver.cu:2352,
    2414-2415) in the viscous half steps: one step of the quadratic
    temperature profile adds dt d/dy(k dT/dy) to E and rho_e.  FP32: the
    increment is ~1e-4 of E, so the state's rounding bounds it to ~1e-3."""
    announce(f"heat conduction {precision} 48x64 1 step")
    p, sv, si = viscous_and_inviscid(solver, C.heat_conduction_case, precision)
    eE, ee, other = C.heat_conduction_error(p, sv, si)
    print(f"  dE {eE:.1e}, d rho_e {ee:.1e}, other fields {other:.1e}", flush=True)
    tol = 2e-3 if precision == "fp64" else 5e-3
    assert eE < tol, eE
    assert ee < tol, ee
    assert other < 1e-3, other


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_dilatational_stress_closed_form(solver, precision):
    """tau_xx = mu (2 du/dx - (2/3) div v) (solver.cu:2395-2396): the viscous
    increment of E and rho_e on the isotropic expansion is dt (4/3) mu a^2
    (twice that with 1/3 for 2/3).  FP32: the increment is ~3e-5 of E, so
    single cells carry ~2e-3 of rounding; the block mean is tighter."""
    announce(f"isotropic expansion {precision} 64x64 1 step")
    p, sv, si = viscous_and_inviscid(solver, C.expansion_case, precision)
    mE, me, worst, mom = C.expansion_error(p, sv, si)
    print(f"  mean dE {mE:.1e}, mean d rho_e {me:.1e}, worst cell {worst:.1e}, momentum {mom:.1e}", flush=True)
    assert mE < 1e-2, mE
    assert me < 1e-2, me
    assert worst < (1e-2 if precision == "fp64" else 3e-2), worst
    assert mom < (1e-3 if precision == "fp64" else 1e-2), mom
