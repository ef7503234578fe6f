"""Batched engine on the GPU: a case advanced in a mix of uneven cases gives
bitwise the state it gets alone in the plain solver; cases finish at their own
step counts; admitting and retiring cases mid-run leaves the others unchanged."""
import numpy as np
import pytest

import fvm_amd
import oracle as O
from fvm_amd import params as P
from fvm_amd.batch import BatchEngine

pytestmark = pytest.mark.gpu


def small_mix(n, seed):
    rng = np.random.default_rng(seed)
    return [P.mixed_case(rng, n_lo=64 * 64, n_hi=160 * 96, nx_range=(64, 160), ny_range=(64, 96)) for _ in range(n)]


def alone(p, nsteps):
    s = fvm_amd.CFDSolver(0)
    s.initialize(p)
    q = p.copy()
    s.advance(q, nsteps)
    out = s.state()
    s.close()
    return out


@pytest.mark.parametrize("nsteps", [30, 70])
def test_batch_equals_alone_bitwise(nsteps):
    """30 steps replay the 16- and 2-step graphs, 70 steps the 64- and 2-step
    graphs, in the batch's persistent grids and in each solo run."""
    cases = small_mix(10, seed=11)
    print(f"\n[case] batch of {len(cases)} mixed cases, {nsteps} steps vs each alone", flush=True)
    b = BatchEngine(0, mem_budget_bytes=2 << 30)
    ids = [b.add(p) for p in cases]
    b.advance(nsteps)
    for cid, p in zip(ids, cases):
        n, t, fin = b.status(cid)
        assert n == nsteps
        got = b.state(cid)
        assert np.isfinite(got).all()
        assert np.array_equal(got, alone(p, nsteps)), f"case {cid} differs from its solo run"
    b.close()


def test_batch_step_budgets_and_admission():
    cases = small_mix(6, seed=12)
    print("\n[case] per-case step budgets, mid-run admission and removal", flush=True)
    b = BatchEngine(0, mem_budget_bytes=2 << 30)
    budgets = [7, 20, 13, 20, 5, 20]
    ids = [b.add(p, steps_total=k) for p, k in zip(cases[:5], budgets[:5])]
    b.advance(10)
    done = {cid: b.status(cid) for cid in ids}
    assert done[ids[0]][0] == 7 and done[ids[0]][2]
    assert done[ids[4]][0] == 5 and done[ids[4]][2]
    assert done[ids[1]][0] == 10 and not done[ids[1]][2]
    b.remove(ids[0])
    late = b.add(cases[5], steps_total=budgets[5])     # admitted while others are mid-run
    b.advance(20)
    assert b.status(ids[1])[0] == 20 and b.status(ids[2])[0] == 13
    assert b.status(late)[0] == 20
    for cid, p, k in [(ids[1], cases[1], 20), (ids[2], cases[2], 13), (ids[4], cases[4], 5), (late, cases[5], 20)]:
        assert np.array_equal(b.state(cid), alone(p, k)), f"case {cid}"
    assert b.active() == 0
    b.close()


def test_budget_rejects_oversized_case():
    p = P.euler_bow_shock(nx=512, ny=256)
    need = BatchEngine.case_bytes(p)
    b = BatchEngine(0, mem_budget_bytes=need // 2)
    with pytest.raises(fvm_amd.FvmError):
        b.add(p)
    b.close()


def mix_sample(cases):
    """32 cases of the 512-case mix that span it: for each of the five shapes,
    viscous and inviscid, the smallest, the median and the largest grid, plus
    the two fastest (highest Mach) cases not yet taken."""
    pick = []
    for shape in range(5):
        for visc in (0, 1):
            group = sorted((k for k, p in enumerate(cases) if p.obstacle_shape == shape and p.enable_viscosity == visc),
                           key=lambda k: cases[k].nx * cases[k].ny)
            for k in (group[0], group[len(group) // 2], group[-1]):
                if k not in pick:
                    pick.append(k)
    for k in sorted(range(len(cases)), key=lambda k: -cases[k].mach):
        if len(pick) >= 32:
            break
        if k not in pick:
            pick.append(k)
    return pick


@pytest.mark.timeout(900)
def test_mix_cases_vs_fp64_oracle():
    """Per-case semantics of the batched step (each case is the reference's
    CFDSolver::step with its own CFL dt, solver.cu:3289-3412): 32 cases of the
    512-case mix (64x64 to 512x256, five shapes, rotated, Mach 0.3-4, viscous
    and inviscid), 200 steps in one batch, each against the FP64 oracle per
    field with the FP32 oracle as the rounding floor.  A branch flip near a
    rough wall state (see parity.field_parity) is judged by restarting the
    FP64 oracle from the batch's own state 2 steps before the end."""
    from parity import field_parity, oracle_state
    cases = P.mixed_batch(512)
    pick = mix_sample(cases)
    sub = [cases[k] for k in pick]
    print(f"\n[case] 32 cases of the 512-case mix, 200 steps, vs FP64 oracle: "
          f"{sum(p.nx * p.ny for p in sub)} cells", flush=True)
    assert len(sub) == 32
    assert {p.obstacle_shape for p in sub} == set(range(5))
    assert {p.enable_viscosity for p in sub} == {0, 1}
    b = BatchEngine(0, mem_budget_bytes=4 << 30, max_cases=64)
    ids = [b.add(p, steps_total=200) for p in sub]
    b.advance(198)
    prev = [b.state(cid) for cid in ids]
    b.advance(2)
    got = [b.state(cid).astype(np.float64) for cid in ids]
    times = [b.status(cid)[1] for cid in ids]
    b.close()
    for k, p, g, g_prev, t in zip(pick, sub, got, prev, times):
        print(f"  case {k}: {p.nx}x{p.ny} shape {p.obstacle_shape} rot {p.obstacle_rotation:.2f} "
              f"Mach {p.mach:.2f} viscous {p.enable_viscosity}", flush=True)
        o = O.Oracle(p, fp64=True)
        o.step(200, p.copy())
        assert t == pytest.approx(o.time()[0], rel=1e-5)
        field_parity(g, o.state(), oracle_state(p, 200), "    ", restart=(p, g_prev, 2))
        o.close()


@pytest.mark.timeout(600)
def test_full_mix_equals_alone_bitwise():
    """All 512 cases of the mix (19M cells) in one batch for 20 steps: every
    case bitwise equal to its solo run through the plain solver."""
    cases = P.mixed_batch(512)
    print(f"\n[case] the 512-case mix, 20 steps, vs each case alone", flush=True)
    b = BatchEngine(0, mem_budget_bytes=4 << 30, max_cases=512)
    ids = [b.add(p) for p in cases]
    b.advance(20)
    s = fvm_amd.CFDSolver(0)
    for cid, p in zip(ids, cases):
        got = b.state(cid)
        s.initialize(p)
        s.advance(p.copy(), 20)
        assert np.array_equal(got, s.state()), f"case {cid} differs from its solo run"
    s.close()
    b.close()
