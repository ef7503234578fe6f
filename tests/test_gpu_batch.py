"""Batched engine on the GPU: a case advanced in a mix of uneven cases gives
bitwise the state it gets alone in the plain solver; cases finish at their own
step counts; admitting and retiring cases mid-run leaves the others unchanged."""
import numpy as np
import pytest

import fvm_amd
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


def test_batch_equals_alone_bitwise():
    cases = small_mix(10, seed=11)
    print(f"\n[case] batch of {len(cases)} mixed cases, 30 steps vs each alone", flush=True)
    b = BatchEngine(0, mem_budget_bytes=2 << 30)
    ids = [b.add(p) for p in cases]
    b.advance(30)
    for cid, p in zip(ids, cases):
        n, t, fin = b.status(cid)
        assert n == 30
        got = b.state(cid)
        assert np.isfinite(got).all()
        assert np.array_equal(got, alone(p, 30)), f"case {cid} differs from its solo run"
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
