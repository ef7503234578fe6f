"""Served (continuous-batching) loop on the GPU: requests arriving while
others run are admitted, complete at their own step counts, and every result
equals the same case run alone (bitwise)."""
import numpy as np
import pytest

import fvm_amd
from fvm_amd import params as P
from fvm_amd.serve import Request, SolveServer

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_runs():
    rng = np.random.default_rng(21)
    reqs = []
    for k in range(8):
        p = P.mixed_case(rng, n_lo=64 * 64, n_hi=192 * 96, nx_range=(64, 192), ny_range=(64, 96))
        reqs.append(Request(k, arrival=0.02 * k, params=p, steps=int(rng.integers(15, 60))))
    print(f"\n[case] served: {len(reqs)} staggered requests, steps 15-60", flush=True)
    srv = SolveServer(0, mem_budget_bytes=1 << 30, max_active=4, chunk_steps=7, keep_results=True)
    stats = srv.run(reqs)
    srv.close()
    assert stats["completed"] == len(reqs)
    s = fvm_amd.CFDSolver(0)
    for r in stats["done"]:
        s.initialize(r.params)
        q = r.params.copy()
        s.advance(q, r.steps)
        assert np.array_equal(r.result, s.state()), f"request {r.rid}"
        assert r.completed >= r.admitted >= r.arrival
    s.close()
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def test_admission_respects_budget():
    p = P.euler_bow_shock(nx=128, ny=64, radius=8.0, center=(32.0, 32.0))
    p.dt_mode = P.DT_DEVICE_CFL
    need = fvm_amd.batch.BatchEngine.case_bytes(p)
    reqs = [Request(k, 0.0, p.copy(), 10) for k in range(5)]
    srv = SolveServer(0, mem_budget_bytes=int(need * 2.5), max_active=16, chunk_steps=5)
    stats = srv.run(reqs)
    srv.close()
    assert stats["completed"] == 5
