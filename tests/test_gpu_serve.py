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


@pytest.mark.timeout(120)
def test_request_larger_than_the_budget_is_refused():
    """A request that would not fit even in an empty engine is refused at
    once (before, the loop waited for it forever when no time limit was
    given); the requests around it are served."""
    small = P.euler_bow_shock(nx=96, ny=48, radius=6.0, center=(24.0, 24.0))
    small.dt = 0.1
    big = P.euler_bow_shock(nx=1024, ny=512, radius=32.0, center=(256.0, 256.0))
    big.dt = 0.1
    budget = int((fvm_amd.batch.BatchEngine.case_bytes(small) + small.nx * small.ny * 2) * 3)
    assert fvm_amd.batch.BatchEngine.case_bytes(big) > budget
    reqs = [Request(0, 0.0, small.copy(), 10), Request(1, 0.0, big.copy(), 10), Request(2, 0.01, small.copy(), 10)]
    print("\n[case] served: a request larger than the whole budget between two small ones", flush=True)
    srv = SolveServer(0, mem_budget_bytes=budget, max_active=4, chunk_steps=5)
    stats = srv.run(reqs)
    srv.close()
    assert stats["completed"] == 2 and stats["rejected"] == 1
    assert sorted(r.rid for r in stats["done"]) == [0, 2]


@pytest.mark.timeout(900)
def test_served_large_requests_vs_fp64_oracle():
    """Requests at the served configuration's large end (1024x512 to
    2048x1024, drawn as the served stream draws them) arrive while smaller
    ones run; each large result is compared with the FP64 oracle per field,
    with the FP32 oracle as the rounding floor, and each small one with its
    solo run (bitwise)."""
    import oracle as O
    from parity import field_parity, oracle_state
    rng = np.random.default_rng(31)
    reqs = []
    sizes = [(1024, 512), (1536, 768), (2048, 1024)]
    for k, (nx, ny) in enumerate(sizes):
        p = P.mixed_case(rng, n_lo=nx * ny, n_hi=nx * ny, nx_range=(nx, nx), ny_range=(ny, ny))
        reqs.append(Request(k, arrival=0.002 * k, params=p, steps=int(rng.integers(60, 101))))
    for k in range(3, 9):
        p = P.mixed_case(rng, n_lo=128 * 64, n_hi=256 * 128, nx_range=(128, 256), ny_range=(64, 128))
        reqs.append(Request(k, arrival=0.001 * k, params=p, steps=int(rng.integers(20, 120))))
    print(f"\n[case] served: 3 large requests {sizes} among 6 small ones", flush=True)
    srv = SolveServer(0, mem_budget_bytes=8 << 30, max_active=16, chunk_steps=10, keep_results=True)
    stats = srv.run(reqs)
    srv.close()
    assert stats["completed"] == len(reqs)
    done = {r.rid: r for r in stats["done"]}
    # the large ones overlapped on the device with each other and with small ones
    big = [done[k] for k in range(3)]
    assert max(r.admitted for r in big) < min(r.completed for r in big)
    s = fvm_amd.CFDSolver(0)
    for r in stats["done"]:
        p = r.params
        print(f"  request {r.rid}: {p.nx}x{p.ny} shape {p.obstacle_shape} Mach {p.mach:.2f} "
              f"viscous {p.enable_viscosity} {r.steps} steps", flush=True)
        if r.rid < 3:
            o = O.Oracle(p, fp64=True)
            o.step(r.steps, p.copy())
            field_parity(r.result.astype(np.float64), o.state(), oracle_state(p, r.steps), "    ")
            o.close()
        else:
            s.initialize(p)
            s.advance(p.copy(), r.steps)
            assert np.array_equal(r.result, s.state()), f"request {r.rid}"
    s.close()
