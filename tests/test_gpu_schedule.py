"""Viscous step schedules on the GPU: the first half step as a launch of its
own (split) or recomputed inside stage 1 on a wider halo (fused) must give
bitwise the same state (the cell arithmetic is shared), in the plain solver at
the autotuned and at explicit tiles, through graph replay and direct launches,
in the batched engine over a viscous/inviscid mix, and where the fused form is
not built (FP64) the request falls back to the split launches.  The split
schedule itself is checked against the oracle by the parity tests."""
import numpy as np
import pytest

import fvm_amd
from fvm_amd import params as P
from fvm_amd.batch import BatchEngine

pytestmark = pytest.mark.gpu


def si_mach3(nx=256, ny=128):
    """The reference defaults: SI Mach-3 viscous circle."""
    p = P.default_params()
    p.nx, p.ny = nx, ny
    p.enable_viscosity = 1
    return P.compute_derived(p)


def run(p, schedule, nsteps, tile=None, direct=False, precision=None):
    q = p.copy()
    q.dt_mode = P.DT_DEVICE_CFL
    if precision is not None:
        q.precision = precision
    s = fvm_amd.CFDSolver(0)
    if tile:
        s.set_tile(*tile)
    s.set_viscous_schedule(schedule)
    s.initialize(q)
    # timing mode launches every kernel directly (no graph)
    s.enable_timing(direct)
    s.advance(q, nsteps)
    names = set(s.kernel_timings()) if direct else set()
    out = s.state64() if precision == P.PRECISION_FP64 else s.state()
    t, n = s.time()
    s.close()
    return out, names, (t, n)


def shaped(p, shape):
    p.obstacle_shape = shape
    return p


def bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


CASES = [
    ("re100_256x128_auto", lambda: P.viscous_cylinder(nx=256, ny=128, diameter=16.0, reynolds=100.0), None),
    ("re100_256x128_32x8", lambda: P.viscous_cylinder(nx=256, ny=128, diameter=16.0, reynolds=100.0), (32, 8)),
    ("re200_1024x512_64x7", lambda: P.viscous_cylinder(nx=1024, ny=512, diameter=32.0, reynolds=200.0), (64, 7)),
    ("si_mach3_256x128_64x16", lambda: si_mach3(), (64, 16)),
    ("ragged_200x90_32x16", lambda: P.viscous_cylinder(nx=200, ny=90, diameter=14.0, reynolds=150.0), (32, 16)),
    ("ragged_200x90_32x7", lambda: P.viscous_cylinder(nx=200, ny=90, diameter=14.0, reynolds=150.0), (32, 7)),
    ("ragged_200x90_32x3", lambda: P.viscous_cylinder(nx=200, ny=90, diameter=14.0, reynolds=150.0), (32, 3)),
]


@pytest.mark.parametrize("name,build,tile", CASES, ids=[c[0] for c in CASES])
def test_fused_schedule_equals_split_bitwise(name, build, tile):
    p = build()
    print(f"\n[case] {name}: 200 steps split vs fused (graph) and fused (direct)", flush=True)
    a, _, ta = run(p, P.SCHEDULE_SPLIT, 200, tile)
    b, _, tb = run(p, P.SCHEDULE_FUSED, 200, tile)
    c, names, tc = run(p, P.SCHEDULE_FUSED, 200, tile, direct=True)
    assert np.isfinite(a).all()
    assert np.array_equal(bits(a), bits(b)), f"max |diff| {np.abs(a - b).max()}"
    assert np.array_equal(bits(a), bits(c))
    assert ta == tb == tc   # clock and step counter advanced by the same dt
    # the fused launches ran: no separate first half step
    assert "stage1_fused" in names and "diffuse_half1" not in names, names


def test_fused_schedule_in_a_batch_equals_split_bitwise():
    mix = [P.viscous_cylinder(nx=256, ny=128, diameter=16.0, reynolds=100.0), P.euler_bow_shock(),
           shaped(P.viscous_cylinder(nx=192, ny=96, diameter=12.0, reynolds=150.0), P.SHAPE_STAR),
           shaped(P.viscous_cylinder(nx=320, ny=160, diameter=20.0, reynolds=300.0), P.SHAPE_TRIANGLE),
           P.euler_bow_shock(nx=160, ny=96, radius=10.0, center=(40.0, 48.0))]
    print(f"\n[case] batch of {len(mix)} (3 viscous), 150 steps split vs fused", flush=True)
    out = []
    for sched in (P.SCHEDULE_SPLIT, P.SCHEDULE_FUSED):
        e = BatchEngine(0, mem_budget_bytes=2 << 30, max_cases=8)
        e.set_viscous_schedule(sched)
        ids = [e.add(q) for q in mix]
        e.advance(150)
        out.append([e.state(i) for i in ids])
        e.close()
    for k, (x, y) in enumerate(zip(*out)):
        assert np.isfinite(x).all()
        assert np.array_equal(bits(x), bits(y)), f"case {k}"


def test_fused_request_falls_back_where_not_built():
    """FP64 has no fused stage 1: the request runs the split launches."""
    p = P.viscous_cylinder(nx=192, ny=96, diameter=12.0, reynolds=100.0)
    print("\n[case] FP64 viscous 192x96 with the fused schedule requested", flush=True)
    a, _, _ = run(p, P.SCHEDULE_SPLIT, 40, precision=P.PRECISION_FP64)
    b, names, _ = run(p, P.SCHEDULE_FUSED, 40, direct=True, precision=P.PRECISION_FP64)
    assert np.array_equal(bits(a), bits(b))
    assert "diffuse_half1" in names and "stage1_fused" not in names, names


def test_schedule_rejects_unknown_values():
    s = fvm_amd.CFDSolver(0)
    with pytest.raises(fvm_amd.FvmError):
        s.set_viscous_schedule(7)
    s.close()
