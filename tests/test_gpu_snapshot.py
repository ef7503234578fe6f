"""State snapshots through the C ABI: a run saved, reloaded in a fresh
solver (or the batched engine) and continued is bitwise the run that was not
interrupted; the host reader sees the saved state."""
import numpy as np
import pytest

import fvm_amd
from fvm_amd import params as P
from fvm_amd import snapshot
from fvm_amd.batch import BatchEngine

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


def case(kind):
    if kind == "viscous":
        p = P.viscous_cylinder(nx=192, ny=96, diameter=16.0, reynolds=100.0)
    else:
        p = P.euler_bow_shock(nx=192, ny=96, radius=14.0, center=(60.0, 48.0))
        if kind == "kahan":
            p.precision = P.PRECISION_FP32_KAHAN
        elif kind == "fp64":
            p.precision = P.PRECISION_FP64
    p.dt_mode = P.DT_DEVICE_CFL
    return p


@pytest.mark.parametrize("kind", ["viscous", "inviscid", "kahan", "fp64"])
def test_snapshot_resume_is_bitwise(tmp_path, kind):
    p = case(kind)
    announce(f"snapshot resume {kind} 192x96, 61 + 61 steps")
    path = str(tmp_path / "run.fvmsnap")
    a = fvm_amd.CFDSolver(0)
    a.initialize(p)
    a.advance(p.copy(), 61)
    a.save_snapshot(path)
    saved = a.state64() if kind == "fp64" else a.state()
    t_saved = a.time()
    a.advance(p.copy(), 61)
    ref = a.state64()
    t_ref = a.time()
    a.close()

    s = snapshot.read(path)
    assert np.array_equal(s.state, saved)
    assert (s.header.t, s.header.steps) == t_saved
    key = ("nx", "ny", "precision", "dt_mode", "enable_viscosity", "obstacle_x", "obstacle_r", "mach", "noise_seed")
    assert [getattr(s.params, k) for k in key] == [getattr(p, k) for k in key]

    b = fvm_amd.CFDSolver(0)
    q = b.load_snapshot(path)
    assert bytes(q) == bytes(s.params)
    b.advance(q, 61)
    assert np.array_equal(b.state64(), ref)
    assert b.time() == t_ref
    b.close()


def test_snapshot_moves_a_case_into_the_batch(tmp_path):
    p = case("viscous")
    announce("plain solver snapshot -> batched engine, 40 + 40 steps")
    path = str(tmp_path / "case.fvmsnap")
    a = fvm_amd.CFDSolver(0)
    a.initialize(p)
    a.advance(p.copy(), 40)
    a.save_snapshot(path)
    a.advance(p.copy(), 40)
    ref = a.state()
    a.close()
    e = BatchEngine(0, mem_budget_bytes=1 << 30, max_cases=4)
    other = e.add(P.euler_bow_shock(nx=128, ny=64, radius=10.0, center=(40.0, 32.0)), steps_total=0)
    cid = e.load(path, steps_total=40)
    e.advance(40)
    e.synchronize()
    n, _, fin = e.status(cid)
    assert fin and n == 40
    assert np.array_equal(e.state(cid), ref)
    path2 = str(tmp_path / "back.fvmsnap")
    e.save(cid, path2)
    assert np.array_equal(snapshot.read(path2).state, ref)
    e.remove(other)
    e.close()


def test_snapshot_rejects_a_bad_file(tmp_path):
    path = tmp_path / "bad.fvmsnap"
    path.write_bytes(b"not a snapshot" * 40)
    s = fvm_amd.CFDSolver(0)
    with pytest.raises(fvm_amd.FvmError):
        s.load_snapshot(str(path))
    s.close()
