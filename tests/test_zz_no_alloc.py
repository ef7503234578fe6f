"""fvm_conservation_sums_static, the entry point that allocates nothing,
through the C ABI: on a state copied into device memory that this test owns,
its totals equal fvm_conservation_sums of the solver holding that state, bit
for bit, on two calls in a row (the second with a smaller grid) and for an
FP64 state; sizes beyond its maximum and bad arguments are rejected."""
import ctypes

import numpy as np
import pytest

import fvm_amd
from fvm_amd import params as P
from hip_mem import D2H, H2D, hip_runtime

pytestmark = pytest.mark.gpu


FVM_ERR_INVALID_ARGUMENT = -1   # fvm.h


def solver_state(s, p, steps, fp64):
    s.initialize(p)
    s.advance(p.copy(), steps)
    st = s.state64() if fp64 else s.state()
    return np.ascontiguousarray(st), s.conservation_sums()


def static_sums(hip, d_state, d_out, st):
    assert hip.hipMemcpy(d_state, st.ctypes.data, st.nbytes, H2D) == 0
    ny, nx = st.shape[1:]
    rc = fvm_amd.lib().fvm_conservation_sums_static(d_state, st.itemsize, nx, ny, d_out, None)
    assert rc == 0, fvm_amd.lib().fvm_last_error()
    out = np.zeros(4, dtype=np.float64)
    assert hip.hipMemcpy(out.ctypes.data, d_out, 32, D2H) == 0   # null-stream copy: after the kernels
    return out


def test_static_sums_equal_the_solver_entry_point():
    hip = hip_runtime()
    d_state, d_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_state), 5 * 160 * 96 * 8) == 0
    assert hip.hipMalloc(ctypes.byref(d_out), 64) == 0
    try:
        with fvm_amd.CFDSolver(0) as s:
            cases = [("viscous 160x96 FP32", P.viscous_cylinder(nx=160, ny=96, diameter=12.0, reynolds=100.0), 12, False),
                     ("inviscid 64x48 FP32 (smaller)", P.euler_bow_shock(nx=64, ny=48, radius=6.0, center=(16.0, 24.0)), 7, False),
                     ("inviscid 96x40 FP64", P.euler_bow_shock(nx=96, ny=40, radius=6.0, center=(24.0, 20.0)), 5, True)]
            for label, p, steps, fp64 in cases:
                print(f"\n[case] static sums: {label}, {steps} steps", flush=True)
                p.dt = 0.1
                if fp64:
                    p.precision = P.PRECISION_FP64
                st, ref = solver_state(s, p, steps, fp64)
                got = static_sums(hip, d_state, d_out, st)
                print(f"  static {got}, solver {ref}", flush=True)
                assert np.array_equal(got, ref), (got, ref)
                assert np.allclose(got, st[:4].astype(np.float64).reshape(4, -1).sum(axis=1), rtol=1e-12)
    finally:
        hip.hipFree(d_state)
        hip.hipFree(d_out)


def test_static_sums_reject_bad_arguments():
    """Host-side checks only: no kernel is launched for these."""
    L = fvm_amd.lib()
    fake = ctypes.c_void_p(0x1000)   # never dereferenced: every call below is rejected first
    assert L.fvm_conservation_sums_static(None, 4, 8, 8, fake, None) == FVM_ERR_INVALID_ARGUMENT
    assert L.fvm_conservation_sums_static(fake, 2, 8, 8, fake, None) == FVM_ERR_INVALID_ARGUMENT
    assert L.fvm_conservation_sums_static(fake, 4, 0, 8, fake, None) == FVM_ERR_INVALID_ARGUMENT
    assert L.fvm_conservation_sums_static(fake, 4, 65536, 65536, fake, None) == FVM_ERR_INVALID_ARGUMENT
    assert b"FVM_STATIC_MAX_CELLS" in L.fvm_last_error()
