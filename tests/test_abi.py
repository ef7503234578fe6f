"""The C ABI library loads without a GPU and exports every entry point that
include/fvm/fvm.h declares; the ctypes mirror of fvm_params matches the C
layout; host-side parameter helpers match the reference's computeDerived."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import fvm_amd
from fvm_amd import params as P
from fvm_amd.lib import EXPORTED, LIB_PATH

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "fvm", "fvm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(EXPORTED) == declared_functions()


def test_params_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "fvm/fvm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(fvm_params), offsetof(fvm_params, gamma),
         offsetof(fvm_params, flux_scheme), offsetof(fvm_params, noise_seed), offsetof(fvm_params, reserved));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(REPO, "include"), c, "-o", exe])
        vals = [int(x) for x in subprocess.check_output([exe]).split()]
    F = P.FvmParams
    assert vals == [ctypes.sizeof(F), F.gamma.offset, F.flux_scheme.offset, F.noise_seed.offset, F.reserved.offset]


def test_default_and_derived_params_match_reference_formulas():
    lib = fvm_amd.lib()
    p = P.FvmParams()
    lib.fvm_params_default(ctypes.byref(p))
    assert (p.nx, p.ny, p.mach, p.T_inf, p.p_inf, p.cfl, p.cfl_visc) == (512, 256, 3.0, 300.0, 101325.0, 0.5, P.f32(0.4))
    lib.fvm_params_compute_derived(ctypes.byref(p))
    q = P.compute_derived(P.default_params())
    for f in ("rho_inf", "c_inf", "u_inf", "dx", "dy", "obstacle_y"):
        assert getattr(p, f) == getattr(q, f), f
    # rho = p/(R T), c = sqrt(gamma R T), common.h:95-97
    assert p.rho_inf == pytest.approx(101325.0 / (287.05 * 300.0), rel=1e-6)
    assert p.c_inf == pytest.approx((1.4 * 287.05 * 300.0) ** 0.5, rel=1e-6)


def test_create_without_gpu_fails_cleanly():
    """No device in the build container: fvm_create reports an error, no crash."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = fvm_amd.lib().fvm_create(0, ctypes.byref(h))
    assert rc != 0
    assert fvm_amd.lib().fvm_last_error()
