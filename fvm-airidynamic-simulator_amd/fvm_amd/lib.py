"""Loader of libfvm_amd.so and ctypes signatures of the C ABI (include/fvm/fvm.h).

The product path is the HIP library; there is no CPU fallback.  Loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

from .params import FvmParams

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfvm_amd.so")

FVM_OK = 0


class FvmError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes)
_P = ctypes.POINTER(FvmParams)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_szp = ctypes.POINTER(ctypes.c_size_t)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

SIGNATURES = [
    ("fvm_abi_version", _i, []),
    ("fvm_last_error", ctypes.c_char_p, []),
    ("fvm_params_default", None, [_P]),
    ("fvm_params_compute_derived", None, [_P]),
    ("fvm_create", _i, [_i, ctypes.POINTER(_vp)]),
    ("fvm_destroy", _i, [_vp]),
    ("fvm_initialize", _i, [_vp, _P]),
    ("fvm_resize", _i, [_vp, _i, _i]),
    ("fvm_reset", _i, [_vp, _P]),
    ("fvm_update_obstacle", _i, [_vp, _P]),
    ("fvm_step", _i, [_vp, _P]),
    ("fvm_advance", _i, [_vp, _P, _i]),
    ("fvm_prepare_graphs", _i, [_vp, _P]),
    ("fvm_compute_stable_dt", _i, [_vp, _P, _fp]),
    ("fvm_queue_dt", _i, [_vp, _P]),
    ("fvm_read_dt", _i, [_vp, _P, _fp]),
    ("fvm_get_cell_types", _i, [_vp, _u8p]),
    ("fvm_get_sdf", _i, [_vp, _fp]),
    ("fvm_field_to_device", _i, [_vp, _i, _vp]),
    ("fvm_field_to_host", _i, [_vp, _i, _fp]),
    ("fvm_get_state", _i, [_vp, _fp]),
    ("fvm_set_state", _i, [_vp, _fp]),
    ("fvm_get_state_f64", _i, [_vp, _dp]),
    ("fvm_conservation_sums", _i, [_vp, _dp]),
    ("fvm_conservation_sums_static", _i, [_vp, _i, _i, _i, _vp, _vp]),
    ("fvm_forces", _i, [_vp, _dp]),
    ("fvm_get_time", _i, [_vp, _dp, _i64p]),
    ("fvm_save_snapshot", _i, [_vp, ctypes.c_char_p]),
    ("fvm_load_snapshot", _i, [_vp, ctypes.c_char_p, _P]),
    ("fvm_max_temperature", _i, [_vp, _fp]),
    ("fvm_max_mach", _i, [_vp, _fp]),
    ("fvm_generate_vector_arrows", _i, [_vp, _vp, _i, _i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.POINTER(_i)]),
    ("fvm_memory_usage", _i, [_vp, _szp]),
    ("fvm_gpu_memory_info", _i, [_i, _szp, _szp]),
    ("fvm_synchronize", _i, [_vp]),
    ("fvm_get_size", _i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    ("fvm_set_tile", _i, [_vp, _i, _i]),
    ("fvm_set_viscous_schedule", _i, [_vp, _i]),
    ("fvm_use_graph", _i, [_vp, _i]),
    ("fvm_fused_schedule_supported", _i, [_i, _i, _i, ctypes.POINTER(_i)]),
    ("fvm_enable_timing", _i, [_vp, _i]),
    ("fvm_kernel_timing", _i, [_vp, _i, ctypes.c_char_p, _i, _dp, _i64p]),
    ("fvm_batch_create", _i, [_i, ctypes.c_size_t, _i, ctypes.POINTER(_vp)]),
    ("fvm_batch_destroy", _i, [_vp]),
    ("fvm_batch_add_case", _i, [_vp, _P, ctypes.c_int32, _i32p]),
    ("fvm_batch_case_bytes", _i, [_P, _szp]),
    ("fvm_batch_remove_case", _i, [_vp, ctypes.c_int32]),
    ("fvm_batch_advance", _i, [_vp, ctypes.c_int32]),
    ("fvm_batch_prepare_graphs", _i, [_vp]),
    ("fvm_batch_case_status", _i, [_vp, ctypes.c_int32, _i32p, _dp, _i32p]),
    ("fvm_batch_get_state", _i, [_vp, ctypes.c_int32, _fp]),
    ("fvm_batch_get_cell_types", _i, [_vp, ctypes.c_int32, _u8p]),
    ("fvm_batch_forces", _i, [_vp, ctypes.c_int32, _dp]),
    ("fvm_batch_save_case", _i, [_vp, ctypes.c_int32, ctypes.c_char_p]),
    ("fvm_batch_load_case", _i, [_vp, ctypes.c_char_p, ctypes.c_int32, _i32p]),
    ("fvm_batch_active_cases", _i, [_vp, _i32p]),
    ("fvm_batch_bytes_in_use", _i, [_vp, _szp]),
    ("fvm_batch_synchronize", _i, [_vp]),
    ("fvm_batch_set_tile", _i, [_vp, _i, _i]),
    ("fvm_batch_set_viscous_schedule", _i, [_vp, _i]),
    ("fvm_batch_use_graph", _i, [_vp, _i]),
]

EXPORTED = [s[0] for s in SIGNATURES]


def lib() -> ctypes.CDLL:
    """The HIP library; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(the FVM step has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, where: str):
    if rc != FVM_OK:
        msg = lib().fvm_last_error()
        raise FvmError(rc, where, msg.decode() if msg else "")
