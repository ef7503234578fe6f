"""State snapshot files (`fvm_save_snapshot` / `fvm_load_snapshot`, include/fvm/fvm.h).

Layout, little-endian: a 256-byte `fvm_snapshot_header`, the `fvm_params`
struct, five state planes (rho, rho_u, rho_v, E, rho_e; nx*ny values each,
row-major idx = j*nx + i; FP32 or FP64 as the header says) and, for the
compensated FP32 variant, five FP32 carry planes.  The header keeps the
device time, the step count and the CFL maxima of the state, so a loaded
snapshot continues bitwise like the run that wrote it.  `read` parses a file
on the host (no GPU needed), e.g. to inspect or compare results.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from .params import FvmParams

MAGIC = b"FVMSNAP\x00"
VERSION = 1


class SnapshotHeader(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 8),
        ("version", ctypes.c_uint32),
        ("header_bytes", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("params_bytes", ctypes.c_uint32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("elem_bytes", ctypes.c_uint32),
        ("kahan", ctypes.c_uint32),
        ("t", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("max_speed_bits", ctypes.c_uint32),
        ("max_nu_bits", ctypes.c_uint32),
        ("max_speed64", ctypes.c_double),
        ("max_nu64", ctypes.c_double),
        ("reserved", ctypes.c_uint8 * 176),
    ]


@dataclasses.dataclass
class Snapshot:
    header: SnapshotHeader
    params: FvmParams
    state: np.ndarray            # (5, ny, nx), float32 or float64
    carries: np.ndarray | None   # (5, ny, nx) float32, compensated variant only

    @property
    def max_speed(self) -> float:
        return float(np.array([self.header.max_speed_bits], dtype=np.uint32).view(np.float32)[0])

    @property
    def max_nu(self) -> float:
        return float(np.array([self.header.max_nu_bits], dtype=np.uint32).view(np.float32)[0])


def read(path: str) -> Snapshot:
    with open(path, "rb") as f:
        raw = f.read()
    h = SnapshotHeader.from_buffer_copy(raw[:ctypes.sizeof(SnapshotHeader)])
    if bytes(raw[:8]) != MAGIC or h.version != VERSION or h.header_bytes != ctypes.sizeof(SnapshotHeader):
        raise ValueError(f"{path}: not an fvm snapshot of version {VERSION}")
    if h.params_bytes != ctypes.sizeof(FvmParams):
        raise ValueError(f"{path}: parameter block of {h.params_bytes} bytes, this library has {ctypes.sizeof(FvmParams)}")
    off = h.header_bytes
    p = FvmParams.from_buffer_copy(raw[off:off + h.params_bytes])
    off += h.params_bytes
    n = h.nx * h.ny
    dt = np.float32 if h.elem_bytes == 4 else np.float64
    need = off + 5 * n * h.elem_bytes + (5 * n * 4 if h.kahan else 0)
    if len(raw) != need:
        raise ValueError(f"{path}: {len(raw)} bytes, the header implies {need}")
    state = np.frombuffer(raw, dtype=dt, count=5 * n, offset=off).reshape(5, h.ny, h.nx).copy()
    off += 5 * n * h.elem_bytes
    carries = None
    if h.kahan:
        carries = np.frombuffer(raw, dtype=np.float32, count=5 * n, offset=off).reshape(5, h.ny, h.nx).copy()
    return Snapshot(h, p, state, carries)
