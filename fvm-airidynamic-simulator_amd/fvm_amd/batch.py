"""Batched engine: many independent cases advanced together (fvm_batch_* C ABI).

Every kernel of a step is one launch over a (case, tile) work list of all
active cases; each case keeps its own CFL dt on the device; the step pair is
replayed from a captured HIP graph.  A case's arithmetic does not depend on
the other cases, so a case alone and in the mix give bitwise the same result.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import params as P
from .lib import check, lib


class BatchEngine:
    def __init__(self, device: int = 0, mem_budget_bytes: int = 8 << 30, max_cases: int = 4096):
        h = ctypes.c_void_p()
        check(lib().fvm_batch_create(device, ctypes.c_size_t(mem_budget_bytes), max_cases, ctypes.byref(h)),
              "fvm_batch_create")
        self._h = h
        self._shape = {}

    def close(self):
        if self._h:
            lib().fvm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def case_bytes(p: P.FvmParams) -> int:
        b = ctypes.c_size_t()
        check(lib().fvm_batch_case_bytes(ctypes.byref(p), ctypes.byref(b)), "fvm_batch_case_bytes")
        return b.value

    def add(self, p: P.FvmParams, steps_total: int = 0) -> int:
        cid = ctypes.c_int32()
        check(lib().fvm_batch_add_case(self._h, ctypes.byref(p), steps_total, ctypes.byref(cid)), "fvm_batch_add_case")
        self._shape[cid.value] = (p.ny, p.nx)
        return cid.value

    def remove(self, cid: int):
        check(lib().fvm_batch_remove_case(self._h, cid), "fvm_batch_remove_case")
        self._shape.pop(cid, None)

    def advance(self, nsteps: int):
        check(lib().fvm_batch_advance(self._h, nsteps), "fvm_batch_advance")

    def prepare_graphs(self):
        """Capture the step graphs of the current cases without stepping."""
        check(lib().fvm_batch_prepare_graphs(self._h), "fvm_batch_prepare_graphs")

    def status(self, cid: int, with_time: bool = True):
        """(steps done, simulated time, finished).  The step count and the
        finished flag are host-side; with_time=False skips the device read of
        the clock (time is then None) and so waits for nothing."""
        n = ctypes.c_int32()
        t = ctypes.c_double()
        f = ctypes.c_int32()
        check(lib().fvm_batch_case_status(self._h, cid, ctypes.byref(n), ctypes.byref(t) if with_time else None,
                                          ctypes.byref(f)), "fvm_batch_case_status")
        return n.value, (t.value if with_time else None), bool(f.value)

    def state(self, cid: int) -> np.ndarray:
        ny, nx = self._shape[cid]
        out = np.empty((5, ny, nx), dtype=np.float32)
        check(lib().fvm_batch_get_state(self._h, cid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
              "fvm_batch_get_state")
        return out

    def save(self, cid: int, path: str):
        check(lib().fvm_batch_save_case(self._h, cid, str(path).encode()), "fvm_batch_save_case")

    def load(self, path: str, steps_total: int = 0) -> int:
        """Admit a case from a snapshot file; it runs steps_total more steps."""
        from .snapshot import read
        cid = ctypes.c_int32()
        check(lib().fvm_batch_load_case(self._h, str(path).encode(), steps_total, ctypes.byref(cid)),
              "fvm_batch_load_case")
        h = read(path).header
        self._shape[cid.value] = (h.ny, h.nx)
        return cid.value

    def forces(self, cid: int) -> np.ndarray:
        out = np.zeros(3, dtype=np.float64)
        check(lib().fvm_batch_forces(self._h, cid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "fvm_batch_forces")
        return out

    def cell_types(self, cid: int) -> np.ndarray:
        ny, nx = self._shape[cid]
        out = np.empty((ny, nx), dtype=np.uint8)
        check(lib().fvm_batch_get_cell_types(self._h, cid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
              "fvm_batch_get_cell_types")
        return out

    def active(self) -> int:
        n = ctypes.c_int32()
        check(lib().fvm_batch_active_cases(self._h, ctypes.byref(n)), "fvm_batch_active_cases")
        return n.value

    def bytes_in_use(self) -> int:
        b = ctypes.c_size_t()
        check(lib().fvm_batch_bytes_in_use(self._h, ctypes.byref(b)), "fvm_batch_bytes_in_use")
        return b.value

    def synchronize(self):
        check(lib().fvm_batch_synchronize(self._h), "fvm_batch_synchronize")

    def set_tile(self, tx: int, ty: int):
        check(lib().fvm_batch_set_tile(self._h, tx, ty), "fvm_batch_set_tile")

    def set_viscous_schedule(self, schedule: int):
        check(lib().fvm_batch_set_viscous_schedule(self._h, int(schedule)), "fvm_batch_set_viscous_schedule")

    def use_graph(self, on: bool):
        check(lib().fvm_batch_use_graph(self._h, 1 if on else 0), "fvm_batch_use_graph")
