"""Python mirror of the reference's ``CFDSolver`` interface over the C ABI.

Method names and argument meaning follow src/include/solver.cuh:22-80; every
call goes to libfvm_amd.so (hand-written HIP for gfx950).  Errors raise
:class:`FvmError` (the reference exits the process, solver.cuh:8-19).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import params as P
from .lib import check, lib


class CFDSolver:
    """One case on one GPU (solver.cuh:22).  State stays in HBM between calls."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().fvm_create(device, ctypes.byref(h)), "fvm_create")
        self._h = h
        self._nx = self._ny = 0
        self.device = device

    def close(self):
        if self._h:
            lib().fvm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- lifecycle (solver.cu:3102-3196)
    def initialize(self, params: P.FvmParams):
        check(lib().fvm_initialize(self._h, ctypes.byref(params)), "fvm_initialize")
        self._nx, self._ny = params.nx, params.ny

    def resize(self, nx: int, ny: int):
        check(lib().fvm_resize(self._h, nx, ny), "fvm_resize")
        if nx > 0 and ny > 0:
            self._nx, self._ny = nx, ny

    def reset(self, params: P.FvmParams):
        check(lib().fvm_reset(self._h, ctypes.byref(params)), "fvm_reset")
        self._nx, self._ny = params.nx, params.ny

    def updateObstacleGeometry(self, params: P.FvmParams):
        check(lib().fvm_update_obstacle(self._h, ctypes.byref(params)), "fvm_update_obstacle")

    # -- stepping (solver.cu:3289-3412)
    def step(self, params: P.FvmParams):
        check(lib().fvm_step(self._h, ctypes.byref(params)), "fvm_step")

    def advance(self, params: P.FvmParams, nsteps: int):
        check(lib().fvm_advance(self._h, ctypes.byref(params), int(nsteps)), "fvm_advance")

    def prepare_graphs(self, params: P.FvmParams):
        """Capture the step graphs advance() replays, without stepping."""
        check(lib().fvm_prepare_graphs(self._h, ctypes.byref(params)), "fvm_prepare_graphs")

    # -- CFL dt (solver.cu:3206-3278)
    def computeStableTimeStep(self, params: P.FvmParams) -> float:
        dt = ctypes.c_float()
        check(lib().fvm_compute_stable_dt(self._h, ctypes.byref(params), ctypes.byref(dt)), "fvm_compute_stable_dt")
        return dt.value

    def queueTimeStepComputation(self, params: P.FvmParams):
        check(lib().fvm_queue_dt(self._h, ctypes.byref(params)), "fvm_queue_dt")

    def readTimeStepResult(self, params: P.FvmParams) -> float:
        dt = ctypes.c_float()
        check(lib().fvm_read_dt(self._h, ctypes.byref(params), ctypes.byref(dt)), "fvm_read_dt")
        return dt.value

    # -- readback (solver.cu:3417-3671)
    def getCellTypes(self) -> np.ndarray:
        out = np.empty((self._ny, self._nx), dtype=np.uint8)
        check(lib().fvm_get_cell_types(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "fvm_get_cell_types")
        return out

    def getSDF(self) -> np.ndarray:
        out = np.empty((self._ny, self._nx), dtype=np.float32)
        check(lib().fvm_get_sdf(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "fvm_get_sdf")
        return out

    def fieldToDevice(self, field: int, dev_ptr: int):
        check(lib().fvm_field_to_device(self._h, field, ctypes.c_void_p(dev_ptr)), "fvm_field_to_device")

    def field(self, field: int) -> np.ndarray:
        out = np.empty((self._ny, self._nx), dtype=np.float32)
        check(lib().fvm_field_to_host(self._h, field, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "fvm_field_to_host")
        return out

    def getMaxTemperature(self) -> float:
        v = ctypes.c_float()
        check(lib().fvm_max_temperature(self._h, ctypes.byref(v)), "fvm_max_temperature")
        return v.value

    def getMaxMach(self) -> float:
        v = ctypes.c_float()
        check(lib().fvm_max_mach(self._h, ctypes.byref(v)), "fvm_max_mach")
        return v.value

    def generateVectorArrows(self, dev_vertex_data: int, max_vertices: int, step: int, u_inf: float,
                             max_arrow_length: float, arrow_head_angle: float, arrow_head_length: float) -> int:
        """generateVectorArrows (solver.cu:3809-3845): GL_LINES vertices
        (x, y, r, g, b) of the velocity arrows into caller device memory;
        returns the vertex count."""
        n = ctypes.c_int()
        check(lib().fvm_generate_vector_arrows(self._h, ctypes.c_void_p(dev_vertex_data), int(max_vertices), int(step),
                                               u_inf, max_arrow_length, arrow_head_angle, arrow_head_length,
                                               ctypes.byref(n)), "fvm_generate_vector_arrows")
        return n.value

    def getSimulationMemoryUsage(self) -> int:
        v = ctypes.c_size_t()
        check(lib().fvm_memory_usage(self._h, ctypes.byref(v)), "fvm_memory_usage")
        return v.value

    @staticmethod
    def getGPUMemoryInfo(device: int = 0):
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().fvm_gpu_memory_info(device, ctypes.byref(free), ctypes.byref(total)), "fvm_gpu_memory_info")
        return free.value, total.value

    def getNx(self) -> int:
        return self._nx

    def getNy(self) -> int:
        return self._ny

    # -- additions
    def state(self) -> np.ndarray:
        """Conserved state (rho, rho_u, rho_v, E, rho_e) as float32 [5, ny, nx]."""
        out = np.empty((5, self._ny, self._nx), dtype=np.float32)
        check(lib().fvm_get_state(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "fvm_get_state")
        return out

    def state64(self) -> np.ndarray:
        out = np.empty((5, self._ny, self._nx), dtype=np.float64)
        check(lib().fvm_get_state_f64(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fvm_get_state_f64")
        return out

    def set_state(self, s: np.ndarray):
        s = np.ascontiguousarray(s, dtype=np.float32).reshape(5, self._ny, self._nx)
        check(lib().fvm_set_state(self._h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "fvm_set_state")

    def conservation_sums(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.float64)
        check(lib().fvm_conservation_sums(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fvm_conservation_sums")
        return out

    def forces(self) -> np.ndarray:
        """Pressure force on the obstacle per unit depth: (Fx, Fy, wall faces)."""
        out = np.zeros(3, dtype=np.float64)
        check(lib().fvm_forces(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fvm_forces")
        return out

    def save_snapshot(self, path: str):
        check(lib().fvm_save_snapshot(self._h, str(path).encode()), "fvm_save_snapshot")

    def load_snapshot(self, path: str) -> P.FvmParams:
        """Initialise from a snapshot file; returns its parameters."""
        p = P.FvmParams()
        check(lib().fvm_load_snapshot(self._h, str(path).encode(), ctypes.byref(p)), "fvm_load_snapshot")
        self._nx, self._ny = p.nx, p.ny
        return p

    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib().fvm_get_time(self._h, ctypes.byref(t), ctypes.byref(n)), "fvm_get_time")
        return t.value, n.value

    def synchronize(self):
        check(lib().fvm_synchronize(self._h), "fvm_synchronize")

    def set_tile(self, tx: int, ty: int):
        check(lib().fvm_set_tile(self._h, tx, ty), "fvm_set_tile")

    def set_viscous_schedule(self, schedule: int):
        """P.SCHEDULE_AUTO (default), P.SCHEDULE_SPLIT or P.SCHEDULE_FUSED (the
        first viscous half step inside stage 1); all give bitwise the same state."""
        check(lib().fvm_set_viscous_schedule(self._h, int(schedule)), "fvm_set_viscous_schedule")

    def use_graph(self, on: bool):
        """Replay captured graphs in advance() (default) or launch directly."""
        check(lib().fvm_use_graph(self._h, 1 if on else 0), "fvm_use_graph")

    @staticmethod
    def fused_schedule_supported(tile_x: int, tile_y: int, precision: int = P.PRECISION_FP32) -> bool:
        v = ctypes.c_int()
        check(lib().fvm_fused_schedule_supported(tile_x, tile_y, precision, ctypes.byref(v)),
              "fvm_fused_schedule_supported")
        return bool(v.value)

    def enable_timing(self, on: bool = True):
        check(lib().fvm_enable_timing(self._h, 1 if on else 0), "fvm_enable_timing")

    def kernel_timings(self) -> dict:
        out = {}
        i = 0
        buf = ctypes.create_string_buffer(64)
        while True:
            avg = ctypes.c_double()
            n = ctypes.c_int64()
            rc = lib().fvm_kernel_timing(self._h, i, buf, 64, ctypes.byref(avg), ctypes.byref(n))
            if rc != 0:
                break
            out[buf.value.decode()] = (avg.value, n.value)
            i += 1
        return out
