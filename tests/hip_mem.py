"""Device buffers owned by a test, from the HIP runtime libfvm_amd itself
loaded (the same library and process state), through ctypes."""
import ctypes

import fvm_amd

H2D, D2H = 1, 2


def hip_runtime():
    fvm_amd.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 not loaded by libfvm_amd"
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class DeviceBuffer:
    """nbytes of device memory; freed on close / exit."""

    def __init__(self, nbytes):
        self.hip = hip_runtime()
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        assert a.nbytes <= self.nbytes and a.flags["C_CONTIGUOUS"]
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, a):
        assert a.nbytes <= self.nbytes and a.flags["C_CONTIGUOUS"]
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    @property
    def address(self):
        return self.ptr.value

    def close(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
