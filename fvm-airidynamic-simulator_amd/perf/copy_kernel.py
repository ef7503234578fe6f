#!/usr/bin/env python3
"""A plain device-to-device copy of known size (the runtime's copy kernel),
used to calibrate rocprofv3 FETCH_SIZE / WRITE_SIZE against known bytes:
factor = bytes copied / counter value of one copy dispatch."""
import ctypes
import sys

NBYTES = int(sys.argv[1]) if len(sys.argv) > 1 else 256 << 20
REPS = int(sys.argv[2]) if len(sys.argv) > 2 else 8

hip = ctypes.CDLL("libamdhip64.so")
a, b = ctypes.c_void_p(), ctypes.c_void_p()
assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(NBYTES)) == 0
assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(NBYTES)) == 0
assert hip.hipMemset(a, 1, ctypes.c_size_t(NBYTES)) == 0
for _ in range(REPS):
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(NBYTES), 3) == 0  # hipMemcpyDeviceToDevice
assert hip.hipDeviceSynchronize() == 0
print(f"copied {NBYTES} bytes x {REPS}")
