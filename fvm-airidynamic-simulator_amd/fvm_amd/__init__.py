"""fvm_amd -- MI355X-native 2D compressible FVM time step (host-side Python mirror).

The compute path is libfvm_amd.so (hand-written HIP for gfx950) behind the C
ABI in include/fvm/fvm.h; this package only binds it with ctypes and builds
the synthetic workloads.
"""
from . import params  # noqa: F401
from .lib import FvmError, LIB_PATH, lib  # noqa: F401
from .solver import CFDSolver  # noqa: F401
from . import batch, diagnostics, serve, snapshot  # noqa: F401
