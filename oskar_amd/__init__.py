"""MI355X-native OSKAR visibility hot path.

The product is the C-ABI shared library ``liboskar_amd.so`` (host C calling
hand-written gfx950 HIP kernels), declared in ``include/*.h``.  This package
loads it with ctypes and adds thin Python helpers for tests and benchmarks.
Loading fails loudly when the library has not been built: there is no
Python or CPU fallback for any hot-path function.
"""
from .capi import lib, load, LIB_PATH  # noqa: F401
from .mem import (  # noqa: F401
    Mem, Jones, Telescope, Station, StationWork, check_status,
    OSKAR_CPU, OSKAR_GPU, OSKAR_SINGLE, OSKAR_DOUBLE, OSKAR_COMPLEX,
    OSKAR_MATRIX, OSKAR_INT, OskarError,
)
