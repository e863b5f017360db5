"""MI355X implementation of QUICHOTTE's self-consistent NEGF path.

The compute path is the gfx950 library lib/libqs_hip.so behind the C ABI of
include/qs_hip.h; this module binds it with ctypes.  There is no CPU fallback:
if the library is missing or no GPU is visible, the calls raise.
"""
from .engine import (QS_ERRORS, QsError, ScfEngine, ScfProblem, ScfResult, device_memory, footprint_bytes,  # noqa: F401
                     load_library)
