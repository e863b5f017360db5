"""MI355X implementation of QUICHOTTE's self-consistent NEGF path (and of
SANCHO's order-2/3 self-energies).

The compute path is the gfx950 library lib/libqs_hip.so behind the C ABI of
include/qs_hip.h; this module binds it with ctypes.  There is no CPU fallback:
if the library is missing or no GPU is visible, the calls raise.
"""
from .engine import (QS_ERRORS, DeviceBuffer, QsError, SanchoSelfEnergies, ScfEngine, ScfGroup, ScfProblem,  # noqa: F401
                     ScfResult, device_memory, footprint_bytes, load_library, sancho_self_energies_static)
