"""drifter_amd: MI355X-native DG right-hand side for the DRIFTER ocean model.

The compute path is libdrifter_hip.so (hand-written HIP for gfx950) behind the
C ABI in include/drifter_hip/drf.h; this package holds its ctypes binding, a
mirror of the reference's solver interfaces, synthetic inputs and the build.
"""
from . import _lib  # noqa: F401
from ._lib import DrfError, load  # noqa: F401

__all__ = ["load", "DrfError"]
