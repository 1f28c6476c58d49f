"""MI355X-native spectral shallow-water hot path (SHT pair, spectral operators,
RK4 tendency loop, REXI banded solves).

The compute lives in libsphswe.so (HIP, gfx950) behind the C ABI declared in
include/sphswe.h; `capi` is its ctypes binding.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
