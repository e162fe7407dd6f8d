"""meturay_amd — MI355X-native wavefront path-tracing core.

The compute path is libmrh.so (hand-written HIP kernels for gfx950 behind the
C-ABI in include/mrh.h); this package only binds it.  See native.py.
"""
from . import native  # noqa: F401

__all__ = ["native"]
