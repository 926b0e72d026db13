"""fastmulticolsort_amd -- MI355X-native multi-column sorting path.

Hand-written HIP kernels for gfx950 behind a C ABI (include/fmcs.h,
libfmcs.so), with the reference's composer / plan API on the host.
"""
from . import _lib  # noqa: F401

__all__ = ["engine", "datagen", "planner", "learned", "columnar", "query", "autotune"]
