"""MI355X-native (gfx950) OFRR projection pipeline.

The compute path is libofrr.so (hand-written HIP, C-ABI in include/ofrr.h);
this package binds it for Python.  Importing the package does not touch the
GPU; the first Context() loads the library and fails loudly when it is absent.
"""
from ._lib import BF16, F16, F32, F64, FP8, LIB_PATH, load  # noqa: F401

__all__ = ["api", "load", "LIB_PATH", "F64", "F32", "F16", "BF16", "FP8"]
