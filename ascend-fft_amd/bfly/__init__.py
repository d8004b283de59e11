"""bfly -- MI355X (gfx950) FFT library: Python face of libbfly_hip.so.

Importing this package needs only ctypes and numpy; the torch entry points
(`bfly.ops`) import torch lazily.
"""
from ._lib import (BflyError, ERR_SHAPE, KIND_C2C, KIND_C2R, KIND_R2C, HIP_LIB, STUB_LIB,
                   SYMBOLS, hip, stub)

__all__ = ["BflyError", "ERR_SHAPE", "KIND_C2C", "KIND_R2C", "KIND_C2R", "HIP_LIB",
           "STUB_LIB", "SYMBOLS", "hip", "stub"]
