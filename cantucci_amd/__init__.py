"""cantucci_amd: MI355X-native Mandelbulb distance estimation behind
cantucci's `Shape` trait surface.

The compute path is the gfx950 HIP library `libcantucci_hip.so` (C ABI in
include/cantucci_hip.h); this package is its host-side binding.
"""
from ._lib import CthError, LIB_PATH  # noqa: F401  (raises if the library is missing)
from .shape import Context, Mandelbulb, MandelbulbTrig, Shape, Sphere, default_context  # noqa: F401

__all__ = ["CthError", "Context", "Mandelbulb", "MandelbulbTrig", "Shape", "Sphere", "default_context", "LIB_PATH"]
