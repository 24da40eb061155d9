"""coffe_amd: MI355X-native spherical-Bessel integral engine behind COFFE's API.

The compute path is libcoffe_engine.so (hand-written HIP kernels for gfx950 behind
a C-ABI, see include/coffe_engine.h).  Importing the package does not touch the
GPU; the first engine call loads the library and fails loudly if it is missing.
"""
from ._lib import EngineError, load_library  # noqa: F401

__all__ = ["EngineError", "load_library"]
