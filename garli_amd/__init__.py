"""MI355X-native GARLI likelihood engine.

The product is libgle.so (garli_amd/csrc, C ABI in include/gle.h): hand-written
gfx950 HIP kernels behind GARLI's tree-scoring surface.  This package holds its
ctypes bindings (engine), the seeded workload simulator (sim) and the build
script (build).
"""
from . import engine  # noqa: F401
from .engine import Instance, Tree, PREC_FP32, PREC_FP64  # noqa: F401
