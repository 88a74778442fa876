"""MI355X (gfx950) implementation of gpu_ppm's pulse-profile hot path.

The compute path is libppm.so (HIP kernels + C++20 host behind the C ABI in
include/ppm.h); this package is its Python mirror for tests and benchmarks.
"""
from . import abi, scenarios  # noqa: F401
from .lib import Context, PpmError, load  # noqa: F401

__all__ = ["abi", "scenarios", "Context", "PpmError", "load"]
