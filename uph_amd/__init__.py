"""MI355X-native SPH hot path of Uph: HIP kernels for gfx950 behind a C ABI.

Modules:
  scenes  -- scene builders and the reference's 12-column deck format
  solver  -- ctypes mirror of the C ABI (SearchNeighbors, the stepping API)
  build   -- in-tree build of libuph_gpu.so
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "solver", "build"]
