"""MI355X (gfx950) implementation of COOLFluiD 3's UFEM incompressible Navier-Stokes hot path.

The compute path is the native library libcf3hip.so (HIP kernels behind the C ABI of
include/cf3hip.h); this package holds its ctypes binding (cabi), the synthetic workloads (meshgen,
cases) and the build recipe (build).
"""
from . import cabi, cases, meshgen  # noqa: F401

__all__ = ["cabi", "cases", "meshgen"]
