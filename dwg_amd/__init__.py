"""dwg_amd — MI355X-native DWG (diffusing winding gradients) reconstruction.

The compute path is libdwg_hip.so (HIP kernels for gfx950 behind the C ABI
in include/dwg_hip.h); this package holds the Python mirror of that ABI
(``gpu``), the DWG driver (``pipeline``), the seeded workload generators
(``shapes``) and quality metrics (``metrics``).
"""
from . import _lib, shapes  # noqa: F401

__all__ = ["_lib", "shapes"]
