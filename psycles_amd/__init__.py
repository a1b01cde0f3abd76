"""psycles_amd: MI355X-native render path for Psycles scenes.

The product is libpsx.so (include/psx.h): HIP kernels for gfx950 plus the
host scene compiler and session runtime.  This package holds its ctypes
binding (abi), the Python mirror of the reference render contract (render)
and the synthetic benchmark scenes (scenes).
"""
from . import abi  # noqa: F401
