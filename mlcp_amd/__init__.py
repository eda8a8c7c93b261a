"""MI355X-native intertwined power/ALS path of MLCP.

The product is libmlcp_hip.so (HIP kernels for gfx950 behind the C-ABI in
include/mlcp_hip.h).  This package holds its build recipe, the ctypes
binding, the term-list/CP-block host types and the host-side mirror of the
reference's Fortran GPU wrappers and solver cycle.
"""
from .hoplist import CPBlock, HopList  # noqa: F401
