"""MI355X-native ILU0-preconditioned BiCGSTAB for block-sparse systems.

The compute path is ``libfpga_amd.so`` (hand-written HIP for gfx950 behind
the C ABI of include/bda_bicgstab.h); this package is a thin host mirror of
that interface.  Importing it fails if the library is not built.
"""
from ._lib import LIB_PATH, lib  # noqa: F401  (raises if the library is missing)
from .solver import (BdaError, BicgstabSolver, BlockSystem, GridSpec, SolveResult,  # noqa: F401
                     KIND_LOWER, KIND_SPMV, KIND_UPPER, KIND_VEC, colour_pattern, debug_level, generate,
                     set_debug_level)
from . import configs  # noqa: F401

__all__ = ["BicgstabSolver", "BdaError", "BlockSystem", "GridSpec", "SolveResult", "generate",
           "colour_pattern", "set_debug_level", "debug_level", "configs", "lib", "LIB_PATH"]
