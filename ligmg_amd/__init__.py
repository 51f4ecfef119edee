"""MI355X-native graph-Laplacian algebraic multigrid (LigMG setup + solve).

The solver runs entirely in the HIP library ``libligmg_amd.so`` (C ABI in
include/ligmg_amd.h); this package is its Python binding plus the seeded
graph generators.
"""
from . import graphs  # noqa: F401
from ._ffi import (KSP_CG, KSP_FCG, PREC_FP64, PREC_MIXED, SMOOTHER_CHEBYSHEV,  # noqa: F401
                   SMOOTHER_JACOBI)
from .solver import (ConvergenceInfo, LigmgError, Options, Solver,  # noqa: F401
                     aggregate_with_strength, default_options, ligmg, trim_memory)
