"""oxiblas_mi355x -- MI355X (gfx950) implementation of OxiBLAS.jl's level-3
hot path: gemm!, the blocked LU with partial pivoting and its solves, the
mixed-precision LU with FP64 refinement (IR / GMRES-IR) and the adjoints of
X = A \\ B and gemm!.

Every computation runs in liboxiblas_mi355x.so (hand-written HIP kernels for
CDNA4) through its C ABI; this package is the host-side mirror of the
reference's API used by tests and benchmarks.
"""
from ._lib import load, handle, Handle, OxiError, LIB_PATH, SIGNATURES  # noqa: F401
from .device import DMatrix, as_dmatrix  # noqa: F401
from .linalg import (  # noqa: F401
    gemm_, trsm_, laswp_, lu, lu_, lu_factor, LU, ldiv, ldiv_, solve, solve_transposed,
    transpose, adjoint, SingularException, DimensionMismatch, det, logdet, logabsdet, inv, factorize,
)
from .expert import (  # noqa: F401
    backward_error, inverse_norm_estimate, reciprocal_condition, refine_solve, expert_solve,
)
from .mixed import (  # noqa: F401
    mixed_solve, equilibrate, sgemm_, gemm_lowp_update_, lu_lowp, LUFactor32, MixedSolveResult, precision_code,
)
from .blas3 import DBatch, gemm_strided_batched_, gemm_batched_, syrk_, herk_, hgemm_  # noqa: F401
from .cholesky import Cholesky, PosDefException, cholesky, cholesky_, cholesky_factor  # noqa: F401
from .pullbacks import ldiv_pullback, rrule_ldiv, gemm_pullback, rrule_mul  # noqa: F401

__all__ = [
    "load", "handle", "Handle", "OxiError", "DMatrix", "as_dmatrix",
    "gemm_", "trsm_", "laswp_", "lu", "lu_", "lu_factor", "LU", "ldiv", "ldiv_", "solve", "solve_transposed",
    "transpose", "adjoint", "SingularException", "DimensionMismatch",
    "mixed_solve", "equilibrate", "sgemm_", "gemm_lowp_update_", "lu_lowp", "LUFactor32", "MixedSolveResult",
    "precision_code", "ldiv_pullback", "rrule_ldiv", "gemm_pullback", "rrule_mul",
    "DBatch", "gemm_strided_batched_", "gemm_batched_", "syrk_", "herk_", "hgemm_",
    "Cholesky", "PosDefException", "cholesky", "cholesky_", "cholesky_factor",
    "det", "logdet", "logabsdet", "inv", "factorize",
    "backward_error", "inverse_norm_estimate", "reciprocal_condition", "refine_solve", "expert_solve",
]
