"""Mixed-precision LU solve, equilibration and the low-precision building
blocks, over the C ABI.

===================================  ==========================================
here                                 reference
===================================  ==========================================
``mixed_solve(A, b; ...)``           ``mixed_solve`` (src/lapack/factorizations.jl:419-427,
                                     src/lapack/extended_lapack.jl:175-226)
``equilibrate(A; radix)``            ``equilibrate`` (src/lapack/dense_utilities.jl:191-212)
``sgemm_``                           ``gemm!`` for Float32 (src/blas/gemm.jl:625-703)
``lu_lowp`` / ``LUFactor32``         ``lu_factor`` of the Float32 low matrix
                                     (extended_lapack.jl:193), trailing products in
                                     BF16 / FP16 / FP8 on top of the reference
``gemm_lowp_update_``                the low-precision trailing product on its own
===================================  ==========================================

``lowtype`` selects the precision of the factorization's trailing products:
"float32" (the reference's default low type for Float64), "bfloat16",
"float16" or "float8_e4m3fn" (torch dtypes are accepted too); the panel, the
triangular solves and the factor stay FP32, the residual and X stay FP64.
"""
import ctypes as C
import re
from collections import namedtuple

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix
from .linalg import SingularException, DimensionMismatch, _trans

EPS = float(np.finfo(np.float64).eps)

_PREC_NAMES = {
    "float32": _lib.OXI_PREC_FP32, "fp32": _lib.OXI_PREC_FP32,
    "bfloat16": _lib.OXI_PREC_BF16, "bf16": _lib.OXI_PREC_BF16,
    "float16": _lib.OXI_PREC_FP16, "fp16": _lib.OXI_PREC_FP16, "half": _lib.OXI_PREC_FP16,
    "float8_e4m3fn": _lib.OXI_PREC_FP8, "fp8": _lib.OXI_PREC_FP8, "e4m3": _lib.OXI_PREC_FP8,
}
_PREC_LABEL = {_lib.OXI_PREC_FP32: "Float32", _lib.OXI_PREC_BF16: "BFloat16",
               _lib.OXI_PREC_FP16: "Float16", _lib.OXI_PREC_FP8: "Float8_E4M3FN"}
_SOLVERS = {"ir": 0, "gmres": 1, "gmres_ir": 1, "auto": 2}
_PIVOTS = {"partial": 0, "rbt": 1, "tournament": 2}
_REASONS = ("converged", "iteration_limit", "nonfinite_iteration", "working_factor_fallback")

MixedSolveResult = namedtuple(
    "MixedSolveResult",
    "x converged iterations residual backward_error metric low_type fallback factorization finite reason "
    "normwise_backward_error inner_iterations solver factor_ms refine_ms")
Equilibration = namedtuple("Equilibration", "matrix row_scale column_scale")


def precision_code(lowtype):
    """Map a low type (name or torch dtype) to OXI_PREC_*."""
    if lowtype is None:
        return _lib.OXI_PREC_FP32
    if isinstance(lowtype, int):
        if lowtype in _PREC_LABEL:
            return lowtype
        raise ValueError(f"unknown precision code {lowtype}")
    name = str(lowtype).replace("torch.", "").lower()
    if name not in _PREC_NAMES:
        raise ValueError("low precision type must be float32, bfloat16, float16 or float8_e4m3fn")
    return _PREC_NAMES[name]


def _h():
    return _lib.handle()


def _singular_from(h):
    m = re.search(r"U\((\d+),", h.last_error())
    return SingularException(int(m.group(1)) if m else 0)


def mixed_solve(A, b, maxiter=10, rtol=1e-13, lowtype=None, fallback=True, factorization="lu", solver="ir",
                gmres_restart=0, threaded=True, pivot="partial"):
    """mixed_solve(A, b; maxiter=10, rtol=1e-13, lowtype, fallback=true,
    factorization=:lu) (the Float64 method, src/lapack/factorizations.jl:425-427,
    over src/lapack/extended_lapack.jl:175-226).

    Radix equilibration, a low-precision LU, then refinement in FP64 until the
    componentwise backward error is <= rtol.  solver "ir" is the reference's
    classical iteration; "gmres" runs GMRES-IR (the low factors as an FP64
    preconditioner) for condition numbers where the classical iteration
    stalls; "auto" switches to GMRES-IR once the error stops halving.
    pivot "partial" factors with partial pivoting as the reference; "rbt"
    applies a random butterfly transform and factors without pivoting (no
    device-wide pivot search per column; refinement restores the accuracy).
    "rbt" applies to the Float32 factor; other low types, n not a multiple of
    4 or a zero pivot use partial pivoting.  "tournament" chooses each
    32-column leaf's pivots by a tournament over its row blocks (one exchange
    per leaf instead of one per column; panels above 16384 rows pivot
    partially).  Neither option has a reference counterpart.  Returns the reference's NamedTuple fields (x, converged, iterations,
    residual, backward_error, metric, low_type, fallback, factorization,
    finite, reason) plus the GPU's normwise error, GMRES iteration count and
    device timings."""
    if factorization not in ("lu", "cholesky"):
        raise NotImplementedError("the GPU mixed solve implements factorization=:lu and :cholesky")
    if not fallback:
        raise NotImplementedError("the GPU mixed solve always falls back to the working-precision factor")
    if maxiter < 0 or not np.isfinite(rtol) or rtol < 0:
        raise ValueError("invalid mixed solve options")
    if factorization == "cholesky":
        return _mixed_cholesky(A, b, maxiter, rtol, lowtype)
    if solver not in _SOLVERS:
        raise ValueError("solver must be ir, gmres or auto")
    if pivot not in _PIVOTS:
        raise ValueError("pivot must be partial, rbt or tournament")
    prec = precision_code(lowtype)
    vec = not isinstance(b, DMatrix) and np.ndim(b) == 1
    Ad, Bd = as_dmatrix(A), as_dmatrix(b)
    if Ad.m != Ad.n:
        raise DimensionMismatch("mixed factorization dimensions differ")
    if Bd.m != Ad.m:
        raise ValueError("invalid mixed solve options")
    n, nrhs = Ad.n, Bd.n
    Xd = DMatrix.empty(n, nrhs, device=Ad.buf.device) if n and nrhs else DMatrix.zeros(n, nrhs)
    info = _lib.IRInfo()
    h = _h()
    rc = h._L.oxi_dgesv_mixed(h.ptr, prec, _SOLVERS[solver], n, nrhs, Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, Xd.ptr, Xd.ld,
                              int(maxiter), float(rtol), int(gmres_restart), _PIVOTS[pivot], C.byref(info))
    if rc == _lib.OXI_ERR_NOT_FINITE:
        raise ValueError("finite mixed input required")
    if rc == _lib.OXI_ERR_SINGULAR:
        raise _singular_from(h)
    if rc == _lib.OXI_ERR_OVERFLOW:
        raise OverflowError(h.last_error())
    h.check(rc, "mixed_solve")
    if isinstance(b, DMatrix):
        x = Xd
    else:
        x = Xd.numpy()
        if vec:
            x = x[:, 0].copy()
    berr = info.backward_error
    return MixedSolveResult(
        x=x, converged=bool(info.converged), iterations=int(info.iterations), residual=info.residual_norm,
        backward_error=berr, metric="componentwise_backward_error", low_type=_PREC_LABEL[prec],
        fallback=bool(info.fallback), factorization="lu", finite=info.reason != 2,
        reason=_REASONS[info.reason], normwise_backward_error=info.normwise_backward_error,
        inner_iterations=int(info.inner_iterations), solver="gmres" if info.solver else "ir",
        factor_ms=info.factor_ms, refine_ms=info.refine_ms)


def _mixed_cholesky(A, b, maxiter, rtol, lowtype):
    """mixed_solve(A, b; factorization=:cholesky): one global scale, the FP32
    Cholesky (FP64 on PosDefException), classical refinement
    (src/lapack/extended_lapack.jl:175-226)."""
    from .cholesky import PosDefException
    prec = precision_code(lowtype)
    if prec != _lib.OXI_PREC_FP32:
        raise ValueError("the Cholesky mixed solve takes Float32 low factors")
    vec = not isinstance(b, DMatrix) and np.ndim(b) == 1
    Ad, Bd = as_dmatrix(A), as_dmatrix(b)
    if Ad.m != Ad.n:
        raise DimensionMismatch("mixed factorization dimensions differ")
    if Bd.m != Ad.m:
        raise ValueError("invalid mixed solve options")
    n, nrhs = Ad.n, Bd.n
    Xd = DMatrix.empty(n, nrhs, device=Ad.buf.device) if n and nrhs else DMatrix.zeros(n, nrhs)
    info = _lib.IRInfo()
    h = _h()
    rc = h._L.oxi_dposv_mixed(h.ptr, prec, n, nrhs, Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, Xd.ptr, Xd.ld, int(maxiter),
                              float(rtol), C.byref(info))
    if rc == _lib.OXI_ERR_NOT_FINITE:
        raise ValueError("finite mixed input required")
    if rc == _lib.OXI_ERR_NOT_POSDEF:
        m = re.search(r"A\((\d+),", h.last_error())
        raise PosDefException(int(m.group(1)) if m else 0)
    if rc == _lib.OXI_ERR_OVERFLOW:
        raise OverflowError(h.last_error())
    h.check(rc, "mixed_solve")
    if isinstance(b, DMatrix):
        x = Xd
    else:
        x = Xd.numpy()
        if vec:
            x = x[:, 0].copy()
    return MixedSolveResult(
        x=x, converged=bool(info.converged), iterations=int(info.iterations), residual=info.residual_norm,
        backward_error=info.backward_error, metric="componentwise_backward_error", low_type=_PREC_LABEL[prec],
        fallback=bool(info.fallback), factorization="cholesky", finite=info.reason != 2,
        reason=_REASONS[info.reason], normwise_backward_error=info.normwise_backward_error,
        inner_iterations=0, solver="ir", factor_ms=info.factor_ms, refine_ms=info.refine_ms)


def equilibrate(A, radix=False):
    """equilibrate(A; radix=false) -> (matrix, row_scale, column_scale)
    (src/lapack/dense_utilities.jl:191-212)."""
    import torch
    Ad = as_dmatrix(A)
    m, n = Ad.m, Ad.n
    E = DMatrix.empty(m, n, device=Ad.buf.device)
    rows = torch.empty(max(m, 1), dtype=torch.float64, device=Ad.buf.device)
    cols = torch.empty(max(n, 1), dtype=torch.float64, device=Ad.buf.device)
    h = _h()
    rc = h._L.oxi_dequilibrate(h.ptr, m, n, Ad.ptr, Ad.ld, E.ptr, E.ld, rows.data_ptr(), cols.data_ptr(),
                               1 if radix else 0)
    if rc == _lib.OXI_ERR_NOT_FINITE:
        raise ValueError("equilibrate: non-finite entry")
    h.check(rc, "equilibrate")
    if isinstance(A, DMatrix):
        return Equilibration(E, rows[:m], cols[:n])
    return Equilibration(E.numpy(), rows[:m].cpu().numpy(), cols[:n].cpu().numpy())


# ---------------------------------------------------------------------------
# FP32 and low-precision building blocks
# ---------------------------------------------------------------------------
def sgemm_(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """gemm! for Float32 device matrices (src/blas/gemm.jl:625-703)."""
    ta, tb = _trans(transA), _trans(transB)
    am, ak = (A.n, A.m) if ta != "N" else (A.m, A.n)
    bk, bn = (B.n, B.m) if tb != "N" else (B.m, B.n)
    if ak != bk or (C_.m, C_.n) != (am, bn):
        raise DimensionMismatch("GEMM dimensions differ")
    h = _h()
    rc = h._L.oxi_sgemm(h.ptr, ta.encode(), tb.encode(), C_.m, C_.n, ak, float(alpha), A.ptr, A.ld, B.ptr, B.ld,
                        float(beta), C_.ptr, C_.ld)
    h.check(rc, "gemm!")
    return C_


def gemm_lowp_update_(C_, A, B, lowtype="bfloat16", alpha=-1.0):
    """C <- C + alpha * round(A) * round(B) with A, B rounded to `lowtype`
    (FP8 with one power-of-two scale per 256-row block of A and 256-column
    block of B), FP32 accumulation.  All three are FP32 device matrices."""
    if A.n != B.m or (C_.m, C_.n) != (A.m, B.n):
        raise DimensionMismatch("GEMM dimensions differ")
    prec = precision_code(lowtype)
    h = _h()
    rc = h._L.oxi_gemm_lowp_update(h.ptr, prec, A.m, B.n, A.n, float(alpha), A.ptr, A.ld, B.ptr, B.ld, C_.ptr,
                                   C_.ld)
    h.check(rc, "gemm_lowp_update")
    return C_


class LUFactor32:
    """FP32 factors of ``lu_lowp`` (LUFactor{Float32} fields: factors, pivots, info)."""

    def __init__(self, factors, ipiv_dev, info, lowtype):
        self.factors = factors
        self.ipiv_dev = ipiv_dev
        self.info = int(info)
        self.lowtype = lowtype

    @property
    def pivots(self):
        return self.ipiv_dev.cpu().numpy() if self.ipiv_dev.numel() else np.zeros(0, dtype=np.int64)

    def solve(self, B, trans="N"):
        """FP32 solve (ldiv! on a copy) with these factors."""
        t = _trans(trans)
        Bd = as_dmatrix(B, dtype=np.float32) if not isinstance(B, DMatrix) else B.copy()
        if Bd.m != self.factors.n:
            raise DimensionMismatch("solve dimensions differ")
        h = _h()
        rc = h._L.oxi_sgetrs(h.ptr, t.encode(), self.factors.n, Bd.n, self.factors.ptr, self.factors.ld,
                             self.ipiv_dev.data_ptr(), Bd.ptr, Bd.ld)
        if rc == _lib.OXI_ERR_SINGULAR:
            raise SingularException(self.info or 0)
        h.check(rc, "solve")
        if isinstance(B, DMatrix):
            return Bd
        out = Bd.numpy()
        return out[:, 0].copy() if np.ndim(B) == 1 else out


def lu_lowp(A, lowtype="float32", check=True, blocksize=0, pivot="partial"):
    """FP32 blocked LU with partial pivoting whose trailing products run in
    `lowtype`; A is copied (a float32 device matrix or any host array).
    pivot "tournament" chooses each 32-column leaf's pivots by a tournament
    over its row blocks (oxi_sgetrf_tournament; no reference counterpart)."""
    if pivot not in ("partial", "tournament"):
        raise ValueError("pivot must be partial or tournament")
    import torch
    Ad = A.copy() if isinstance(A, DMatrix) else DMatrix.from_numpy(A, dtype=np.float32)
    q = min(Ad.m, Ad.n)
    ipiv = torch.zeros(max(q, 1), dtype=torch.int64, device=Ad.buf.device)
    info = C.c_int64(0)
    h = _h()
    fn = h._L.oxi_sgetrf_lowp if pivot == "partial" else h._L.oxi_sgetrf_tournament
    rc = fn(h.ptr, precision_code(lowtype), Ad.m, Ad.n, Ad.ptr, Ad.ld, ipiv.data_ptr(), int(blocksize), C.byref(info))
    h.check(rc, "lu")
    if check and info.value != 0:
        raise SingularException(info.value)
    return LUFactor32(Ad, ipiv[:q], info.value, lowtype)
