"""Expert drivers and error estimates over the C ABI.

=====================================  ==============================================
here                                   reference
=====================================  ==============================================
``backward_error(A, x, b)``            ``backward_error`` (src/lapack/dense_utilities.jl:214-228)
``inverse_norm_estimate(F, n)``        ``inverse_norm_estimate`` (src/lapack/expert_refinement.jl:45-60)
``reciprocal_condition(A)``            ``reciprocal_condition`` (src/lapack/expert_refinement.jl:61-66)
``refine_solve(A, F, b)``              ``refine_solve`` (src/lapack/expert_refinement.jl:89-104)
``expert_solve(A, b)``                 ``expert_solve`` (src/lapack/advanced_factorizations.jl:591-645)
=====================================  ==============================================
"""
import ctypes as C
from collections import namedtuple

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix
from .linalg import LU, DimensionMismatch, SingularException, lu

EPS = float(np.finfo(np.float64).eps)

BackwardError = namedtuple("BackwardError", "error residual")
RefineResult = namedtuple("RefineResult",
                          "x converged iterations backward_error residual_norm rcond_estimate forward_error_estimate")
ExpertResult = namedtuple("ExpertResult", "x converged iterations backward_error row_scale column_scale factor")


def _h():
    return _lib.handle()


def _out(Xd, b):
    if isinstance(b, DMatrix):
        return Xd
    X = Xd.numpy()
    return X[:, 0].copy() if np.ndim(b) == 1 else X


def backward_error(A, x, b):
    """(error, residual): R = b - A x and max |r| / (|A||x| + |b|) componentwise
    (dense_utilities.jl:214-228)."""
    Ad, Xd, Bd = as_dmatrix(A), as_dmatrix(x), as_dmatrix(b)
    if Ad.n != Xd.m or (Bd.m, Bd.n) != (Ad.m, Xd.n):
        raise DimensionMismatch("residual dimensions differ")
    R = DMatrix.empty(Bd.m, Bd.n, device=Bd.buf.device) if Bd.m and Bd.n else DMatrix.zeros(Bd.m, Bd.n)
    e = C.c_double(0)
    h = _h()
    rc = h._L.oxi_dbackward_error(h.ptr, Ad.m, Ad.n, Bd.n, Ad.ptr, Ad.ld, Xd.ptr, Xd.ld, Bd.ptr, Bd.ld, R.ptr, R.ld,
                                  C.byref(e))
    h.check(rc, "backward_error")
    return BackwardError(float(e.value), _out(R, b))


def inverse_norm_estimate(F, n=None, maxiter=8):
    """Hager-Higham estimate of ||A^-1||_1 from the LU factors F
    (expert_refinement.jl:45-60); an estimate, not a bound."""
    if not isinstance(F, LU):
        raise TypeError("inverse_norm_estimate takes an LU factorization")
    n = F.factors.n if n is None else int(n)
    if n < 0 or maxiter < 1:
        raise ValueError("invalid inverse norm estimator options")
    if n != F.factors.n:
        raise DimensionMismatch("n differs from the factorization")
    est = C.c_double(0)
    h = _h()
    rc = h._L.oxi_dinverse_norm_estimate(h.ptr, n, F.factors.ptr, F.factors.ld, F.ipiv_dev.data_ptr(), int(maxiter),
                                         C.byref(est))
    if rc == _lib.OXI_ERR_SINGULAR:
        raise SingularException(F.info or 0)
    h.check(rc, "inverse_norm_estimate")
    return float(est.value)


def reciprocal_condition(A, maxiter=8):
    """1 / (||A||_1 est ||A^-1||_1), 0 for a singular A (expert_refinement.jl:61-66)."""
    a = A.numpy() if isinstance(A, DMatrix) else np.asarray(A, dtype=np.float64)
    n = a.shape[0]
    if n == 0:
        return 1.0
    F = lu(A, check=False)
    if F.info != 0:
        return 0.0
    value = np.abs(a).sum(axis=0).max() * inverse_norm_estimate(F, n, maxiter)
    return 1.0 if value == 0 else min(1.0, 1.0 / value)


def refine_solve(A, F, b, maxiter=8, rtol=8 * EPS):
    """refine_solve(A, F, b; maxiter=8, rtol=8eps) (expert_refinement.jl:89-104):
    X = F \\ b refined against A until the componentwise backward error is
    <= rtol; with the (non-certified) condition and forward-error estimates."""
    if not isinstance(F, LU):
        raise TypeError("refine_solve takes the LU factorization of A")
    Ad, Bd = as_dmatrix(A), as_dmatrix(b)
    n = Ad.n
    if Ad.m != n or Bd.m != n or maxiter < 0 or not rtol >= 0:
        raise ValueError("invalid refinement dimensions/options")
    if F.factors.n != n:
        raise DimensionMismatch("factorization size differs")
    F._ready()
    Xd = DMatrix.empty(n, Bd.n, device=Bd.buf.device) if n and Bd.n else DMatrix.zeros(n, Bd.n)
    info = _lib.RefineInfo()
    h = _h()
    rc = h._L.oxi_drefine_solve(h.ptr, n, Bd.n, Ad.ptr, Ad.ld, F.factors.ptr, F.factors.ld, F.ipiv_dev.data_ptr(),
                                Bd.ptr, Bd.ld, Xd.ptr, Xd.ld, int(maxiter), float(rtol), C.byref(info))
    if rc == _lib.OXI_ERR_SINGULAR:
        raise SingularException(F.info or 0)
    h.check(rc, "refine_solve")
    return RefineResult(_out(Xd, b), bool(info.converged), int(info.iterations), info.backward_error,
                        info.residual_norm, info.rcond_estimate, info.forward_error_estimate)


def expert_solve(A, b, maxiter=8, rtol=8 * EPS, equilibrate=True, pivot="partial"):
    """expert_solve(A, b; maxiter=8, rtol=8eps, equilibrate=true)
    (advanced_factorizations.jl:591-645): row and column scaling by the
    largest entries, LU of the scaled matrix, refinement against A.  Returns
    (x, converged, iterations, backward_error, row_scale, column_scale,
    factor) with `factor` the LU of the scaled matrix."""
    import torch
    if pivot != "partial":
        raise NotImplementedError("expert_solve implements pivot=:partial")
    Ad, Bd = as_dmatrix(A), as_dmatrix(b)
    n = Ad.n
    if Ad.m != n or Bd.m != n or maxiter < 0 or not rtol >= 0:
        raise ValueError("invalid expert solve dimensions/options")
    dev = Ad.buf.device
    LUd = DMatrix.empty(n, n, device=dev) if n else DMatrix.zeros(0, 0)
    ipiv = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
    rows = torch.ones(max(n, 1), dtype=torch.float64, device=dev)
    cols = torch.ones(max(n, 1), dtype=torch.float64, device=dev)
    Xd = DMatrix.empty(n, Bd.n, device=dev) if n and Bd.n else DMatrix.zeros(n, Bd.n)
    info = _lib.RefineInfo()
    h = _h()
    rc = h._L.oxi_dexpert_solve(h.ptr, n, Bd.n, Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, Xd.ptr, Xd.ld, int(maxiter),
                                float(rtol), 1 if equilibrate else 0, LUd.ptr, LUd.ld, ipiv.data_ptr(),
                                rows.data_ptr(), cols.data_ptr(), C.byref(info))
    if rc == _lib.OXI_ERR_SINGULAR:
        import re
        m = re.search(r"U\((\d+),", h.last_error())
        raise SingularException(int(m.group(1)) if m else 0)
    h.check(rc, "expert_solve")
    F = LU(LUd, ipiv[:n], 0)
    return ExpertResult(_out(Xd, b), bool(info.converged), int(info.iterations), info.backward_error,
                        rows[:n].cpu().numpy(), cols[:n].cpu().numpy(), F)
