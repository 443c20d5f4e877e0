"""ctypes wrappers over the oracle library ``liboxo.so`` (test infrastructure).

Arrays are numpy, Fortran-ordered (column-major) as the reference's Julia
matrices; pivots are int64, 1-based global rows as the reference's
``pivots::Vector{Int}`` (src/lapack/factorizations.jl:118).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboxo.so")
_lib = None

i64 = C.c_int64
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int64)


def build():
    """Compile liboxo.so from oxo.cpp with the Makefile next to it."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def available():
    return os.path.exists(_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            build()
        L = C.CDLL(_LIB_PATH)
        ch = C.c_char
        L.oxo_dgemm.argtypes = [ch, ch, i64, i64, i64, C.c_double, dp, i64, dp, i64, C.c_double, dp, i64, C.c_int]
        L.oxo_sgemm.argtypes = [ch, ch, i64, i64, i64, C.c_float, fp, i64, fp, i64, C.c_float, fp, i64, C.c_int]
        L.oxo_dtrsm.argtypes = [ch, ch, ch, ch, i64, i64, C.c_double, dp, i64, dp, i64, i64, C.c_int]
        L.oxo_dtrsv.argtypes = [ch, ch, ch, i64, dp, i64, dp]
        L.oxo_dgetrf.argtypes = [i64, i64, dp, i64, ip, i64, C.c_int, ip, C.c_int]
        L.oxo_sgetrf.argtypes = [i64, i64, fp, i64, ip, i64, C.c_int, ip, C.c_int]
        L.oxo_sgetrf_tp.argtypes = [i64, i64, fp, i64, ip, i64, C.c_int, ip, C.c_int]
        L.oxo_dlaswp.argtypes = [i64, dp, i64, i64, i64, ip, C.c_int]
        L.oxo_dgetri.argtypes = [i64, dp, i64, ip, dp, i64, C.c_int]
        L.oxo_dlogabsdet.argtypes = [i64, dp, i64, ip, dp, dp]
        L.oxo_dlogabsdet.restype = None
        L.oxo_dbackward_error.argtypes = [i64, i64, i64, dp, i64, dp, i64, dp, i64, dp, i64]
        L.oxo_dbackward_error.restype = C.c_double
        L.oxo_dinverse_norm_estimate.argtypes = [i64, dp, i64, ip, C.c_int]
        L.oxo_dinverse_norm_estimate.restype = C.c_double
        L.oxo_drefine_solve.argtypes = [i64, i64, dp, i64, dp, i64, ip, dp, i64, dp, i64, C.c_int, C.c_double, dp]
        L.oxo_dexpert_solve.argtypes = [i64, i64, dp, i64, dp, i64, dp, i64, C.c_int, C.c_double, C.c_int, dp, dp,
                                        dp, C.c_int]
        L.oxo_dsyrk.argtypes = [ch, ch, i64, i64, C.c_double, dp, i64, C.c_double, dp, i64]
        L.oxo_dpotrf.argtypes = [i64, dp, i64, i64, ip, C.c_int]
        L.oxo_dpotrs.argtypes = [i64, i64, dp, i64, dp, i64, C.c_int, C.c_int]
        L.oxo_dgetrs.argtypes = [ch, i64, i64, dp, i64, ip, dp, i64, C.c_int, C.c_int]
        L.oxo_sgetrs.argtypes = [ch, i64, i64, fp, i64, ip, fp, i64, C.c_int, C.c_int]
        L.oxo_dequilibrate.argtypes = [i64, i64, dp, i64, dp, dp, C.c_int]
        L.oxo_dequilibrate.restype = None
        L.oxo_refinement_residual.argtypes = [i64, i64, dp, i64, dp, i64, dp, i64, dp, i64]
        L.oxo_refinement_residual.restype = C.c_double
        L.oxo_mixed_solve.argtypes = [i64, i64, dp, i64, dp, i64, dp, i64, C.c_int, C.c_double, C.c_int,
                                      ip, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
        L.oxo_mixed_solve_chol.argtypes = L.oxo_mixed_solve.argtypes
        L.oxo_mixed_solve_chol.restype = C.c_int
        for f in ("oxo_dgemm", "oxo_sgemm", "oxo_dtrsm", "oxo_dtrsv", "oxo_dgetrf", "oxo_sgetrf", "oxo_sgetrf_tp",
                  "oxo_dgetrs", "oxo_sgetrs", "oxo_dlaswp", "oxo_dsyrk", "oxo_dpotrf", "oxo_dpotrs", "oxo_dgetri",
                  "oxo_drefine_solve", "oxo_dexpert_solve", "oxo_mixed_solve", "oxo_max_threads"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def max_threads():
    return lib().oxo_max_threads()


def _f(a, dtype=np.float64):
    """Fortran-ordered array of `dtype` (copy only when needed)."""
    a = np.asarray(a, dtype=dtype)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)


def _ptr(a, t=dp):
    return a.ctypes.data_as(t)


def _ld(a):
    return max(1, a.strides[1] // a.itemsize) if a.shape[1] > 1 else max(1, a.shape[0])


def _check(rc, what):
    if rc < 0:
        raise ValueError(f"{what}: invalid argument {-rc}")
    return rc


def dgemm(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N", threads=1):
    """C <- alpha*op(A)*op(B) + beta*C in place (src/blas/gemm.jl:635-703)."""
    assert C_.flags.f_contiguous and C_.dtype == np.float64
    A = _f(A); B = _f(B)
    m, n = C_.shape
    k = A.shape[0] if transA.upper() in "TC" else A.shape[1]
    rc = lib().oxo_dgemm(transA.encode(), transB.encode(), m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B),
                         beta, _ptr(C_), max(1, m), threads)
    _check(rc, "dgemm")
    return C_


def sgemm(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N", threads=1):
    assert C_.flags.f_contiguous and C_.dtype == np.float32
    A = _f(A, np.float32); B = _f(B, np.float32)
    m, n = C_.shape
    k = A.shape[0] if transA.upper() in "TC" else A.shape[1]
    rc = lib().oxo_sgemm(transA.encode(), transB.encode(), m, n, k, alpha, _ptr(A, fp), _ld(A), _ptr(B, fp),
                         _ld(B), beta, _ptr(C_, fp), max(1, m), threads)
    _check(rc, "sgemm")
    return C_


def dtrsm(B, A, alpha=1.0, side="L", uplo="U", trans="N", unit=False, blocksize=8, threads=1):
    """B <- alpha*op(A)^-1*B (side L) or alpha*B*op(A)^-1 (side R), in place
    (src/blas/level3.jl:518-544).  Returns 0, or the 1-based index of a zero
    diagonal (SingularException) with B untouched."""
    assert B.flags.f_contiguous and B.dtype == np.float64
    A = _f(A)
    m, n = B.shape
    rc = lib().oxo_dtrsm(side.encode(), uplo.encode(), trans.encode(), b"U" if unit else b"N", m, n, alpha,
                         _ptr(A), _ld(A), _ptr(B), max(1, m), blocksize, threads)
    return _check(rc, "dtrsm")


def dtrsv(x, A, uplo="U", trans="N", unit=False):
    """x <- op(A)^-1 x in place (src/blas/level2.jl:436-472)."""
    assert x.flags.c_contiguous and x.dtype == np.float64 and x.ndim == 1
    A = _f(A)
    rc = lib().oxo_dtrsv(uplo.encode(), trans.encode(), b"U" if unit else b"N", x.shape[0], _ptr(A), _ld(A),
                         _ptr(x))
    return _check(rc, "dtrsv")


def dgetrf(A, blocksize=128, check=False, threads=1):
    """Blocked LU with partial pivoting on a copy of A (lu_factor,
    src/lapack/factorizations.jl:17-77).  Returns (factors, ipiv, info)."""
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = F.shape
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int64)
    info = C.c_int64(0)
    rc = lib().oxo_dgetrf(m, n, _ptr(F), max(1, m), _ptr(ipiv, ip), blocksize, 1 if check else 0,
                          C.byref(info), threads)
    _check(rc, "dgetrf")
    return F, ipiv[:min(m, n)], info.value


def sgetrf(A, blocksize=128, check=False, threads=1, tournament=False):
    """FP32 lu_factor; tournament=True: the leaves of oxi_sgetrf_tournament."""
    F = np.array(A, dtype=np.float32, order="F", copy=True)
    m, n = F.shape
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int64)
    info = C.c_int64(0)
    fn = lib().oxo_sgetrf_tp if tournament else lib().oxo_sgetrf
    rc = fn(m, n, _ptr(F, fp), max(1, m), _ptr(ipiv, ip), blocksize, 1 if check else 0, C.byref(info), threads)
    _check(rc, "sgetrf")
    return F, ipiv[:min(m, n)], info.value


def dlaswp(A, k1, k2, ipiv, inverse=False):
    """Interchanges k1..k2 (1-based) of ipiv on every column of A, in place
    (_lu_swap!, src/lapack/factorizations.jl:80-88; reverse order for the
    inverse as _rowswap!, src/api/linalg_api.jl:100-110)."""
    assert A.flags.f_contiguous and A.dtype == np.float64 and A.ndim == 2
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int64)
    rc = lib().oxo_dlaswp(A.shape[1], _ptr(A), max(1, A.shape[0]), k1, k2, _ptr(ipiv, ip), 1 if inverse else 0)
    _check(rc, "dlaswp")
    return A


def dsyrk(C_, A, alpha=1.0, beta=0.0, uplo="U", trans="N"):
    """syrk!(C, A, alpha, beta; uplo, trans) on one triangle, in place
    (src/blas/level3.jl:546-783)."""
    assert C_.flags.f_contiguous and C_.dtype == np.float64
    A = _f(A)
    n = C_.shape[0]
    k = A.shape[0] if trans.upper() in "TC" else A.shape[1]
    rc = lib().oxo_dsyrk(uplo.encode(), trans.encode(), n, k, alpha, _ptr(A), _ld(A), beta, _ptr(C_), max(1, n))
    _check(rc, "dsyrk")
    return C_


def dpotrf(A, leaf=64, threads=1):
    """Cholesky of the lower triangle of a copy of A (_cholesky_factor_info!,
    src/lapack/factorizations.jl:246-286) -> (L, info)."""
    L = np.array(A, dtype=np.float64, order="F", copy=True)
    n = L.shape[0]
    info = C.c_int64(0)
    rc = lib().oxo_dpotrf(n, _ptr(L), max(1, n), leaf, C.byref(info), threads)
    _check(rc, "dpotrf")
    return L, info.value


def dpotrs(L, B, threads=1):
    """solve(F::CholeskyFactor, b) (src/lapack/factorizations.jl:288-300)."""
    vec = np.ndim(B) == 1
    X = _f(np.array(B, dtype=np.float64, copy=True))
    L = _f(L)
    n = L.shape[0]
    rc = lib().oxo_dpotrs(n, X.shape[1], _ptr(L), _ld(L), _ptr(X), max(1, n), 1 if vec else 0, threads)
    _check(rc, "dpotrs")
    return X[:, 0].copy() if vec else X


def hgemm(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """gemm! for Float16 (src/blas/extended_blas.jl:39-88), restated in numpy:
    alpha and beta rounded to Float16, each element accumulated in Float32
    over p = 1..k in order (tile += a*b), then Float16(a*acc + b*Float32(C));
    beta == 0 never reads C.  Returns a new float16 array."""
    f16, f32 = np.float16, np.float32
    a16, b16 = f16(alpha), f16(beta)
    Aop = np.asarray(A, dtype=f16); Bop = np.asarray(B, dtype=f16)
    Aop = Aop.T if transA.upper() in "TC" else Aop
    Bop = Bop.T if transB.upper() in "TC" else Bop
    m, k = Aop.shape
    n = Bop.shape[1]
    if a16 == 0 or k == 0:
        return np.zeros((m, n), f16) if b16 == 0 else (f32(b16) * np.asarray(C_, f16).astype(f32)).astype(f16)
    acc = np.zeros((m, n), f32)
    A32, B32 = Aop.astype(f32), Bop.astype(f32)
    for p in range(k):
        acc = acc + np.outer(A32[:, p], B32[p, :]).astype(f32)
    v = f32(a16) * acc
    if b16 != 0:
        v = v + f32(b16) * np.asarray(C_, f16).astype(f32)
    return v.astype(f16)


def dgetri(F, ipiv, threads=1):
    """native_inverse from LU factors (src/lapack/dense_utilities.jl:14-36)."""
    F = _f(F)
    n = F.shape[0]
    out = np.zeros((n, n), order="F")
    rc = lib().oxo_dgetri(n, _ptr(F), _ld(F), _ptr(np.ascontiguousarray(ipiv, dtype=np.int64), ip), _ptr(out),
                          max(1, n), threads)
    _check(rc, "dgetri")
    if rc > 0:
        raise np.linalg.LinAlgError(f"singular U at {rc}")
    return out


def dlogabsdet(F, ipiv):
    """logabsdet(F::LU) (src/api/linalg_api.jl:185-191) -> (log|det|, sign)."""
    F = _f(F)
    la, sg = C.c_double(0), C.c_double(0)
    lib().oxo_dlogabsdet(F.shape[0], _ptr(F), _ld(F), _ptr(np.ascontiguousarray(ipiv, dtype=np.int64), ip),
                         C.byref(la), C.byref(sg))
    return la.value, sg.value


def backward_error(A, X, B):
    """backward_error(A, x, b) (src/lapack/dense_utilities.jl:214-228) -> (error, residual)."""
    vec = np.ndim(B) == 1
    A = _f(A); Xf = _f(X); Bf = _f(B)
    m = A.shape[0]
    R = np.zeros((m, Bf.shape[1]), order="F")
    e = lib().oxo_dbackward_error(m, A.shape[1], Bf.shape[1], _ptr(A), _ld(A), _ptr(Xf), _ld(Xf), _ptr(Bf), _ld(Bf),
                                  _ptr(R), max(1, m))
    return e, (R[:, 0].copy() if vec else R)


def inverse_norm_estimate(F, ipiv, maxiter=8):
    """inverse_norm_estimate (src/lapack/expert_refinement.jl:45-60)."""
    F = _f(F)
    return lib().oxo_dinverse_norm_estimate(F.shape[0], _ptr(F), _ld(F),
                                            _ptr(np.ascontiguousarray(ipiv, dtype=np.int64), ip), maxiter)


def refine_solve(A, F, ipiv, b, maxiter=8, rtol=8 * np.finfo(float).eps):
    """refine_solve (src/lapack/expert_refinement.jl:89-104) -> dict."""
    vec = np.ndim(b) == 1
    A = _f(A); F = _f(F); B = _f(b)
    n = A.shape[0]
    X = np.zeros((n, B.shape[1]), order="F")
    out = np.zeros(6)
    rc = lib().oxo_drefine_solve(n, B.shape[1], _ptr(A), _ld(A), _ptr(F), _ld(F),
                                 _ptr(np.ascontiguousarray(ipiv, dtype=np.int64), ip), _ptr(B), _ld(B), _ptr(X),
                                 max(1, n), maxiter, rtol, _ptr(out))
    _check(rc, "refine_solve")
    return dict(x=X[:, 0].copy() if vec else X, iterations=int(out[0]), converged=bool(out[1]),
                backward_error=out[2], residual_norm=out[3], rcond_estimate=out[4], forward_error_estimate=out[5])


def expert_solve(A, b, maxiter=8, rtol=8 * np.finfo(float).eps, equilibrate=True, threads=1):
    """expert_solve (src/lapack/advanced_factorizations.jl:591-645) -> dict."""
    vec = np.ndim(b) == 1
    A = _f(A); B = _f(b)
    n = A.shape[0]
    X = np.zeros((n, B.shape[1]), order="F")
    rows = np.ones(max(n, 1)); cols = np.ones(max(n, 1)); out = np.zeros(3)
    rc = lib().oxo_dexpert_solve(n, B.shape[1], _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(X), max(1, n), maxiter, rtol,
                                 1 if equilibrate else 0, _ptr(rows), _ptr(cols), _ptr(out), threads)
    if rc != 0:
        raise np.linalg.LinAlgError(f"singular ({rc})")
    return dict(x=X[:, 0].copy() if vec else X, iterations=int(out[0]), converged=bool(out[1]),
                backward_error=out[2], row_scale=rows[:n], column_scale=cols[:n])


def dgetrs(F, ipiv, B, trans="N", vector=None, threads=1):
    """Solve with stored factors (src/api/linalg_api.jl:168-184); a 1-D B is a
    vector right-hand side (trsv path).  Returns the solution (a new array)."""
    vec = (np.ndim(B) == 1) if vector is None else vector
    X = _f(np.array(B, dtype=np.float64, copy=True))
    F = _f(F)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int64)
    n = F.shape[0]
    rc = lib().oxo_dgetrs(trans.encode(), n, X.shape[1], _ptr(F), _ld(F), _ptr(ipiv, ip), _ptr(X), max(1, n),
                          1 if vec else 0, threads)
    _check(rc, "dgetrs")
    if rc > 0:
        raise np.linalg.LinAlgError(f"singular U at {rc}")
    return X[:, 0].copy() if np.ndim(B) == 1 else X


def sgetrs(F, ipiv, B, trans="N", vector=None, threads=1):
    vec = (np.ndim(B) == 1) if vector is None else vector
    X = _f(np.array(B, dtype=np.float32, copy=True), np.float32)
    F = _f(F, np.float32)
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int64)
    n = F.shape[0]
    rc = lib().oxo_sgetrs(trans.encode(), n, X.shape[1], _ptr(F, fp), _ld(F), _ptr(ipiv, ip), _ptr(X, fp),
                          max(1, n), 1 if vec else 0, threads)
    _check(rc, "sgetrs")
    return X[:, 0].copy() if np.ndim(B) == 1 else X


def dequilibrate(A, radix=True):
    """equilibrate(A; radix) (src/lapack/dense_utilities.jl:191-212) ->
    (scaled matrix, row scales, column scales)."""
    B = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = B.shape
    rows = np.ones(m); cols = np.ones(n)
    lib().oxo_dequilibrate(m, n, _ptr(B), max(1, m), _ptr(rows), _ptr(cols), 1 if radix else 0)
    return B, rows, cols


def refinement_residual(A, X, B):
    """_refinement_residual! (src/lapack/expert_refinement.jl:67-81) ->
    (residual, componentwise backward error)."""
    A = _f(A); X = _f(X); B = _f(B)
    n = A.shape[0]
    R = np.zeros((n, X.shape[1]), order="F")
    e = lib().oxo_refinement_residual(n, X.shape[1], _ptr(A), _ld(A), _ptr(X), _ld(X), _ptr(B), _ld(B),
                                      _ptr(R), max(1, n))
    return R, e


_REASONS = {0: "converged", 1: "iteration_limit", 2: "nonfinite_iteration", 3: "working_factor_fallback"}


def mixed_solve(A, b, maxiter=10, rtol=1e-13, threads=1, factorization="lu"):
    """mixed_solve(A, b) with Float32 factors and Float64 refinement
    (src/lapack/factorizations.jl:425-427, src/lapack/extended_lapack.jl:175-226);
    factorization "lu" (radix equilibration, LU) or "cholesky" (one global
    scale, Cholesky)."""
    A = _f(A); B = _f(b)
    n = A.shape[0]
    X = np.zeros((n, B.shape[1]), order="F")
    it = C.c_int64(0); conv = C.c_int(0); berr = C.c_double(0); fb = C.c_int(0); reason = C.c_int(0)
    f = lib().oxo_mixed_solve_chol if factorization == "cholesky" else lib().oxo_mixed_solve
    rc = f(n, B.shape[1], _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(X), max(1, n), maxiter, rtol,
           threads, C.byref(it), C.byref(conv), C.byref(berr), C.byref(fb), C.byref(reason))
    if rc < 0:
        raise ValueError(f"mixed_solve failed ({rc})")
    x = X[:, 0].copy() if np.ndim(b) == 1 else X
    return dict(x=x, converged=bool(conv.value), iterations=it.value, backward_error=berr.value,
                fallback=bool(fb.value), reason=_REASONS[reason.value], low_type="Float32")


def perm_from_ipiv(ipiv, m):
    """Permutation p (1-based) with A[p,:] == L*U, built by applying the
    interchanges in order (src/lapack/factorizations.jl:72-75)."""
    perm = np.arange(1, m + 1, dtype=np.int64)
    for j, p in enumerate(np.asarray(ipiv, dtype=np.int64)):
        perm[j], perm[p - 1] = perm[p - 1], perm[j]
    return perm
