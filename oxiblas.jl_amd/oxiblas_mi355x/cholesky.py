"""Cholesky factorization and solve over the C ABI.

=============================  ============================================
here                           reference
=============================  ============================================
``cholesky(A; check)``         ``OxiBLAS.cholesky`` (src/api/linalg_api.jl:271-281)
``cholesky_(A; check)``        ``cholesky!`` (src/api/linalg_api.jl:282-296)
``cholesky_factor(A)``         ``cholesky_factor`` (src/lapack/factorizations.jl:206-209)
``Cholesky``                   ``OxiBLAS.Cholesky`` (src/api/linalg_api.jl:221-258)
``F.solve(B)``                 ``solve(F::CholeskyFactor, b)`` (factorizations.jl:288-300)
``PosDefException``            ``LinearAlgebra.PosDefException(info)``
=============================  ============================================
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix
from .linalg import DimensionMismatch, SingularException


class PosDefException(np.linalg.LinAlgError):
    """LinearAlgebra.PosDefException(info): the 1-based column whose pivot is
    not positive (-1: the input is not symmetric)."""

    def __init__(self, info):
        super().__init__(f"PosDefException({info})")
        self.info = int(info)


def _h():
    return _lib.handle()


class Cholesky:
    """A = L L^T with L lower triangular (OxiBLAS.Cholesky wrapping
    CholeskyFactor).  ``factors`` is the device matrix holding L (its strict
    upper triangle zero); ``info`` is 0 or the failing column."""

    def __init__(self, factors, info):
        self.factors = factors
        self.info = int(info)

    @property
    def size(self):
        return self.factors.shape

    @property
    def uplo(self):
        return "L"

    def issuccess(self):
        return self.info == 0

    def _ready(self):
        if self.info != 0:
            raise PosDefException(self.info)    # _chol_ready, linalg_api.jl:239

    @property
    def L(self):
        return np.tril(self.factors.numpy())

    @property
    def U(self):
        return self.L.T

    def ldiv_(self, B):
        """In-place solve A X = B (ldiv!(F, B), linalg_api.jl:240-245)."""
        self._ready()
        host = not isinstance(B, DMatrix)
        vec = host and np.ndim(B) == 1
        Bd = as_dmatrix(B)
        if Bd.m != self.factors.m:
            raise DimensionMismatch("Cholesky solve dimensions differ")
        h = _h()
        rc = h._L.oxi_dpotrs(h.ptr, b"L", self.factors.m, Bd.n, self.factors.ptr, self.factors.ld, Bd.ptr, Bd.ld)
        if rc == _lib.OXI_ERR_SINGULAR:
            raise SingularException(0)
        h.check(rc, "ldiv!")
        if host:
            X = Bd.numpy()
            out = X[:, 0] if vec else X
            if isinstance(B, np.ndarray) and B.flags.writeable and B.dtype == np.float64:
                B[...] = out
                return B
            return out
        return Bd

    def solve(self, B):
        if isinstance(B, DMatrix):
            return self.ldiv_(B.copy())
        return self.ldiv_(np.array(B, dtype=np.float64, copy=True))

    def logabsdet(self):
        """(2 sum log L_ii, 1) (linalg_api.jl:247-251)."""
        self._ready()
        d = np.diag(self.factors.numpy())
        return float(2.0 * np.sum(np.log(d))), 1.0

    def det(self):
        la, s = self.logabsdet()
        return s * np.exp(la)


def _potrf(Ad, check):
    h = _h()
    info = C.c_int64(0)
    rc = h._L.oxi_dpotrf(h.ptr, b"L", Ad.n, Ad.ptr, Ad.ld, C.byref(info))
    h.check(rc, "cholesky")
    if check and info.value != 0:
        raise PosDefException(info.value)
    return Cholesky(Ad, info.value)


def cholesky(A, check=True):
    """OxiBLAS.cholesky(A; check): factors a copy of A.  A host matrix must be
    symmetric (PosDefException(-1) otherwise, linalg_api.jl:263-281); a device
    matrix is read from its lower triangle only."""
    if isinstance(A, DMatrix):
        if A.m != A.n:
            raise DimensionMismatch("matrix is not square")
        return _potrf(A.copy(), check)
    a = np.asarray(A, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise DimensionMismatch("matrix is not square")
    if not np.array_equal(a, a.T):
        if check:
            raise PosDefException(-1)
        return Cholesky(DMatrix.from_numpy(a), -1)
    return _potrf(DMatrix.from_numpy(a), check)


def cholesky_(A, check=True):
    """cholesky!(A): the lower triangle of the device matrix A is overwritten
    by L, its strict upper triangle zeroed, and A becomes F.factors."""
    if not isinstance(A, DMatrix):
        raise TypeError("cholesky_ factors a device matrix in place")
    if A.m != A.n:
        raise DimensionMismatch("matrix is not square")
    return _potrf(A, check)


def cholesky_factor(A):
    """cholesky_factor(A) (factorizations.jl:206-209): lower triangle of a
    copy, PosDefException on failure."""
    Ad = A.copy() if isinstance(A, DMatrix) else DMatrix.from_numpy(np.asarray(A, dtype=np.float64))
    if Ad.m != Ad.n:
        raise DimensionMismatch("matrix is not square")
    return _potrf(Ad, True)
