"""Host-side mirror of OxiBLAS.jl's level-3 hot-path API over the C ABI.

Names, argument meaning and error behaviour follow the reference:

===========================  ==============================================
here                          reference
===========================  ==============================================
``gemm_(C, A, B, a, b, ...)`` ``gemm!`` (src/blas/gemm.jl:625-672)
``trsm_(B, A, a, ...)``       ``trsm!`` (src/blas/level3.jl:518-544)
``lu(A)`` / ``lu_(A)``        ``OxiBLAS.lu`` / ``lu!`` (src/api/linalg_api.jl:217-218)
``lu_factor(A; blocksize)``   ``lu_factor`` (src/lapack/factorizations.jl:17-22)
``LU``                        ``OxiBLAS.LU`` (src/api/linalg_api.jl:145-204)
``ldiv(F, B)``, ``ldiv_``     ``OxiBLAS.ldiv`` / ``ldiv!`` (src/api/linalg_api_solve.jl:133-200)
``F.solve(B)`` (``F \\ B``)    ``ldiv!(F, B)`` on a copy (src/api/linalg_api.jl:168-174)
``transpose(F)``/``adjoint``  ``ldiv!(transpose(F), B)`` (src/api/linalg_api.jl:176-184)
``solve``/``solve_transposed`` src/lapack/factorizations.jl:177-193, expert_refinement.jl:2-23
``SingularException``         ``LinearAlgebra.SingularException(info)``
``DimensionMismatch``         ``DimensionMismatch``
===========================  ==============================================

Operands are column-major FP64 ``DMatrix`` objects in device memory; numpy
arrays are accepted for convenience (copied to the GPU, result copied back).
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix


class SingularException(np.linalg.LinAlgError):
    """LinearAlgebra.SingularException(info): the 1-based column of the first
    zero pivot."""

    def __init__(self, info):
        super().__init__(f"SingularException({info})")
        self.info = int(info)


class DimensionMismatch(ValueError):
    pass


def _trans(t):
    t = t.decode() if isinstance(t, bytes) else str(t)
    if t in ("N", "n", ":N"):
        return "N"
    if t in ("T", "t", ":T"):
        return "T"
    if t in ("C", "c", ":C"):
        return "C"
    raise ValueError("transpose must be N, T or C")  # ArgumentError (src/blas/gemm.jl:156-161)


def _h():
    return _lib.handle()


def _host_in(*xs):
    return any(not isinstance(x, DMatrix) for x in xs)


# ---------------------------------------------------------------------------
# gemm! / trsm!
# ---------------------------------------------------------------------------
def gemm_(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """C <- alpha*op(A)*op(B) + beta*C in place; returns C.

    beta == 0 never reads C, alpha == 0 or k == 0 never reads A or B, an
    operand overlapping C is snapshotted first (src/blas/gemm.jl:625-703)."""
    ta, tb = _trans(transA), _trans(transB)
    host = not isinstance(C_, DMatrix)
    Cd = as_dmatrix(C_)
    Ad, Bd = as_dmatrix(A), as_dmatrix(B)
    am, ak = (Ad.n, Ad.m) if ta != "N" else (Ad.m, Ad.n)
    bk, bn = (Bd.n, Bd.m) if tb != "N" else (Bd.m, Bd.n)
    if ak != bk or (Cd.m, Cd.n) != (am, bn):
        raise DimensionMismatch(f"GEMM: C{Cd.shape}, op(A)({am}, {ak}), op(B)({bk}, {bn})")
    h = _h()
    rc = h._L.oxi_dgemm(h.ptr, ta.encode(), tb.encode(), Cd.m, Cd.n, ak, float(alpha), Ad.ptr, Ad.ld, Bd.ptr, Bd.ld,
                        float(beta), Cd.ptr, Cd.ld)
    h.check(rc, "gemm!")
    if host:
        C_[...] = Cd.numpy()
        return C_
    return Cd


def trsm_(B, A, alpha=1.0, side="L", uplo="U", trans="N", unit=False):
    """B <- alpha*op(A)^-1 B (side 'L') or alpha*B*op(A)^-1 (side 'R') in place
    (src/blas/level3.jl:518-544).  A zero on a non-unit diagonal raises
    SingularException before B is touched."""
    if side not in ("L", "R"):
        raise ValueError("side must be L or R")
    if uplo not in ("U", "L"):
        raise ValueError("uplo must be U or L")
    t = _trans(trans)
    host = not isinstance(B, DMatrix)
    Bd, Ad = as_dmatrix(B), as_dmatrix(A)
    if Ad.m != Ad.n:
        raise DimensionMismatch("triangular matrix must be square")
    if Ad.n != (Bd.m if side == "L" else Bd.n):
        raise DimensionMismatch("TRSM dimensions differ")
    h = _h()
    where = C.c_int64(0)
    rc = h._L.oxi_dtrsm(h.ptr, side.encode(), uplo.encode(), t.encode(), b"U" if unit else b"N", Bd.m, Bd.n,
                        float(alpha), Ad.ptr, Ad.ld, Bd.ptr, Bd.ld, C.byref(where))
    if rc == _lib.OXI_ERR_SINGULAR:
        raise SingularException(where.value)
    h.check(rc, "trsm!")
    if host:
        B[...] = Bd.numpy()
        return B
    return Bd


def laswp_(A, k1, k2, ipiv, inverse=False):
    """Interchanges k1..k2 (1-based, inclusive) of ``ipiv`` on every column of
    the device matrix A, in place; in reverse order when ``inverse``
    (_lu_swap!, src/lapack/factorizations.jl:80-88; _rowswap!,
    src/api/linalg_api.jl:100-110).  ``ipiv`` is a device int64 tensor or a
    host array of 1-based global rows."""
    import torch
    if not isinstance(A, DMatrix):
        raise TypeError("laswp_ permutes a device matrix in place")
    if isinstance(ipiv, torch.Tensor):
        pv = ipiv.to(device=A.buf.device, dtype=torch.int64).contiguous()
    else:
        pv = torch.from_numpy(np.ascontiguousarray(ipiv, dtype=np.int64)).to(A.buf.device)
    if k2 > pv.numel():
        raise ValueError("laswp: k2 beyond the pivot vector")
    h = _h()
    rc = h._L.oxi_dlaswp(h.ptr, A.n, A.ptr, A.ld, int(k1), int(k2), pv.data_ptr(), 1 if inverse else 0)
    h.check(rc, "laswp")
    return A


# ---------------------------------------------------------------------------
# LU
# ---------------------------------------------------------------------------
class LU:
    """Dense LU with row pivoting: ``F.L @ F.U == A[F.p - 1, :]``
    (OxiBLAS.LU wrapping LUFactor, src/api/linalg_api.jl:145-204).

    The factors and pivots stay in device memory (``factors``, ``ipiv_dev``)
    so that solves and adjoints reuse them."""

    def __init__(self, factors, ipiv_dev, info):
        self.factors = factors      # DMatrix, unit-lower L below the diagonal, U on and above
        self.ipiv_dev = ipiv_dev    # torch int64 tensor, 1-based global rows
        self.info = int(info)
        self._ipiv = None
        self._plan = None           # solve plan of these factors, made by the second solve
        self._solved = False        # a solve has run on these factors

    def _solve_plan(self):
        """The factors' solve plan (device permutation and the inverses of
        the diagonal blocks of both triangles), made once."""
        if self._plan is None:
            h = _h()
            p = C.c_void_p()
            rc = h._L.oxi_dgetrs_plan_create(h.ptr, self.factors.m, self.factors.ptr, self.factors.ld,
                                             self.ipiv_dev.data_ptr(), C.byref(p))
            h.check(rc, "ldiv!")
            self._plan = (h, p)
        return self._plan[1]

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None:
            try:
                plan[0]._L.oxi_dgetrs_plan_destroy(plan[0].ptr, plan[1])
            except Exception:
                pass

    # properties of the reference (propertynames: L, U, p, P; private: factors, ipiv, info)
    @property
    def size(self):
        return self.factors.shape

    @property
    def ipiv(self):
        if self._ipiv is None:
            self._ipiv = self.ipiv_dev.cpu().numpy().astype(np.int64)
        return self._ipiv

    @property
    def p(self):
        """Permutation (1-based) from the interchanges (src/lapack/factorizations.jl:72-75)."""
        m = self.factors.m
        perm = np.arange(1, m + 1, dtype=np.int64)
        for j, q in enumerate(self.ipiv):
            perm[j], perm[q - 1] = perm[q - 1], perm[j]
        return perm

    @property
    def P(self):
        p = self.p
        n = len(p)
        P = np.zeros((n, n))
        P[np.arange(n), p - 1] = 1.0
        return P

    @property
    def L(self):
        a = self.factors.numpy()
        m, n = a.shape
        k = min(m, n)
        L = np.tril(a[:, :k], -1)
        L[np.arange(k), np.arange(k)] = 1.0
        return L

    @property
    def U(self):
        a = self.factors.numpy()
        k = min(a.shape)
        return np.triu(a[:k, :])

    def issuccess(self):
        return self.info == 0

    def _ready(self):
        m, n = self.factors.shape
        if m != n:
            raise DimensionMismatch("matrix is not square")
        if self.info != 0:
            raise SingularException(self.info)  # src/api/linalg_api.jl:163-167

    def ldiv_(self, B, trans="N"):
        """In-place solve op(A) X = B (ldiv!(F, B)).  A complex host B on these
        real factors is solved as its real and imaginary parts
        (_apply_split!, src/api/linalg_api.jl:67-78)."""
        self._ready()
        if isinstance(B, np.ndarray) and np.iscomplexobj(B):
            X = self.solve(B.real, trans) + 1j * self.solve(B.imag, trans)
            if B.flags.writeable:
                B[...] = X
                return B
            return X
        t = _trans(trans)
        host = not isinstance(B, DMatrix)
        vec = host and np.ndim(B) == 1
        Bd = as_dmatrix(B)
        if Bd.m != self.factors.m:
            raise DimensionMismatch(f"factorization has {self.factors.m} rows, right-hand side has {Bd.m}")
        h = _h()
        if self._plan is None and not self._solved:
            # the first solve: oxi_dgetrs prepares in the handle's workspace;
            # the plan, which keeps that preparation for later solves in
            # buffers of its own, is made from the second solve on
            self._solved = True
            rc = h._L.oxi_dgetrs(h.ptr, t.encode(), Bd.m, Bd.n, self.factors.ptr, self.factors.ld,
                                 self.ipiv_dev.data_ptr(), Bd.ptr, Bd.ld)
        else:
            plan = self._solve_plan()
            rc = h._L.oxi_dgetrs_plan(h.ptr, plan, t.encode(), Bd.n, Bd.ptr, Bd.ld)
        h.check(rc, "ldiv!")
        if host:
            X = Bd.numpy()
            out = X[:, 0] if vec else X
            if isinstance(B, np.ndarray) and B.flags.writeable and B.dtype == np.float64:
                B[...] = out
                return B
            return out
        return Bd

    def solve(self, B, trans="N"):
        """``F \\ B``: a new solution, B untouched."""
        if isinstance(B, DMatrix):
            return self.ldiv_(B.copy(), trans)
        if np.iscomplexobj(B):
            return self.ldiv_(np.array(B, dtype=np.complex128, copy=True), trans)
        return self.ldiv_(np.array(B, dtype=np.float64, copy=True), trans)

    def transpose(self):
        return _TransposedLU(self, "T")

    def adjoint(self):
        return _TransposedLU(self, "C")

    @property
    def T(self):
        return self.transpose()

    def logabsdet(self):
        """(log|det|, sign) (src/api/linalg_api.jl:185-191): one device
        reduction over U's diagonal and the interchanges.  A singular factor
        gives (-inf, 0.0), as the reference (no exception)."""
        m, n = self.factors.shape
        if m != n:
            raise DimensionMismatch("matrix is not square")
        h = _h()
        la, sg = C.c_double(0), C.c_double(0)
        rc = h._L.oxi_dlogabsdet(h.ptr, n, self.factors.ptr, self.factors.ld, self.ipiv_dev.data_ptr(), C.byref(la),
                                 C.byref(sg))
        h.check(rc, "logabsdet")
        return float(la.value), float(sg.value)

    def logdet(self):
        la, sg = self.logabsdet()
        if sg < 0:
            raise ValueError("logdet: negative determinant")   # DomainError
        return la if sg > 0 else -np.inf

    def det(self):
        if self.info != 0:
            return 0.0
        la, s = self.logabsdet()
        return s * np.exp(la)

    def inv(self):
        """inv(A) from these factors: inv(U) inv(L) P (native_inverse,
        src/lapack/dense_utilities.jl:14-36) as a new device matrix."""
        self._ready()
        n = self.factors.n
        out = DMatrix.empty(n, n, device=self.factors.buf.device)
        h = _h()
        rc = h._L.oxi_dgetri(h.ptr, n, self.factors.ptr, self.factors.ld, self.ipiv_dev.data_ptr(), out.ptr, out.ld)
        if rc == _lib.OXI_ERR_SINGULAR:
            raise SingularException(self.info or 0)
        h.check(rc, "inv")
        return out


class _TransposedLU:
    def __init__(self, F, trans):
        self.parent = F
        self.trans = trans

    def ldiv_(self, B):
        return self.parent.ldiv_(B, self.trans)

    def solve(self, B):
        return self.parent.solve(B, self.trans)


def transpose(F):
    return F.transpose()


def adjoint(F):
    return F.adjoint()


def _getrf(Ad, check, blocksize):
    import torch
    h = _h()
    q = min(Ad.m, Ad.n)
    ipiv = torch.zeros(max(q, 1), dtype=torch.int64, device=Ad.buf.device)
    info = C.c_int64(0)
    rc = h._L.oxi_dgetrf(h.ptr, Ad.m, Ad.n, Ad.ptr, Ad.ld, ipiv.data_ptr(), int(blocksize), C.byref(info))
    h.check(rc, "lu")
    F = LU(Ad, ipiv[:q], info.value)
    if check and info.value != 0:
        raise SingularException(info.value)
    return F


def lu(A, check=True, blocksize=0):
    """OxiBLAS.lu(A; check): factors a copy of A (src/api/linalg_api.jl:217).
    A device matrix is read once by oxi_dgetrf_copy, whose change of layout
    writes the copy; a host matrix is copied to the device and factored
    there."""
    if not isinstance(A, DMatrix):
        return _getrf(as_dmatrix(A), check, blocksize)
    import torch
    h = _h()
    F = DMatrix.empty(A.m, A.n, device=A.buf.device)
    q = min(A.m, A.n)
    ipiv = torch.zeros(max(q, 1), dtype=torch.int64, device=A.buf.device)
    info = C.c_int64(0)
    rc = h._L.oxi_dgetrf_copy(h.ptr, A.m, A.n, A.ptr, A.ld, F.ptr, F.ld, ipiv.data_ptr(), int(blocksize),
                              C.byref(info))
    h.check(rc, "lu")
    Fl = LU(F, ipiv[:q], info.value)
    if check and info.value != 0:
        raise SingularException(info.value)
    return Fl


def lu_(A, check=True, blocksize=0):
    """OxiBLAS.lu!(A; check): A is overwritten and becomes F.factors
    (src/api/linalg_api.jl:218)."""
    if not isinstance(A, DMatrix):
        raise TypeError("lu_ factors a device matrix in place")
    return _getrf(A, check, blocksize)


def lu_factor(A, blocksize=128, check=True):
    """lu_factor(A; blocksize=128, check=true) (src/lapack/factorizations.jl:17-22)."""
    return lu(A, check=check, blocksize=blocksize)


def ldiv(F, B):
    """OxiBLAS.ldiv(F, B): a new solution (src/api/linalg_api_solve.jl:188-199)."""
    return F.solve(B)


def ldiv_(*args):
    """ldiv!(F, B) in place, or ldiv!(Y, F, B): Y <- F \\ B
    (src/api/linalg_api_solve.jl:133-178)."""
    if len(args) == 2:
        F, B = args
        return F.ldiv_(B)
    Y, F, B = args
    if isinstance(Y, DMatrix):
        Bd = as_dmatrix(B)
        if Y.shape != Bd.shape:
            raise DimensionMismatch("ldiv!: Y and B differ in size")
        if Y is not B:
            Y.copy_from(Bd)
        return F.ldiv_(Y)
    Y[...] = F.solve(B)
    return Y


def solve(F, b):
    """solve(F::LUFactor, b) (src/lapack/factorizations.jl:177-193)."""
    return F.solve(b)


def solve_transposed(F, b, trans="C"):
    """solve_transposed(F, b; trans) (src/lapack/expert_refinement.jl:2-23)."""
    t = _trans(trans)
    if t == "N":
        return F.solve(b)
    return F.solve(b, t)


# ---------------------------------------------------------------------------
# what is built on lu (src/api/linalg_api.jl:1101-1143, dense_utilities.jl)
# ---------------------------------------------------------------------------
def det(A):
    """OxiBLAS.det(A) = det(lu(A; check=false))."""
    return A.det() if isinstance(A, LU) else lu(A, check=False).det()


def logabsdet(A):
    """OxiBLAS.logabsdet(A) = logabsdet(lu(A; check=false))."""
    return A.logabsdet() if isinstance(A, LU) else lu(A, check=False).logabsdet()


def logdet(A):
    return A.logdet() if isinstance(A, LU) else lu(A, check=False).logdet()


def inv(A):
    """OxiBLAS.inv(A) = native_inverse(A) (dense_utilities.jl:14-36); a host
    matrix gives a host matrix."""
    if isinstance(A, LU):
        return A.inv()
    F = lu(A, check=True)
    X = F.inv()
    return X if isinstance(A, DMatrix) else X.numpy()


def factorize(A):
    """factorize(A) (src/api/linalg_api.jl:1101-1113) for a square matrix: a
    symmetric matrix with a positive real diagonal is tried with Cholesky
    (check=false) and kept when that succeeds; otherwise LU.  (The reference
    turns a symmetric indefinite matrix into Bunch-Kaufman and a rectangular
    one into pivoted QR, which are outside this library: those raise or give
    LU.)"""
    from .cholesky import cholesky
    a = A.numpy() if isinstance(A, DMatrix) else np.asarray(A, dtype=np.float64)
    m, n = a.shape
    if m != n:
        raise NotImplementedError("factorize of a rectangular matrix (pivoted QR) is not implemented")
    if np.array_equal(a, a.T) and np.all(np.diag(a) > 0) and np.all(np.isfinite(np.diag(a))):
        Fc = cholesky(A if isinstance(A, DMatrix) else a, check=False)
        if Fc.issuccess():
            return Fc
    return lu(A)
