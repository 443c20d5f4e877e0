"""Batched GEMM and the symmetric rank-k update over the C ABI.

===============================================  ==========================================
here                                             reference
===============================================  ==========================================
``gemm_strided_batched_(C, A, B, a, b, ...)``    ``gemm_strided_batched!`` (src/blas/batched.jl:80-100)
``gemm_batched_(Cs, As, Bs, a, b, ...)``         ``gemm_batched!`` (src/blas/batched.jl:49-78)
``syrk_(C, A, a, b; uplo, trans)``               ``syrk!`` (src/blas/level3.jl:779)
``herk_(C, A, a, b; uplo, trans)``               ``herk!`` (src/blas/level3.jl:781), real data
``hgemm_(C, A, B, a, b; transA, transB)``        ``gemm!`` for Float16 (src/blas/extended_blas.jl:39-88)
===============================================  ==========================================
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix
from .linalg import DimensionMismatch, _trans


def _h():
    return _lib.handle()


class DBatch:
    """`batch` column-major FP64 matrices of one shape in one device buffer,
    matrix z at element offset z * stride (a 3-D Julia array, dimension three
    the batch index)."""
    __slots__ = ("buf", "m", "n", "ld", "stride", "batch")

    def __init__(self, buf, m, n, batch, ld=None, stride=None):
        self.buf = buf
        self.m, self.n, self.batch = int(m), int(n), int(batch)
        self.ld = int(ld if ld is not None else max(1, m))
        self.stride = int(stride if stride is not None else self.ld * self.n)

    @classmethod
    def from_numpy(cls, a):
        """a: host array of shape (m, n, batch)."""
        import torch
        a = np.asarray(a, dtype=np.float64)
        m, n, b = a.shape
        host = np.ascontiguousarray(np.transpose(a, (2, 1, 0)))   # [z][j][i]
        buf = torch.from_numpy(host.reshape(-1) if host.size else np.zeros(1)).to("cuda")
        return cls(buf, m, n, b)

    def numpy(self):
        host = self.buf.cpu().numpy()
        out = np.empty((self.m, self.n, self.batch))
        for z in range(self.batch):
            blk = host[z * self.stride: z * self.stride + self.ld * self.n].reshape(self.n, self.ld)[:, :self.m]
            out[:, :, z] = blk.T
        return out

    @property
    def ptr(self):
        return self.buf.data_ptr()


def gemm_strided_batched_(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """C[:,:,z] <- alpha op(A[:,:,z]) op(B[:,:,z]) + beta C[:,:,z] for every z,
    one launch (gemm_strided_batched!, src/blas/batched.jl:80-100)."""
    ta, tb = _trans(transA), _trans(transB)
    host = not isinstance(C_, DBatch)
    Cd = C_ if not host else DBatch.from_numpy(C_)
    Ad = A if isinstance(A, DBatch) else DBatch.from_numpy(A)
    Bd = B if isinstance(B, DBatch) else DBatch.from_numpy(B)
    if not (Cd.batch == Ad.batch == Bd.batch):
        raise DimensionMismatch("batch counts differ")
    am, ak = (Ad.n, Ad.m) if ta != "N" else (Ad.m, Ad.n)
    bk, bn = (Bd.n, Bd.m) if tb != "N" else (Bd.m, Bd.n)
    if ak != bk or (Cd.m, Cd.n) != (am, bn):
        raise DimensionMismatch("batched GEMM dimensions differ")
    h = _h()
    rc = h._L.oxi_dgemm_strided_batched(h.ptr, ta.encode(), tb.encode(), Cd.m, Cd.n, ak, float(alpha), Ad.ptr, Ad.ld,
                                        Ad.stride, Bd.ptr, Bd.ld, Bd.stride, float(beta), Cd.ptr, Cd.ld, Cd.stride,
                                        Cd.batch)
    h.check(rc, "gemm_strided_batched!")
    if host:
        C_[...] = Cd.numpy()
        return C_
    return Cd


def gemm_batched_(Cs, As, Bs, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """Cs[i] <- alpha op(As[i]) op(Bs[i]) + beta Cs[i] for lists of device
    matrices (gemm_batched!, src/blas/batched.jl:49-78).  Members of one shape
    and leading dimensions run in one launch; a batch of several shapes runs
    one launch per shape."""
    ta, tb = _trans(transA), _trans(transB)
    if not (len(Cs) == len(As) == len(Bs)):
        raise DimensionMismatch("batch counts differ")
    groups = {}
    for i, (Cm, Am, Bm) in enumerate(zip(Cs, As, Bs)):
        if not all(isinstance(x, DMatrix) for x in (Cm, Am, Bm)):
            raise TypeError("gemm_batched_ takes device matrices")
        am, ak = (Am.n, Am.m) if ta != "N" else (Am.m, Am.n)
        bk, bn = (Bm.n, Bm.m) if tb != "N" else (Bm.m, Bm.n)
        if ak != bk or (Cm.m, Cm.n) != (am, bn):
            raise DimensionMismatch(f"batched GEMM dimensions differ at member {i}")
        groups.setdefault((Cm.m, Cm.n, ak, Am.ld, Bm.ld, Cm.ld), []).append(i)
    # every output disjoint from every other, across the groups too
    spans = sorted((c.ptr, c.ptr + 8 * ((c.n - 1) * c.ld + c.m)) for c in Cs if c.m and c.n)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        if b0 < a1:
            raise ValueError("batch outputs must be disjoint")
    h = _h()
    for (m, n, k, lda, ldb, ldc), idx in groups.items():
        nb = len(idx)
        pa = (C.c_void_p * nb)(*[As[i].ptr for i in idx])
        pb = (C.c_void_p * nb)(*[Bs[i].ptr for i in idx])
        pc = (C.c_void_p * nb)(*[Cs[i].ptr for i in idx])
        rc = h._L.oxi_dgemm_batched(h.ptr, ta.encode(), tb.encode(), m, n, k, float(alpha), C.cast(pa, C.c_void_p),
                                    lda, C.cast(pb, C.c_void_p), ldb, float(beta), C.cast(pc, C.c_void_p), ldc, nb)
        h.check(rc, "gemm_batched!")
    return Cs


def syrk_(C_, A, alpha=1.0, beta=0.0, uplo="U", trans="N"):
    """C[tri] <- alpha op(A) op(A)^T + beta C[tri] on the `uplo` triangle, the
    other untouched (syrk!, src/blas/level3.jl:779; _rankk!, :692-777)."""
    if uplo not in ("U", "L"):
        raise ValueError("uplo must be U or L")
    t = _trans(trans)
    if t == "C":
        raise ValueError("symmetric rank-k accepts N or T")   # level3.jl:695
    return _rankk(C_, A, alpha, beta, uplo, t)


def herk_(C_, A, alpha=1.0, beta=0.0, uplo="U", trans="N"):
    """herk! for real data (src/blas/level3.jl:781): the Hermitian rank-k
    update accepts N or C."""
    if uplo not in ("U", "L"):
        raise ValueError("uplo must be U or L")
    t = _trans(trans)
    if t == "T":
        raise ValueError("Hermitian rank-k accepts N or C")   # level3.jl:694
    return _rankk(C_, A, alpha, beta, uplo, t)


def _rankk(C_, A, alpha, beta, uplo, t):
    host = not isinstance(C_, DMatrix)
    Cd, Ad = as_dmatrix(C_), as_dmatrix(A)
    if Cd.m != Cd.n:
        raise DimensionMismatch("C must be square")
    n = Cd.n
    k = Ad.n if t == "N" else Ad.m
    if (Ad.m if t == "N" else Ad.n) != n:
        raise DimensionMismatch("rank-k dimensions differ")
    h = _h()
    rc = h._L.oxi_dsyrk(h.ptr, uplo.encode(), t.encode(), n, k, float(alpha), Ad.ptr, Ad.ld, float(beta), Cd.ptr,
                        Cd.ld)
    h.check(rc, "syrk!")
    if host:
        C_[...] = Cd.numpy()
        return C_
    return Cd


def hgemm_(C_, A, B, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """gemm! for Float16 (src/blas/extended_blas.jl:39-88): Float16 operands
    and output, Float32 accumulation over all of K, one rounding to Float16
    (alpha and beta rounded to Float16 first).  Host float16 arrays or
    float16 device matrices."""
    ta, tb = _trans(transA), _trans(transB)
    host = not isinstance(C_, DMatrix)
    Cd = C_ if not host else DMatrix.from_numpy(C_, dtype=np.float16)
    Ad = A if isinstance(A, DMatrix) else DMatrix.from_numpy(A, dtype=np.float16)
    Bd = B if isinstance(B, DMatrix) else DMatrix.from_numpy(B, dtype=np.float16)
    am, ak = (Ad.n, Ad.m) if ta != "N" else (Ad.m, Ad.n)
    bk, bn = (Bd.n, Bd.m) if tb != "N" else (Bd.m, Bd.n)
    if ak != bk or (Cd.m, Cd.n) != (am, bn):
        raise DimensionMismatch("GEMM dimensions differ")
    h = _h()
    rc = h._L.oxi_hgemm(h.ptr, ta.encode(), tb.encode(), Cd.m, Cd.n, ak, float(alpha), Ad.ptr, Ad.ld, Bd.ptr, Bd.ld,
                        float(beta), Cd.ptr, Cd.ld)
    h.check(rc, "gemm!")
    if host:
        C_[...] = Cd.numpy()
        return C_
    return Cd
