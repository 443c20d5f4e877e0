"""Reverse-mode rules for the hot path, over the C ABI.

The reference ships no AD rules; these are the ChainRules-style ``rrule``s a
user of its API would attach to ``X = A \\ B`` and ``gemm!``, computed on the
GPU from the reference's own building blocks: the transposed LU solve
``ldiv!(transpose(F), .)`` (src/api/linalg_api.jl:176-184,
src/lapack/expert_refinement.jl:2-23) and ``gemm!`` (src/blas/gemm.jl:625-703).

    X = A \\ B:   Bbar = A^-T Xbar,  Abar = -Bbar X^T
    C = alpha op(A) op(B) + beta C0:
                 op(A)bar = alpha Cbar op(B)^T,  op(B)bar = alpha op(A)^T Cbar,
                 C0bar = beta Cbar
"""
import ctypes as C

import numpy as np

from . import _lib
from .device import DMatrix, as_dmatrix
from .linalg import LU, SingularException, DimensionMismatch, _trans, gemm_, lu


def ldiv_pullback(F, X, Xbar, need_A=True):
    """(Abar, Bbar) of X = A \\ B from the stored factors F (an ``LU``), the
    solution X and its cotangent Xbar.  Abar is None when need_A is False."""
    if not isinstance(F, LU):
        raise TypeError("ldiv_pullback needs the LU factors of A")
    F._ready()
    host = not (isinstance(X, DMatrix) and isinstance(Xbar, DMatrix))
    Xd, Xbd = as_dmatrix(X), as_dmatrix(Xbar)
    n = F.factors.n
    if F.factors.m != n or Xd.m != n or Xbd.shape != Xd.shape:
        raise DimensionMismatch("adjoint dimensions differ")
    nrhs = Xd.n
    dev = F.factors.buf.device
    Bbar = DMatrix.empty(n, nrhs, device=dev) if nrhs else DMatrix.zeros(n, 0, device=dev)
    Abar = DMatrix.empty(n, n, device=dev) if need_A else None
    h = _lib.handle()
    rc = h._L.oxi_dgesv_adjoint(h.ptr, n, nrhs, F.factors.ptr, F.factors.ld, F.ipiv_dev.data_ptr(), Xd.ptr, Xd.ld,
                                Xbd.ptr, Xbd.ld, Bbar.ptr, Bbar.ld, Abar.ptr if Abar is not None else None,
                                Abar.ld if Abar is not None else 1)
    if rc == _lib.OXI_ERR_SINGULAR:
        raise SingularException(F.info or 0)
    h.check(rc, "ldiv pullback")
    if host:
        return (Abar.numpy() if Abar is not None else None), Bbar.numpy()
    return Abar, Bbar


def rrule_ldiv(A, B):
    """X = A \\ B and its pullback Xbar -> (Abar, Bbar); the factors stay on the
    device between the two."""
    F = lu(A)
    X = F.solve(B)
    vec = not isinstance(B, DMatrix) and np.ndim(B) == 1

    def pullback(Xbar):
        Xb = np.asarray(Xbar).reshape(-1, 1) if vec else Xbar
        Xm = np.asarray(X).reshape(-1, 1) if vec else X
        Abar, Bbar = ldiv_pullback(F, Xm, Xb)
        if vec:
            Bbar = Bbar[:, 0].copy()
        return Abar, Bbar

    return X, pullback


def gemm_pullback(A, B, Cbar, alpha=1.0, beta=0.0, transA="N", transB="N"):
    """(Abar, Bbar, C0bar) of C = alpha op(A) op(B) + beta C0, each a gemm!."""
    ta, tb = _trans(transA), _trans(transB)
    host = not all(isinstance(x, DMatrix) for x in (A, B, Cbar))
    Ad, Bd, Cb = as_dmatrix(A), as_dmatrix(B), as_dmatrix(Cbar)
    dev = Cb.buf.device
    am, ak = (Ad.n, Ad.m) if ta != "N" else (Ad.m, Ad.n)
    bk, bn = (Bd.n, Bd.m) if tb != "N" else (Bd.m, Bd.n)
    if ak != bk or (Cb.m, Cb.n) != (am, bn):
        raise DimensionMismatch("GEMM dimensions differ")
    Abar = DMatrix.zeros(Ad.m, Ad.n, device=dev)
    Bbar = DMatrix.zeros(Bd.m, Bd.n, device=dev)
    flip = {"N": "T", "T": "N", "C": "N"}
    if ta == "N":   # Abar = alpha Cbar op(B)^T
        gemm_(Abar, Cb, Bd, alpha, 0.0, "N", flip[tb])
    else:           # Abar = (alpha Cbar op(B)^T)^T = alpha op(B) Cbar^T
        gemm_(Abar, Bd, Cb, alpha, 0.0, tb, "T")
    if tb == "N":   # Bbar = alpha op(A)^T Cbar
        gemm_(Bbar, Ad, Cb, alpha, 0.0, flip[ta], "N")
    else:           # Bbar = alpha Cbar^T op(A)
        gemm_(Bbar, Cb, Ad, alpha, 0.0, "T", ta)
    C0bar = None
    if beta != 0.0:
        C0bar = Cb.copy()   # beta * Cbar: gemm! with k = 0 scales C only
        gemm_(C0bar, DMatrix.zeros(Cb.m, 0, device=dev), DMatrix.zeros(0, Cb.n, device=dev), 0.0, float(beta))
    if host:
        return Abar.numpy(), Bbar.numpy(), (C0bar.numpy() if C0bar is not None else None)
    return Abar, Bbar, C0bar


def rrule_mul(A, B, alpha=1.0, transA="N", transB="N"):
    """C = alpha op(A) op(B) and its pullback Cbar -> (Abar, Bbar)."""
    ta, tb = _trans(transA), _trans(transB)
    Ad, Bd = as_dmatrix(A), as_dmatrix(B)
    am = Ad.n if ta != "N" else Ad.m
    bn = Bd.m if tb != "N" else Bd.n
    Cd = DMatrix.zeros(am, bn, device=Ad.buf.device)
    gemm_(Cd, Ad, Bd, alpha, 0.0, ta, tb)
    host = not (isinstance(A, DMatrix) and isinstance(B, DMatrix))

    def pullback(Cbar):
        Ab, Bb, _ = gemm_pullback(Ad, Bd, as_dmatrix(Cbar), alpha, 0.0, ta, tb)
        return (Ab.numpy(), Bb.numpy()) if host else (Ab, Bb)

    return (Cd.numpy() if host else Cd), pullback
