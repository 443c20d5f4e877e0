"""Column-major FP64 matrices in device memory.

A ``DMatrix`` is the device counterpart of the reference's ``Matrix{Float64}``
or strided view with unit row stride (src/blas/gemm.jl:360-361): element
(i, j) lives at ``base[i + j*ld]``.  The storage is a flat torch tensor on the
GPU (torch is used for allocation and streams only; every computation goes
through liboxiblas_mi355x.so).
"""
import numpy as np


def _torch():
    import torch
    return torch


class DMatrix:
    __slots__ = ("buf", "offset", "m", "n", "ld")

    def __init__(self, buf, m, n, ld=None, offset=0):
        self.buf = buf          # 1-D torch tensor (float64 or float32) on the GPU
        self.m = int(m)
        self.n = int(n)
        self.ld = int(ld if ld is not None else max(1, m))
        self.offset = int(offset)
        if self.ld < max(1, self.m):
            raise ValueError("leading dimension smaller than the row count")
        need = self.offset + (self.ld * (self.n - 1) + self.m if self.m and self.n else 0)
        if need > buf.numel():
            raise ValueError("view exceeds its storage")

    # -- construction -----------------------------------------------------
    @classmethod
    def empty(cls, m, n, ld=None, dtype=None, device=None):
        torch = _torch()
        ld = int(ld if ld is not None else max(1, m))
        dtype = dtype or torch.float64
        buf = torch.empty(max(1, ld * n), dtype=dtype, device=device or "cuda")
        return cls(buf, m, n, ld)

    @classmethod
    def zeros(cls, m, n, ld=None, dtype=None, device=None):
        d = cls.empty(m, n, ld, dtype, device)
        d.buf.zero_()
        return d

    @classmethod
    def from_numpy(cls, a, ld=None, dtype=np.float64, device=None):
        """Copy a host matrix (any order) to a new column-major device matrix."""
        torch = _torch()
        a = np.asarray(a, dtype=dtype)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        m, n = a.shape
        ld = int(ld if ld is not None else max(1, m))
        host = np.zeros((n, ld), dtype=dtype)  # row j of `host` = column j
        host[:, :m] = a.T
        buf = torch.from_numpy(host.reshape(-1)).to(device or "cuda")
        if buf.numel() == 0:
            buf = torch.zeros(1, dtype=buf.dtype, device=buf.device)
        return cls(buf, m, n, ld)

    def numpy(self):
        """Copy back to a Fortran-ordered host array."""
        torch = _torch()
        if self.m == 0 or self.n == 0:
            return np.zeros((self.m, self.n), dtype=self.np_dtype, order="F")
        span = self.ld * (self.n - 1) + self.m
        flat = self.buf[self.offset:self.offset + span]
        padded = torch.empty(self.ld * self.n, dtype=self.buf.dtype, device=self.buf.device)
        padded[:span] = flat
        host = padded.cpu().numpy().reshape(self.n, self.ld)[:, :self.m]
        return np.asfortranarray(host.T)

    # -- views -------------------------------------------------------------
    def view(self, rows, cols):
        """Contiguous sub-block A[rows, cols] (Python slices, step 1)."""
        r0, r1, _ = rows.indices(self.m)
        c0, c1, _ = cols.indices(self.n)
        return DMatrix(self.buf, max(0, r1 - r0), max(0, c1 - c0), self.ld, self.offset + r0 + c0 * self.ld)

    def copy(self):
        out = DMatrix.empty(self.m, self.n, dtype=self.buf.dtype, device=self.buf.device)
        if self.m and self.n:
            src = self.buf[self.offset:self.offset + self.ld * (self.n - 1) + self.m]
            if self.ld == self.m:
                out.buf[: self.m * self.n].copy_(src)
            else:
                full = _torch().empty(self.ld * self.n, dtype=self.buf.dtype, device=self.buf.device)
                full[: src.numel()] = src
                out.buf[: self.m * self.n].view(self.n, self.m).copy_(full.view(self.n, self.ld)[:, : self.m])
        return out

    def strided(self):
        """The (n, m) torch view whose row j is column j of this matrix."""
        return self.buf.as_strided((self.n, self.m), (self.ld, 1), self.buf.storage_offset() + self.offset)

    def copy_from(self, src):
        """Device copy of src (same shape and dtype) into this matrix."""
        if src.shape != self.shape:
            raise ValueError("copy_from: shapes differ")
        if self.m and self.n:
            self.strided().copy_(src.strided())
        return self

    @property
    def ptr(self):
        return self.buf.data_ptr() + self.offset * self.buf.element_size()

    @property
    def shape(self):
        return (self.m, self.n)

    @property
    def np_dtype(self):
        torch = _torch()
        return {torch.float64: np.float64, torch.float32: np.float32, torch.float16: np.float16}[self.buf.dtype]

    def __repr__(self):
        return f"DMatrix({self.m}x{self.n}, ld={self.ld}, {self.buf.dtype})"


def as_dmatrix(x, dtype=np.float64):
    if isinstance(x, DMatrix):
        return x
    return DMatrix.from_numpy(x, dtype=dtype)
