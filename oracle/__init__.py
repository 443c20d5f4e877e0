"""CPU oracle for the OxiBLAS.jl level-3 hot path (test infrastructure only).

``oracle/oxo.cpp`` restates the reference's algorithms in C++; this module
loads the built ``liboxo.so`` with ctypes and wraps it for numpy arrays in
column-major (Fortran) order.  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, as the checker and the timed CPU
baseline.  The product library never imports it.
"""
from .oxo import (  # noqa: F401
    available, build, lib, max_threads,
    dgemm, sgemm, dtrsm, dtrsv, dgetrf, sgetrf, dgetrs, sgetrs, dlaswp,
    dsyrk, dpotrf, dpotrs, dgetri, dlogabsdet, backward_error, inverse_norm_estimate, refine_solve,
    expert_solve, hgemm,
    dequilibrate, refinement_residual, mixed_solve, perm_from_ipiv,
)
