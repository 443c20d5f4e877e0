"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

Every LU fixture is cross-checked against LAPACK dgetrf (scipy) before it is
written: same pivots (as test/lapack/lapack_dense.jl:12-14 asserts for the
reference) except where a case pins the reference's own tie and NaN rules
(_lu_pivot, src/lapack/factorizations.jl:89-107), which LAPACK's idamax
shares for ties.  GEMM fixtures are cross-checked against numpy.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

SEED = 0x0B1A5


def gemm_cases(rng):
    out = {}
    cases = [(65, 67, 49, "N", "N", 0.7, -0.4), (33, 13, 35, "T", "N", 0.7, -0.4), (17, 19, 23, "N", "T", 1.0, 0.0),
             (7, 9, 11, "T", "T", -1.0, 1.0), (130, 66, 129, "N", "N", -1.0, 1.0), (1, 1, 1, "N", "N", 2.0, 0.5)]
    for i, (m, n, k, ta, tb, al, be) in enumerate(cases):
        A = rng.standard_normal((m, k) if ta == "N" else (k, m))
        B = rng.standard_normal((k, n) if tb == "N" else (n, k))
        C0 = np.asfortranarray(rng.standard_normal((m, n)))
        C = C0.copy(order="F")
        O.dgemm(C, A, B, al, be, ta, tb)
        ref = al * ((A if ta == "N" else A.T) @ (B if tb == "N" else B.T)) + be * C0
        assert np.allclose(C, ref, rtol=3e-12, atol=3e-12 * k)
        out.update({f"c{i}_A": A, f"c{i}_B": B, f"c{i}_C0": C0, f"c{i}_C": C,
                    f"c{i}_meta": np.array([m, n, k, ord(ta), ord(tb)], dtype=np.int64),
                    f"c{i}_ab": np.array([al, be])})
    out["count"] = np.array(len(cases))
    return out


def lu_cases(rng):
    out = {}
    cases = []
    cases.append(("gaussian64", rng.standard_normal((64, 64)), 32, True))
    cases.append(("gaussian_rect_33x20", rng.standard_normal((33, 20)), 16, True))
    cases.append(("gaussian_rect_20x33", rng.standard_normal((20, 33)), 16, True))
    A = rng.standard_normal((40, 40)); A[:, 22] = 0.0
    cases.append(("zero_column_23", A, 128, True))
    G = rng.standard_normal((80, 80)) @ np.diag(10.0 ** np.linspace(0, -12, 80))
    cases.append(("graded80", G, 32, True))
    # exact ties in the first column: the first row of largest magnitude wins
    T = rng.standard_normal((24, 24)); T[:, 0] = 0.0; T[5, 0] = 3.0; T[11, 0] = -3.0; T[17, 0] = 3.0
    cases.append(("tie_first_column", T, 8, True))
    # NaN below the diagonal never wins; the pivot is the largest number
    Nn = rng.standard_normal((16, 16)); Nn[3, 0] = np.nan; Nn[9, 0] = 7.0
    cases.append(("nan_below_diagonal", Nn, 8, False))
    for i, (name, A, bs, lapack_check) in enumerate(cases):
        F, ipiv, info = O.dgetrf(A, bs, check=False)
        if lapack_check:
            _, lp = sl.lu_factor(A, check_finite=False)
            assert np.array_equal(ipiv, lp[:len(ipiv)] + 1), name
        out.update({f"c{i}_A": A, f"c{i}_F": F, f"c{i}_ipiv": ipiv, f"c{i}_info": np.array(info),
                    f"c{i}_bs": np.array(bs), f"c{i}_name": np.array(name)})
        if A.shape[0] == A.shape[1] and info == 0:
            b = rng.standard_normal(A.shape[0]); B = rng.standard_normal((A.shape[0], 3))
            out.update({f"c{i}_b": b, f"c{i}_x": O.dgetrs(F, ipiv, b), f"c{i}_B": B,
                        f"c{i}_X": O.dgetrs(F, ipiv, B), f"c{i}_xT": O.dgetrs(F, ipiv, b, "T")})
    out["count"] = np.array(len(cases))
    return out


def main():
    rng = np.random.default_rng(SEED)
    np.savez_compressed(os.path.join(HERE, "gemm_cases.npz"), **gemm_cases(rng))
    np.savez_compressed(os.path.join(HERE, "lu_cases.npz"), **lu_cases(rng))
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
