"""The committed fixtures (tests/golden/make_golden.py) are reproduced by the
oracle bit for bit, and pin the pivot rules of _lu_pivot
(src/lapack/factorizations.jl:89-107) on exact ties and NaN entries."""
import os

import numpy as np

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def test_gemm_fixtures_reproduced():
    z = load("gemm_cases.npz")
    for i in range(int(z["count"])):
        m, n, k, ta, tb = (int(v) for v in z[f"c{i}_meta"])
        al, be = z[f"c{i}_ab"]
        C = np.asfortranarray(z[f"c{i}_C0"].copy())
        O.dgemm(C, z[f"c{i}_A"], z[f"c{i}_B"], al, be, chr(ta), chr(tb))
        assert np.array_equal(C, z[f"c{i}_C"])


def test_lu_fixtures_reproduced():
    z = load("lu_cases.npz")
    for i in range(int(z["count"])):
        F, ipiv, info = O.dgetrf(z[f"c{i}_A"], int(z[f"c{i}_bs"]), check=False)
        assert np.array_equal(F, z[f"c{i}_F"], equal_nan=True), str(z[f"c{i}_name"])
        assert np.array_equal(ipiv, z[f"c{i}_ipiv"]) and info == int(z[f"c{i}_info"])


def test_pivot_rules_on_ties_and_nan():
    z = load("lu_cases.npz")
    names = [str(z[f"c{i}_name"]) for i in range(int(z["count"]))]
    tie = names.index("tie_first_column")
    assert z[f"c{tie}_ipiv"][0] == 6       # rows 6, 12, 18 hold |3|: the first one wins
    nan = names.index("nan_below_diagonal")
    assert z[f"c{nan}_ipiv"][0] == 10      # NaN at row 4 never beats the 7.0 at row 10
    zc = names.index("zero_column_23")
    assert int(z[f"c{zc}_info"]) == 23
