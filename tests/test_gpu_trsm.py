"""trsm! on the GPU (oxi_dtrsm through the C ABI) against the oracle's
restatement of src/blas/level3.jl:518-544, every side / uplo / trans / diag,
widths that hit each leaf variant (8, 16, 32, 64 and others) and the
recursive split above 64."""
import itertools

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def tri(rng, n, uplo, unit):
    T = rng.standard_normal((n, n)) / max(n, 1) + 2.0 * np.eye(n)
    T = np.triu(T) if uplo == "U" else np.tril(T)
    return np.asfortranarray(T)   # a unit solve ignores the diagonal


@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product("LR", "UL", "NT", "NU")))
def test_trsm_all_variants_vs_oracle(gpu, rng, side, uplo, trans, diag):
    for n, other in ((1, 5), (5, 3), (8, 70), (16, 33), (32, 129), (33, 7), (64, 300), (100, 65), (129, 40),
                     (256, 1000)):
        T = tri(rng, n, uplo, diag == "U")
        m, k = (n, other) if side == "L" else (other, n)
        B = np.asfortranarray(rng.standard_normal((m, k)))
        alpha = 0.75
        X = gpu.trsm_(B.copy(order="F"), T, alpha, side, uplo, trans, diag == "U")
        Xo = B.copy(order="F")
        assert O.dtrsm(Xo, T, alpha, side, uplo, trans, diag == "U") == 0
        scale = np.abs(Xo).max() + 1e-300
        assert np.abs(X - Xo).max() <= 1e-12 * scale * max(n, 1), (side, uplo, trans, diag, n, other)


def test_trsm_right_tall(gpu, rng):
    # the shape of U12 = A12 * (L11^T)^-1 in getrf's working layout: many rows, 256 columns
    n, rows = 256, 5000
    T = tri(rng, n, "U", True)
    B = np.asfortranarray(rng.standard_normal((rows, n)))
    X = gpu.trsm_(B.copy(order="F"), T, 1.0, "R", "U", "N", True)
    Xo = B.copy(order="F")
    assert O.dtrsm(Xo, T, 1.0, "R", "U", "N", True) == 0
    assert np.abs(X - Xo).max() <= 1e-11 * np.abs(Xo).max()


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product("UL", "NT", "NU")))
def test_trsm_right_tall_all_variants(gpu, rng, uplo, trans, diag):
    """Right-side solves with 1500 rows: when op(T) is upper and n a multiple
    of 64 they invert op(T)'s 64 x 64 diagonal blocks and replace each
    64-column sweep by an in-place block product; n = 100 and the lower op(T)
    keep the column sweeps."""
    for n in (64, 192, 320, 100):
        T = tri(rng, n, uplo, diag == "U")
        B = np.asfortranarray(rng.standard_normal((1500, n)))
        X = gpu.trsm_(B.copy(order="F"), T, -0.5, "R", uplo, trans, diag == "U")
        Xo = B.copy(order="F")
        assert O.dtrsm(Xo, T, -0.5, "R", uplo, trans, diag == "U") == 0
        scale = np.abs(Xo).max() + 1e-300
        assert np.abs(X - Xo).max() <= 1e-12 * scale * n, (uplo, trans, diag, n)


def test_trsm_singular_reported_before_write(gpu, rng):
    T = tri(rng, 20, "U", False)
    T[7, 7] = 0.0
    B = np.asfortranarray(rng.standard_normal((20, 3)))
    B0 = B.copy()
    with pytest.raises(gpu.SingularException) as e:
        gpu.trsm_(B, T, 1.0, "L", "U", "N", False)
    assert e.value.info == 8
    assert np.array_equal(B, B0)
