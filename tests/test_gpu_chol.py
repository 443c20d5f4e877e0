"""Cholesky factorization and solve on the GPU (oxi_dpotrf / oxi_dpotrs
through the C ABI) against the oracle and LAPACK dpotrf, following
test/lapack/factorizations.jl and test/api/linalg_api.jl."""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def spd(rng, n, shift=None):
    G = rng.standard_normal((n, n))
    return G @ G.T + (n if shift is None else shift) * np.eye(n)


@pytest.mark.parametrize("n", [1, 2, 17, 63, 64, 65, 130, 257, 1000, 2048])
def test_potrf_vs_oracle_and_lapack(gpu, rng, n):
    A = spd(rng, n)
    F = gpu.cholesky(A)
    L = F.factors.numpy()
    assert F.issuccess() and np.array_equal(np.triu(L, 1), np.zeros((n, n)))
    Lo, info = O.dpotrf(A, threads=8)
    assert info == 0
    np.testing.assert_allclose(L, Lo, rtol=0, atol=64 * n * EPS * np.abs(Lo).max())
    Lref = sl.cholesky(A, lower=True)
    np.testing.assert_allclose(L, Lref, rtol=0, atol=64 * n * EPS * np.abs(Lref).max())
    assert np.linalg.norm(L @ L.T - A) <= 30 * n * EPS * np.linalg.norm(A)


def test_potrf_reads_lower_triangle_only(gpu, rng):
    n = 300
    A = spd(rng, n)
    Au = A.copy(); Au[np.triu_indices(n, 1)] = np.nan
    F1 = gpu.cholesky_factor(gpu.DMatrix.from_numpy(Au))
    F2 = gpu.cholesky(A)
    assert np.array_equal(F1.factors.numpy(), F2.factors.numpy())


def test_potrf_info_and_exceptions(gpu, rng):
    ox = gpu
    # the last three run the blocked form (256-column blocks): a failing column
    # at the end and the start of a block, and inside a later one
    for n, j in ((10, 4), (100, 70), (300, 1), (300, 150), (700, 640), (1024, 256), (1024, 257), (1500, 1300)):
        A = spd(rng, n)
        A[j - 1, j - 1] = -1e3 * n
        with pytest.raises(ox.PosDefException) as e:
            ox.cholesky(A)
        assert e.value.info == j
        F = ox.cholesky(A, check=False)
        assert not F.issuccess() and F.info == j
        assert O.dpotrf(A)[1] == j
        with pytest.raises(ox.PosDefException):
            F.solve(np.ones(n))
    A = spd(rng, 9); A[5, 5] = np.nan
    # a NaN pivot stops the factorization at its column; on a host matrix the
    # symmetry check (LinearAlgebra.ishermitian compares the diagonal too)
    # already fails with PosDefException(-1)
    assert ox.cholesky(ox.DMatrix.from_numpy(A), check=False).info == 6
    assert ox.cholesky(A, check=False).info == -1
    N = rng.standard_normal((5, 5))
    with pytest.raises(ox.PosDefException) as e:
        ox.cholesky(N)                                   # not symmetric: PosDefException(-1)
    assert e.value.info == -1


def test_cholesky_in_place_on_view(gpu, rng):
    ox = gpu
    n = 200
    A = spd(rng, n)
    Big = rng.standard_normal((n + 5, n + 3))
    Big[2:n + 2, 1:n + 1] = A
    Bd = ox.DMatrix.from_numpy(Big)
    V = Bd.view(slice(2, n + 2), slice(1, n + 1))
    F = ox.cholesky_(V)
    assert F.factors is V
    after = Bd.numpy()
    assert np.array_equal(after[:2], Big[:2]) and np.array_equal(after[:, :1], Big[:, :1])
    assert np.array_equal(after[n + 2:], Big[n + 2:]) and np.array_equal(after[:, n + 1:], Big[:, n + 1:])
    L = after[2:n + 2, 1:n + 1]
    assert np.linalg.norm(L @ L.T - A) <= 30 * n * EPS * np.linalg.norm(A)


@pytest.mark.parametrize("n", [1, 9, 129, 1000])
def test_potrs_vs_oracle(gpu, rng, n):
    A = spd(rng, n)
    F = gpu.cholesky(A)
    L = F.factors.numpy()
    b = rng.standard_normal(n); B = rng.standard_normal((n, 7)); B64 = rng.standard_normal((n, 64))
    for rhs in (b, B, B64):
        x = F.solve(rhs)
        np.testing.assert_allclose(x, O.dpotrs(L, rhs), rtol=1e-10, atol=1e-12 * np.abs(x).max())
        assert np.linalg.norm(A @ x - rhs) <= 30 * n * EPS * np.linalg.norm(A) * np.linalg.norm(x)
    la, s = F.logabsdet()
    ld = np.linalg.slogdet(A)
    assert s == 1.0 and la == pytest.approx(ld[1], rel=1e-12)


def test_cholesky_n8192(gpu):
    """A larger factorization (8192): the reconstruction bound, checked on the
    GPU with torch as the checker."""
    import torch
    ox = gpu
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(11)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = G @ G.T + n * torch.eye(n, dtype=torch.float64, device="cuda")
    Ad = ox.DMatrix(A.reshape(-1).clone(), n, n)          # symmetric: row-major == column-major
    F = ox.cholesky_(Ad)
    Lt = Ad.buf.view(n, n)                                # Lt[j, i] = L(i, j)
    L = Lt.T
    err = torch.linalg.matrix_norm(L @ L.T - A).item() / (torch.linalg.matrix_norm(A).item() * n * EPS)
    assert err <= 30
    assert torch.count_nonzero(torch.triu(L, 1)).item() == 0
    # 64 right-hand sides (the chained solve): normwise backward error
    B = torch.randn(64, n, dtype=torch.float64, device="cuda", generator=g)   # B[c, i] = column c
    X = F.solve(ox.DMatrix(B.reshape(-1).clone(), n, 64))
    Xt = X.buf[:n * 64].view(64, n)
    R = A @ Xt.T - B.T
    berr = torch.linalg.matrix_norm(R).item() / (torch.linalg.matrix_norm(A).item() *
                                                 torch.linalg.matrix_norm(Xt).item() * n * EPS)
    assert berr <= 30, berr
