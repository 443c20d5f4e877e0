"""Blocked LU with partial pivoting on the GPU (oxi_dgetrf / oxi_dgetrs through
the C ABI) against the oracle and LAPACK, following test/lapack/lapack_dense.jl,
test/lapack/factorizations.jl and test/api/linalg_api.jl."""
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
LU_CASES = ((1, 1, 32), (5, 5, 1), (17, 17, 16), (33, 20, 16), (20, 33, 16), (64, 64, 32), (130, 97, 32),
            (97, 130, 64), (257, 257, 32), (257, 257, 64))  # test/lapack/lapack_dense.jl:9


def bwd(E, A):
    """_bwd of test/lapack/lapack_dense.jl:4."""
    return np.linalg.norm(E) / (max(np.linalg.norm(A), np.finfo(float).tiny) * max(1, *A.shape) * EPS)


def factor(ox, A, bs=0, check=False):
    F = ox.lu(ox.DMatrix.from_numpy(A), check=check, blocksize=bs)
    return F, F.factors.numpy(), F.ipiv, F.info


def recon(Fh, ipiv, m, n):
    k = min(m, n)
    L = np.tril(Fh[:, :k], -1) + np.eye(m, k)
    U = np.triu(Fh[:k, :])
    return L @ U, O.perm_from_ipiv(ipiv, m)


@pytest.mark.parametrize("m,n,bs", LU_CASES)
def test_reference_cases(gpu, rng, m, n, bs):
    A = rng.standard_normal((m, n))
    F, Fh, ipiv, info = factor(gpu, A, bs)
    assert info == 0
    LU, p = recon(Fh, ipiv, m, n)
    assert bwd(LU - A[p - 1], A) <= 30
    # identical pivots to the oracle and to LAPACK getrf (lapack_dense.jl:12-14)
    _, oipiv, _ = O.dgetrf(A, bs)
    assert np.array_equal(ipiv, oipiv)
    _, lp = sl.lu_factor(A, check_finite=False)
    assert np.array_equal(ipiv, lp[:min(m, n)] + 1)


def test_blocksizes_and_shapes(gpu, rng):
    # test/lapack/factorizations.jl:2-11
    for (m, n) in ((1, 1), (9, 9), (13, 7), (7, 13), (67, 65), (65, 67)):
        for block in (1, 7, 32):
            A = rng.standard_normal((m, n))
            F, Fh, ipiv, info = factor(gpu, A, block)
            LU, p = recon(Fh, ipiv, m, n)
            np.testing.assert_allclose(LU, A[p - 1], rtol=3e-11, atol=3e-12 * max(1, n))
            _, oipiv, _ = O.dgetrf(A, block)
            assert np.array_equal(ipiv, oipiv)


def test_golden_fixtures(gpu):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "lu_cases.npz"), allow_pickle=False)
    for i in range(int(z["count"])):
        name = str(z[f"c{i}_name"])
        A = z[f"c{i}_A"]
        F, Fh, ipiv, info = factor(gpu, A, int(z[f"c{i}_bs"]))
        assert np.array_equal(ipiv, z[f"c{i}_ipiv"]), name
        assert info == int(z[f"c{i}_info"]), name
        if not np.isnan(A).any():
            m, n = A.shape
            LU, p = recon(Fh, ipiv, m, n)
            assert bwd(LU - A[p - 1], A) <= 30, name
            tol = 30 * max(m, n) * EPS * np.linalg.norm(A)
            assert np.linalg.norm(Fh - z[f"c{i}_F"]) <= tol * 1e3, name
        if f"c{i}_x" in z.files:
            np.testing.assert_allclose(F.solve(z[f"c{i}_b"]), z[f"c{i}_x"], rtol=1e-9, atol=1e-10)
            np.testing.assert_allclose(F.solve(z[f"c{i}_B"]), z[f"c{i}_X"], rtol=1e-9, atol=1e-10)
            np.testing.assert_allclose(F.transpose().solve(z[f"c{i}_b"]), z[f"c{i}_xT"], rtol=1e-9, atol=1e-10)


def test_singular_and_hard_inputs(gpu, rng):
    ox = gpu
    A = rng.standard_normal((40, 40)); A[:, 22] = 0
    with pytest.raises(ox.SingularException):
        ox.lu(A)
    F, Fh, ipiv, info = factor(ox, A)
    assert info == 23 and not F.issuccess()
    LU, p = recon(Fh, ipiv, 40, 40)
    assert bwd(LU - A[p - 1], A) <= 30
    F, Fh, ipiv, info = factor(ox, np.zeros((3, 3)))
    assert info == 1
    with pytest.raises(ox.SingularException):
        F.solve(np.ones(3))
    G = rng.standard_normal((80, 80)) @ np.diag(10.0 ** np.linspace(0, -12, 80))
    F, Fh, ipiv, info = factor(ox, G)
    LU, p = recon(Fh, ipiv, 80, 80)
    assert bwd(LU - G[p - 1], G) <= 30
    N = rng.standard_normal((60, 60)); N[:, -1] = N[:, 0] + EPS * rng.standard_normal(60)
    F, Fh, ipiv, info = factor(ox, N)
    LU, p = recon(Fh, ipiv, 60, 60)
    assert bwd(LU - N[p - 1], N) <= 30


@pytest.mark.parametrize("n", [600, 1000, 2500, 4096])
def test_multi_workgroup_panels(gpu, rng, n):
    """Panels taller than 256 rows run the multi-workgroup leaf (device-wide
    pivot exchange): pivots equal the oracle's, reconstruction <= 30, and a
    second run is bitwise identical (lapack_dense.jl:33-38)."""
    A = rng.standard_normal((n, n))
    F, Fh, ipiv, info = factor(gpu, A, 0)
    assert info == 0
    _, oipiv, _ = O.dgetrf(A, 128, threads=16)
    assert np.array_equal(ipiv, oipiv)
    LU, p = recon(Fh, ipiv, n, n)
    assert bwd(LU - A[p - 1], A) <= 30
    F2, Fh2, ipiv2, _ = factor(gpu, A, 0)
    assert np.array_equal(Fh, Fh2) and np.array_equal(ipiv, ipiv2)
    A[:, n // 2] = 0
    F, Fh, ipiv, info = factor(gpu, A, 0)
    assert info == n // 2 + 1


def test_solves(gpu, rng):
    ox = gpu
    for n in (1, 9, 64, 129, 300):
        A = rng.standard_normal((n, n))
        F = ox.lu(A)
        c = np.linalg.cond(A)
        b = rng.standard_normal(n); B = rng.standard_normal((n, 3)); B64 = rng.standard_normal((n, 64))
        assert bwd(A @ F.solve(b) - b, A) <= 30 * c
        assert bwd(A @ F.solve(B) - B, A) <= 30 * c
        assert bwd(A @ F.solve(B64) - B64, A) <= 30 * c
        assert bwd(A.T @ F.transpose().solve(b) - b, A) <= 30 * c
        assert bwd(A.T @ ox.solve_transposed(F, B, "C") - B, A) <= 30 * c
        # the solve forms agree (test/api/linalg_api_solve.jl:12-20)
        x1 = F.solve(b); x2 = ox.ldiv(F, b); x3 = ox.solve(F, b); x4 = ox.ldiv_(F, b.copy())
        y = np.zeros(n); ox.ldiv_(y, F, b)
        for x in (x2, x3, x4, y):
            np.testing.assert_allclose(x, x1, rtol=1e-12, atol=1e-13)
    with pytest.raises(ox.DimensionMismatch):
        ox.lu(rng.standard_normal((3, 3))).solve(np.ones(4))


def test_api_properties(gpu, rng):
    """test/api/linalg_api.jl:23-52."""
    ox = gpu
    n = 9
    A = rng.standard_normal((n, n)); b = rng.standard_normal(n); B = rng.standard_normal((n, 3))
    F = ox.lu(A)
    assert isinstance(F, ox.LU) and F.size == (n, n) and F.issuccess()
    np.testing.assert_allclose(F.L @ F.U, A[F.p - 1], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(F.P @ A, F.L @ F.U, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(F.solve(b), np.linalg.solve(A, b), rtol=1e-10)
    np.testing.assert_allclose(F.solve(B), np.linalg.solve(A, B), rtol=1e-10)
    np.testing.assert_allclose(ox.adjoint(F).solve(b), np.linalg.solve(A.T, b), rtol=1e-10)
    assert F.det() == pytest.approx(np.linalg.det(A), rel=1e-10)
    la, s = F.logabsdet()
    sd, ld = np.linalg.slogdet(A)
    assert la == pytest.approx(ld, rel=1e-12) and s == sd
    S = A.copy(); S[:, 1] = 0
    G = ox.lu(S, check=False)
    assert not G.issuccess() and G.det() == 0.0
    with pytest.raises(ox.SingularException):
        G.solve(b)


def test_lu_in_place_on_views(gpu, rng):
    """lu!(X) on a view with lda > m (test/api/linalg_api.jl:54-68): the view
    becomes the factors, the storage around it is untouched."""
    ox = gpu
    for n in (9, 300):
        b = rng.standard_normal(n)
        Big = rng.standard_normal((n + 7, n + 5))
        Bd = ox.DMatrix.from_numpy(Big)
        V = Bd.view(slice(3, n + 3), slice(2, n + 2))
        A = Big[3:n + 3, 2:n + 2].copy()
        F = ox.lu_(V)
        assert F.factors is V
        np.testing.assert_allclose(F.solve(b), np.linalg.solve(A, b), rtol=1e-9)
        after = Bd.numpy()
        assert np.array_equal(after[:3, :], Big[:3, :]) and np.array_equal(after[:, :2], Big[:, :2])
        assert np.array_equal(after[n + 3:, :], Big[n + 3:, :]) and np.array_equal(after[:, n + 2:], Big[:, n + 2:])
        LU, p = recon(after[3:n + 3, 2:n + 2], F.ipiv, n, n)
        assert bwd(LU - A[p - 1], A) <= 30


def test_full_size_16384(gpu):
    """n = 16384, as the benchmark: size-independent properties checked on
    the GPU in FP64 (torch is the checker here, the factorization is ours):
    normwise ||PA - LU|| / (||A|| n eps) <= 30 and a solve with backward error
    ||Ax - b|| / (||A|| ||x|| n eps) <= 30."""
    import torch
    ox = gpu
    n = 16384
    g = torch.Generator(device="cuda").manual_seed(0x0B1A5)
    At = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)   # At[j] = column j of A
    A = ox.DMatrix(At.reshape(-1).clone(), n, n)
    F = ox.lu(A, check=True)
    assert F.info == 0
    Ft = F.factors.buf.view(n, n)          # Ft[j, i] = factor(i, j)
    Lt = torch.tril(Ft.T, -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    Ut = torch.triu(Ft.T)
    p = torch.from_numpy(F.p - 1).cuda()
    PA = At.T[p]
    err = torch.linalg.matrix_norm(PA - Lt @ Ut).item() / (torch.linalg.matrix_norm(At).item() * n * EPS)
    assert err <= 30
    del Lt, Ut, PA
    b = torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g)
    x = F.solve(ox.DMatrix(b.reshape(-1).clone(), n, 1))
    xt = x.buf[:n].view(n, 1)
    r = At.T @ xt - b
    berr = torch.linalg.norm(r).item() / (torch.linalg.matrix_norm(At).item() * torch.linalg.norm(xt).item() * n * EPS)
    assert berr <= 30
