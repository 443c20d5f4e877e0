"""What is built on the LU factors, on the GPU through the C ABI, against the
oracle and closed forms: inverse (native_inverse), log|det|, complex
right-hand sides (_apply_split!), backward error, the Hager-Higham estimate,
refine_solve, expert_solve and factorize (src/lapack/dense_utilities.jl,
src/lapack/expert_refinement.jl, src/lapack/advanced_factorizations.jl:591-645,
src/api/linalg_api.jl:67-78,1101-1143)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.mark.parametrize("n", [1, 2, 65, 300, 1024])
def test_inverse_vs_oracle(gpu, rng, n):
    A = rng.standard_normal((n, n))
    F = gpu.lu(A)
    X = F.inv().numpy()
    Xo = O.dgetri(F.factors.numpy(), F.ipiv, threads=8)
    c = np.linalg.cond(A)
    np.testing.assert_allclose(X, Xo, rtol=0, atol=64 * n * EPS * c * np.abs(Xo).max())
    assert np.linalg.norm(A @ X - np.eye(n)) <= 30 * n * EPS * c
    np.testing.assert_allclose(gpu.inv(A), X, rtol=0, atol=0)
    S = A.copy(); S[:, n // 2] = 0
    with pytest.raises(gpu.SingularException):
        gpu.inv(S)


def test_logabsdet_det(gpu, rng):
    for n in (1, 7, 300, 2000):
        A = rng.standard_normal((n, n))
        F = gpu.lu(A)
        la, s = F.logabsdet()
        lo, so = O.dlogabsdet(F.factors.numpy(), F.ipiv)
        sn, ln = np.linalg.slogdet(A)
        assert s == so == sn and la == pytest.approx(lo, rel=1e-12) and la == pytest.approx(ln, rel=1e-10)
        assert gpu.logabsdet(A) == (la, s)
    A = rng.standard_normal((9, 9))
    assert gpu.det(A) == pytest.approx(np.linalg.det(A), rel=1e-10)
    S = A.copy(); S[:, 3] = 0
    G = gpu.lu(S, check=False)
    assert G.logabsdet() == (-np.inf, 0.0) and G.det() == 0.0 and gpu.det(S) == 0.0
    P = np.eye(6)[[1, 0, 2, 3, 5, 4]]
    assert gpu.logabsdet(P) == (0.0, 1.0)


def test_complex_rhs_on_real_factor(gpu, rng):
    # _apply_split!: a complex right-hand side on real factors is solved as
    # its real and imaginary parts (test/api/linalg_api.jl:23-88)
    n = 200
    A = rng.standard_normal((n, n))
    F = gpu.lu(A)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    for rhs in (b, B):
        x = F.solve(rhs)
        assert np.iscomplexobj(x)
        np.testing.assert_allclose(A @ x, rhs, rtol=1e-9, atol=1e-10)
        xt = F.transpose().solve(rhs)
        np.testing.assert_allclose(A.T @ xt, rhs, rtol=1e-9, atol=1e-10)
    y = b.copy()
    F.ldiv_(y)
    np.testing.assert_allclose(A @ y, b, rtol=1e-9, atol=1e-10)


def test_backward_error_vs_oracle(gpu, rng):
    for (m, n, k) in ((1, 1, 1), (30, 20, 3), (1000, 700, 1), (4096, 4096, 5)):
        A = rng.standard_normal((m, n)); X = rng.standard_normal((n, k)); B = rng.standard_normal((m, k))
        e, R = gpu.backward_error(A, X, B)
        eo, Ro = O.backward_error(A, X, B)
        np.testing.assert_allclose(R, Ro, rtol=1e-12, atol=1e-12 * np.abs(Ro).max())
        assert e == pytest.approx(eo, rel=1e-10)
    Ai = np.diag(np.arange(1.0, 6.0)); xi = np.arange(1.0, 6.0)
    assert gpu.backward_error(Ai, xi, Ai @ xi).error == 0.0
    assert gpu.backward_error(np.array([[np.inf, 0.0], [0.0, 1.0]]), np.ones(2), np.ones(2)).error == np.inf


def test_inverse_norm_estimate_vs_oracle(gpu, rng):
    for n in (1, 20, 300, 2000):
        A = rng.standard_normal((n, n))
        F = gpu.lu(A)
        est = gpu.inverse_norm_estimate(F)
        eo = O.inverse_norm_estimate(F.factors.numpy(), F.ipiv)
        assert est == pytest.approx(eo, rel=1e-8)
        if n <= 300:
            true = np.abs(np.linalg.inv(A)).sum(axis=0).max()
            assert est <= true * (1 + 1e-8) and est >= true / 10
    rc = gpu.reciprocal_condition(A)
    assert 0.1 / np.linalg.cond(A, 1) <= rc <= 1.0 / np.linalg.cond(A, 1) * (1 + 1e-8)


@pytest.mark.parametrize("n", [1, 50, 700])
def test_refine_solve_vs_oracle(gpu, rng, n):
    A = rng.standard_normal((n, n)) * np.logspace(-5, 5, n)[:, None]
    b = rng.standard_normal(n); B = rng.standard_normal((n, 3))
    F = gpu.lu(A)
    for rhs in (b, B):
        r = gpu.refine_solve(A, F, rhs)
        o = O.refine_solve(A, F.factors.numpy(), F.ipiv, rhs)
        assert r.converged and o["converged"] and r.backward_error <= 8 * EPS
        assert abs(r.iterations - o["iterations"]) <= 1
        c = np.linalg.cond(A, 1)
        np.testing.assert_allclose(r.x, o["x"], rtol=0, atol=1e-14 * c * np.abs(o["x"]).max())
        assert r.rcond_estimate == pytest.approx(o["rcond_estimate"], rel=1e-6)
        assert r.residual_norm <= 10 * max(o["residual_norm"], 1e-300) + 1e-12 * np.abs(rhs).max()


@pytest.mark.parametrize("n", [1, 50, 700])
def test_expert_solve_vs_oracle(gpu, rng, n):
    A = rng.standard_normal((n, n)) * np.logspace(-8, 8, n)[:, None] * np.logspace(3, -3, n)[None, :]
    b = rng.standard_normal(n)
    r = gpu.expert_solve(A, b)
    o = O.expert_solve(A, b)
    assert r.converged and o["converged"] and r.backward_error <= 8 * EPS
    assert np.array_equal(r.row_scale, o["row_scale"])
    np.testing.assert_allclose(r.column_scale, o["column_scale"], rtol=1e-15, atol=0)
    c = np.linalg.cond(A, 1)
    np.testing.assert_allclose(r.x, o["x"], rtol=0, atol=1e-14 * c * np.abs(o["x"]).max())
    assert isinstance(r.factor, gpu.LU) and r.factor.factors.shape == (n, n)
    r2 = gpu.expert_solve(A, b, equilibrate=False)
    assert r2.converged and np.all(r2.row_scale == 1.0)
    Z = rng.standard_normal((6, 6)); Z[3] = 0
    with pytest.raises(gpu.SingularException) as e:
        gpu.expert_solve(Z, np.ones(6))
    assert e.value.info == 4


def test_factorize_selects(gpu, rng):
    n = 100
    G = rng.standard_normal((n, n))
    S = G @ G.T + n * np.eye(n)
    F = gpu.factorize(S)
    assert isinstance(F, gpu.Cholesky)
    np.testing.assert_allclose(S @ F.solve(np.ones(n)), np.ones(n), rtol=1e-10)
    Fi = gpu.factorize(G + G.T)             # symmetric indefinite: an LU here
    assert isinstance(Fi, gpu.LU)
    Fg = gpu.factorize(G)
    assert isinstance(Fg, gpu.LU)
    np.testing.assert_allclose(G @ Fg.solve(np.ones(n)), np.ones(n), rtol=1e-9)
