"""The benchmark configurations at their sizes on the GPU: the mixed
LU + refinement sweep (A = U diag(sigma) V^T, Haar U and V, geometric sigma,
kappa in {1e2, 1e4, 1e6, 1e8}, BF16 / FP16 / FP8 / FP32 trailing products,
n = 8192 and one n = 16384 case) with the convergence contract of
test/lapack/factorizations.jl:57-63 at the target's backward error
1e-15 n, and the adjoints of X = A \\ B (n = 8192, 64 right-hand sides,
32 central finite differences, h = 1e-6) and of gemm! (4096^3)."""
import math

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def haar_q(n, g):
    Q, R = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    return Q * torch.sign(torch.diagonal(R))


@pytest.fixture(scope="module")
def haar8192():
    g = torch.Generator(device="cuda").manual_seed(0x0B1A5)
    return haar_q(8192, g), haar_q(8192, g)


def conditioned_from(U, V, kappa):
    n = U.shape[0]
    s = kappa ** (-torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
    return (U * s) @ V.T


# Which (low type, solver) pairs must reach the target 1e-15 n at a given
# kappa, as measured on MI355X at n = 8192 (A = U diag(sigma) V^T):
#  - FP32 factors: classical refinement and GMRES-IR up to kappa = 1e6 (at
#    1e8, kappa u32 ~ 6 and neither converges);
#  - BF16 / FP16 trailing products (FP32 panel): GMRES-IR at kappa = 1e2 only;
#    the factor's error grows with the ~30 rounded trailing updates of an
#    n = 8192 matrix, and from kappa = 1e4 on the preconditioned system is too
#    far from the identity for restarted GMRES;
#  - FP8 (e4m3, per-block scales) trailing products: none.
# Every run must report honestly: converged == (backward error <= rtol), and
# a consistent reason.
MUST = {("float32", "ir"): 1e6, ("float32", "gmres"): 1e6, ("bfloat16", "gmres"): 1e2,
        ("float16", "gmres"): 1e2, ("float8_e4m3fn", "gmres"): 0.0}


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6, 1e8])
def test_mixed_sweep_n8192(gpu, haar8192, kappa):
    n = 8192
    U, V = haar8192
    At = conditioned_from(U, V, kappa)
    g = torch.Generator(device="cuda").manual_seed(7)
    xt = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    b = At @ xt
    A = gpu.DMatrix(At.T.contiguous().reshape(-1), n, n)
    bd = gpu.DMatrix(b.clone(), n, 1)
    rtol = 1e-15 * n
    results = {}
    for (low, solver), kmax in MUST.items():
        # a run expected to stall is bounded (its report is what is checked)
        maxiter = 30 if kappa <= kmax else 6
        r = gpu.mixed_solve(A, bd, lowtype=low, solver=solver, maxiter=maxiter, rtol=rtol, gmres_restart=50)
        x = r.x.buf[:n]
        fwd = (torch.linalg.norm(x - xt) / torch.linalg.norm(xt)).item()
        print(f"n={n} kappa={kappa:g} {low}/{solver}: converged={r.converged} berr={r.backward_error:.2e} "
              f"fwd={fwd:.2e} it={r.iterations}/{r.inner_iterations} reason={r.reason} "
              f"factor {r.factor_ms:.1f} ms refine {r.refine_ms:.1f} ms")
        results[(low, solver)] = (r, fwd)
    for (low, solver), (r, fwd) in results.items():
        # honest reporting in every case
        assert r.converged == (r.backward_error <= rtol)
        assert r.reason in ("converged", "iteration_limit", "nonfinite_iteration", "working_factor_fallback")
        assert (r.reason == "converged") == (r.converged and not r.fallback)
        if kappa <= MUST[(low, solver)]:
            assert r.converged, (low, solver, kappa, r.backward_error)
            # a backward-stable solve: the forward error is within the conditioning
            assert fwd <= 100 * kappa * EPS * math.sqrt(n)
    del A, bd, At, b


def test_mixed_n16384(gpu):
    n = 16384
    g = torch.Generator(device="cuda").manual_seed(0x0B1A5)
    U = haar_q(n, g)
    s = 1e6 ** (-torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
    At = (U * s) @ haar_q(n, g).T
    del U
    xt = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    b = At @ xt
    A = gpu.DMatrix(At.T.contiguous().reshape(-1), n, n)
    bd = gpu.DMatrix(b.clone(), n, 1)
    for low in ("float32", "bfloat16", "float16"):
        r = gpu.mixed_solve(A, bd, lowtype=low, solver="gmres", maxiter=30 if low == "float32" else 4,
                            rtol=1e-15 * n, gmres_restart=50)
        fwd = (torch.linalg.norm(r.x.buf[:n] - xt) / torch.linalg.norm(xt)).item()
        print(f"n={n} kappa=1e6 {low}/gmres: converged={r.converged} berr={r.backward_error:.2e} fwd={fwd:.2e} "
              f"it={r.iterations}/{r.inner_iterations} factor {r.factor_ms:.1f} ms refine {r.refine_ms:.1f} ms")
        assert r.converged == (r.backward_error <= 1e-15 * n)
        if low == "float32":
            assert r.converged
            assert fwd <= 100 * 1e6 * EPS * math.sqrt(n)
    # the FP32 factor with tournament pivoting (pivot 2) meets the same bar
    r = gpu.mixed_solve(A, bd, lowtype="float32", solver="gmres", pivot="tournament", maxiter=30,
                        rtol=1e-15 * n, gmres_restart=50)
    fwd = (torch.linalg.norm(r.x.buf[:n] - xt) / torch.linalg.norm(xt)).item()
    print(f"n={n} kappa=1e6 float32/gmres/tournament: converged={r.converged} berr={r.backward_error:.2e} "
          f"fwd={fwd:.2e} it={r.iterations}/{r.inner_iterations} factor {r.factor_ms:.1f} ms refine {r.refine_ms:.1f} ms")
    assert r.converged and not r.fallback and fwd <= 100 * 1e6 * EPS * math.sqrt(n)


def test_adjoint_ldiv_n8192(gpu):
    """X = A \\ B, A = N(0,1) + 8192 I, B 8192 x 64: Bbar = A^-T Xbar by the
    transposed solve on the stored factors against the oracle's transposed
    solve, Abar = -Bbar X^T against exact sampled dot products, and 32 central
    finite differences of <Xbar, A \\ B>."""
    n, k = 8192, 64
    rng = np.random.default_rng(0x0B1A5)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    B = rng.standard_normal((n, k)); Xbar = rng.standard_normal((n, k))
    X, back = gpu.rrule_ldiv(A, B)
    Abar, Bbar = back(Xbar)
    F, ipiv, info = O.dgetrf(A, 128, threads=16)
    Bo = O.dgetrs(F, ipiv, Xbar, "T", threads=16)
    np.testing.assert_allclose(Bbar, Bo, rtol=0, atol=1e-12 * np.abs(Bo).max())
    idx = rng.integers(0, n, (64, 2))
    for i, j in idx:
        exact = -math.fsum(Bbar[i, c] * X[j, c] for c in range(k))
        assert abs(Abar[i, j] - exact) <= 1e-13 * math.fsum(abs(Bbar[i, c] * X[j, c]) for c in range(k)) + 1e-300
    # finite differences on the device (the perturbed systems are solved by
    # the forward path, independent of the pullback)
    At = torch.from_numpy(np.asfortranarray(A).reshape(-1, order="F")).cuda()     # column-major
    Bt = torch.from_numpy(np.asfortranarray(B).reshape(-1, order="F")).cuda()
    Xbt = torch.from_numpy(np.asfortranarray(Xbar).reshape(-1, order="F")).cuda()
    Abt = torch.from_numpy(np.asfortranarray(Abar).reshape(-1, order="F")).cuda()
    Bbt = torch.from_numpy(np.asfortranarray(Bbar).reshape(-1, order="F")).cuda()
    h = 1e-6
    g = torch.Generator(device="cuda").manual_seed(5)

    def f(dA, dB, sgn):
        Ap = gpu.DMatrix(At + sgn * h * dA, n, n)
        Xp = gpu.lu_(Ap).solve(gpu.DMatrix(Bt + sgn * h * dB, n, k))
        return float(torch.sum(Xbt * Xp.buf[:n * k]).item())
    for _ in range(32):
        dA = torch.randn(n * n, dtype=torch.float64, device="cuda", generator=g)
        dB = torch.randn(n * k, dtype=torch.float64, device="cuda", generator=g)
        fd = (f(dA, dB, 1) - f(dA, dB, -1)) / (2 * h)
        an = float(torch.sum(Abt * dA).item() + torch.sum(Bbt * dB).item())
        assert fd == pytest.approx(an, rel=1e-5, abs=1e-7)


def test_gemm_adjoints_4096(gpu):
    """The gemm! pullback at m = n = k = 4096: Abar = Cbar B^T, Bbar = A^T Cbar
    against exact (fsum) dot products on sampled entries."""
    n = 4096
    rng = np.random.default_rng(11)
    A = rng.standard_normal((n, n)); B = rng.standard_normal((n, n)); Cbar = rng.standard_normal((n, n))
    C, back = gpu.rrule_mul(A, B, 1.0, "N", "N")
    Abar, Bbar = back(Cbar)
    for i, j in rng.integers(0, n, (48, 2)):
        ea = math.fsum(Cbar[i, c] * B[j, c] for c in range(n))
        eb = math.fsum(A[c, i] * Cbar[c, j] for c in range(n))
        ec = math.fsum(A[i, c] * B[c, j] for c in range(n))
        for got, ex, terms in ((Abar[i, j], ea, Cbar[i] * B[j]), (Bbar[i, j], eb, A[:, i] * Cbar[:, j]),
                               (C[i, j], ec, A[i] * B[:, j])):
            assert abs(got - ex) <= n * EPS * math.fsum(np.abs(terms))
