"""Mixed-precision LU with FP64 refinement, equilibration, the FP32 and
low-precision building blocks, the single-RHS triangular solves and the
adjoints, on the GPU through the C ABI, against the oracle and closed forms
(test/lapack/factorizations.jl:57-63 for the mixed-solve contract)."""
import numpy as np
import pytest
import torch

import oracle as O
from test_oracle import exact_radix_scale, radix_boundary_values

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def solve16(gpu, A, b, **kw):
    """mixed_solve run to the internal defaults of _mixed_precision_impl
    (maxiter 20, rtol 16 eps: src/lapack/extended_lapack.jl:175) unless given."""
    kw.setdefault("maxiter", 20)
    kw.setdefault("rtol", 16 * EPS)
    return gpu.mixed_solve(A, b, **kw)


def haar(rng, n):
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    return Q * np.sign(np.diag(R))


def conditioned(rng, n, kappa):
    """A = U diag(sigma) V^T, sigma_i = kappa^(-(i-1)/(n-1)) (Haar U, V)."""
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    return (haar(rng, n) * s) @ haar(rng, n).T


# ---------------------------------------------------------------------------
# equilibrate (src/lapack/dense_utilities.jl:191-212): bit-exact vs the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("radix", [True, False])
@pytest.mark.parametrize("m,n", [(1, 1), (50, 37), (37, 50), (300, 300), (600, 257)])
def test_equilibrate_bit_exact(gpu, rng, m, n, radix):
    A = rng.standard_normal((m, n)) * 2.0 ** rng.uniform(-40, 40, (m, 1)) * 2.0 ** rng.uniform(-40, 40, (1, n))
    if m > 3:
        A[2] = 0.0
    E, r, c = gpu.equilibrate(A, radix=radix)
    Eo, ro, co = O.dequilibrate(A, radix=radix)
    assert np.array_equal(r, ro) and np.array_equal(c, co)
    assert np.array_equal(E, Eo)


def test_equilibrate_radix_boundaries(gpu):
    v = radix_boundary_values()
    A = np.column_stack([v, v[::-1]])
    _, r, c = gpu.equilibrate(A, radix=True)
    _, ro, co = O.dequilibrate(A, radix=True)
    assert np.array_equal(r, ro) and np.array_equal(c, co)
    assert all(r[i] == exact_radix_scale(max(abs(A[i, 0]), abs(A[i, 1]))) for i in range(len(v)))


def test_equilibrate_rejects_non_finite(gpu, rng):
    A = rng.standard_normal((20, 20)); A[4, 5] = np.inf
    with pytest.raises(ValueError):
        gpu.equilibrate(A, radix=True)


# ---------------------------------------------------------------------------
# single right-hand-side solves (the chained trsv kernel) in both directions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 7, 127, 128, 129, 255, 257, 1000, 2050])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_vector_solves_vs_oracle(gpu, rng, n, nrhs):
    A = rng.standard_normal((n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, nrhs))
    F = gpu.lu(A)
    LU, ipiv = F.factors.numpy(), F.ipiv
    for trans in ("N", "T"):
        X = F.solve(B, trans) if trans == "N" else F.transpose().solve(B)
        Xo = O.dgetrs(LU, ipiv, B, trans=trans)
        np.testing.assert_allclose(X, Xo, rtol=1e-11, atol=1e-12 * np.abs(Xo).max())
        R = (A if trans == "N" else A.T) @ X - B
        assert np.linalg.norm(R) <= 30 * n * EPS * np.linalg.norm(A) * np.linalg.norm(X) + 1e-300


def test_vector_solve_padded_leading_dimension(gpu, rng):
    n = 301   # odd: no vector loads; ld padded
    A = rng.standard_normal((n, n)) + 30 * np.eye(n)
    Ad = gpu.DMatrix.from_numpy(A, ld=n + 5)
    F = gpu.lu_(Ad)
    b = rng.standard_normal(n)
    x = F.solve(b)
    np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-10)
    xt = F.transpose().solve(b)
    np.testing.assert_allclose(A.T @ xt, b, rtol=1e-10, atol=1e-10)


# ---------------------------------------------------------------------------
# FP32 factor and solve; low-precision trailing product
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [5, 64, 300, 1000])
def test_fp32_lu_matches_oracle(gpu, rng, n):
    A = rng.standard_normal((n, n)).astype(np.float32)
    F = gpu.lu_lowp(A, "float32")
    Fo, ipo, info = O.sgetrf(A)
    assert F.info == 0 and info == 0
    assert np.array_equal(F.pivots, ipo)
    Fh = F.factors.numpy()
    # both factors are Float32 eliminations with the same pivots in different
    # summation orders: they agree to a few n*eps32 of the factor's size, and
    # each reconstructs P*A to the bound of test/lapack/lapack_dense.jl:11
    # taken at Float32 eps
    e32 = float(np.finfo(np.float32).eps)
    assert np.abs(Fh.astype(np.float64) - Fo).max() <= 4 * n * e32 * max(1.0, np.abs(Fo).max())
    L = np.tril(Fh.astype(np.float64), -1) + np.eye(n); U = np.triu(Fh.astype(np.float64))
    p = O.perm_from_ipiv(F.pivots, n)
    A64 = A.astype(np.float64)
    assert np.linalg.norm(L @ U - A64[p - 1]) <= 30 * n * e32 * np.linalg.norm(A64)
    b = rng.standard_normal(n).astype(np.float32)
    x = F.solve(b)
    xo = O.sgetrs(Fo, ipo, b)
    np.testing.assert_allclose(x, xo, rtol=2e-3, atol=2e-3 * np.abs(xo).max())


def _round(x, dt):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(torch.float64).numpy()


def _fp8_block_round(X, axis):
    """Power-of-two scale per 256 rows (axis 0) / columns (axis 1) with
    max/scale <= 448, then e4m3 round-to-nearest-even; returns (values, scales)."""
    X = X.astype(np.float32)
    out = np.empty(X.shape, dtype=np.float64)
    nb = -(-X.shape[axis] // 256)
    scales = np.empty(nb)
    for b in range(nb):
        sl = (slice(b * 256, (b + 1) * 256), slice(None)) if axis == 0 else (slice(None), slice(b * 256, (b + 1) * 256))
        blk = X[sl]
        m = np.float32(np.abs(blk).max())
        s = 1.0
        if m > 0:
            s = float(np.ldexp(1.0, int(np.frexp(np.float32(m / np.float32(448.0)))[1])))
        scales[b] = s
        out[sl] = _round(blk * np.float32(1.0 / s), torch.float8_e4m3fn) * s
    return out, scales


@pytest.mark.parametrize("n", [64, 300, 1000, 3000])
def test_tournament_lu_matches_oracle(gpu, rng, n):
    # oxi_sgetrf_tournament against the oracle's restatement (tp_leaf): the
    # same pivots, factors to a few n*eps32, and the reconstruction bound
    A = rng.standard_normal((n, n)).astype(np.float32)
    F = gpu.lu_lowp(A, "float32", pivot="tournament")
    Fo, ipo, info = O.sgetrf(A, tournament=True, threads=8)
    assert F.info == 0 and info == 0
    Fh = F.factors.numpy()
    e32 = float(np.finfo(np.float32).eps)
    # the first leaf sees A itself, the same operations in both: its pivots
    # (a tournament over the waves of up to 12 workgroups here) agree
    # exactly; later leaves see trailing updates rounded in different orders,
    # and every round of the tournament is a partial-pivoting search in which
    # the top two candidates of a column can lie closer than that rounding
    # (seen at n = 1000)
    assert np.array_equal(F.pivots[:32], ipo[:32])
    if n <= 300:
        assert np.array_equal(F.pivots, ipo)
        assert np.abs(Fh.astype(np.float64) - Fo).max() <= 4 * n * e32 * max(1.0, np.abs(Fo).max())
    L = np.tril(Fh.astype(np.float64), -1) + np.eye(n); U = np.triu(Fh.astype(np.float64))
    p = O.perm_from_ipiv(F.pivots, n)
    A64 = A.astype(np.float64)
    assert np.linalg.norm(L @ U - A64[p - 1]) <= 30 * n * e32 * np.linalg.norm(A64)
    # a zero column is reported at its step, as by the partial-pivoting leaf
    Z = A.copy(); Z[:, min(n - 1, 40)] = 0
    Fz = gpu.lu_lowp(Z, "float32", pivot="tournament", check=False)
    assert Fz.info == O.sgetrf(Z, tournament=True)[2] == min(n - 1, 40) + 1


@pytest.mark.parametrize("m", [40, 64, 257, 1000, 4097, 8192, 8193, 12345, 16384])
def test_tournament_leaf_bitwise(gpu, rng, m):
    # one leaf (m x 32): every round of the tournament -- the waves, the
    # workgroup, one to three tree levels over up to 32 workgroups -- on the
    # same data as the oracle, so the pivots and the factors agree bit for bit
    A = rng.standard_normal((m, 32)).astype(np.float32)
    A[rng.integers(0, m, 8), rng.integers(0, 32, 8)] *= 64.0      # some dominant entries
    F = gpu.lu_lowp(A, "float32", pivot="tournament")
    Fo, ipo, info = O.sgetrf(A, tournament=True)
    assert F.info == info == 0
    assert np.array_equal(F.pivots, ipo)
    assert np.array_equal(F.factors.numpy(), Fo)


def test_tournament_lu_n8192_and_solve(gpu):
    """Larger panels (512-row blocks beyond 8192 rows do not occur; 32 blocks
    of 256 rows per leaf at the top): reconstruction on the GPU in FP64, and
    the mixed solve with tournament pivoting to the same accuracy."""
    import torch
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(7)
    A = torch.randn(n, n, dtype=torch.float32, device="cuda", generator=g)
    Ad = gpu.DMatrix(A.T.contiguous().reshape(-1), n, n)     # column-major float32
    F = gpu.lu_lowp(Ad, "float32", pivot="tournament")
    Ft = F.factors.buf[: n * n].view(n, n).T.double()          # column-major -> (row, col)
    L = torch.tril(Ft, -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    U = torch.triu(Ft)
    p = torch.from_numpy(O.perm_from_ipiv(F.pivots, n) - 1).to("cuda")
    e32 = float(np.finfo(np.float32).eps)
    A64 = A.double()
    err = torch.linalg.matrix_norm(L @ U - A64[p]).item() / (n * e32 * torch.linalg.matrix_norm(A64).item())
    assert err <= 30
    rng = np.random.default_rng(3)
    B = rng.standard_normal((2048, 2048)) + 2048 * np.eye(2048)
    b = rng.standard_normal(2048)
    r = gpu.mixed_solve(B, b, lowtype="float32", pivot="tournament", maxiter=10, rtol=16 * EPS)
    assert r.converged and r.backward_error <= 16 * EPS and not r.fallback


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_mixed_solve_tournament_conditioned(gpu, rng, kappa):
    n = 1024
    A = conditioned(rng, n, kappa); xt = rng.standard_normal(n); b = A @ xt
    for solver in ("ir", "gmres"):
        r = gpu.mixed_solve(A, b, lowtype="float32", solver=solver, pivot="tournament", maxiter=30, rtol=16 * EPS)
        assert r.converged and r.backward_error <= 16 * EPS and not r.fallback
        assert np.linalg.norm(r.x - xt) / np.linalg.norm(xt) <= 50 * kappa * EPS


@pytest.mark.parametrize("prec", ["bfloat16", "float16", "float8_e4m3fn"])
@pytest.mark.parametrize("M,N,K", [(300, 260, 200), (512, 640, 256), (33, 17, 5)])
def test_lowp_update_vs_emulation(gpu, rng, prec, M, N, K):
    A = (rng.standard_normal((M, K)) * np.exp(rng.uniform(-2, 2, (M, 1)))).astype(np.float32)
    B = (rng.standard_normal((K, N)) * np.exp(rng.uniform(-2, 2, (1, N)))).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    alpha = -1.0
    Cd = gpu.DMatrix.from_numpy(C0, dtype=np.float32)
    gpu.gemm_lowp_update_(Cd, gpu.DMatrix.from_numpy(A, dtype=np.float32), gpu.DMatrix.from_numpy(B, dtype=np.float32),
                          prec, alpha)
    C = Cd.numpy().astype(np.float64)
    if prec == "float8_e4m3fn":
        Ar, sa = _fp8_block_round(A, 0)
        Br, sb = _fp8_block_round(B, 1)
    else:
        dt = torch.bfloat16 if prec == "bfloat16" else torch.float16
        Ar, Br = _round(A, dt), _round(B, dt)
    ref = C0.astype(np.float64) + alpha * (Ar @ Br)
    bound = 4 * (K + 8) * 2.0 ** -24 * (np.abs(Ar) @ np.abs(Br) + np.abs(C0)) + 1e-30
    if prec == "float8_e4m3fn":
        # e4m3 subnormals (below 2^-6 after scaling): one subnormal step
        # (2^-9 of the block scale) per operand, for a conversion that treats
        # them differently from torch's
        ra = np.repeat(sa, 256)[:M][:, None]; rb = np.repeat(sb, 256)[:N][None, :]
        bound = bound + 2.0 ** -9 * (ra * np.abs(Br).sum(0)[None, :] + rb * np.abs(Ar).sum(1)[:, None])
    assert np.all(np.abs(C - ref) <= bound), np.max(np.abs(C - ref) / bound)
    # the rounding matters: FP32 operands would be outside the bound somewhere
    exact = C0.astype(np.float64) + alpha * (A.astype(np.float64) @ B.astype(np.float64))
    assert np.any(np.abs(exact - ref) > bound)


@pytest.mark.parametrize("prec", ["bfloat16", "float16", "float8_e4m3fn"])
def test_lowp_update_aligned_and_unaligned_c_agree(gpu, rng, prec):
    """Tiles inside an aligned C (ld % 4 == 0, 16-byte base) take 16-byte C
    accesses; the same update on an unaligned C (odd ld, or a base one float
    off) takes per-element accesses.  Both must give the same bits."""
    M, N, K = 640, 520, 256
    A = gpu.DMatrix.from_numpy(rng.standard_normal((M, K)).astype(np.float32), dtype=np.float32)
    B = gpu.DMatrix.from_numpy((rng.standard_normal((K, N)) * 4).astype(np.float32), dtype=np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    outs = []
    for ld, off in ((M, 0), (M + 1, 0), (M + 4, 1)):
        buf = torch.zeros(off + ld * N, dtype=torch.float32, device="cuda")
        buf[off:].view(N, ld)[:, :M].copy_(torch.from_numpy(C0.T.copy()))
        Cd = gpu.DMatrix(buf, M, N, ld, off)
        gpu.gemm_lowp_update_(Cd, A, B, prec, -0.75)
        outs.append(Cd.numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert not np.array_equal(outs[0], C0)


@pytest.mark.parametrize("prec", ["bfloat16", "float16", "float8_e4m3fn"])
def test_lowp_lu_is_a_good_preconditioner(gpu, rng, prec):
    n = 1024
    A = (rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)).astype(np.float32)
    F = gpu.lu_lowp(A, prec)
    F32 = gpu.lu_lowp(A, "float32")
    assert np.array_equal(F.pivots[:32], F32.pivots[:32])   # the first panel is exact FP32
    b = rng.standard_normal(n).astype(np.float32)
    x = F.solve(b)
    rel = np.linalg.norm(A.astype(np.float64) @ x - b) / np.linalg.norm(b)
    assert rel < {"bfloat16": 0.1, "float16": 0.02, "float8_e4m3fn": 0.5}[prec]


# ---------------------------------------------------------------------------
# mixed_solve (src/lapack/extended_lapack.jl:175-226)
# ---------------------------------------------------------------------------
def test_mixed_solve_defaults_match_oracle(gpu, rng):
    # the Float64 method's defaults maxiter = 10, rtol = 1e-13
    # (src/lapack/factorizations.jl:425), on the device and in the oracle
    for n in (1, 17, 67, 300):
        A = rng.standard_normal((n, n)) + n * np.eye(n); b = rng.standard_normal(n)
        r = gpu.mixed_solve(A, b)
        o = O.mixed_solve(A, b)
        assert r.converged == o["converged"] and r.converged and r.backward_error <= 1e-13
        assert abs(r.iterations - o["iterations"]) <= 1 and r.iterations <= 10
        np.testing.assert_allclose(r.x, o["x"], rtol=1e-11, atol=1e-13)
    # an iteration cap the solve cannot meet is reported, as in the oracle
    A = conditioned(rng, 256, 1e6); b = rng.standard_normal(256)
    r = gpu.mixed_solve(A, b, maxiter=1, rtol=1e-300)
    o = O.mixed_solve(A, b, maxiter=1, rtol=1e-300)
    assert not r.converged and not o["converged"] and r.iterations == o["iterations"] == 1
    assert r.reason == o["reason"] == "iteration_limit"


def test_mixed_solve_contract(gpu, rng):
    # test/lapack/factorizations.jl:57-63
    for n in (1, 17, 67, 300):
        A = rng.standard_normal((n, n)) + n * np.eye(n); b = rng.standard_normal(n)
        r = solve16(gpu, A, b)
        assert r.converged and r.iterations <= 10 and not r.fallback and r.reason == "converged"
        assert r.backward_error <= 16 * EPS and r.finite and r.low_type == "Float32"
        np.testing.assert_allclose(A @ r.x, b, rtol=2e-12, atol=2e-12)
        o = O.mixed_solve(A, b, maxiter=20, rtol=16 * EPS)
        assert o["converged"] and abs(o["iterations"] - r.iterations) <= 2
        np.testing.assert_allclose(r.x, o["x"], rtol=1e-12, atol=1e-14)
        scale = np.linalg.norm(A) * np.linalg.norm(r.x) + np.linalg.norm(b)
        assert abs(r.residual - np.linalg.norm(b - A @ r.x)) <= 1e-14 * scale


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
def test_mixed_solve_fp32_ir_conditioned(gpu, rng, kappa):
    n = 512
    A = conditioned(rng, n, kappa); xt = rng.standard_normal(n); b = A @ xt
    r = solve16(gpu, A, b, lowtype="float32", solver="ir")
    assert r.converged and r.backward_error <= 16 * EPS
    assert np.linalg.norm(r.x - xt) / np.linalg.norm(xt) <= 50 * kappa * EPS
    # the componentwise error agrees with the oracle's evaluation of the same x
    _, e = O.refinement_residual(A, r.x, b)
    assert r.backward_error <= max(4 * e, 4 * EPS) and e <= 16 * EPS


@pytest.mark.parametrize("prec,kappa", [("bfloat16", 1e2), ("float16", 1e2), ("float16", 1e4), ("float8_e4m3fn", 1e1)])
def test_mixed_solve_gmres_lowp(gpu, rng, prec, kappa):
    n = 768
    A = conditioned(rng, n, kappa); xt = rng.standard_normal(n); b = A @ xt
    r = solve16(gpu, A, b, lowtype=prec, solver="gmres", maxiter=20, gmres_restart=50)
    assert r.converged and r.backward_error <= 16 * EPS, r
    assert r.solver == "gmres" and r.inner_iterations >= 1
    assert np.linalg.norm(r.x - xt) / np.linalg.norm(xt) <= 100 * kappa * EPS


def test_mixed_solve_auto_switches(gpu, rng):
    n = 512
    # FP8 factors (u = 2^-4) at kappa = 10: the classical iteration cannot halve the error
    A = conditioned(rng, n, 1e1); b = rng.standard_normal(n)
    r = solve16(gpu, A, b, lowtype="float8_e4m3fn", solver="auto", maxiter=30, gmres_restart=50)
    assert r.converged and r.solver == "gmres"
    A = conditioned(rng, n, 1e4)
    r2 = solve16(gpu, A, b, lowtype="float32", solver="auto")
    assert r2.converged and r2.solver == "ir"


def test_mixed_solve_multiple_rhs(gpu, rng):
    n = 400
    A = conditioned(rng, n, 1e3)
    for nrhs in (3, 9):
        B = rng.standard_normal((n, nrhs))
        r = solve16(gpu, A, B)
        assert r.converged and r.x.shape == (n, nrhs)
        _, e = O.refinement_residual(A, r.x, B)
        assert e <= 16 * EPS
        r = solve16(gpu, A, B, lowtype="bfloat16", solver="gmres", maxiter=20)
        assert r.converged


# ---------------------------------------------------------------------------
# mixed_solve with two-level random butterflies and a factor without pivoting
# (pivot="rbt"): a wrong transform or a mismatch between the transform of the
# matrix and of the corrections would leave the refinement without a
# contraction, so convergence to 16 eps pins the whole chain
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 8, 64, 300, 301, 1024])
def test_mixed_solve_rbt_contract(gpu, rng, n):
    A = rng.standard_normal((n, n)) + n * np.eye(n); b = rng.standard_normal(n)
    r = solve16(gpu, A, b, pivot="rbt")
    assert r.converged and not r.fallback and r.backward_error <= 16 * EPS and r.iterations <= 10, r
    o = O.mixed_solve(A, b, maxiter=20, rtol=16 * EPS)
    np.testing.assert_allclose(r.x, o["x"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("prec,solver,kappa", [("float32", "gmres", 1e6), ("float32", "ir", 1e3),
                                               ("float32", "auto", 1e5),
                                               # below Float32 the option keeps partial pivoting
                                               ("bfloat16", "gmres", 1e2), ("float16", "gmres", 1e3)])
def test_mixed_solve_rbt_conditioned(gpu, rng, prec, solver, kappa):
    n = 1024
    A = conditioned(rng, n, kappa); xt = rng.standard_normal(n); b = A @ xt
    r = solve16(gpu, A, b, lowtype=prec, solver=solver, maxiter=30, gmres_restart=50, pivot="rbt")
    assert r.converged and r.backward_error <= 16 * EPS and not r.fallback, r
    assert np.linalg.norm(r.x - xt) / np.linalg.norm(xt) <= 100 * kappa * EPS
    _, e = O.refinement_residual(A, r.x, b)
    assert e <= 16 * EPS


def test_mixed_solve_rbt_zero_diagonal(gpu, rng):
    # every diagonal entry is zero: LU without pivoting of A itself stops at
    # the first step; after the butterflies the factor needs no interchange
    n = 512
    A = 3.0 * np.roll(np.eye(n), 1, axis=1) + 0.1 * rng.standard_normal((n, n))
    np.fill_diagonal(A, 0.0)
    b = rng.standard_normal(n)
    r = solve16(gpu, A, b, pivot="rbt", solver="gmres")
    assert r.converged and r.backward_error <= 16 * EPS, r


def test_mixed_solve_rbt_multiple_rhs_and_determinism(gpu, rng):
    n = 512
    A = conditioned(rng, n, 1e3)
    for nrhs in (3, 9):
        B = rng.standard_normal((n, nrhs))
        r = solve16(gpu, A, B, pivot="rbt")
        assert r.converged and r.x.shape == (n, nrhs)
        _, e = O.refinement_residual(A, r.x, B)
        assert e <= 16 * EPS
        r2 = solve16(gpu, A, B, pivot="rbt")
        assert np.array_equal(r.x, r2.x)
    with pytest.raises(ValueError):
        solve16(gpu, A, B, pivot="complete")


def test_mixed_solve_fallback_and_errors(gpu, rng):
    # singular after rounding to Float32: the FP64 factor is used
    n = 64
    A = np.eye(n) * 3.0
    A[10:12, 10:12] = [[1.0, 1.0], [1.0, 1.0 + 1e-12]]
    b = rng.standard_normal(n)
    r = solve16(gpu, A, b)
    assert r.fallback and r.reason == "working_factor_fallback" and r.converged
    _, e = O.refinement_residual(A, r.x, b)     # kappa ~ 4e12: a backward-error check
    assert e <= 16 * EPS
    # exactly singular in both precisions
    S = rng.standard_normal((n, n)); S[:, 7] = 0
    with pytest.raises(gpu.SingularException):
        solve16(gpu, S, b)
    bad = rng.standard_normal((n, n)); bad[3, 3] = np.nan
    with pytest.raises(ValueError):
        solve16(gpu, bad, b)
    with pytest.raises(ValueError):
        solve16(gpu, A, np.full(n, np.inf))
    with pytest.raises(ValueError):
        solve16(gpu, A, b, maxiter=-1)
    with pytest.raises(gpu.DimensionMismatch):
        solve16(gpu, rng.standard_normal((4, 5)), np.ones(4))
    # maxiter = 0: the first low-precision solve only, reported as not converged
    r0 = solve16(gpu, conditioned(rng, 128, 1e6), np.ones(128), maxiter=0)
    assert r0.iterations == 0 and (r0.converged == (r0.backward_error <= 16 * EPS))
    # empty
    r = solve16(gpu, np.zeros((0, 0)), np.zeros(0))
    assert r.converged and r.x.shape == (0,)


def test_mixed_solve_deterministic(gpu, rng):
    n = 1000
    A = conditioned(rng, n, 1e4); b = rng.standard_normal(n)
    r1 = solve16(gpu, A, b, lowtype="float16", solver="gmres")
    r2 = solve16(gpu, A, b, lowtype="float16", solver="gmres")
    assert np.array_equal(r1.x, r2.x) and r1.backward_error == r2.backward_error


def test_mixed_solve_large(gpu, rng):
    n = 4096
    A = rng.standard_normal((n, n)); xt = rng.standard_normal(n); b = A @ xt
    for prec, solver in (("float32", "ir"), ("bfloat16", "gmres"), ("float16", "auto")):
        r = solve16(gpu, A, b, lowtype=prec, solver=solver, maxiter=30)
        assert r.converged and r.backward_error <= 16 * EPS, (prec, solver, r.backward_error)


# ---------------------------------------------------------------------------
# adjoints: closed forms and central finite differences
# ---------------------------------------------------------------------------
def test_ldiv_pullback_closed_form_and_fd(gpu, rng):
    n, k = 200, 3
    A = rng.standard_normal((n, n)) + 2 * np.sqrt(n) * np.eye(n); B = rng.standard_normal((n, k))
    X, back = gpu.rrule_ldiv(A, B)
    Xbar = rng.standard_normal((n, k))
    Abar, Bbar = back(Xbar)
    Bref = np.linalg.solve(A.T, Xbar)
    np.testing.assert_allclose(Bbar, Bref, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(Abar, -Bref @ X.T, rtol=1e-10, atol=1e-12)
    h = 1e-6
    f = lambda A_, B_: float(np.sum(Xbar * np.linalg.solve(A_, B_)))   # noqa: E731
    for _ in range(4):
        dA = rng.standard_normal((n, n)); dB = rng.standard_normal((n, k))
        fd = (f(A + h * dA, B + h * dB) - f(A - h * dA, B - h * dB)) / (2 * h)
        an = np.sum(Abar * dA) + np.sum(Bbar * dB)
        assert fd == pytest.approx(an, rel=1e-6, abs=1e-8)
    # vector right-hand side
    b = rng.standard_normal(n)
    x, back = gpu.rrule_ldiv(A, b)
    Ab, bb = back(np.ones(n))
    np.testing.assert_allclose(bb, np.linalg.solve(A.T, np.ones(n)), rtol=1e-10)
    assert Ab.shape == (n, n) and bb.shape == (n,)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_pullback(gpu, rng, ta, tb):
    m, n, k = 70, 50, 40
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    op = lambda X, t: X if t == "N" else X.T   # noqa: E731
    C, back = gpu.rrule_mul(A, B, 1.5, ta, tb)
    np.testing.assert_allclose(C, 1.5 * op(A, ta) @ op(B, tb), rtol=1e-12, atol=1e-12)
    Cbar = rng.standard_normal((m, n))
    Abar, Bbar = back(Cbar)
    np.testing.assert_allclose(op(Abar, ta), 1.5 * Cbar @ op(B, tb).T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(op(Bbar, tb), 1.5 * op(A, ta).T @ Cbar, rtol=1e-12, atol=1e-12)
    Ab, Bb, C0b = gpu.gemm_pullback(A, B, Cbar, 1.5, -0.5, ta, tb)
    np.testing.assert_allclose(C0b, -0.5 * Cbar, rtol=0, atol=0)


# ---------------------------------------------------------------------------
# mixed_solve(A, b; factorization=:cholesky) (src/lapack/extended_lapack.jl:175-226)
# ---------------------------------------------------------------------------
def spd_conditioned(rng, n, kappa):
    U = haar(rng, n)
    s = kappa ** (-np.arange(n) / max(n - 1, 1))
    A = (U * s) @ U.T
    return (A + A.T) / 2


@pytest.mark.parametrize("n", [1, 17, 300, 1024])
def test_mixed_cholesky_vs_oracle(gpu, rng, n):
    G = rng.standard_normal((n, n))
    A = G @ G.T + n * np.eye(n); b = rng.standard_normal(n)
    r = gpu.mixed_solve(A, b, factorization="cholesky")
    o = O.mixed_solve(A, b, factorization="cholesky", threads=8)
    assert r.converged and o["converged"] and r.factorization == "cholesky" and not r.fallback
    assert abs(r.iterations - o["iterations"]) <= 1 and r.backward_error <= 1e-13
    np.testing.assert_allclose(r.x, o["x"], rtol=1e-11, atol=1e-13 * np.abs(o["x"]).max())


@pytest.mark.parametrize("kappa", [1e2, 1e5])
def test_mixed_cholesky_conditioned(gpu, rng, kappa):
    n = 2048
    A = spd_conditioned(rng, n, kappa); xt = rng.standard_normal(n); b = A @ xt
    r = solve16(gpu, A, b, factorization="cholesky")
    assert r.converged and r.backward_error <= 16 * EPS
    _, e = O.refinement_residual(A, r.x, b)
    assert e <= 16 * EPS
    assert np.linalg.norm(r.x - xt) / np.linalg.norm(xt) <= 50 * kappa * EPS


def test_mixed_cholesky_fallback_and_errors(gpu, rng):
    n = 64
    A = np.eye(n) * 3.0
    A[10:12, 10:12] = [[1.0, 1.0], [1.0, 1.0 + 1e-12]]     # not positive definite once rounded to Float32
    b = rng.standard_normal(n)
    r = solve16(gpu, A, b, factorization="cholesky")
    assert r.fallback and r.reason == "working_factor_fallback" and r.converged
    I = np.eye(n); I[5, 5] = -1.0                           # indefinite in both precisions
    with pytest.raises(gpu.PosDefException) as e:
        gpu.mixed_solve(I, b, factorization="cholesky")
    assert e.value.info == 6
    bad = A.copy(); bad[3, 3] = np.nan
    with pytest.raises(ValueError):
        gpu.mixed_solve(bad, b, factorization="cholesky")
    with pytest.raises(ValueError):
        gpu.mixed_solve(A, b, factorization="cholesky", lowtype="bfloat16")
