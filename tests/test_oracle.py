"""Pins the CPU oracle (oracle/oxo.cpp) before it is trusted as the checker.

The reference ships no golden vectors; its own tests compare against
LinearAlgebra/LAPACK computed at test time (test/testsetup.jl:1-3).  The same
comparisons pin the oracle here: numpy/OpenBLAS products, LAPACK dgetrf pivots
(test/lapack/lapack_dense.jl:12-14), scipy triangular solves, the normwise
bounds of the reference's tests, exact integer arithmetic and the reference's
no-read / singular / convergence contracts.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

EPS = np.finfo(np.float64).eps
GEMM_SHAPES = ((0, 0, 0), (0, 3, 7), (4, 0, 5), (3, 4, 0), (1, 1, 1), (2, 3, 5),
               (7, 9, 11), (8, 6, 4), (17, 19, 23), (33, 13, 35), (65, 67, 49))  # test/blas/gemm.jl:2-3


def op(a, t):
    return a if t == "N" else a.T


def bwd(E, A):
    """_bwd of test/lapack/lapack_dense.jl:4."""
    return np.linalg.norm(E) / (max(np.linalg.norm(A), np.finfo(float).tiny) * max(1, *A.shape) * EPS)


# ---------------------------------------------------------------------------
# gemm
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ta", "NTC")
@pytest.mark.parametrize("tb", "NTC")
def test_gemm_shapes_vs_numpy(rng, ta, tb):
    for (m, n, k) in GEMM_SHAPES:
        A = rng.standard_normal((m, k) if ta == "N" else (k, m))
        B = rng.standard_normal((k, n) if tb == "N" else (n, k))
        C = np.asfortranarray(rng.standard_normal((m, n)))
        old = C.copy()
        expected = 0.7 * (op(A, ta) @ op(B, tb)) - 0.4 * old
        O.dgemm(C, A, B, 0.7, -0.4, ta, tb)
        # the reference's tolerance (test/blas/gemm.jl:13, tol(Float64)=3e-12)
        np.testing.assert_allclose(C, expected, rtol=3e-12, atol=3e-12 * max(1, k))
        # normwise bound k*eps*|A||B| (+ the beta*C rounding)
        if m and n:
            err = np.linalg.norm(C - expected)
            bound = (k + 2) * EPS * (0.7 * np.linalg.norm(A) * np.linalg.norm(B) + 0.4 * np.linalg.norm(old))
            assert err <= bound


def test_gemm_integer_inputs_bit_exact(rng):
    # small integers: every product and partial sum is exact, so any blocking
    # gives the exact result -- a wrong index or transposition cannot hide
    for (m, n, k) in ((65, 67, 49), (600, 300, 500), (5, 7, 900), (16, 6, 448), (17, 7, 449)):
        for ta in "NT":
            for tb in "NT":
                A = rng.integers(-8, 9, (m, k) if ta == "N" else (k, m)).astype(float)
                B = rng.integers(-8, 9, (k, n) if tb == "N" else (n, k)).astype(float)
                C = np.asfortranarray(rng.integers(-8, 9, (m, n)).astype(float))
                expected = 2.0 * (op(A, ta) @ op(B, tb)) - 3.0 * C
                O.dgemm(C, A, B, 2.0, -3.0, ta, tb, threads=3)
                assert np.array_equal(C, expected)


def test_gemm_no_read_rules(rng):
    # beta == 0 never reads C (test/blas/gemm.jl:50-52)
    A = rng.standard_normal((17, 13)); B = rng.standard_normal((13, 19))
    C = np.asfortranarray(np.full((17, 19), np.nan))
    O.dgemm(C, A, B, 1.0, 0.0)
    np.testing.assert_allclose(C, A @ B, rtol=3e-12, atol=3e-12 * 13)
    # alpha == 0 never reads A or B (:45-49)
    for beta in (0.0, 1.0, 0.5):
        C = np.asfortranarray(np.full((9, 7), 2.0))
        O.dgemm(C, np.full((9, 3), np.nan), np.full((3, 7), np.nan), 0.0, beta)
        assert np.array_equal(C, np.full((9, 7), 2.0 * beta))
    # k == 0 scales C only, even with alpha NaN (:53-55)
    C = np.asfortranarray(np.full((5, 7), 4.0))
    O.dgemm(C, np.zeros((5, 0)), np.zeros((0, 7)), np.nan, 0.5)
    assert np.array_equal(C, np.full((5, 7), 2.0))


def test_gemm_threaded_equals_sequential(rng):
    # beta in {0, 1}: the worker grid changes no element's arithmetic
    A = rng.standard_normal((700, 900)); B = rng.standard_normal((900, 650))
    for beta in (0.0, 1.0):
        C0 = np.asfortranarray(rng.standard_normal((700, 650)))
        C1 = C0.copy(order="F")
        O.dgemm(C0, A, B, -1.0, beta, threads=1)
        O.dgemm(C1, A, B, -1.0, beta, threads=4)
        assert np.array_equal(C0, C1)


def test_gemm_invalid_arguments():
    C = np.asfortranarray(np.zeros((4, 4)))
    with pytest.raises(ValueError):
        O.dgemm(C, np.zeros((4, 4)), np.zeros((4, 4)), transA="X")


# ---------------------------------------------------------------------------
# triangular solves
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [0, 1, 17, 35, 65])  # test/blas/level3_blocked.jl:21-38
def test_trsm_all_variants_vs_scipy(rng, n):
    for side in "LR":
        for uplo in "UL":
            for trans in "NTC":
                for unit in (False, True):
                    A = rng.standard_normal((n, n)) / max(n, 1) + 2 * np.eye(n)  # well-conditioned, also unit
                    B = np.asfortranarray(rng.standard_normal((n, 5) if side == "L" else (5, n)))
                    X = B.copy(order="F")
                    rc = O.dtrsm(X, A, 1.5, side, uplo, trans, unit)
                    assert rc == 0
                    T = np.triu(A) if uplo == "U" else np.tril(A)
                    if unit:
                        np.fill_diagonal(T, 1.0)
                    opT = T if trans == "N" else T.T
                    if side == "L":
                        ref = 1.5 * np.linalg.solve(opT, B) if n else B
                        back = opT @ X
                    else:
                        ref = 1.5 * np.linalg.solve(opT.T, B.T).T if n else B
                        back = X @ opT
                    if n:
                        # normwise backward error of the solve, and the forward
                        # error against numpy within the conditioning
                        res = np.linalg.norm(back - 1.5 * B)
                        assert res <= 10 * n * EPS * np.linalg.norm(T) * np.linalg.norm(X)
                        np.testing.assert_allclose(X, ref, rtol=1e-9, atol=1e-10)


def test_trsm_singular_reported_before_write(rng):
    A = np.triu(rng.standard_normal((20, 20))) + 5 * np.eye(20)
    A[7, 7] = 0.0
    B = np.asfortranarray(rng.standard_normal((20, 3)))
    old = B.copy(order="F")
    rc = O.dtrsm(B, A, 1.0, "L", "U", "N", False)
    assert rc == 8 and np.array_equal(B, old)   # SingularException(8), B untouched


def test_trsv_vs_scipy(rng):
    for n in (1, 3, 4, 5, 17, 64, 101):
        for uplo in "UL":
            for trans in "NT":
                for unit in (False, True):
                    A = rng.standard_normal((n, n)) + 2 * n * np.eye(n)
                    x = rng.standard_normal(n)
                    y = x.copy()
                    assert O.dtrsv(y, A, uplo, trans, unit) == 0
                    ref = sl.solve_triangular(A, x, lower=(uplo == "L"), trans=0 if trans == "N" else 1,
                                              unit_diagonal=unit)
                    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-13)


# ---------------------------------------------------------------------------
# LU
# ---------------------------------------------------------------------------
LU_CASES = ((1, 1, 32), (5, 5, 1), (17, 17, 16), (33, 20, 16), (20, 33, 16), (64, 64, 32), (130, 97, 32),
            (97, 130, 64), (257, 257, 32), (257, 257, 64))  # test/lapack/lapack_dense.jl:9


def reconstruct(F, ipiv, m, n):
    k = min(m, n)
    L = np.tril(F[:, :k], -1) + np.eye(m, k)
    U = np.triu(F[:k, :])
    return L @ U, O.perm_from_ipiv(ipiv, m)


@pytest.mark.parametrize("m,n,bs", LU_CASES)
def test_getrf_matches_lapack_pivots_and_reconstructs(rng, m, n, bs):
    A = rng.standard_normal((m, n))
    F, ipiv, info = O.dgetrf(A, bs)
    assert info == 0
    LU, p = reconstruct(F, ipiv, m, n)
    assert bwd(LU - A[p - 1], A) <= 30
    # the same partial-pivot choices as LAPACK getrf (lapack_dense.jl:12-14)
    _, lapack_piv = sl.lu_factor(A, check_finite=False)
    assert np.array_equal(ipiv, lapack_piv[:min(m, n)] + 1)


def test_getrf_blocked_shapes(rng):
    # test/lapack/factorizations.jl:2-11
    for (m, n) in ((0, 0), (1, 1), (9, 9), (13, 7), (7, 13), (67, 65), (65, 67)):
        for block in (1, 7, 32):
            A = rng.standard_normal((m, n))
            F, ipiv, info = O.dgetrf(A, block)
            if m and n:
                LU, p = reconstruct(F, ipiv, m, n)
                np.testing.assert_allclose(LU, A[p - 1], rtol=3e-11, atol=3e-12 * max(1, n))


def test_getrf_singular_and_hard_inputs(rng):
    A = rng.standard_normal((40, 40)); A[:, 22] = 0
    F, ipiv, info = O.dgetrf(A, 128, check=False)
    assert info == 23          # lapack_dense.jl:24-27
    LU, p = reconstruct(F, ipiv, 40, 40)
    assert bwd(LU - A[p - 1], A) <= 30
    F, ipiv, info = O.dgetrf(np.zeros((3, 3)), check=False)
    assert info == 1           # test/lapack/factorizations.jl:13-15
    G = rng.standard_normal((80, 80)) @ np.diag(10.0 ** np.linspace(0, -12, 80))
    F, ipiv, info = O.dgetrf(G)
    LU, p = reconstruct(F, ipiv, 80, 80)
    assert bwd(LU - G[p - 1], G) <= 30
    N = rng.standard_normal((60, 60)); N[:, -1] = N[:, 0] + EPS * rng.standard_normal(60)
    F, ipiv, info = O.dgetrf(N, check=False)
    LU, p = reconstruct(F, ipiv, 60, 60)
    assert bwd(LU - N[p - 1], N) <= 30


def test_getrf_threaded_bitwise_equal(rng):
    # lapack_dense.jl:33-38
    A = rng.standard_normal((600, 600))
    F1, p1, i1 = O.dgetrf(A, threads=1)
    F2, p2, i2 = O.dgetrf(A, threads=4)
    assert np.array_equal(F1, F2) and np.array_equal(p1, p2)
    A[:, 299] = 0
    F1, p1, i1 = O.dgetrf(A, check=False, threads=1)
    F2, p2, i2 = O.dgetrf(A, check=False, threads=4)
    assert i1 == i2 == 300 and np.array_equal(F1, F2)


def test_getrs_solves(rng):
    for n in (1, 9, 64, 129):
        A = rng.standard_normal((n, n))
        F, ipiv, info = O.dgetrf(A, 32)
        c = np.linalg.cond(A)
        b = rng.standard_normal(n); B = rng.standard_normal((n, 3))
        assert bwd(A @ O.dgetrs(F, ipiv, b) - b, A) <= 30 * c
        assert bwd(A @ O.dgetrs(F, ipiv, B) - B, A) <= 30 * c
        assert bwd(A.T @ O.dgetrs(F, ipiv, b, "T") - b, A) <= 30 * c
        assert bwd(A.T @ O.dgetrs(F, ipiv, B, "C") - B, A) <= 30 * c
        # the matrix and vector forms agree (test/api/linalg_api_solve.jl:12-20)
        np.testing.assert_allclose(O.dgetrs(F, ipiv, b), O.dgetrs(F, ipiv, b.reshape(-1, 1))[:, 0], rtol=1e-10,
                                   atol=1e-12)


# ---------------------------------------------------------------------------
# equilibration, residual, mixed solve
# ---------------------------------------------------------------------------
def test_equilibrate_radix_properties(rng):
    A = rng.standard_normal((50, 40)) * np.logspace(-30, 30, 50)[:, None]
    B, r, c = O.dequilibrate(A, radix=True)
    # exact powers of two, and the scaling is exact
    assert np.all(np.frexp(r)[0] == 0.5) and np.all(np.frexp(c)[0] == 0.5)
    assert np.array_equal(B, A / r[:, None] / c[None, :])
    rowscaled = A / r[:, None]
    assert np.all(np.abs(rowscaled).max(1) < 2 ** 0.5 + 1e-12) and np.all(np.abs(rowscaled).max(1) >= 2 ** -0.5)
    assert np.all(np.abs(B).max(0) < 2 ** 0.5 + 1e-12) and np.all(np.abs(B).max(0) >= 2 ** -0.5)
    # a zero row keeps scale 1
    Z = A.copy(); Z[3] = 0
    _, r2, _ = O.dequilibrate(Z)
    assert r2[3] == 1.0


def exact_radix_scale(v):
    """2^round(log2 v) with log2 correctly rounded to double (80-digit Decimal)
    and ties to even, clamped to [exponent(floatmin), exponent(floatmax)]:
    the closed form of equilibrate(A; radix=true)'s scale."""
    from decimal import Decimal, localcontext
    import math
    with localcontext() as ctx:
        ctx.prec = 80
        L = float(Decimal(v).ln() / Decimal(2).ln())
    return math.ldexp(1.0, min(1023, max(-1022, round(L))))


def radix_boundary_values():
    """Values whose log2 sits at or next to a half-integer (sqrt(2) 2^k and its
    neighbours, where round-half-even decides), powers of two, subnormals."""
    s = 1.4142135623730951
    mant = (s, np.nextafter(s, 0.0), np.nextafter(s, 2.0), np.nextafter(np.nextafter(s, 0.0), 0.0), 1.0, 1.5,
            np.nextafter(2.0, 0.0))
    ks = list(range(-40, 41)) + [-1030, -1022, -1000, -700, 500, 1000, 1020, 1023]
    vals = [m * 2.0 ** k for k in ks for m in mant if np.isfinite(m * 2.0 ** k) and m * 2.0 ** k > 0]
    return np.array(vals + [5e-324, 1e-310, 2.2250738585072014e-308, 1.7976931348623157e308])


def test_equilibrate_radix_rounding_is_exact():
    v = radix_boundary_values()
    _, r, c = O.dequilibrate(v.reshape(-1, 1), radix=True)
    expect = np.array([exact_radix_scale(x) for x in v])
    bad = np.nonzero(r != expect)[0]
    assert bad.size == 0, [(v[i], r[i], expect[i]) for i in bad[:5]]
    # a plausible mistake -- rounding log2 half away from zero -- is caught
    from decimal import Decimal, localcontext, ROUND_HALF_UP
    with localcontext() as ctx:
        ctx.prec = 80
        half_up = [int((Decimal(float(Decimal(x).ln() / Decimal(2).ln()))).quantize(Decimal(1), ROUND_HALF_UP))
                   for x in v if x > 1e-300]
    assert any(2.0 ** e != exact_radix_scale(x) for e, x in zip(half_up, [x for x in v if x > 1e-300]))
    assert c[0] == exact_radix_scale(np.max(v / r))


def test_refinement_residual(rng):
    A = rng.standard_normal((30, 30)); X = rng.standard_normal((30, 2)); B = rng.standard_normal((30, 2))
    R, e = O.refinement_residual(A, X, B)
    np.testing.assert_allclose(R, B - A @ X, rtol=1e-12, atol=1e-12)
    den = np.abs(A) @ np.abs(X) + np.abs(B)
    assert e == pytest.approx(np.max(np.abs(R) / den), rel=1e-12)
    # an exact solution of an integer system has zero residual and zero error
    Ai = np.diag(np.arange(1.0, 6.0)); xi = np.arange(1.0, 6.0); bi = Ai @ xi
    R, e = O.refinement_residual(Ai, xi, bi)
    assert np.all(R == 0) and e == 0.0


def test_mixed_solve_contract(rng):
    # test/lapack/factorizations.jl:57-63
    for n in (1, 17, 67):
        A = rng.standard_normal((n, n)) + n * np.eye(n); b = rng.standard_normal(n)
        r = O.mixed_solve(A, b)
        assert r["converged"] and r["iterations"] <= 10 and not r["fallback"]
        np.testing.assert_allclose(A @ r["x"], b, rtol=2e-12, atol=2e-12)
    # singular after rounding to Float32: falls back to the working precision
    A = np.array([[1.0, 1.0], [1.0, 1.0 + 1e-12]]); b = np.array([1.0, 2.0])
    r = O.mixed_solve(A, b)
    assert r["fallback"] and r["reason"] == "working_factor_fallback"
    np.testing.assert_allclose(A @ r["x"], b, rtol=1e-9)


# ---------------------------------------------------------------------------
# Float32 oracle functions (the checker of the FP32 factor of the mixed solve)
# ---------------------------------------------------------------------------
EPS32 = float(np.finfo(np.float32).eps)


@pytest.mark.parametrize("m,n,bs", LU_CASES + ((600, 600, 128),))
def test_sgetrf_matches_lapack_sgetrf(rng, m, n, bs):
    # Float32 lu_factor (src/lapack/factorizations.jl:17-77 at T = Float32):
    # the pivots of LAPACK sgetrf (scipy on float32 input) and the
    # reconstruction bound of test/lapack/lapack_dense.jl:11 at Float32 eps
    A = rng.standard_normal((m, n)).astype(np.float32)
    F, ipiv, info = O.sgetrf(A, bs)
    assert info == 0 and F.dtype == np.float32
    k = min(m, n)
    L = np.tril(F[:, :k].astype(np.float64), -1) + np.eye(m, k)
    U = np.triu(F[:k, :].astype(np.float64))
    p = O.perm_from_ipiv(ipiv, m)
    r = np.linalg.norm(L @ U - A[p - 1].astype(np.float64)) / (np.linalg.norm(A) * max(m, n) * EPS32)
    assert r <= 30
    lapack_lu, lapack_piv = sl.lu_factor(A, check_finite=False)
    assert lapack_lu.dtype == np.float32            # really sgetrf
    assert np.array_equal(ipiv, lapack_piv[:k] + 1)
    # the factors agree with sgetrf's to the rounding of a Float32 elimination
    assert np.abs(F.astype(np.float64) - lapack_lu).max() <= 64 * max(m, n) * EPS32 * np.abs(lapack_lu).max()


@pytest.mark.parametrize("n", [5, 33, 64])
def test_sgetrf_tournament_one_wave_is_partial_pivoting(rng, n):
    # the tournament leaf of oxi_sgetrf_tournament: with the rows of one wave
    # (at most 64) the candidates are the partial-pivoting pivots themselves
    # and the elimination is partial pivoting's, operation for operation, so
    # the factors are sgetrf's (pinned to LAPACK above) bit for bit
    A = rng.standard_normal((n, n)).astype(np.float32)
    F1, p1, _ = O.sgetrf(A)
    F2, p2, _ = O.sgetrf(A, tournament=True)
    assert np.array_equal(p1, p2) and np.array_equal(F1, F2)
    assert np.array_equal(p2, sl.lu_factor(A, check_finite=False)[1][:n] + 1)


def lapack_pivot_rows(A, rows, w=32):
    """The rows (positions, ascending) that LAPACK sgetrf picks, in order, as
    the first min(w, len(rows)) pivots of A[rows, :w]: partial pivoting over
    the set, ties to the first row in position order."""
    rows = np.sort(np.asarray(rows))
    _, bp = sl.lu_factor(A[rows, :w], check_finite=False)
    order = np.arange(len(rows))
    for j, q in enumerate(bp):
        order[[j, q]] = order[[q, j]]
    return rows[order[:min(w, len(rows))]]


def tournament_pivot_rows(A, w=32):
    """The first leaf's pivot rows of the tournament, with LAPACK for every
    partial-pivoting step: each wave's rows (wave v of workgroup g: rows
    g*R + s*256 + v*64 + l), then the union of a workgroup's wave candidates,
    then unions of four consecutive sets until one is left."""
    m = A.shape[0]
    R = 512 if m > 8192 else 256
    sets = []
    for g0 in range(0, m, R):
        cands = []
        for v in range(4):
            rows = [g0 + s * 256 + v * 64 + l for s in range(R // 256) for l in range(64)]
            rows = [r for r in rows if r < m]
            if rows:
                cands.extend(lapack_pivot_rows(A, rows, w))
        sets.append(lapack_pivot_rows(A, cands, w))
    while len(sets) > 1:
        sets = [lapack_pivot_rows(A, np.concatenate(sets[g:g + 4]), w) for g in range(0, len(sets), 4)]
    return lapack_pivot_rows(A, sets[0], w)


@pytest.mark.parametrize("n", [200, 1000, 2000])
def test_sgetrf_tournament_first_leaf_and_reconstruction(rng, n):
    # several waves and workgroups: the first leaf's pivot rows are those of
    # the same tournament run with LAPACK sgetrf for every partial-pivoting
    # step (columns 0..31), and the whole factor reconstructs P*A to the
    # Float32 bound of test/lapack/lapack_dense.jl:11
    A = rng.standard_normal((n, n)).astype(np.float32)
    F, ipiv, info = O.sgetrf(A, tournament=True, threads=8)
    assert info == 0
    p = O.perm_from_ipiv(ipiv, n) - 1
    assert np.array_equal(np.sort(p), np.arange(n))
    assert np.array_equal(p[:32], tournament_pivot_rows(A))
    L = np.tril(F.astype(np.float64), -1) + np.eye(n)
    U = np.triu(F.astype(np.float64))
    A64 = A.astype(np.float64)
    assert np.linalg.norm(L @ U - A64[p]) <= 30 * n * EPS32 * np.linalg.norm(A64)
    assert not np.array_equal(p, O.perm_from_ipiv(O.sgetrf(A)[1], n) - 1)   # not partial pivoting


def test_sgetrf_tournament_tree_is_not_flat(rng):
    # the tree matters: a flat tournament (partial pivoting over all the
    # wave candidates at once) picks different first-leaf rows on this matrix,
    # so a dropped tree level would fail the test above
    n = 2000
    A = rng.standard_normal((n, n)).astype(np.float32)
    flat = []
    for g0 in range(0, n, 256):
        for v in range(4):
            rows = [r for r in range(g0 + v * 64, g0 + v * 64 + 64) if r < n]
            flat.extend(lapack_pivot_rows(A, rows))
    assert not np.array_equal(lapack_pivot_rows(A, flat), tournament_pivot_rows(A))


def test_sgetrf_singular_column_info(rng):
    A = rng.standard_normal((40, 40)).astype(np.float32)
    A[:, 17] = 0
    F, ipiv, info = O.sgetrf(A, 16, check=False)
    assert info == 18


@pytest.mark.parametrize("n", [1, 9, 64, 129, 300])
def test_sgetrs_backward_error(rng, n):
    # Float32 solve(F, b) (src/lapack/factorizations.jl:177-193): normwise
    # backward error at Float32 eps, both transposes, vector and matrix forms,
    # and agreement with LAPACK sgetrs on sgetrf's own factors
    A = rng.standard_normal((n, n)).astype(np.float32)
    F, ipiv, _ = O.sgetrf(A, 32)
    b = rng.standard_normal(n).astype(np.float32)
    B = rng.standard_normal((n, 4)).astype(np.float32)
    A64 = A.astype(np.float64)
    for trans, opA in (("N", A64), ("T", A64.T)):
        for rhs in (b, B):
            x = O.sgetrs(F, ipiv, rhs, trans)
            assert x.dtype == np.float32
            x64 = x.astype(np.float64)
            res = np.linalg.norm(opA @ x64 - rhs.astype(np.float64))
            assert res <= 30 * n * EPS32 * np.linalg.norm(A64) * np.linalg.norm(x64) + 1e-30
            ref = sl.lu_solve((F, ipiv - 1), rhs, trans=0 if trans == "N" else 1, check_finite=False)
            assert ref.dtype == np.float32
            cond = np.linalg.cond(A64)
            np.testing.assert_allclose(x64, ref.astype(np.float64), rtol=0,
                                       atol=64 * n * EPS32 * cond * np.abs(ref).max())


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_sgemm_vs_float64_product(rng, ta, tb):
    # Float32 gemm! (src/blas/gemm.jl:625-703) on the 16x16 Float32 tile
    # (edges at 15, 16, 17, 31, 33 rows/columns) and across the kc = 896 block
    # boundary of K: normwise error <= k eps32 |A||B| against the exact
    # (float64) product, and bit-exact on small-integer inputs
    for (m, n, k) in ((15, 16, 17), (16, 16, 16), (17, 31, 33), (33, 17, 895), (31, 33, 897), (100, 70, 1800)):
        A = rng.standard_normal((m, k) if ta == "N" else (k, m)).astype(np.float32)
        B = rng.standard_normal((k, n) if tb == "N" else (n, k)).astype(np.float32)
        C0 = rng.standard_normal((m, n)).astype(np.float32)
        C = np.asfortranarray(C0.copy())
        O.sgemm(C, A, B, 1.25, -0.5, ta, tb)
        exact = 1.25 * (op(A, ta).astype(np.float64) @ op(B, tb).astype(np.float64)) - 0.5 * C0.astype(np.float64)
        err = np.linalg.norm(C.astype(np.float64) - exact)
        bound = (k + 2) * EPS32 * (1.25 * np.linalg.norm(A.astype(np.float64)) * np.linalg.norm(B.astype(np.float64))
                                   + 0.5 * np.linalg.norm(C0.astype(np.float64)))
        assert err <= bound
        Ai = rng.integers(-8, 9, A.shape).astype(np.float32)
        Bi = rng.integers(-8, 9, B.shape).astype(np.float32)
        Ci = np.asfortranarray(rng.integers(-8, 9, (m, n)).astype(np.float32))
        expect = 2.0 * (op(Ai, ta).astype(np.float64) @ op(Bi, tb).astype(np.float64)) - 3.0 * Ci.astype(np.float64)
        O.sgemm(Ci, Ai, Bi, 2.0, -3.0, ta, tb, threads=3)
        assert np.array_equal(Ci.astype(np.float64), expect)


def test_sgemm_no_read_rules(rng):
    A = rng.standard_normal((17, 13)).astype(np.float32); B = rng.standard_normal((13, 19)).astype(np.float32)
    C = np.asfortranarray(np.full((17, 19), np.nan, dtype=np.float32))
    O.sgemm(C, A, B, 1.0, 0.0)
    assert np.all(np.isfinite(C))
    C = np.asfortranarray(np.full((9, 7), 2.0, dtype=np.float32))
    O.sgemm(C, np.full((9, 3), np.nan, np.float32), np.full((3, 7), np.nan, np.float32), 0.0, 0.5)
    assert np.array_equal(C, np.full((9, 7), 1.0, dtype=np.float32))


def test_dlaswp_closed_form(rng):
    # interchanges in order give A[p, :] with p = perm_from_ipiv (the
    # permutation of factorizations.jl:72-75); in reverse order its inverse
    m, ncols = 700, 5
    ipiv = np.array([rng.integers(j, m) + 1 for j in range(m)], dtype=np.int64)
    A = np.asfortranarray(rng.standard_normal((m, ncols)))
    p = O.perm_from_ipiv(ipiv, m)
    fwd = O.dlaswp(A.copy(order="F"), 1, m, ipiv)
    assert np.array_equal(fwd, A[p - 1])
    back = O.dlaswp(fwd.copy(order="F"), 1, m, ipiv, inverse=True)
    assert np.array_equal(back, A)
    # a sub-range k1..k2 equals the permutation of those steps alone
    sub = ipiv.copy(); sub[:99] = np.arange(1, 100); sub[400:] = np.arange(401, m + 1)
    part = O.dlaswp(A.copy(order="F"), 100, 400, ipiv)
    assert np.array_equal(part, A[O.perm_from_ipiv(sub, m) - 1])
    # a dropped step or a reversed order is caught
    assert not np.array_equal(O.dlaswp(A.copy(order="F"), 2, m, ipiv), fwd)
    assert not np.array_equal(O.dlaswp(A.copy(order="F"), 1, m, ipiv, inverse=True), fwd)
    with pytest.raises(ValueError):
        O.dlaswp(A.copy(order="F"), 1, 3, np.array([1, 2, m + 1]))


# ---------------------------------------------------------------------------
# syrk / Cholesky (src/blas/level3.jl:546-783, src/lapack/factorizations.jl:206-300)
# ---------------------------------------------------------------------------
def spd(rng, n, shift=None):
    G = rng.standard_normal((n, n))
    return G @ G.T + (n if shift is None else shift) * np.eye(n)


@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
def test_syrk_triangle_vs_numpy(rng, uplo, trans):
    for (n, k) in ((0, 3), (1, 1), (5, 0), (17, 9), (40, 65)):
        A = rng.standard_normal((n, k) if trans == "N" else (k, n))
        C0 = np.asfortranarray(rng.standard_normal((n, n)))
        C = C0.copy(order="F")
        O.dsyrk(C, A, 0.7, -0.3, uplo, trans)
        P = A @ A.T if trans == "N" else A.T @ A
        full = 0.7 * P - 0.3 * C0
        tri = np.triu if uplo == "U" else np.tril
        other = np.tril(np.ones((n, n)), -1) if uplo == "U" else np.triu(np.ones((n, n)), 1)
        np.testing.assert_allclose(tri(C), tri(full), rtol=1e-12, atol=1e-12 * max(1, k))
        assert np.array_equal(C[other.astype(bool)], C0[other.astype(bool)])   # the other triangle untouched
    # beta == 0 never reads C; integers are exact
    A = rng.integers(-5, 6, (30, 12)).astype(float)
    C = np.asfortranarray(np.full((30, 30), np.nan))
    O.dsyrk(C, A, 2.0, 0.0, "L")
    assert np.array_equal(np.tril(C), np.tril(2.0 * A @ A.T))
    assert np.all(np.isnan(C[np.triu_indices(30, 1)]))


@pytest.mark.parametrize("n", [1, 2, 17, 64, 65, 130, 257])
def test_potrf_vs_lapack(rng, n):
    A = spd(rng, n)
    L, info = O.dpotrf(A)
    assert info == 0
    assert np.array_equal(np.triu(L, 1), np.zeros((n, n)))
    # LAPACK dpotrf (scipy) and the reconstruction bound
    Lref = sl.cholesky(A, lower=True)
    np.testing.assert_allclose(L, Lref, rtol=0, atol=64 * n * EPS * np.abs(Lref).max())
    assert np.linalg.norm(L @ L.T - A) <= 30 * n * EPS * np.linalg.norm(A)
    # only the lower triangle is read
    Au = A.copy(); Au[np.triu_indices(n, 1)] = np.nan
    L2, info2 = O.dpotrf(Au)
    assert info2 == 0 and np.array_equal(L2, L)


def test_potrf_info_first_failing_column(rng):
    # a non-positive leading minor at column j: info == j (LAPACK's order k);
    # scipy names the same order
    for n, j in ((10, 4), (100, 70), (200, 1), (300, 150)):
        A = spd(rng, n)
        A[j - 1, j - 1] = -1e3 * n
        L, info = O.dpotrf(A)
        assert info == j
        with pytest.raises(np.linalg.LinAlgError, match=f"^{j}-th leading minor"):
            sl.cholesky(A, lower=True)
    A = spd(rng, 9); A[5, 5] = np.nan
    assert O.dpotrf(A)[1] == 6


def test_potrs_vs_scipy(rng):
    for n in (1, 9, 70, 200):
        A = spd(rng, n)
        L, _ = O.dpotrf(A)
        b = rng.standard_normal(n); B = rng.standard_normal((n, 5))
        for rhs in (b, B):
            x = O.dpotrs(L, rhs)
            ref = sl.cho_solve((L, True), rhs)
            np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-12)
            assert np.linalg.norm(A @ x - rhs) <= 30 * n * EPS * np.linalg.norm(A) * np.linalg.norm(x)


# ---------------------------------------------------------------------------
# what is built on LU: inverse, log|det|, backward error, estimates, drivers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 17, 64, 150])
def test_getri_vs_numpy_inv(rng, n):
    A = rng.standard_normal((n, n))
    F, ipiv, info = O.dgetrf(A)
    X = O.dgetri(F, ipiv)
    np.testing.assert_allclose(X, np.linalg.inv(A), rtol=0, atol=1e3 * n * EPS * np.linalg.cond(A) * np.abs(X).max())
    assert np.linalg.norm(A @ X - np.eye(n)) <= 30 * n * EPS * np.linalg.cond(A)
    # integer unimodular matrix: the inverse is an exact integer matrix
    U = np.triu(rng.integers(-2, 3, (6, 6)).astype(float), 1) + np.eye(6)
    L = np.tril(rng.integers(-2, 3, (6, 6)).astype(float), -1) + np.eye(6)
    M = (L @ U)[[3, 1, 5, 0, 2, 4]]
    F, ipiv, _ = O.dgetrf(M)
    np.testing.assert_allclose(O.dgetri(F, ipiv), np.round(np.linalg.inv(M)), rtol=0, atol=1e-9)
    S = A.copy(); S[:, 0] = 0
    F, ipiv, info = O.dgetrf(S)
    with pytest.raises(np.linalg.LinAlgError):
        O.dgetri(F, ipiv)


def test_logabsdet_vs_slogdet(rng):
    for n in (1, 5, 40, 200):
        A = rng.standard_normal((n, n))
        F, ipiv, _ = O.dgetrf(A)
        la, s = O.dlogabsdet(F, ipiv)
        sn, ln = np.linalg.slogdet(A)
        assert s == sn and la == pytest.approx(ln, rel=1e-12, abs=1e-12)
    # a permutation matrix: log|det| 0, sign the parity
    P = np.eye(7)[[1, 0, 3, 2, 4, 6, 5]]
    F, ipiv, _ = O.dgetrf(P)
    assert O.dlogabsdet(F, ipiv) == (0.0, -1.0)
    S = rng.standard_normal((9, 9)); S[:, 4] = 0
    F, ipiv, _ = O.dgetrf(S)
    la, s = O.dlogabsdet(F, ipiv)
    assert la == -np.inf and s == 0.0


def test_backward_error_closed_form(rng):
    A = rng.standard_normal((30, 20)); X = rng.standard_normal((20, 3)); B = rng.standard_normal((30, 3))
    e, R = O.backward_error(A, X, B)
    np.testing.assert_allclose(R, B - A @ X, rtol=1e-12, atol=1e-13)
    assert e == pytest.approx(np.max(np.abs(B - A @ X) / (np.abs(A) @ np.abs(X) + np.abs(B))), rel=1e-12)
    # exact solution of an integer system: zero; 0/0 = 0; a non-finite denominator: Inf
    Ai = np.diag(np.arange(1.0, 6.0)); xi = np.arange(1.0, 6.0)
    assert O.backward_error(Ai, xi, Ai @ xi)[0] == 0.0
    assert O.backward_error(np.zeros((2, 2)), np.zeros(2), np.zeros(2))[0] == 0.0
    assert O.backward_error(np.array([[np.inf, 0.0], [0.0, 1.0]]), np.ones(2), np.ones(2))[0] == np.inf
    assert O.backward_error(np.array([[np.nan]]), np.ones(1), np.ones(1))[0] == np.inf   # NaN denominator


def test_inverse_norm_estimate_vs_exact(rng):
    # Hager's estimate is a lower bound of ||A^-1||_1 that is exact on most
    # small matrices; it is never above the true norm
    exact_hits = 0
    for n in (1, 4, 20, 60, 120):
        for _ in range(3):
            A = rng.standard_normal((n, n))
            F, ipiv, _ = O.dgetrf(A)
            est = O.inverse_norm_estimate(F, ipiv)
            true = np.abs(np.linalg.inv(A)).sum(axis=0).max()
            assert est <= true * (1 + 1e-10) and est >= true / 10
            exact_hits += est == pytest.approx(true, rel=1e-10)
    assert exact_hits >= 8


def test_refine_and_expert_solve_contract(rng):
    for n in (1, 10, 80):
        A = rng.standard_normal((n, n)) * np.logspace(-6, 6, n)[:, None]   # badly scaled rows
        b = rng.standard_normal(n)
        F, ipiv, _ = O.dgetrf(A)
        r = O.refine_solve(A, F, ipiv, b)
        assert r["converged"] and r["backward_error"] <= 8 * EPS and r["iterations"] <= 8
        c1 = np.linalg.cond(A, 1)
        xs = np.linalg.solve(A, b)
        np.testing.assert_allclose(r["x"], xs, rtol=0, atol=1e-14 * c1 * np.abs(xs).max())
        assert 0.1 / c1 <= r["rcond_estimate"] <= 1.0 / c1 * (1 + 1e-8)
        e = O.expert_solve(A, b)
        assert e["converged"] and e["backward_error"] <= 8 * EPS
        np.testing.assert_allclose(e["x"], xs, rtol=0, atol=1e-14 * c1 * np.abs(xs).max())
        np.testing.assert_allclose(e["row_scale"], np.abs(A).max(axis=1), rtol=0, atol=0)
    Z = rng.standard_normal((5, 5)); Z[2] = 0
    with pytest.raises(np.linalg.LinAlgError):
        O.expert_solve(Z, np.ones(5))


def test_mixed_solve_cholesky_contract(rng):
    # mixed_solve(A, b; factorization=:cholesky): test/lapack/factorizations.jl:57-63
    for n in (1, 17, 80):
        A = spd(rng, n); b = rng.standard_normal(n)
        r = O.mixed_solve(A, b, factorization="cholesky")
        assert r["converged"] and r["iterations"] <= 10 and not r["fallback"]
        np.testing.assert_allclose(A @ r["x"], b, rtol=2e-12, atol=2e-12 * np.abs(b).max())
    # positive definite in Float64, not after rounding to Float32: the Float64 factor
    A = np.array([[1.0, 1.0], [1.0, 1.0 + 1e-12]]); b = np.array([1.0, 2.0])
    r = O.mixed_solve(A, b, factorization="cholesky")
    assert r["fallback"] and r["reason"] == "working_factor_fallback"
    np.testing.assert_allclose(A @ r["x"], b, rtol=1e-9)
    with pytest.raises(ValueError):
        O.mixed_solve(np.array([[1.0, 2.0], [2.0, 1.0]]), b, factorization="cholesky")   # indefinite


def test_hgemm_restatement(rng):
    # small integers: every product and partial sum is exact in Float32, the
    # single rounding to Float16 is the only one, so the result equals the
    # exact product rounded once -- a dropped beta term or a transposed
    # operand cannot hide
    for ta in "NT":
        for tb in "NT":
            m, n, k = 17, 9, 40
            A = rng.integers(-4, 5, (m, k) if ta == "N" else (k, m)).astype(np.float16)
            B = rng.integers(-4, 5, (k, n) if tb == "N" else (n, k)).astype(np.float16)
            C = rng.integers(-8, 9, (m, n)).astype(np.float16)
            got = O.hgemm(C, A, B, 0.5, -2.0, ta, tb)
            exact = 0.5 * (op(A, ta).astype(np.float64) @ op(B, tb).astype(np.float64)) - 2.0 * C.astype(np.float64)
            assert np.array_equal(got, exact.astype(np.float16))
    # alpha, beta rounded to Float16 first; beta == 0 never reads C
    A = rng.standard_normal((5, 3)).astype(np.float16); B = rng.standard_normal((3, 4)).astype(np.float16)
    got = O.hgemm(np.full((5, 4), np.nan, np.float16), A, B, 1.0 / 3.0, 0.0)
    exact = np.float32(np.float16(1.0 / 3.0)) * (A.astype(np.float32) @ B.astype(np.float32))
    assert np.all(np.isfinite(got)) and np.allclose(got.astype(np.float32), exact, rtol=2e-3, atol=1e-3)
