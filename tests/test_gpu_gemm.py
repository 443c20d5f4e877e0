"""FP64 GEMM on the GPU (oxi_dgemm through the C ABI) against the oracle.

Follows test/blas/gemm.jl: every N/T/C combination over its shapes, the
no-read rules, aliasing snapshots and views with guard regions; plus exact
integer inputs (bit-identical results), the golden fixtures and the full-size
shapes of the benchmark (8192^3 and the 16384 x 16384 x 256 trailing update)
checked entry-wise on a random sample.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
SHAPES = ((0, 0, 0), (0, 3, 7), (4, 0, 5), (3, 4, 0), (1, 1, 1), (2, 3, 5),
          (7, 9, 11), (8, 6, 4), (17, 19, 23), (33, 13, 35), (65, 67, 49))  # test/blas/gemm.jl:2-3


def op(a, t):
    return a if t == "N" else a.T


def run(ox, C0, A, B, al, be, ta, tb):
    Cd = ox.DMatrix.from_numpy(C0)
    ox.gemm_(Cd, ox.DMatrix.from_numpy(A), ox.DMatrix.from_numpy(B), al, be, ta, tb)
    return Cd.numpy()


@pytest.mark.parametrize("ta", "NTC")
@pytest.mark.parametrize("tb", "NTC")
def test_shapes_against_oracle(gpu, rng, ta, tb):
    for (m, n, k) in SHAPES + ((129, 130, 257), (300, 129, 512), (128, 128, 16), (257, 1, 300), (1, 300, 77)):
        A = rng.standard_normal((m, k) if ta == "N" else (k, m))
        B = rng.standard_normal((k, n) if tb == "N" else (n, k))
        C0 = np.asfortranarray(rng.standard_normal((m, n)))
        got = run(gpu, C0, A, B, 0.7, -0.4, ta, tb)
        ref = C0.copy(order="F")
        O.dgemm(ref, A, B, 0.7, -0.4, ta, tb)
        # the reference's tolerance (test/blas/gemm.jl:13) and the normwise bound
        np.testing.assert_allclose(got, ref, rtol=3e-12, atol=3e-12 * max(1, k))
        if m and n:
            bound = 2 * (k + 2) * EPS * (0.7 * np.linalg.norm(A) * np.linalg.norm(B) + 0.4 * np.linalg.norm(C0))
            assert np.linalg.norm(got - ref) <= bound


def test_integer_inputs_bit_exact(gpu, rng):
    for (m, n, k) in ((65, 67, 49), (600, 300, 500), (5, 7, 900), (129, 257, 130)):
        for ta in "NT":
            for tb in "NT":
                A = rng.integers(-8, 9, (m, k) if ta == "N" else (k, m)).astype(float)
                B = rng.integers(-8, 9, (k, n) if tb == "N" else (n, k)).astype(float)
                C0 = np.asfortranarray(rng.integers(-8, 9, (m, n)).astype(float))
                got = run(gpu, C0, A, B, 2.0, -3.0, ta, tb)
                assert np.array_equal(got, 2.0 * (op(A, ta) @ op(B, tb)) - 3.0 * C0)


def test_no_read_rules(gpu, rng):
    ox = gpu
    A = rng.standard_normal((17, 13)); B = rng.standard_normal((13, 19))
    got = run(ox, np.full((17, 19), np.nan), A, B, 1.0, 0.0, "N", "N")
    np.testing.assert_allclose(got, A @ B, rtol=3e-12, atol=3e-12 * 13)     # beta == 0 never reads C
    for beta in (0.0, 1.0, 0.5):
        got = run(ox, np.full((9, 7), 2.0), np.full((9, 3), np.nan), np.full((3, 7), np.nan), 0.0, beta, "N", "N")
        assert np.array_equal(got, np.full((9, 7), 2.0 * beta))              # alpha == 0 never reads A, B
    got = run(ox, np.full((5, 7), 4.0), np.zeros((5, 0)), np.zeros((0, 7)), np.nan, 0.5, "N", "N")
    assert np.array_equal(got, np.full((5, 7), 2.0))                         # k == 0


def test_aliasing_snapshots(gpu, rng):
    ox = gpu
    # C is A, C is B (test/blas/gemm.jl:56-62)
    A = rng.standard_normal((17, 17)); B = rng.standard_normal((17, 17))
    Ad = ox.DMatrix.from_numpy(A); Bd = ox.DMatrix.from_numpy(B)
    ox.gemm_(Ad, Ad, Bd, 1.0, 0.3)
    np.testing.assert_allclose(Ad.numpy(), A @ B + 0.3 * A, rtol=3e-12, atol=3e-12 * 17)
    A2 = Ad.numpy()
    ox.gemm_(Bd, Ad, Bd)
    np.testing.assert_allclose(Bd.numpy(), A2 @ B, rtol=3e-12, atol=3e-12 * 17)
    # three overlapping views of one matrix (:63-66)
    X = rng.standard_normal((20, 20))
    Xd = ox.DMatrix.from_numpy(X)
    C = Xd.view(slice(1, 18), slice(1, 18)); Av = Xd.view(slice(0, 17), slice(0, 17)); Bv = Xd.view(slice(2, 19), slice(2, 19))
    expected = X[0:17, 0:17] @ X[2:19, 2:19] + X[1:18, 1:18]
    ox.gemm_(C, Av, Bv, 1.0, 1.0)
    np.testing.assert_allclose(Xd.numpy()[1:18, 1:18], expected, rtol=3e-12, atol=3e-12 * 17)


def test_views_and_guard_regions(gpu, rng):
    ox = gpu
    # unit row stride with padded leading dimension; unaligned first row (test/blas/gemm.jl:28-33)
    A = rng.standard_normal((17, 15)); B = rng.standard_normal((15, 19))
    storage = np.full((23, 22), 88.0)
    Sd = ox.DMatrix.from_numpy(storage)
    C = Sd.view(slice(1, 18), slice(2, 21))
    ox.gemm_(C, ox.DMatrix.from_numpy(A), ox.DMatrix.from_numpy(B))
    S = Sd.numpy()
    np.testing.assert_allclose(S[1:18, 2:21], A @ B, rtol=3e-12, atol=3e-12 * 15)
    mask = np.ones_like(S, dtype=bool); mask[1:18, 2:21] = False
    assert np.all(S[mask] == 88.0)
    # operands with a leading dimension larger than their rows
    Ab = ox.DMatrix.from_numpy(A, ld=40); Bb = ox.DMatrix.from_numpy(B, ld=33)
    Cd = ox.DMatrix.from_numpy(np.zeros((17, 19)), ld=21)
    ox.gemm_(Cd, Ab, Bb)
    np.testing.assert_allclose(Cd.numpy(), A @ B, rtol=3e-12, atol=3e-12 * 15)


def test_argument_validation(gpu):
    ox = gpu
    C = ox.DMatrix.from_numpy(np.zeros((64, 64))); A = ox.DMatrix.from_numpy(np.ones((64, 64)))
    with pytest.raises(ox.DimensionMismatch):
        ox.gemm_(C, A, ox.DMatrix.from_numpy(np.zeros((63, 64))))
    with pytest.raises(ValueError):
        ox.gemm_(C, A, A, transA="bad")
    assert np.all(C.numpy() == 0)


def test_golden_fixtures(gpu):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "gemm_cases.npz"), allow_pickle=False)
    for i in range(int(z["count"])):
        m, n, k, ta, tb = (int(v) for v in z[f"c{i}_meta"])
        al, be = z[f"c{i}_ab"]
        got = run(gpu, z[f"c{i}_C0"], z[f"c{i}_A"], z[f"c{i}_B"], al, be, chr(ta), chr(tb))
        np.testing.assert_allclose(got, z[f"c{i}_C"], rtol=3e-12, atol=3e-12 * max(1, k))


def test_run_to_run_bitwise(gpu, rng):
    A = rng.standard_normal((700, 900)); B = rng.standard_normal((900, 650)); C0 = rng.standard_normal((700, 650))
    r1 = run(gpu, C0, A, B, -1.0, 1.0, "N", "N")
    r2 = run(gpu, C0, A, B, -1.0, 1.0, "N", "N")
    assert np.array_equal(r1, r2)


def sampled_check(rows_a, cols_b, c0, got, al, be):
    """Entry-wise check on a random sample against exact dot products:
    |C_ij - ref_ij| <= (k+2) eps (|alpha| sum|a||b| + |beta c|).
    rows_a[s] = row i_s of op(A), cols_b[s] = column j_s of op(B)."""
    prod = (rows_a.astype(np.longdouble) * cols_b.astype(np.longdouble)).sum(axis=1)
    ref = al * prod.astype(np.float64) + be * c0
    k = rows_a.shape[1]
    tol = (k + 2) * EPS * (abs(al) * np.sum(np.abs(rows_a) * np.abs(cols_b), axis=1) + np.abs(be * c0))
    bad = np.abs(got - ref) > tol
    assert not bad.any(), (np.flatnonzero(bad)[:5], got[bad][:5], ref[bad][:5])


@pytest.mark.parametrize("shape", [(8192, 8192, 8192), (16384, 16384, 256)])
def test_full_size_shapes(gpu, shape):
    """The benchmark shapes, alpha=1, beta=-1 (the update form of the GEMM
    workload), all four transpose variants, data generated on the GPU and
    300 random entries checked against exact dot products."""
    ox = gpu
    import torch
    m, n, k = shape
    g = torch.Generator(device="cuda").manual_seed(0x0B1A5 + k)
    C0 = torch.randn(m * n, dtype=torch.float64, device="cuda", generator=g)
    rng = np.random.default_rng(k)
    ii = torch.from_numpy(rng.integers(0, m, 300)).cuda(); jj = torch.from_numpy(rng.integers(0, n, 300)).cuda()
    c0 = C0.view(n, m)[jj, ii].cpu().numpy()
    for ta in "NT":
        for tb in "NT":
            A = torch.randn(m * k, dtype=torch.float64, device="cuda", generator=g)   # storage of A (or A^T)
            B = torch.randn(k * n, dtype=torch.float64, device="cuda", generator=g)
            Ad = ox.DMatrix(A, m, k) if ta == "N" else ox.DMatrix(A, k, m)
            Bd = ox.DMatrix(B, k, n) if tb == "N" else ox.DMatrix(B, n, k)
            Cd = ox.DMatrix(C0.clone(), m, n)
            ox.gemm_(Cd, Ad, Bd, 1.0, -1.0, ta, tb)
            torch.cuda.synchronize()
            # column-major storage viewed as [col][row]
            Am = A.view(k, m) if ta == "N" else A.view(m, k)     # Am[p, i] = A[i, p] ('N') ; Am[i, p] = A[p, i] ('T')
            rows_a = (Am[:, ii].T if ta == "N" else Am[ii, :]).cpu().numpy()
            Bm = B.view(n, k) if tb == "N" else B.view(k, n)     # Bm[j, p] = B[p, j] ('N') ; Bm[p, j] = B[j, p] ('T')
            cols_b = (Bm[jj, :] if tb == "N" else Bm[:, jj].T).cpu().numpy()
            got = Cd.buf.view(n, m)[jj, ii].cpu().numpy()
            sampled_check(rows_a, cols_b, c0, got, 1.0, -1.0)
            del A, B, Cd


# Shapes that select each GEMM kernel (the tile rule: 128 x 128 on the 16x16x4
# kernel when that grid has >= 448 workgroups, 64 x 64 when its grid has
# >= 384, else 32 x 32; a side of at most 32 takes a 32-wide tile), once with
# every tile inside C and K a multiple of the slab depth (the unchecked
# interior path) and once with edges in M, N and K.
TILE_SHAPES = [
    ("128x128 interior", 2048, 4096, 256), ("128x128 edges", 2100, 4100, 300),
    ("64x64 interior", 1600, 1088, 96), ("64x64 edges", 1630, 1100, 70),
    ("32x32 interior", 640, 320, 64), ("32x32 edges", 650, 333, 37),
    ("32x64 (32 rows)", 32, 30000, 40), ("64x32 (20 columns)", 30000, 20, 40),
]


@pytest.mark.parametrize("name,m,n,k", TILE_SHAPES, ids=[s[0] for s in TILE_SHAPES])
def test_tile_kernels_every_entry(gpu, rng, name, m, n, k):
    """Every entry of C for each kernel path, all four transpose pairs,
    beta = 0 (C never read) and beta != 0, against numpy's product within the
    normwise bound of the shape tests."""
    ox = gpu
    for ta in "NT":
        for tb in "NT":
            A = rng.standard_normal((m, k) if ta == "N" else (k, m))
            B = rng.standard_normal((k, n) if tb == "N" else (n, k))
            for al, be in ((1.0, 0.0), (-1.0, 0.75)):
                C0 = np.asfortranarray(rng.standard_normal((m, n))) if be else np.full((m, n), np.nan, order="F")
                got = run(ox, C0, A, B, al, be, ta, tb)
                ref = al * (op(A, ta) @ op(B, tb)) + (be * C0 if be else 0.0)
                bound = 2 * (k + 2) * EPS * (abs(al) * np.linalg.norm(A) * np.linalg.norm(B) +
                                             (abs(be) * np.linalg.norm(C0) if be else 0.0))
                assert np.isfinite(got).all(), (name, ta, tb, be)
                assert np.linalg.norm(got - ref) <= bound, (name, ta, tb, be)
                np.testing.assert_allclose(got, ref, rtol=3e-12, atol=3e-12 * k)


def test_gemm_512_cube_beta_zero(gpu, rng):
    """The 512^3 FP64 product C = A*B (beta = 0, C overwritten: its NaN
    contents are never read), all transposes, against the oracle within the
    reference's tolerance (test/blas/gemm.jl:13) and k*eps*|A||B|."""
    n = 512
    for ta in "NT":
        for tb in "NT":
            A = rng.standard_normal((n, n)); B = rng.standard_normal((n, n))
            got = run(gpu, np.full((n, n), np.nan), A, B, 1.0, 0.0, ta, tb)
            ref = np.asfortranarray(np.full((n, n), np.nan))
            O.dgemm(ref, A, B, 1.0, 0.0, ta, tb, threads=8)
            assert np.all(np.isfinite(got))
            np.testing.assert_allclose(got, ref, rtol=3e-12, atol=3e-12 * n)
            assert np.linalg.norm(got - ref) <= n * EPS * np.linalg.norm(A) * np.linalg.norm(B)


# ---------------------------------------------------------------------------
# batched products (src/blas/batched.jl) and the rank-k update (level3.jl:546-783)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_strided_batched_vs_oracle(gpu, rng, ta, tb):
    for (m, n, k, b) in ((17, 19, 23, 5), (64, 64, 64, 33), (130, 70, 257, 3), (1, 300, 40, 2), (200, 200, 16, 40)):
        A = rng.standard_normal(((m, k) if ta == "N" else (k, m)) + (b,))
        B = rng.standard_normal(((k, n) if tb == "N" else (n, k)) + (b,))
        C0 = rng.standard_normal((m, n, b))
        got = gpu.gemm_strided_batched_(C0.copy(), A, B, 0.7, -0.4, ta, tb)
        for z in range(b):
            ref = np.asfortranarray(C0[:, :, z].copy())
            O.dgemm(ref, A[:, :, z], B[:, :, z], 0.7, -0.4, ta, tb)
            np.testing.assert_allclose(got[:, :, z], ref, rtol=3e-12, atol=3e-12 * k)


def test_strided_batched_rules(gpu, rng):
    ox = gpu
    A = rng.standard_normal((9, 7, 4)); B = rng.standard_normal((7, 5, 4))
    got = ox.gemm_strided_batched_(np.full((9, 5, 4), np.nan), A, B, 1.0, 0.0)   # beta 0 never reads C
    assert np.all(np.isfinite(got))
    got = ox.gemm_strided_batched_(np.full((9, 5, 4), 3.0), np.full((9, 7, 4), np.nan), np.full((7, 5, 4), np.nan),
                                   0.0, 0.5)
    assert np.array_equal(got, np.full((9, 5, 4), 1.5))                       # alpha 0 never reads A, B
    # overlapping outputs are rejected before any write
    Cd = ox.DBatch.from_numpy(np.zeros((9, 5, 4)))
    Ad, Bd = ox.DBatch.from_numpy(A), ox.DBatch.from_numpy(B)
    bad = ox.DBatch(Cd.buf, 9, 5, 4, stride=20)
    with pytest.raises(ValueError):
        ox.gemm_strided_batched_(bad, Ad, Bd)
    assert np.array_equal(Cd.numpy(), np.zeros((9, 5, 4)))
    with pytest.raises(ox.DimensionMismatch):
        ox.gemm_strided_batched_(np.zeros((9, 5, 3)), A, B)


def test_strided_batched_many_products(gpu, rng):
    # more products than one grid dimension holds, 4 x 4 products
    b = 70000
    A = rng.standard_normal((4, 4, b)); B = rng.standard_normal((4, 4, b))
    got = gpu.gemm_strided_batched_(np.zeros((4, 4, b)), A, B)
    np.testing.assert_allclose(got, np.einsum("ikz,kjz->ijz", A, B), rtol=1e-12, atol=1e-12)


def test_pointer_batched_mixed_shapes(gpu, rng):
    ox = gpu
    shapes = [(17, 19, 23), (64, 64, 64), (17, 19, 23), (5, 300, 7), (64, 64, 64), (64, 64, 64)]
    As = [rng.standard_normal((m, k)) for (m, n, k) in shapes]
    Bs = [rng.standard_normal((n, k)) for (m, n, k) in shapes]
    Cs = [rng.standard_normal((m, n)) for (m, n, k) in shapes]
    Ad = [ox.DMatrix.from_numpy(a) for a in As]
    Bd = [ox.DMatrix.from_numpy(b) for b in Bs]
    Cd = [ox.DMatrix.from_numpy(c) for c in Cs]
    ox.gemm_batched_(Cd, Ad, Bd, 1.5, 0.5, "N", "T")
    for i, (m, n, k) in enumerate(shapes):
        ref = np.asfortranarray(Cs[i].copy())
        O.dgemm(ref, As[i], Bs[i], 1.5, 0.5, "N", "T")
        np.testing.assert_allclose(Cd[i].numpy(), ref, rtol=3e-12, atol=3e-12 * k)
    # an input aliasing another member's output is read before that output is written
    X = ox.DMatrix.from_numpy(rng.standard_normal((8, 8)))
    Y = ox.DMatrix.from_numpy(rng.standard_normal((8, 8)))
    x0, y0 = X.numpy(), Y.numpy()
    ox.gemm_batched_([X, Y], [Y, X], [Y, X])
    np.testing.assert_allclose(X.numpy(), y0 @ y0, rtol=1e-12)
    np.testing.assert_allclose(Y.numpy(), x0 @ x0, rtol=1e-12)
    with pytest.raises(ValueError):
        ox.gemm_batched_([X, X], [Y, Y], [Y, Y])


@pytest.mark.parametrize("uplo", "UL")
@pytest.mark.parametrize("trans", "NT")
def test_syrk_vs_oracle(gpu, rng, uplo, trans):
    for (n, k) in ((1, 1), (17, 9), (64, 64), (130, 257), (300, 40), (1000, 128), (3000, 48)):   # 3000: 128 x 128 tiles
        A = rng.standard_normal((n, k) if trans == "N" else (k, n))
        C0 = np.asfortranarray(rng.standard_normal((n, n)))
        got = gpu.syrk_(C0.copy(order="F"), A, 0.7, -0.3, uplo, trans)
        ref = C0.copy(order="F")
        O.dsyrk(ref, A, 0.7, -0.3, uplo, trans)
        np.testing.assert_allclose(got, ref, rtol=3e-12, atol=3e-12 * k)
        mask = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
        assert np.array_equal(got[mask], C0[mask])         # the other triangle is untouched
    # NaN in the other triangle stays NaN on the 128 x 128 tiles as well
    n = 3000
    A = rng.standard_normal((n, 40) if trans == "N" else (40, n))
    C0 = np.asfortranarray(rng.standard_normal((n, n)))
    mask = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
    C0[mask] = np.nan
    got = gpu.syrk_(C0.copy(order="F"), A, -1.0, 1.0, uplo, trans)
    assert np.all(np.isnan(got[mask])) and np.all(np.isfinite(got[~mask]))
    A = rng.standard_normal((40, 10))
    got = gpu.syrk_(np.full((40, 40), np.nan), A, 1.0, 0.0, uplo)   # beta 0 never reads C
    tri = np.triu_indices(40) if uplo == "U" else np.tril_indices(40)
    assert np.all(np.isfinite(got[tri]))
    got = gpu.syrk_(np.full((6, 6), 2.0), np.full((6, 3), np.nan), 0.0, 0.5, uplo)   # alpha 0 never reads A
    expect = np.full((6, 6), 2.0)
    t6 = np.triu_indices(6) if uplo == "U" else np.tril_indices(6)
    expect[t6] = 1.0
    assert np.array_equal(got, expect)
    with pytest.raises(ValueError):
        gpu.syrk_(np.zeros((4, 4)), np.zeros((4, 2)), trans="C")


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_sgemm_vs_oracle(gpu, rng, ta, tb):
    # Float32 gemm! (src/blas/gemm.jl:625-703) on every tile path: 32 x 32 and
    # 64 x 64 (small grids), 128 x 128 interior tiles (K a multiple of 32),
    # edge tiles and K not a multiple of 32; normwise error <= k eps32 |A||B|
    # against the float64 product, and the oracle's Float32 gemm! beside it
    e32 = float(np.finfo(np.float32).eps)
    for (m, n, k) in ((17, 9, 40), (300, 260, 200), (2816, 2816, 128), (2900, 2830, 96), (2816, 2944, 45)):
        A = rng.standard_normal((m, k) if ta == "N" else (k, m)).astype(np.float32)
        B = rng.standard_normal((k, n) if tb == "N" else (n, k)).astype(np.float32)
        C0 = rng.standard_normal((m, n)).astype(np.float32)
        Ad = gpu.DMatrix.from_numpy(A, dtype=np.float32)
        Bd = gpu.DMatrix.from_numpy(B, dtype=np.float32)
        for beta in (0.0, -0.5):
            Cd = gpu.DMatrix.from_numpy(C0 if beta else np.full((m, n), np.nan, np.float32), dtype=np.float32)
            gpu.sgemm_(Cd, Ad, Bd, 1.25, beta, ta, tb)
            got = Cd.numpy().astype(np.float64)
            opA = (A if ta == "N" else A.T).astype(np.float64)
            opB = (B if tb == "N" else B.T).astype(np.float64)
            exact = 1.25 * opA @ opB + beta * C0.astype(np.float64)
            err = np.linalg.norm(got - exact)
            assert err <= 2 * k * e32 * (1.25 * np.linalg.norm(opA) * np.linalg.norm(opB) + abs(beta) * np.linalg.norm(C0))
            if m * n <= 300 * 300:
                Co = np.asfortranarray(C0.copy())
                O.sgemm(Co, A, B, 1.25, beta, ta, tb)
                np.testing.assert_allclose(got, Co, rtol=0, atol=4 * k * e32 * np.abs(exact).max())


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_hgemm_vs_oracle(gpu, rng, ta, tb):
    # gemm! for Float16 (src/blas/extended_blas.jl:39-88)
    for (m, n, k) in ((1, 1, 1), (17, 9, 40), (130, 70, 257), (256, 256, 64)):
        A = (rng.standard_normal((m, k) if ta == "N" else (k, m)) / 4).astype(np.float16)
        B = (rng.standard_normal((k, n) if tb == "N" else (n, k)) / 4).astype(np.float16)
        C = rng.standard_normal((m, n)).astype(np.float16)
        got = gpu.hgemm_(C.copy(), A, B, 0.75, -0.5, ta, tb)
        ref = O.hgemm(C, A, B, 0.75, -0.5, ta, tb)
        # both round once to Float16 from Float32 sums that differ in order:
        # at most one Float16 ulp apart, plus the Float32 reassociation
        tol = np.abs(ref).astype(np.float64) * 2.0 ** -10 + k * 2.0 ** -23 * (
            np.abs(op(A, ta)).astype(np.float64) @ np.abs(op(B, tb)).astype(np.float64)) + 2.0 ** -24
        assert np.all(np.abs(got.astype(np.float64) - ref.astype(np.float64)) <= tol)
        Ai = rng.integers(-4, 5, A.shape).astype(np.float16)
        Bi = rng.integers(-4, 5, B.shape).astype(np.float16)
        Ci = rng.integers(-8, 9, C.shape).astype(np.float16)
        assert np.array_equal(gpu.hgemm_(Ci.copy(), Ai, Bi, 0.5, -2.0, ta, tb), O.hgemm(Ci, Ai, Bi, 0.5, -2.0, ta, tb))
    got = gpu.hgemm_(np.full((7, 5), np.nan, np.float16), A[:7] if ta == "N" else A[:, :7],
                     B[:, :5] if tb == "N" else B[:5], 1.0, 0.0, ta, tb)
    assert np.all(np.isfinite(got))                       # beta == 0 never reads C
