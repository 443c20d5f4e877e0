// oracle/oxo.cpp -- CPU restatement of OxiBLAS.jl's level-3 hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
// implementation: only tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py may load it.  The product library never links or calls it.
//
// Every function restates the reference algorithm it names (paths relative to
// the OxiBLAS.jl source tree, revision 0.6.0-dev).  Matrices are column-major,
// indices 0-based here; pivots are stored 1-based with global row numbers as
// the reference's `pivots::Vector{Int}` (src/lapack/factorizations.jl:118).
//
// Parity status per function (how each one is pinned; see tests/test_oracle.py):
//   gemm      -- numpy matmul within k*eps*|A||B|, bit-exact on integer-valued
//                inputs, no-read rules with NaN fills, all N/T/C shapes of
//                test/blas/gemm.jl:2-14.
//   trsm/trsv -- scipy.linalg.solve_triangular for every side/uplo/trans/diag,
//                round trips B == op(A)*X, singular diagonals reported first.
//   getrf     -- pivots identical to LAPACK dgetrf (scipy) as the reference
//                asserts (test/lapack/lapack_dense.jl:12-14), normwise
//                reconstruction <= 30 (lapack_dense.jl:4,11), info of a zero
//                column (lapack_dense.jl:24-27, test/lapack/factorizations.jl:13-15).
//   getrs     -- backward error <= 30*cond(A) for N/T solves, solve forms agree.
//   equilibrate, residual, mixed_solve -- closed forms (power-of-two scales,
//                numpy residuals) and the convergence contract of
//                test/lapack/factorizations.jl:57-63.
// Bit-level agreement with the reference itself is "parity unpinned": the
// reference is Julia and cannot run in this environment.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

namespace {

// ---------------------------------------------------------------------------
// Operand views.  `tr` = transposed access (T and C coincide for real data).
// ---------------------------------------------------------------------------
template <class T>
struct CView {  // read-only view of op(X)
    const T* p; i64 ld; bool tr;
    inline T operator()(i64 i, i64 j) const { return tr ? p[j + i * ld] : p[i + j * ld]; }
};

// Register tile per element type, as the reference picks for AVX-512
// (src/utils/hardware.jl:26-48): Float64 (W,R,N)=(8,2,6) -> 16x6,
// Float32 (16,1,16) -> 16x16.
template <class T> struct Shape;
template <> struct Shape<double> { static constexpr int MR = 16, NR = 6; };
template <> struct Shape<float>  { static constexpr int MR = 16, NR = 16; };

// GemmConfig defaults (src/blas/gemm.jl:8-21).
template <class T> struct Cfg {
    static constexpr i64 mc = 576;
    static constexpr i64 kc = sizeof(T) < 8 ? 896 : 448;
    static constexpr i64 nc = 2048;
    static constexpr i64 parallel_threshold = 8000000;
    static constexpr i64 small_threshold = 8;
};

// _scale! (src/blas/gemm.jl:236-246)
template <class T>
void scale(T* C, i64 ldc, i64 m, i64 n, T beta) {
    if (beta == T(1)) return;
    for (i64 j = 0; j < n; ++j)
        for (i64 i = 0; i < m; ++i) {
            T& c = C[i + j * ldc];
            c = (beta == T(0)) ? T(0) : c * beta;
        }
}

// _small_gemm! (src/blas/gemm.jl:260-284): four-row register strips.
template <class T>
void small_gemm(T* C, i64 ldc, const CView<T>& A, const CView<T>& B, i64 m, i64 n, i64 k, T alpha, T beta) {
    i64 stop = m - (m % 4);
    for (i64 j = 0; j < n; ++j) {
        for (i64 row = 0; row < stop; row += 4) {
            T a = 0, b = 0, c = 0, d = 0;
            for (i64 p = 0; p < k; ++p) {
                T z = B(p, j);
                a = std::fma(A(row, p), z, a); b = std::fma(A(row + 1, p), z, b);
                c = std::fma(A(row + 2, p), z, c); d = std::fma(A(row + 3, p), z, d);
            }
            T acc[4] = {a, b, c, d};
            for (int r = 0; r < 4; ++r) {
                T& out = C[row + r + j * ldc];
                // alpha*s+beta*C is two rounded products and a rounded sum
                out = (beta == T(0)) ? alpha * acc[r] : alpha * acc[r] + beta * out;
            }
        }
        for (i64 row = stop; row < m; ++row) {
            T s = 0;
            for (i64 p = 0; p < k; ++p) s = std::fma(A(row, p), B(p, j), s);
            T& out = C[row + j * ldc];
            out = (beta == T(0)) ? alpha * s : alpha * s + beta * out;
        }
    }
}

// _pack_a! / _pack_b! (src/blas/gemm.jl:314-329): micro-panel order, zero tails.
template <class T, int MR>
void pack_a(T* dst, const CView<T>& A, i64 ic, i64 pc, i64 ib, i64 kb) {
    for (i64 ii = 0; ii < ib; ii += MR)
        for (i64 p = 0; p < kb; ++p)
            for (int i = 0; i < MR; ++i)
                dst[ii * kb + p * MR + i] = (ii + i < ib) ? A(ic + ii + i, pc + p) : T(0);
}
template <class T, int NR>
void pack_b(T* dst, const CView<T>& B, i64 pc, i64 jc, i64 kb, i64 jb) {
    for (i64 jj = 0; jj < jb; jj += NR)
        for (i64 p = 0; p < kb; ++p)
            for (int j = 0; j < NR; ++j)
                dst[jj * kb + p * NR + j] = (jj + j < jb) ? B(pc + p, jc + jj + j) : T(0);
}

// _simd_micro! (src/blas/simd.jl:151-199): one fused multiply-add per lane and
// K step, alpha*acc at the store, fma(beta, C, alpha*acc) when beta != 0.
template <class T, int MR, int NR>
void micro(T* c, i64 ldc, const T* a, const T* b, i64 k, T alpha, T beta) {
    T acc[NR][MR];
    for (int j = 0; j < NR; ++j) for (int i = 0; i < MR; ++i) acc[j][i] = 0;
    for (i64 p = 0; p < k; ++p) {
        const T* ap = a + p * MR; const T* bp = b + p * NR;
        for (int j = 0; j < NR; ++j) {
            T bb = bp[j];
            for (int i = 0; i < MR; ++i) acc[j][i] = std::fma(ap[i], bb, acc[j][i]);
        }
    }
    for (int j = 0; j < NR; ++j)
        for (int i = 0; i < MR; ++i) {
            T v = alpha * acc[j][i];
            T& out = c[i + j * ldc];
            out = (beta == T(0)) ? v : std::fma(beta, out, v);
        }
}

// _packed_region! (src/blas/gemm.jl:414-485): jc -> pc -> pack B -> ic ->
// pack A -> micro tiles; beta only with the first K block (bp), edge tiles go
// through a scratch tile and are added as v + bp*C (unfused).
template <class T>
void packed_region(T* C, i64 ldc, const CView<T>& A, const CView<T>& B, i64 k, T alpha, T beta,
                   i64 r0, i64 r1, i64 c0, i64 c1, T* pa, T* pb) {
    constexpr int MR = Shape<T>::MR, NR = Shape<T>::NR;
    T edge[MR * NR];
    for (i64 jc = c0; jc < c1; jc += Cfg<T>::nc) {
        i64 jb = std::min<i64>(Cfg<T>::nc, c1 - jc);
        for (i64 pc = 0; pc < k; pc += Cfg<T>::kc) {
            i64 kb = std::min<i64>(Cfg<T>::kc, k - pc);
            pack_b<T, NR>(pb, B, pc, jc, kb, jb);
            T bp = (pc == 0) ? beta : T(1);
            for (i64 ic = r0; ic < r1; ic += Cfg<T>::mc) {
                i64 ib = std::min<i64>(Cfg<T>::mc, r1 - ic);
                pack_a<T, MR>(pa, A, ic, pc, ib, kb);
                for (i64 jj = 0; jj < jb; jj += NR)
                    for (i64 ii = 0; ii < ib; ii += MR) {
                        const T* aa = pa + ii * kb;
                        const T* bb = pb + jj * kb;
                        i64 mi = std::min<i64>(MR, ib - ii), nj = std::min<i64>(NR, jb - jj);
                        if (mi == MR && nj == NR) {
                            micro<T, MR, NR>(C + (ic + ii) + (jc + jj) * ldc, ldc, aa, bb, kb, alpha, bp);
                        } else {
                            micro<T, MR, NR>(edge, MR, aa, bb, kb, alpha, T(0));
                            for (i64 j = 0; j < nj; ++j)
                                for (i64 i = 0; i < mi; ++i) {
                                    T v = edge[i + j * MR];
                                    T& out = C[(ic + ii + i) + (jc + jj + j) * ldc];
                                    out = (bp == T(0)) ? v : v + bp * out;
                                }
                        }
                    }
            }
        }
    }
}

// _gemm_workers / _worker_grid (src/blas/gemm.jl:532-560).
int gemm_workers(double work, int cost, int limit) {
    if (limit < 2 || work < double(Cfg<double>::parallel_threshold)) return 1;
    double w = std::floor(work * cost / 6e6);
    return int(std::max(2.0, std::min(double(limit), w)));
}
void worker_grid(int nt, i64 m, i64 n, i64 rt, i64 ct, int& pr_out, int& pc_out) {
    double target = std::sqrt(double(nt) * m / n);
    int best_r = 1, best_c = int(std::min<i64>(nt, ct)); double score = -INFINITY;
    for (int pr = 1; pr <= std::min<i64>(nt, rt); ++pr) {
        int pc = int(std::min<i64>(nt / pr, ct));
        int w = pr * pc; if (w == 0) continue;
        double s = double(w) - 1.5 * std::fabs(std::log(double(pr) / target));
        if (s > score) { best_r = pr; best_c = pc; score = s; }
    }
    pr_out = best_r; pc_out = best_c;
}

// _run_packed! (src/blas/gemm.jl:561-612): 2-D grid of output blocks, one per
// worker, each with its own packing panels.
template <class T>
void run_packed(T* C, i64 ldc, const CView<T>& A, const CView<T>& B, i64 m, i64 n, i64 k, T alpha, T beta, int threads) {
    constexpr int MR = Shape<T>::MR, NR = Shape<T>::NR;
    i64 acap = (std::min<i64>(m, Cfg<T>::mc) + 15) * std::min<i64>(k, Cfg<T>::kc);
    i64 bcap = (std::min<i64>(n, Cfg<T>::nc) + 15) * std::min<i64>(k, Cfg<T>::kc);
    int nt = gemm_workers(double(m) * n * k, sizeof(T) / 4, threads);
    i64 rt = (m + MR - 1) / MR, ct = (n + NR - 1) / NR;
    int pr = 1, pcn = 1;
    if (nt > 1) { worker_grid(int(std::min<i64>(nt, rt * ct)), m, n, rt, ct, pr, pcn); }
    nt = pr * pcn;
    if (nt <= 1) {
        std::vector<T> pa(acap), pb(bcap);
        packed_region(C, ldc, A, B, k, alpha, beta, 0, m, 0, n, pa.data(), pb.data());
        return;
    }
#pragma omp parallel for schedule(static, 1) num_threads(nt)
    for (int worker = 0; worker < nt; ++worker) {
        int i = worker / pcn, j = worker % pcn;
        i64 r0 = (i64(i) * rt / pr) * MR, r1 = std::min<i64>(m, (i64(i + 1) * rt / pr) * MR);
        i64 q0 = (i64(j) * ct / pcn) * NR, q1 = std::min<i64>(n, (i64(j + 1) * ct / pcn) * NR);
        if (r0 >= r1 || q0 >= q1) continue;
        std::vector<T> pa(acap), pb(bcap);
        packed_region(C, ldc, A, B, k, alpha, beta, r0, r1, q0, q1, pa.data(), pb.data());
    }
}

bool ranges_overlap(const void* a, size_t abytes, const void* b, size_t bbytes) {
    const char* x = (const char*)a; const char* y = (const char*)b;
    return x < y + bbytes && y < x + abytes;
}

// gemm! -> _gemm_ops! -> _gemm_run! (src/blas/gemm.jl:635-703): dimension and
// option checks before any write, m==0||n==0 returns, alpha==0||k==0 scales C
// only (A and B never read), beta==0 never reads C, operands overlapping C are
// copied first, the small path for max(m,n)<=8 && k<=512.
template <class T>
int gemm(char ta, char tb, i64 m, i64 n, i64 k, T alpha, const T* A, i64 lda, const T* B, i64 ldb,
         T beta, T* C, i64 ldc, int threads) {
    auto kind = [](char t) { return (t == 'N' || t == 'n') ? 0 : (t == 'T' || t == 't' || t == 'C' || t == 'c') ? 1 : -1; };
    int ka = kind(ta), kb = kind(tb);
    if (ka < 0) return -1;
    if (kb < 0) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (lda < std::max<i64>(1, ka ? k : m)) return -8;
    if (ldb < std::max<i64>(1, kb ? n : k)) return -10;
    if (ldc < std::max<i64>(1, m)) return -13;
    if (m == 0 || n == 0) return 0;
    if (alpha == T(0) || k == 0) { scale(C, ldc, m, n, beta); return 0; }
    // overlap snapshot
    size_t cb = size_t((n - 1) * ldc + m) * sizeof(T);
    i64 ar = ka ? k : m, ac = ka ? m : k, br = kb ? n : k, bc = kb ? k : n;
    size_t ab = size_t((ac - 1) * lda + ar) * sizeof(T), bb = size_t((bc - 1) * ldb + br) * sizeof(T);
    std::vector<T> acopy, bcopy;
    if (ranges_overlap(C, cb, A, ab)) {
        acopy.resize(size_t(ar) * ac);
        for (i64 j = 0; j < ac; ++j) for (i64 i = 0; i < ar; ++i) acopy[i + j * ar] = A[i + j * lda];
        A = acopy.data(); lda = ar;
    }
    if (ranges_overlap(C, cb, B, bb)) {
        bcopy.resize(size_t(br) * bc);
        for (i64 j = 0; j < bc; ++j) for (i64 i = 0; i < br; ++i) bcopy[i + j * br] = B[i + j * ldb];
        B = bcopy.data(); ldb = br;
    }
    CView<T> Av{A, lda, ka != 0}, Bv{B, ldb, kb != 0};
    bool small = Cfg<T>::small_threshold > 0 && std::max(m, n) <= Cfg<T>::small_threshold && k <= 512;
    if (small) { small_gemm(C, ldc, Av, Bv, m, n, k, alpha, beta); return 0; }
    run_packed(C, ldc, Av, Bv, m, n, k, alpha, beta, threads);
    return 0;
}

// ---------------------------------------------------------------------------
// Triangular solves.
// ---------------------------------------------------------------------------
// _trsm_leaf_generic! (src/blas/level3.jl:309-336) and _tri_left4!
// (src/blas/triangular_microkernels.jl:31-59): both evaluate, per element,
// s = B[i,j]; s = fma(-A[i,p], B[p,j], s) over p in order; s/A[i,i] when not
// unit.  The four-column leaf only changes which columns share a pass, so the
// per-element arithmetic below is the one both leaves perform.
template <class T>
void trsm_leaf(T* Bp, i64 ldb, i64 m, i64 n, const CView<T>& A, bool left, bool upper, bool unit) {
    if (left) {
        for (i64 j = 0; j < n; ++j) {
            T* b = Bp + j * ldb;
            if (upper) {
                for (i64 i = m - 1; i >= 0; --i) {
                    T s = b[i];
                    for (i64 p = i + 1; p < m; ++p) s = std::fma(-A(i, p), b[p], s);
                    b[i] = unit ? s : s / A(i, i);
                }
            } else {
                for (i64 i = 0; i < m; ++i) {
                    T s = b[i];
                    for (i64 p = 0; p < i; ++p) s = std::fma(-A(i, p), b[p], s);
                    b[i] = unit ? s : s / A(i, i);
                }
            }
        }
    } else {
        // right side: contiguous column updates (level3.jl:319-333)
        auto step = [&](i64 j) {
            if (!unit) { T d = A(j, j); for (i64 i = 0; i < m; ++i) Bp[i + j * ldb] /= d; }
            i64 lo = upper ? j + 1 : 0, hi = upper ? n : j;
            for (i64 p = lo; p < hi; ++p) {
                T a = -A(j, p);
                for (i64 i = 0; i < m; ++i) Bp[i + p * ldb] = std::fma(Bp[i + j * ldb], a, Bp[i + p * ldb]);
            }
        };
        if (upper) for (i64 j = 0; j < n; ++j) step(j);
        else for (i64 j = n - 1; j >= 0; --j) step(j);
    }
}

// op(A) sub-block view helper: rows r0.., cols c0.. of op(A).
template <class T>
CView<T> sub(const CView<T>& A, i64 r0, i64 c0) {
    CView<T> s = A;
    s.p = A.tr ? A.p + c0 + r0 * A.ld : A.p + r0 + c0 * A.ld;
    return s;
}

// _trsm_blocks_left! / _trsm_blocks_right! (src/blas/level3.jl:338-376):
// halve the triangle until it is at most `bs`, GEMM updates in between.
template <class T>
void trsm_blocks(T* B, i64 ldb, i64 m, i64 n, const CView<T>& A, i64 na, bool left, bool upper, bool unit, i64 bs,
                 int threads) {
    if (na <= bs) { trsm_leaf(B, ldb, m, n, A, left, upper, unit); return; }
    i64 h = na / 2;
    CView<T> A1 = A, A2 = sub(A, h, h);
    if (left) {
        T* B1 = B; T* B2 = B + h;
        if (upper) {
            trsm_blocks(B2, ldb, m - h, n, A2, na - h, left, upper, unit, bs, threads);
            CView<T> A12 = sub(A, 0, h);
            // C=B1 (h x n), A=A12 (h x (na-h)), B=B2 ((na-h) x n)
            gemm<T>(A12.tr ? 'T' : 'N', 'N', h, n, na - h, T(-1), A12.p, A12.ld, B2, ldb, T(1), B1, ldb, threads);
            trsm_blocks(B1, ldb, h, n, A1, h, left, upper, unit, bs, threads);
        } else {
            trsm_blocks(B1, ldb, h, n, A1, h, left, upper, unit, bs, threads);
            CView<T> A21 = sub(A, h, 0);
            gemm<T>(A21.tr ? 'T' : 'N', 'N', na - h, n, h, T(-1), A21.p, A21.ld, B1, ldb, T(1), B2, ldb, threads);
            trsm_blocks(B2, ldb, m - h, n, A2, na - h, left, upper, unit, bs, threads);
        }
    } else {
        T* B1 = B; T* B2 = B + h * ldb;
        if (upper) {
            trsm_blocks(B1, ldb, m, h, A1, h, left, upper, unit, bs, threads);
            CView<T> A12 = sub(A, 0, h);
            gemm<T>('N', A12.tr ? 'T' : 'N', m, na - h, h, T(-1), B1, ldb, A12.p, A12.ld, T(1), B2, ldb, threads);
            trsm_blocks(B2, ldb, m, na - h, A2, na - h, left, upper, unit, bs, threads);
        } else {
            trsm_blocks(B2, ldb, m, na - h, A2, na - h, left, upper, unit, bs, threads);
            CView<T> A21 = sub(A, h, 0);
            gemm<T>('N', A21.tr ? 'T' : 'N', m, h, na - h, T(-1), B2, ldb, A21.p, A21.ld, T(1), B1, ldb, threads);
            trsm_blocks(B1, ldb, m, h, A1, h, left, upper, unit, bs, threads);
        }
    }
}

// trsm! (src/blas/level3.jl:518-544, _trsm_trans! :485-516, _trsm_compute!
// :453-467): transposition flips the triangle, a zero on a non-unit diagonal
// is reported (1-based) before B is touched, alpha scales B first.
template <class T>
int trsm(char side, char uplo, char trans, char diag, i64 m, i64 n, T alpha, const T* A, i64 lda, T* B, i64 ldb,
         i64 bs, int threads) {
    bool left = side == 'L' || side == 'l';
    if (!left && !(side == 'R' || side == 'r')) return -1;
    bool upper = uplo == 'U' || uplo == 'u';
    if (!upper && !(uplo == 'L' || uplo == 'l')) return -2;
    int tk = (trans == 'N' || trans == 'n') ? 0 : (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c') ? 1 : -1;
    if (tk < 0) return -3;
    bool unit = diag == 'U' || diag == 'u';
    if (!unit && !(diag == 'N' || diag == 'n')) return -4;
    if (m < 0) return -5;
    if (n < 0) return -6;
    i64 na = left ? m : n;
    if (lda < std::max<i64>(1, na)) return -9;
    if (ldb < std::max<i64>(1, m)) return -11;
    if (bs <= 0) return -12;
    if (alpha == T(0) || m == 0 || n == 0) { scale(B, ldb, m, n, alpha); return 0; }
    CView<T> Av{A, lda, tk != 0};
    if (tk) upper = !upper;
    if (!unit)
        for (i64 i = 0; i < na; ++i)
            if (Av(i, i) == T(0)) return int(i + 1);  // SingularException(i+1), nothing written
    scale(B, ldb, m, n, alpha);
    trsm_blocks(B, ldb, m, n, Av, na, left, upper, unit, bs, threads);
    return 0;
}

// trsv! (src/blas/level2.jl:436-472): column form with four fused columns for
// an untransposed operand (_trsv_columns! :408-435), row dot products for a
// transposed one.
template <class T>
void trsv_diag(T* x, const CView<T>& A, bool unit, bool upper, i64 lo, i64 hi) {
    if (upper) {
        for (i64 jj = hi; jj >= lo; --jj) {
            T v = unit ? x[jj] : x[jj] / A(jj, jj); x[jj] = v;
            for (i64 ii = lo; ii < jj; ++ii) x[ii] = std::fma(A(ii, jj), -v, x[ii]);
        }
    } else {
        for (i64 jj = lo; jj <= hi; ++jj) {
            T v = unit ? x[jj] : x[jj] / A(jj, jj); x[jj] = v;
            for (i64 ii = jj + 1; ii <= hi; ++ii) x[ii] = std::fma(A(ii, jj), -v, x[ii]);
        }
    }
}
template <class T>
int trsv(char uplo, char trans, char diag, i64 n, const T* A, i64 lda, T* x) {
    bool upper = uplo == 'U' || uplo == 'u';
    if (!upper && !(uplo == 'L' || uplo == 'l')) return -1;
    int tk = (trans == 'N' || trans == 'n') ? 0 : (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c') ? 1 : -1;
    if (tk < 0) return -2;
    bool unit = diag == 'U' || diag == 'u';
    if (!unit && !(diag == 'N' || diag == 'n')) return -3;
    if (n < 0) return -4;
    if (lda < std::max<i64>(1, n)) return -6;
    CView<T> Av{A, lda, tk != 0};
    if (tk) upper = !upper;
    if (!unit)
        for (i64 i = 0; i < n; ++i)
            if (Av(i, i) == T(0)) return int(i + 1);
    if (!tk) {
        if (upper) {
            i64 j = n - 1;
            while (j >= 3) {
                trsv_diag(x, Av, unit, true, j - 3, j);
                T c1 = -x[j - 3], c2 = -x[j - 2], c3 = -x[j - 1], c4 = -x[j];
                for (i64 i = 0; i < j - 3; ++i)
                    x[i] = std::fma(Av(i, j - 3), c1, std::fma(Av(i, j - 2), c2, std::fma(Av(i, j - 1), c3, std::fma(Av(i, j), c4, x[i]))));
                j -= 4;
            }
            if (j >= 0) trsv_diag(x, Av, unit, true, 0, j);
        } else {
            i64 j = 0;
            while (j + 3 < n) {
                trsv_diag(x, Av, unit, false, j, j + 3);
                T c1 = -x[j], c2 = -x[j + 1], c3 = -x[j + 2], c4 = -x[j + 3];
                for (i64 i = j + 4; i < n; ++i)
                    x[i] = std::fma(Av(i, j + 3), c4, std::fma(Av(i, j + 2), c3, std::fma(Av(i, j + 1), c2, std::fma(Av(i, j), c1, x[i]))));
                j += 4;
            }
            if (j < n) trsv_diag(x, Av, unit, false, j, n - 1);
        }
        return 0;
    }
    if (upper) {
        for (i64 i = n - 1; i >= 0; --i) {
            T s = 0;
            for (i64 j = i + 1; j < n; ++j) s = std::fma(Av(i, j), x[j], s);
            s = x[i] - s; x[i] = unit ? s : s / Av(i, i);
        }
    } else {
        for (i64 i = 0; i < n; ++i) {
            T s = 0;
            for (i64 j = 0; j < i; ++j) s = std::fma(Av(i, j), x[j], s);
            s = x[i] - s; x[i] = unit ? s : s / Av(i, i);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// LU with partial pivoting.
// ---------------------------------------------------------------------------
// _lu_swap! (src/lapack/factorizations.jl:80-88): interchanges of steps
// j1..j2 (0-based step indices, pivots 1-based global rows minus `poff`) on a
// column range, column by column.
template <class T>
void lu_swap(T* F, i64 ldf, const i64* piv, i64 poff, i64 j1, i64 j2, i64 c0, i64 c1) {
    for (i64 col = c0; col < c1; ++col)
        for (i64 j = j1; j <= j2; ++j) {
            i64 p = piv[j] - 1 - poff;
            if (p != j) std::swap(F[j + col * ldf], F[p + col * ldf]);
        }
}

// _lu_pivot (src/lapack/factorizations.jl:89-107): largest |value| at or
// below row j; strict '>' keeps the first row on ties.
template <class T>
i64 lu_pivot(const T* P, i64 ldp, i64 mp, i64 j) {
    i64 p = j; T best = std::fabs(P[j + j * ldp]);
    for (i64 i = j + 1; i < mp; ++i) {
        T v = std::fabs(P[i + j * ldp]);
        if (v > best) { best = v; p = i; }
    }
    return p;
}

struct LuState { i64* pivots; i64 info; bool check; bool stop; int threads; bool tournament = false; };

// Tournament pivoting for a leaf (oxi_sgetrf_tournament, the FP32 factor of
// the mixed solve with pivot 2; no reference counterpart).  tp_gepp: partial
// pivoting over a set of rows (row-major, 32 values each, positions id) for
// columns 0..w-1, with ties to the lowest position and NaN never winning;
// returns the pivot positions in order and, with prow, each pivot row as
// partial pivoting leaves it (multipliers left of its column, U from it on).
// A zero pivot leaves its column and updates nothing.
template <class T>
static T tp_key(T v) {
    T a = std::fabs(v);
    return a != a ? T(-1) : a;
}
template <class T>
static void tp_gepp(std::vector<T>& V, const std::vector<i64>& id, int w, std::vector<i64>& piv,
                    std::vector<T>* prow) {
    const size_t nr = id.size();
    std::vector<char> live(nr, 1);
    piv.clear();
    for (int c = 0; c < w; ++c) {
        i64 best = -1, bid = std::numeric_limits<i64>::max();
        T bk = T(-2);
        for (size_t r = 0; r < nr; ++r)
            if (live[r]) {
                const T k = tp_key(V[r * 32 + c]);
                if (k > bk || (k == bk && id[r] < bid)) { bk = k; bid = id[r]; best = i64(r); }
            }
        if (best < 0) break;
        piv.push_back(id[best]);
        live[best] = 0;
        const T* u = &V[best * 32];
        if (prow) prow->insert(prow->end(), u, u + 32);
        const T d = u[c];
        if (d == T(0)) continue;
        for (size_t r = 0; r < nr; ++r)
            if (live[r]) {
                T* v = &V[r * 32];
                const T l = v[c] / d;
                v[c] = l;
                for (int k = c + 1; k < 32; ++k) v[k] = std::fma(-l, u[k], v[k]);
            }
    }
}

// The tournament leaf on the mp x np block P (np <= 32).  The rows sit in
// workgroups of R rows (R = 256, 512 beyond 8192 rows), each of four waves;
// wave v of workgroup g holds rows g*R + s*256 + v*64 + l (s < R/256,
// l < 64).  Candidates by partial pivoting inside each wave's rows, then by
// partial pivoting over the union of a workgroup's wave candidates, then over
// the union of the candidates of four consecutive workgroups (of four
// consecutive groups, and so on) until one set is left; the leaf's pivots are
// partial pivoting's over that set, the pivot rows as it leaves them, and
// every other row is eliminated with them.  Rows are stored in the order the
// interchanges give, pivots as for the partial-pivoting leaf.
template <class T>
void tp_leaf(T* P, i64 ldp, i64 mp, i64 np, i64 offset, LuState& st) {
    const int w = int(np);
    const i64 R = mp > 8192 ? 512 : 256;
    auto load = [&](i64 r, T* dst) {
        for (int c = 0; c < 32; ++c) dst[c] = c < w ? P[r + c * ldp] : T(0);
    };
    auto pivots_of = [&](const std::vector<i64>& rows, std::vector<i64>& pv, std::vector<T>* prow) {
        std::vector<T> V(rows.size() * 32);
        for (size_t k = 0; k < rows.size(); ++k) load(rows[k], &V[k * 32]);
        tp_gepp(V, rows, w, pv, prow);
    };
    std::vector<std::vector<i64>> sets;
    for (i64 g0 = 0; g0 < mp; g0 += R) {
        std::vector<i64> uni, pv;
        for (int v = 0; v < 4; ++v) {
            std::vector<i64> rows;
            for (i64 s = 0; s < R / 256; ++s)
                for (int l = 0; l < 64; ++l) {
                    const i64 r = g0 + s * 256 + v * 64 + l;
                    if (r < mp) rows.push_back(r);
                }
            if (rows.empty()) continue;
            pivots_of(rows, pv, nullptr);
            uni.insert(uni.end(), pv.begin(), pv.end());
        }
        pivots_of(uni, pv, nullptr);
        sets.push_back(pv);
    }
    while (sets.size() > 1) {
        std::vector<std::vector<i64>> next;
        for (size_t g = 0; g < sets.size(); g += 4) {
            std::vector<i64> uni, pv;
            for (size_t i = g; i < std::min(sets.size(), g + 4); ++i) uni.insert(uni.end(), sets[i].begin(), sets[i].end());
            pivots_of(uni, pv, nullptr);
            next.push_back(pv);
        }
        sets.swap(next);
    }
    std::vector<T> prow;
    std::vector<i64> pid;
    pivots_of(sets[0], pid, &prow);
    for (int c = 0; c < w; ++c)
        if (prow[size_t(c) * 32 + c] == T(0)) {
            if (st.info == 0) st.info = offset + c + 1;
            if (st.check) { st.stop = true; return; }
            break;
        }
    // the interchanges: position of every row as the steps move it
    std::vector<i64> perm(mp), where(mp);
    for (i64 i = 0; i < mp; ++i) perm[i] = where[i] = i;
    for (int c = 0; c < w; ++c) {
        const i64 q = where[pid[c]];
        st.pivots[offset + c] = offset + q + 1;
        const i64 a = perm[c], b = perm[q];
        std::swap(perm[c], perm[q]);
        where[a] = q; where[b] = c;
    }
    std::vector<T> out(size_t(mp) * 32);
    for (i64 i = 0; i < mp; ++i) {
        T* o = &out[size_t(i) * 32];
        if (i < w) { std::copy(&prow[size_t(i) * 32], &prow[size_t(i) * 32] + 32, o); continue; }
        load(perm[i], o);
        for (int c = 0; c < w; ++c) {
            const T* u = &prow[size_t(c) * 32];
            const T d = u[c];
            if (d == T(0)) continue;
            const T l = o[c] / d;
            o[c] = l;
            for (int k = c + 1; k < 32; ++k) o[k] = std::fma(-l, u[k], o[k]);
        }
    }
    for (i64 i = 0; i < mp; ++i)
        for (int c = 0; c < w; ++c) P[i + c * ldp] = out[size_t(i) * 32 + c];
}

// _lu_panel! (src/lapack/factorizations.jl:108-152): recursive getrf2 with a
// column-by-column leaf of up to 32 columns (division by the pivot).
template <class T>
void lu_panel(T* P, i64 ldp, i64 mp, i64 np, i64 offset, LuState& st) {
    if (st.stop) return;
    if (np <= 32 && st.tournament && mp >= np) { tp_leaf(P, ldp, mp, np, offset, st); return; }
    if (np <= 32) {
        for (i64 j = 0; j < std::min(mp, np); ++j) {
            i64 p = lu_pivot(P, ldp, mp, j);
            st.pivots[offset + j] = offset + p + 1;
            if (p != j)
                for (i64 col = 0; col < np; ++col) std::swap(P[j + col * ldp], P[p + col * ldp]);
            T d = P[j + j * ldp];
            if (d == T(0)) {
                if (st.info == 0) st.info = offset + j + 1;
                if (st.check) { st.stop = true; return; }
                continue;
            }
            for (i64 i = j + 1; i < mp; ++i) P[i + j * ldp] /= d;
            for (i64 col = j + 1; col < np; ++col) {
                T u = P[j + col * ldp];
                for (i64 i = j + 1; i < mp; ++i) P[i + col * ldp] = std::fma(-P[i + j * ldp], u, P[i + col * ldp]);
            }
        }
        return;
    }
    i64 h = np / 2;
    lu_panel(P, ldp, mp, h, offset, st);
    if (st.stop) return;
    lu_swap(P, ldp, st.pivots + offset, offset, 0, h - 1, h, np);
    trsm<T>('L', 'L', 'N', 'U', h, np - h, T(1), P, ldp, P + h * ldp, ldp, 8, st.threads);
    if (mp > h) {
        gemm<T>('N', 'N', mp - h, np - h, h, T(-1), P + h, ldp, P + h * ldp, ldp, T(1), P + h + h * ldp, ldp, st.threads);
        lu_panel(P + h + h * ldp, ldp, mp - h, np - h, offset + h, st);
        if (st.stop) return;
        lu_swap(P, ldp, st.pivots + offset, offset, h, std::min(mp, np) - 1, 0, h);
    }
}

// _lu_factor! (src/lapack/factorizations.jl:30-77): panels of `blocksize`
// columns, interchanges applied to both sides, TRSM on the U12 block and the
// trailing GEMM update.  Look-ahead only reorders independent work, so the
// sequential order below gives the same factors (lapack_dense.jl:33-38).
template <class T>
int getrf(i64 m, i64 n, T* F, i64 ldf, i64* pivots, i64 bs, int check, i64* info, int threads,
          bool tournament = false) {
    if (m < 0) return -1;
    if (n < 0) return -2;
    if (ldf < std::max<i64>(1, m)) return -4;
    if (bs <= 0) return -6;
    i64 q = std::min(m, n);
    LuState st{pivots, 0, check != 0, false, threads, tournament};
    for (i64 first = 0; first < q; first += bs) {
        i64 last = std::min(q, first + bs) - 1;  // inclusive
        lu_panel(F + first + first * ldf, ldf, m - first, last - first + 1, first, st);
        if (st.stop) break;
        // the pivots hold global rows: swap on F with offset 0
        lu_swap(F, ldf, pivots, 0, first, last, 0, first);
        lu_swap(F, ldf, pivots, 0, first, last, last + 1, n);
        if (last + 1 >= n) continue;
        trsm<T>('L', 'L', 'N', 'U', last - first + 1, n - last - 1, T(1), F + first + first * ldf, ldf,
                F + first + (last + 1) * ldf, ldf, 8, threads);
        if (last + 1 >= m) continue;
        gemm<T>('N', 'N', m - last - 1, n - last - 1, last - first + 1, T(-1), F + (last + 1) + first * ldf, ldf,
                F + first + (last + 1) * ldf, ldf, T(1), F + (last + 1) + (last + 1) * ldf, ldf, threads);
    }
    *info = st.info;
    return 0;
}

// ldiv!(F::LU, B) and its transpose (src/api/linalg_api.jl:168-184), with
// _rowswap! (:100-110); a vector right-hand side goes to trsv!, a matrix to
// trsm! (:112-113).
template <class T>
int getrs(char trans, i64 n, i64 nrhs, const T* LU, i64 ldlu, const i64* pivots, T* B, i64 ldb, int vector_rhs,
          int threads) {
    int tk = (trans == 'N' || trans == 'n') ? 0 : (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c') ? 1 : -1;
    if (tk < 0) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (ldlu < std::max<i64>(1, n)) return -5;
    if (ldb < std::max<i64>(1, n)) return -8;
    if (vector_rhs && nrhs != 1) return -9;
    auto rowswap = [&](bool inverse) {
        for (i64 col = 0; col < nrhs; ++col) {
            T* c = B + col * ldb;
            for (i64 k = 0; k < n; ++k) {
                i64 j = inverse ? n - 1 - k : k;
                i64 q = pivots[j] - 1;
                if (q != j) std::swap(c[j], c[q]);
            }
        }
    };
    auto tri = [&](char uplo, char tr, char diag) {
        if (vector_rhs) return trsv<T>(uplo, tr, diag, n, LU, ldlu, B);
        return trsm<T>('L', uplo, tr, diag, n, nrhs, T(1), LU, ldlu, B, ldb, 8, threads);
    };
    int r;
    if (!tk) {
        rowswap(false);
        if ((r = tri('L', 'N', 'U'))) return r;
        if ((r = tri('U', 'N', 'N'))) return r;
    } else {
        if ((r = tri('U', 'T', 'N'))) return r;
        if ((r = tri('L', 'T', 'U'))) return r;
        rowswap(true);
    }
    return 0;
}

// equilibrate(A; radix=true) (src/lapack/dense_utilities.jl:191-212).
template <class T>
void equilibrate(i64 m, i64 n, T* B, i64 ldb, T* rows, T* cols, int radix) {
    const int emin = std::numeric_limits<T>::min_exponent - 1;   // exponent(floatmin)
    const int emax = std::numeric_limits<T>::max_exponent - 1;   // exponent(floatmax)
    auto sc = [&](T value) -> T {
        if (!radix) return value;
        double l = std::nearbyint(std::log2(double(value)));  // round-half-even as Julia round(Int, x)
        int e = int(std::min<double>(emax, std::max<double>(emin, l)));
        return std::ldexp(T(1), e);
    };
    for (i64 i = 0; i < m; ++i) {
        rows[i] = 1;
        T v = 0;
        for (i64 j = 0; j < n; ++j) v = std::max(v, std::fabs(B[i + j * ldb]));
        if (v > 0) { rows[i] = sc(v); for (i64 j = 0; j < n; ++j) B[i + j * ldb] = B[i + j * ldb] / rows[i]; }
    }
    for (i64 j = 0; j < n; ++j) {
        cols[j] = 1;
        T v = 0;
        for (i64 i = 0; i < m; ++i) v = std::max(v, std::fabs(B[i + j * ldb]));
        if (v > 0) { cols[j] = sc(v); for (i64 i = 0; i < m; ++i) B[i + j * ldb] = B[i + j * ldb] / cols[j]; }
    }
}

// _refinement_residual! (src/lapack/expert_refinement.jl:67-81): residual in
// the working type and the componentwise backward error
// max_i |r_i| / (|A||x| + |b|)_i.
double refinement_residual(i64 n, i64 nrhs, const double* A, i64 lda, const double* X, i64 ldx, const double* B,
                           i64 ldb, double* R, i64 ldr) {
    double err = 0;
    for (i64 j = 0; j < nrhs; ++j)
        for (i64 i = 0; i < n; ++i) {
            double value = B[i + j * ldb], den = std::fabs(B[i + j * ldb]);
            for (i64 k = 0; k < n; ++k) {
                double a = A[i + k * lda], x = X[k + j * ldx];
                double prod = a * x; value = value - prod;
                double absprod = std::fabs(a) * std::fabs(x); den = den + absprod;
            }
            R[i + j * ldr] = value;
            double rel = !std::isfinite(den) ? INFINITY : den == 0 ? (value == 0 ? 0.0 : INFINITY) : std::fabs(value) / den;
            if (std::isnan(rel) || rel > err) err = rel;
        }
    return err;
}

// syrk!/herk! on one triangle (src/blas/level3.jl:546-783) for real data:
// C[i,j] <- alpha * sum_p op(A)[i,p] op(A)[j,p] + beta * C[i,j] on the
// triangle, by the small-case formula of _triangle_leaf! (:555-568): the
// products accumulated in order of p with muladd, then alpha, then beta * C
// (not read when beta == 0); alpha == 0 or k == 0 scales the triangle
// (_triangle_scale!, :546-553).
template <class T>
int syrk(char uplo, char trans, i64 n, i64 k, T alpha, const T* A, i64 lda, T beta, T* C, i64 ldc) {
    const bool upper = uplo == 'U' || uplo == 'u';
    if (!upper && !(uplo == 'L' || uplo == 'l')) return -1;
    const bool tr = trans == 'T' || trans == 't' || trans == 'C' || trans == 'c';
    if (!tr && !(trans == 'N' || trans == 'n')) return -2;
    if (n < 0) return -3;
    if (k < 0) return -4;
    if (lda < std::max<i64>(1, tr ? k : n)) return -7;
    if (ldc < std::max<i64>(1, n)) return -10;
    auto a = [&](i64 i, i64 p) { return tr ? A[p + i * lda] : A[i + p * lda]; };
    for (i64 j = 0; j < n; ++j)
        for (i64 i = upper ? 0 : j; i < (upper ? j + 1 : n); ++i) {
            if (alpha == T(0) || k == 0) {
                C[i + j * ldc] = beta == T(0) ? T(0) : beta * C[i + j * ldc];
                continue;
            }
            T acc = 0;
            for (i64 p = 0; p < k; ++p) acc = std::fma(a(i, p), a(j, p), acc);
            T v = alpha * acc;
            if (beta != T(0)) v += beta * C[i + j * ldc];
            C[i + j * ldc] = v;
        }
    return 0;
}

// _cholesky_recursive! (src/lapack/factorizations.jl:263-286): leaves of at
// most `leaf` columns right-looking (d = sqrt(a_cc), l_ic = a_ic / d,
// a_ij = muladd(-l_ic, l_jc, a_ij)), otherwise L11, L21 = A21 L11^-T (trsm!
// side R, lower, trans C), A22 -= L21 L21^T (herk! lower), L22.  Returns 0
// or the global 1-based column of the first pivot that is not positive and
// finite.
template <class T>
i64 chol_rec(T* L, i64 ldl, i64 n, i64 offset, i64 leaf, int threads) {
    if (n <= leaf) {
        for (i64 col = 0; col < n; ++col) {
            T d = L[col + col * ldl];
            if (!(std::isfinite(d) && d > 0)) return offset + col + 1;
            d = std::sqrt(d);
            L[col + col * ldl] = d;
            for (i64 i = col + 1; i < n; ++i) L[i + col * ldl] /= d;
            for (i64 j = col + 1; j < n; ++j) {
                const T c = L[j + col * ldl];
                for (i64 i = j; i < n; ++i) L[i + j * ldl] = std::fma(-L[i + col * ldl], c, L[i + j * ldl]);
            }
        }
        return 0;
    }
    const i64 h = n / 2;
    i64 info = chol_rec(L, ldl, h, offset, leaf, threads);
    if (info) return info;
    trsm<T>('R', 'L', 'C', 'N', n - h, h, T(1), L, ldl, L + h, ldl, 8, threads);
    syrk<T>('L', 'N', n - h, h, T(-1), L + h, ldl, T(1), L + h + h * ldl, ldl);
    return chol_rec(L + h + h * ldl, ldl, n - h, offset + h, leaf, threads);
}

}  // namespace

extern "C" {

// native_inverse (src/lapack/dense_utilities.jl:14-36) from the factors:
// X = inv(U) by the column formula of _trtri!'s leaf (:94-105, applied to
// the whole triangle: the recursion's TRMM blocking only reorders sums),
// X <- X inv(L) (trsm! side R, lower, unit), columns placed by the
// permutation: result[:, p[k]] = X[:, k].  Returns 0 or the 1-based index
// of a zero pivot (SingularException).
int oxo_dgetri(i64 n, const double* LU, i64 lda, const i64* ipiv, double* out, i64 ldo, int threads) {
    if (n < 0) return -1;
    for (i64 j = 0; j < n; ++j) if (LU[j + j * lda] == 0.0) return int(j + 1);
    std::vector<double> X(size_t(n) * n, 0.0);
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i <= j; ++i) X[i + j * n] = LU[i + j * lda];
    for (i64 j = 0; j < n; ++j) {
        const double d = (X[j + j * n] = 1.0 / X[j + j * n]);
        for (i64 i = 0; i < j; ++i) {
            double sacc = X[i + i * n] * X[i + j * n];
            for (i64 k = i + 1; k < j; ++k) sacc += X[i + k * n] * X[k + j * n];
            X[i + j * n] = -sacc * d;
        }
    }
    if (n > 1) trsm<double>('R', 'L', 'N', 'U', n, n, 1.0, LU, lda, X.data(), n, 8, threads);
    std::vector<i64> perm(n);
    for (i64 i = 0; i < n; ++i) perm[i] = i;
    for (i64 j = 0; j < n; ++j) std::swap(perm[j], perm[ipiv[j] - 1]);
    for (i64 k = 0; k < n; ++k) for (i64 i = 0; i < n; ++i) out[i + perm[k] * ldo] = X[i + k * n];
    return 0;
}

// LinearAlgebra.logabsdet(F::LU) (src/api/linalg_api.jl:185-191): sum of
// log|u_ii| in order, sign = prod sign(u_ii), negated per interchange.
void oxo_dlogabsdet(i64 n, const double* LU, i64 lda, const i64* ipiv, double* logabs, double* sign) {
    double d = 0, sg = 1;
    for (i64 i = 0; i < n; ++i) {
        const double u = LU[i + i * lda];
        d += std::log(std::fabs(u));
        sg *= u > 0 ? 1.0 : u < 0 ? -1.0 : 0.0;
        if (ipiv[i] != i + 1) sg = -sg;
    }
    *logabs = d;
    *sign = sg == 0 ? 0.0 : sg;   // no negative zero
}

// backward_error (src/lapack/dense_utilities.jl:214-228): R = B - A X and the
// largest componentwise ratio (sums in order of k).
double oxo_dbackward_error(i64 m, i64 n, i64 nrhs, const double* A, i64 lda, const double* X, i64 ldx,
                           const double* B, i64 ldb, double* R, i64 ldr) {
    double err = 0;
    for (i64 j = 0; j < nrhs; ++j)
        for (i64 i = 0; i < m; ++i) {
            double value = 0, den = std::fabs(B[i + j * ldb]);
            for (i64 k = 0; k < n; ++k) {
                value += A[i + k * lda] * X[k + j * ldx];
                den += std::fabs(A[i + k * lda]) * std::fabs(X[k + j * ldx]);
            }
            const double r = B[i + j * ldb] - value;
            R[i + j * ldr] = r;
            const double rel = !std::isfinite(den) ? INFINITY : den == 0 ? (r == 0 ? 0.0 : INFINITY) : std::fabs(r) / den;
            if (std::isnan(rel) || std::isnan(err)) err = NAN; else err = std::max(err, rel);
        }
    return err;
}

// inverse_norm_estimate (src/lapack/expert_refinement.jl:45-60) with the LU
// factors: solve(F, x) and solve_transposed(F, signs).
double oxo_dinverse_norm_estimate(i64 n, const double* LU, i64 lda, const i64* ipiv, int maxiter) {
    if (n == 0) return 0.0;
    std::vector<double> x(n, 1.0 / double(n)), y(n), z(n);
    double estimate = 0;
    i64 previous = -1;
    for (int it = 0; it < maxiter; ++it) {
        y = x;
        getrs<double>('N', n, 1, LU, lda, ipiv, y.data(), n, 1, 1);
        double next = 0;
        for (i64 i = 0; i < n; ++i) next += std::fabs(y[i]);
        estimate = std::max(estimate, next);
        for (i64 i = 0; i < n; ++i) z[i] = y[i] == 0 ? 1.0 : (y[i] > 0 ? 1.0 : -1.0);
        getrs<double>('C', n, 1, LU, lda, ipiv, z.data(), n, 1, 1);
        i64 index = 0;
        double zmax = -1, dot = 0;
        for (i64 i = 0; i < n; ++i) {
            if (std::fabs(z[i]) > zmax) { zmax = std::fabs(z[i]); index = i; }
            dot += z[i] * x[i];
        }
        if (zmax <= dot) break;
        if (index == previous) break;
        std::fill(x.begin(), x.end(), 0.0);
        x[index] = 1.0;
        previous = index;
    }
    for (i64 i = 0; i < n; ++i) x[i] = ((i % 2 == 0) ? 1.0 : -1.0) * (1.0 + double(i) / double(std::max<i64>(1, n - 1)));
    getrs<double>('N', n, 1, LU, lda, ipiv, x.data(), n, 1, 1);
    double alt = 0;
    for (i64 i = 0; i < n; ++i) alt += std::fabs(x[i]);
    return std::max(estimate, 2.0 * alt / (3.0 * double(n)));
}

// refine_solve (src/lapack/expert_refinement.jl:89-104).  out: [iterations,
// converged, backward_error, residual_norm, rcond_estimate, forward_error_estimate].
int oxo_drefine_solve(i64 n, i64 nrhs, const double* A, i64 lda, const double* LU, i64 ldlu, const i64* ipiv,
                      const double* B, i64 ldb, double* X, i64 ldx, int maxiter, double rtol, double* out) {
    if (n < 0 || nrhs < 0 || maxiter < 0 || !(rtol >= 0)) return -1;
    const bool vec = nrhs == 1;
    for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] = B[i + j * ldb];
    getrs<double>('N', n, nrhs, LU, ldlu, ipiv, X, ldx, vec, 1);
    std::vector<double> R(size_t(n) * nrhs), C(size_t(n) * nrhs);
    double err = INFINITY;
    int it = 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        err = oxo_dbackward_error(n, n, nrhs, A, lda, X, ldx, B, ldb, R.data(), n);
        it = iteration;
        if (err <= rtol || iteration == maxiter) break;
        C = R;
        getrs<double>('N', n, nrhs, LU, ldlu, ipiv, C.data(), n, vec, 1);
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] += C[i + j * n];
    }
    const double inv_norm = oxo_dinverse_norm_estimate(n, LU, ldlu, ipiv, 8);
    auto norm1 = [&](const double* M, i64 ld, i64 cols) {
        double best = 0;
        for (i64 j = 0; j < cols; ++j) {
            double sacc = 0;
            for (i64 i = 0; i < n; ++i) sacc += std::fabs(M[i + j * ld]);
            best = std::max(best, sacc);
        }
        return best;
    };
    double rn = 0;
    for (double v : R) rn += v * v;
    rn = std::sqrt(rn);
    const double anorm = norm1(A, lda, n), condition = anorm * inv_norm;
    const double xnorm = norm1(X, ldx, nrhs);
    const double delta = xnorm == 0 ? (rn == 0 ? 0.0 : INFINITY) : inv_norm * norm1(R.data(), n, nrhs) / xnorm;
    out[0] = it; out[1] = err <= rtol; out[2] = err; out[3] = rn;
    out[4] = condition == 0 ? 1.0 : std::min(1.0, 1.0 / condition);
    out[5] = delta < 1 ? delta / (1 - delta) : INFINITY;
    return 0;
}

// expert_solve (src/lapack/advanced_factorizations.jl:591-645) with
// equilibration and partial pivoting.  out: [iterations, converged,
// backward_error]; rows, cols: the scales.  Returns 0, or -(i) for a zero row
// i (SingularException(i)), or the getrf info as a positive value.
int oxo_dexpert_solve(i64 n, i64 nrhs, const double* A, i64 lda, const double* B, i64 ldb, double* X, i64 ldx,
                      int maxiter, double rtol, int equil, double* rows, double* cols, double* out, int threads) {
    std::vector<double> S(size_t(n) * n);
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < n; ++i) S[i + j * n] = A[i + j * lda];
    for (i64 i = 0; i < n; ++i) { rows[i] = 1; cols[i] = 1; }
    if (equil) {
        for (i64 i = 0; i < n; ++i) {
            double v = 0;
            for (i64 j = 0; j < n; ++j) v = std::max(v, std::fabs(S[i + j * n]));
            if (v == 0) return -int(i + 1);
            rows[i] = v;
            for (i64 j = 0; j < n; ++j) S[i + j * n] /= v;
        }
        for (i64 j = 0; j < n; ++j) {
            double v = 0;
            for (i64 i = 0; i < n; ++i) v = std::max(v, std::fabs(S[i + j * n]));
            cols[j] = v == 0 ? 1.0 : v;
            for (i64 i = 0; i < n; ++i) S[i + j * n] /= cols[j];
        }
    }
    std::vector<i64> piv(std::max<i64>(n, 1));
    i64 info = 0;
    getrf<double>(n, n, S.data(), n, piv.data(), 128, 0, &info, threads);
    if (info) return int(info);
    const bool vec = nrhs == 1;
    std::vector<double> T(size_t(n) * nrhs), R(size_t(n) * nrhs);
    for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) T[i + j * n] = B[i + j * ldb] / rows[i];
    getrs<double>('N', n, nrhs, S.data(), n, piv.data(), T.data(), n, vec, threads);
    for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] = T[i + j * n] / cols[i];
    double err = INFINITY;
    int it = 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        err = oxo_dbackward_error(n, n, nrhs, A, lda, X, ldx, B, ldb, R.data(), n);
        it = iteration;
        if (err <= rtol || iteration == maxiter) break;
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) T[i + j * n] = R[i + j * n] / rows[i];
        getrs<double>('N', n, nrhs, S.data(), n, piv.data(), T.data(), n, vec, threads);
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] += T[i + j * n] / cols[i];
    }
    out[0] = it; out[1] = err <= rtol; out[2] = err;
    return 0;
}

int oxo_dsyrk(char uplo, char trans, i64 n, i64 k, double alpha, const double* A, i64 lda, double beta, double* C,
              i64 ldc) {
    return syrk<double>(uplo, trans, n, k, alpha, A, lda, beta, C, ldc);
}

// _cholesky_factor_info!(L; blocksize, check=false) (factorizations.jl:246-261):
// the lower triangle of L in place; *info = 0 or the failing column; on
// success the strict upper triangle is zeroed.
int oxo_dpotrf(i64 n, double* L, i64 ldl, i64 leaf, i64* info, int threads) {
    if (n < 0) return -1;
    if (ldl < std::max<i64>(1, n)) return -3;
    if (leaf <= 0) return -4;
    *info = chol_rec<double>(L, ldl, n, 0, leaf, threads);
    if (*info == 0)
        for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < j; ++i) L[i + j * ldl] = 0.0;
    return 0;
}

// solve(F::CholeskyFactor, B) (factorizations.jl:288-300): L then L^T.
int oxo_dpotrs(i64 n, i64 nrhs, const double* L, i64 ldl, double* B, i64 ldb, int vector_rhs, int threads) {
    if (n < 0 || nrhs < 0) return -1;
    if (vector_rhs) {
        int r = trsv<double>('L', 'N', 'N', n, L, ldl, B);
        if (r) return r;
        return trsv<double>('L', 'C', 'N', n, L, ldl, B);
    }
    int r = trsm<double>('L', 'L', 'N', 'N', n, nrhs, 1.0, L, ldl, B, ldb, 8, threads);
    if (r) return r;
    return trsm<double>('L', 'L', 'C', 'N', n, nrhs, 1.0, L, ldl, B, ldb, 8, threads);
}

int oxo_dgemm(char ta, char tb, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
              double beta, double* C, i64 ldc, int threads) {
    return gemm<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, threads);
}
int oxo_sgemm(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda, const float* B, i64 ldb,
              float beta, float* C, i64 ldc, int threads) {
    return gemm<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, threads);
}
int oxo_dtrsm(char side, char uplo, char trans, char diag, i64 m, i64 n, double alpha, const double* A, i64 lda,
              double* B, i64 ldb, i64 bs, int threads) {
    return trsm<double>(side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bs, threads);
}
int oxo_dtrsv(char uplo, char trans, char diag, i64 n, const double* A, i64 lda, double* x) {
    return trsv<double>(uplo, trans, diag, n, A, lda, x);
}
int oxo_dgetrf(i64 m, i64 n, double* A, i64 lda, i64* ipiv, i64 nb, int check, i64* info, int threads) {
    return getrf<double>(m, n, A, lda, ipiv, nb, check, info, threads);
}
int oxo_sgetrf(i64 m, i64 n, float* A, i64 lda, i64* ipiv, i64 nb, int check, i64* info, int threads) {
    return getrf<float>(m, n, A, lda, ipiv, nb, check, info, threads);
}
// the same with tournament-pivoting leaves (tp_leaf)
int oxo_sgetrf_tp(i64 m, i64 n, float* A, i64 lda, i64* ipiv, i64 nb, int check, i64* info, int threads) {
    return getrf<float>(m, n, A, lda, ipiv, nb, check, info, threads, true);
}
// _lu_swap!(F, pivots, k1, k2, cols) (src/lapack/factorizations.jl:80-88) on
// columns 0..ncols-1; inverse != 0 applies the steps in reverse order as
// _rowswap!(X, pivots, true) (src/api/linalg_api.jl:100-110).  k1, k2 1-based.
int oxo_dlaswp(i64 ncols, double* A, i64 lda, i64 k1, i64 k2, const i64* ipiv, int inverse) {
    if (ncols < 0) return -1;
    if (lda < 1) return -3;
    if (k1 < 1) return -4;
    if (k2 < k1 - 1) return -5;
    for (i64 col = 0; col < ncols; ++col)
        for (i64 t = 0; t <= k2 - k1; ++t) {
            const i64 j = inverse ? k2 - 1 - t : k1 - 1 + t;
            const i64 p = ipiv[j] - 1;
            if (p < 0 || p >= lda) return -6;
            if (p != j) std::swap(A[j + col * lda], A[p + col * lda]);
        }
    return 0;
}

int oxo_dgetrs(char trans, i64 n, i64 nrhs, const double* LU, i64 lda, const i64* ipiv, double* B, i64 ldb,
               int vector_rhs, int threads) {
    return getrs<double>(trans, n, nrhs, LU, lda, ipiv, B, ldb, vector_rhs, threads);
}
int oxo_sgetrs(char trans, i64 n, i64 nrhs, const float* LU, i64 lda, const i64* ipiv, float* B, i64 ldb,
               int vector_rhs, int threads) {
    return getrs<float>(trans, n, nrhs, LU, lda, ipiv, B, ldb, vector_rhs, threads);
}
void oxo_dequilibrate(i64 m, i64 n, double* A, i64 lda, double* rows, double* cols, int radix) {
    equilibrate<double>(m, n, A, lda, rows, cols, radix);
}
double oxo_refinement_residual(i64 n, i64 nrhs, const double* A, i64 lda, const double* X, i64 ldx, const double* B,
                               i64 ldb, double* R, i64 ldr) {
    return refinement_residual(n, nrhs, A, lda, X, ldx, B, ldb, R, ldr);
}

// mixed_solve(A, b) for Float64 with Float32 low factors
// (src/lapack/factorizations.jl:425-427 -> src/lapack/extended_lapack.jl:175-226,
// correction step :114-131).  Outputs: X (n x nrhs), iterations, converged,
// backward error, fallback flag, reason code (0 converged, 1 iteration limit,
// 2 non-finite, 3 working-precision fallback).  Returns <0 on invalid input.
int oxo_mixed_solve(i64 n, i64 nrhs, const double* A, i64 lda, const double* B, i64 ldb, double* X, i64 ldx,
                    int maxiter, double rtol, int threads, i64* iterations, int* converged, double* berr,
                    int* fallback, int* reason) {
    if (n < 0 || nrhs < 0 || maxiter < 0 || !std::isfinite(rtol) || rtol < 0) return -1;
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < n; ++i) if (!std::isfinite(A[i + j * lda])) return -2;
    for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) if (!std::isfinite(B[i + j * ldb])) return -2;
    std::vector<double> E(size_t(n) * n), rows(n), cols(n);
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < n; ++i) E[i + j * n] = A[i + j * lda];
    equilibrate<double>(n, n, E.data(), n, rows.data(), cols.data(), 1);
    bool used_fallback = false;
    std::vector<float> low(size_t(n) * n);
    std::vector<double> Fd;
    std::vector<i64> piv(std::max<i64>(n, 1));
    bool overflow = false;
    for (size_t t = 0; t < low.size(); ++t) { low[t] = float(E[t]); if (!std::isfinite(low[t])) overflow = true; }
    i64 info = 0;
    if (!overflow) getrf<float>(n, n, low.data(), n, piv.data(), 128, 1, &info, threads);
    if (overflow || info != 0) {
        used_fallback = true;
        Fd = E;
        getrf<double>(n, n, Fd.data(), n, piv.data(), 128, 1, &info, threads);
        if (info != 0) return -3;  // SingularException(info) from the working-precision factor
    }
    // _mixed_extended_correction: scale by rows, binary magnitude, low solve, scale back by cols
    auto correction = [&](const double* Rin, i64 ldr, double* out, i64 ldo) -> int {
        std::vector<double> scaled(size_t(n) * nrhs);
        double mag = 0;
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) {
            double s = Rin[i + j * ldr] / rows[i]; scaled[i + j * n] = s; mag = std::max(mag, std::fabs(s));
        }
        if (mag == 0) { for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = 0; return 0; }
        if (!std::isfinite(mag)) return -4;
        int e; std::frexp(mag, &e); mag = std::ldexp(1.0, e - 1);  // ldexp(1, exponent(mag))
        if (!used_fallback) {
            std::vector<float> rhs(size_t(n) * nrhs);
            for (size_t t = 0; t < rhs.size(); ++t) { rhs[t] = float(scaled[t] / mag); if (!std::isfinite(rhs[t])) return -4; }
            getrs<float>('N', n, nrhs, low.data(), n, piv.data(), rhs.data(), n, nrhs == 1, threads);
            for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = double(rhs[i + j * n]) * mag / cols[i];
        } else {
            std::vector<double> rhs(size_t(n) * nrhs);
            for (size_t t = 0; t < rhs.size(); ++t) rhs[t] = scaled[t] / mag;
            getrs<double>('N', n, nrhs, Fd.data(), n, piv.data(), rhs.data(), n, nrhs == 1, threads);
            for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = rhs[i + j * n] * mag / cols[i];
        }
        return 0;
    };
    if (correction(B, ldb, X, ldx)) return -4;
    std::vector<double> R(size_t(n) * nrhs), C(size_t(n) * nrhs);
    double err = INFINITY; i64 it = 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        err = refinement_residual(n, nrhs, A, lda, X, ldx, B, ldb, R.data(), n); it = iteration;
        if (err <= rtol || iteration == maxiter || !std::isfinite(err)) break;
        if (correction(R.data(), n, C.data(), n)) return -4;
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] += C[i + j * n];
    }
    bool finite = std::isfinite(err);
    for (i64 j = 0; j < nrhs && finite; ++j) for (i64 i = 0; i < n; ++i) if (!std::isfinite(X[i + j * ldx])) { finite = false; break; }
    *iterations = it; *converged = err <= rtol; *berr = err; *fallback = used_fallback;
    *reason = !finite ? 2 : used_fallback ? 3 : err <= rtol ? 0 : 1;
    return 0;
}

// mixed_solve(A, B; factorization=:cholesky) (src/lapack/extended_lapack.jl:175-226):
// scale = max |a_ij| (1 for a zero matrix), rows = scale, cols = 1; Float32
// Cholesky (chol_rec, leaf 64) of round(A / scale), Float64 on
// PosDefException; the correction step with L and L^T solves.  Returns <0 on
// invalid input, -3 when the Float64 factor fails too (PosDefException).
int oxo_mixed_solve_chol(i64 n, i64 nrhs, const double* A, i64 lda, const double* B, i64 ldb, double* X, i64 ldx,
                         int maxiter, double rtol, int threads, i64* iterations, int* converged, double* berr,
                         int* fallback, int* reason) {
    if (n < 0 || nrhs < 0 || maxiter < 0 || !std::isfinite(rtol) || rtol < 0) return -1;
    double scale = 0;
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < n; ++i) {
        if (!std::isfinite(A[i + j * lda])) return -2;
        scale = std::max(scale, std::fabs(A[i + j * lda]));
    }
    for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) if (!std::isfinite(B[i + j * ldb])) return -2;
    if (scale == 0) scale = 1;
    std::vector<double> E(size_t(n) * n);
    for (i64 j = 0; j < n; ++j) for (i64 i = 0; i < n; ++i) E[i + j * n] = A[i + j * lda] / scale;
    bool used_fallback = false, overflow = false;
    std::vector<float> low(size_t(n) * n);
    for (size_t t = 0; t < low.size(); ++t) { low[t] = float(E[t]); if (!std::isfinite(low[t])) overflow = true; }
    i64 info = overflow ? 1 : chol_rec<float>(low.data(), n, n, 0, 64, threads);
    std::vector<double> Fd;
    if (info != 0) {
        used_fallback = true;
        Fd = E;
        if (chol_rec<double>(Fd.data(), n, n, 0, 64, threads) != 0) return -3;
    }
    auto correction = [&](const double* Rin, i64 ldr, double* out, i64 ldo) -> int {
        std::vector<double> scaled(size_t(n) * nrhs);
        double mag = 0;
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) {
            double sc = Rin[i + j * ldr] / scale; scaled[i + j * n] = sc; mag = std::max(mag, std::fabs(sc));
        }
        if (mag == 0) { for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = 0; return 0; }
        if (!std::isfinite(mag)) return -4;
        int e; std::frexp(mag, &e); mag = std::ldexp(1.0, e - 1);
        if (!used_fallback) {
            std::vector<float> rhs(size_t(n) * nrhs);
            for (size_t t = 0; t < rhs.size(); ++t) { rhs[t] = float(scaled[t] / mag); if (!std::isfinite(rhs[t])) return -4; }
            trsm<float>('L', 'L', 'N', 'N', n, nrhs, 1.0f, low.data(), n, rhs.data(), n, 8, threads);
            trsm<float>('L', 'L', 'C', 'N', n, nrhs, 1.0f, low.data(), n, rhs.data(), n, 8, threads);
            for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = double(rhs[i + j * n]) * mag;
        } else {
            std::vector<double> rhs(size_t(n) * nrhs);
            for (size_t t = 0; t < rhs.size(); ++t) rhs[t] = scaled[t] / mag;
            trsm<double>('L', 'L', 'N', 'N', n, nrhs, 1.0, Fd.data(), n, rhs.data(), n, 8, threads);
            trsm<double>('L', 'L', 'C', 'N', n, nrhs, 1.0, Fd.data(), n, rhs.data(), n, 8, threads);
            for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) out[i + j * ldo] = rhs[i + j * n] * mag;
        }
        return 0;
    };
    if (correction(B, ldb, X, ldx)) return -4;
    std::vector<double> R(size_t(n) * nrhs), C(size_t(n) * nrhs);
    double err = INFINITY; i64 it = 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        err = refinement_residual(n, nrhs, A, lda, X, ldx, B, ldb, R.data(), n); it = iteration;
        if (err <= rtol || iteration == maxiter || !std::isfinite(err)) break;
        if (correction(R.data(), n, C.data(), n)) return -4;
        for (i64 j = 0; j < nrhs; ++j) for (i64 i = 0; i < n; ++i) X[i + j * ldx] += C[i + j * n];
    }
    bool finite = std::isfinite(err);
    *iterations = it; *converged = err <= rtol; *berr = err; *fallback = used_fallback;
    *reason = !finite ? 2 : used_fallback ? 3 : err <= rtol ? 0 : 1;
    return 0;
}

int oxo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
