// FP64 GEMM on CDNA4 matrix cores: C <- alpha*op(A)*op(B) + beta*C.
//
// Replaces gemm! (src/blas/gemm.jl:625-703) and its packed engine
// (_packed_region! :414-485, _simd_micro! src/blas/simd.jl:151-199).  The CPU
// engine packs mc x kc / kc x nc panels into cache-resident buffers; here the
// "packing" is the staging of BK-deep slabs of op(A) and op(B) into LDS with
// coalesced global loads, double-buffered so that the next slab is in flight
// while the matrix cores consume the current one.
//
// Work decomposition: a 256-thread workgroup owns a tile of C, its four
// waves split it 2 x 2.  Large products take 128 x 128 tiles on
// v_mfma_f64_16x16x4 (dgemm_m16_kernel); products whose 128 x 128 grid would
// leave CUs idle take 64 x 64 or 32 x 32 tiles on v_mfma_f64_4x4x4_4b
// (dgemm_mfma_kernel).  Both instructions reach the full FP64 rate on gfx950.  The epilogue applies alpha and beta as the
// reference's full tiles do: v = alpha*acc; C = beta == 0 ? v : fma(beta, C, v)
// (src/blas/simd.jl:190-196); beta == 0 never reads C.
#include "oxi_internal.hpp"

#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace {

constexpr int BK = 32, NT = 256;            // slab depth of the 4x4x4 kernel; threads per workgroup
constexpr int PAD = 1;                 // one double of padding per LDS row

// Load one BK-deep slab of an operand tile (O rows of op(X)) into registers.
//   OUTER_CONTIG: element (o, kk) at X[(o0+o) + (k0+kk)*ldx] (op(A) for A 'N',
//                 op(B) for B 'T'); otherwise at X[(k0+kk) + (o0+o)*ldx].
// Out-of-range elements are zero so partial tiles need no special casing in
// the MFMA loop.
template <bool OUTER_CONTIG, int O, int BK = 32>
__device__ __forceinline__ void load_slab(double (&r)[O * BK / NT], const double* __restrict__ X, i64 ldx, i64 o0,
                                          i64 k0, i64 outer, i64 K, int t) {
    constexpr int LOADS = O * BK / NT;
    if (o0 + O <= outer && k0 + BK <= K) {
        // interior slab: no bounds checks (the common case)
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            int o, kk;
            if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
            else { kk = t % BK; o = t / BK + i * (NT / BK); }
            const i64 go = o0 + o, gk = k0 + kk;
            r[i] = OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx];
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < LOADS; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BK; o = t / BK + i * (NT / BK); }
        i64 go = o0 + o, gk = k0 + kk;
        r[i] = (go < outer && gk < K) ? (OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx]) : 0.0;
    }
}

// The same without bounds checks, for tiles that lie inside the operands
// with K a multiple of the slab depth: no branch in the slab loop, whose
// merge with the guarded path would otherwise leave a wait for all loads
// between the two operands' loads.
template <bool OUTER_CONTIG, int O, int BK>
__device__ __forceinline__ void load_slab_inner(double (&r)[O * BK / NT], const double* __restrict__ X, i64 ldx,
                                                i64 o0, i64 k0, int t) {
    constexpr int LOADS = O * BK / NT;
#pragma unroll
    for (int i = 0; i < LOADS; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BK; o = t / BK + i * (NT / BK); }
        const i64 go = o0 + o, gk = k0 + kk;
        r[i] = OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx];
    }
}

template <bool OUTER_CONTIG, int O, int BK = 32>
__device__ __forceinline__ void store_slab(double* __restrict__ S, const double (&r)[O * BK / NT], int t) {
    constexpr int LOADS = O * BK / NT, LDS = O + PAD;
#pragma unroll
    for (int i = 0; i < LOADS; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BK; o = t / BK + i * (NT / BK); }
        S[kk * LDS + o] = r[i];
    }
}

// Map the linear workgroup id to a (tile_m, tile_n) pair.  Workgroups are
// dispatched round-robin over the 8 XCDs (id % 8); each XCD gets a contiguous
// range of the logical order, and the logical order walks groups of 8 tile
// rows column by column, so the tiles resident on one XCD share A row-panels
// and B column-panels in that XCD's L2.
__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
    const int total = tiles_m * tiles_n;
    const int nx = 8;
    int logical = bid;
    if (total >= nx) {
        int xcd = bid % nx, local = bid / nx, per = total / nx, rem = total % nx;
        logical = xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
    }
    const int G = 8;
    int group_size = G * tiles_n;
    int group = logical / group_size;
    int first_m = group * G;
    int gm = min(tiles_m - first_m, G);
    int in = logical % group_size;
    tm = first_m + in % gm;
    tn = in / gm;
}

// The same for a grid over the tiles of one triangle of a square tile grid
// (T tiles a side, T (T + 1) / 2 workgroups): each XCD gets a contiguous
// range of the lower triangle's tiles in row order (tile row tm holds tiles
// 0 .. tm), so every XCD has the same number of tiles; an upper triangle
// takes the transposed pairs.  (The rectangular grid with the tiles outside
// the triangle leaving at once left some XCDs with almost no tiles.)
__device__ __forceinline__ void tri_tile_of(int bid, int T, bool upper, int& tm, int& tn) {
    const int total = T * (T + 1) / 2;
    const int nx = 8;
    int logical = bid;
    if (total >= nx) {
        int xcd = bid % nx, local = bid / nx, per = total / nx, rem = total % nx;
        logical = xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
    }
    int r = int((__builtin_sqrt(8.0 * double(logical) + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > logical) --r;
    while ((r + 1) * (r + 2) / 2 <= logical) ++r;
    const int c = logical - r * (r + 1) / 2;
    tm = upper ? c : r;
    tn = upper ? r : c;
}

// Launch modes of the tile kernels.  PLAIN: one product.  STRIDED: product
// blockIdx.y of a batch whose operands sit at fixed strides.  PTRS: product
// blockIdx.y of a batch given by device arrays of operand pointers.  LOWER /
// UPPER: one product of which only the lower (row >= column) or upper
// triangle of C is computed and written (syrk!/herk!): tiles outside the
// triangle leave at once, tiles across the diagonal mask their stores.
enum GemmMode { PLAIN = 0, STRIDED = 1, PTRS = 2, LOWER = 3, UPPER = 4 };
struct BatchArg {
    i64 sa, sb, sc;                    // STRIDED: element strides between the products' operands
    const double* const* pa;           // PTRS: device arrays of operand pointers
    const double* const* pb;
    double* const* pc;
};
template <int MODE>
__device__ __forceinline__ void batch_operands(const BatchArg& ba, const double* __restrict__& A,
                                               const double* __restrict__& B, double* __restrict__& C) {
    if (MODE == STRIDED) {
        const i64 z = blockIdx.y;
        A += z * ba.sa; B += z * ba.sb; C += z * ba.sc;
    } else if (MODE == PTRS) {
        const int z = blockIdx.y;
        A = ba.pa[z]; B = ba.pb[z]; C = ba.pc[z];
    }
}
// false when the BM x BN tile at (m0, n0) has no element in the triangle
template <int MODE>
__device__ __forceinline__ bool tile_in_triangle(i64 m0, i64 n0, int BM, int BN) {
    if (MODE == LOWER) return m0 + BM - 1 >= n0;
    if (MODE == UPPER) return n0 + BN - 1 >= m0;
    return true;
}
template <int MODE>
__device__ __forceinline__ bool keep(i64 row, i64 col) {
    if (MODE == LOWER) return row >= col;
    if (MODE == UPPER) return row <= col;
    return true;
}
// The same for element (m0 + r, n0 + c) of a tile, from its offsets r, c in
// the tile and dn = n0 - m0 clamped to +-2^30 (32-bit compares; a clamped
// tile lies wholly on one side of the diagonal, and so do its elements)
__device__ __forceinline__ int diag_offset(i64 m0, i64 n0) {
    const i64 d = n0 - m0;
    return d > (i64(1) << 30) ? (1 << 30) : d < -(i64(1) << 30) ? -(1 << 30) : int(d);
}
template <int MODE>
__device__ __forceinline__ bool keep_local(int r, int c, int dn) {
    if (MODE == LOWER) return r - c >= dn;
    if (MODE == UPPER) return r - c <= dn;
    return true;
}

template <int BM, int BN, bool TA, bool TB, bool BETA0, int MODE = PLAIN>
__global__ __launch_bounds__(NT) void dgemm_mfma_kernel(i64 M, i64 N, i64 K, double alpha,
                                                        const double* __restrict__ A, i64 lda,
                                                        const double* __restrict__ B, i64 ldb, double beta,
                                                        double* __restrict__ C, i64 ldc, int tiles_m, int tiles_n,
                                                        BatchArg ba = BatchArg{}) {
    constexpr int LDS_A = BM + PAD, LDS_B = BN + PAD;
    constexpr int QM = BM / 16, QN = BN / 16;       // 4x4 blocks per lane group, down and across
    __shared__ double As[2][BK * LDS_A];
    __shared__ double Bs[2][BK * LDS_B];

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    if ((MODE == LOWER || MODE == UPPER) && BM == BN) tri_tile_of(blockIdx.x, tiles_m, MODE == UPPER, tm, tn);
    else tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const i64 m0 = i64(tm) * BM, n0 = i64(tn) * BN;
    if (!tile_in_triangle<MODE>(m0, n0, BM, BN)) return;
    batch_operands<MODE>(ba, A, B, C);

    // op(A) rows contiguous when A is 'N'; op(B) columns contiguous when B is 'T'
    constexpr bool A_OC = !TA, B_OC = TB;

    // v_mfma_f64_4x4x4_4b runs four independent 4x4x4 products, one per lane
    // group g = (lane % 16) / 4 (measured on gfx950): A_g[i][k] at lane
    // 16k + 4g + i, B_g[k][j] at lane 16k + 4g + j, D_g[i][j] at lane
    // 16i + 4g + j.  The four waves split the tile 2 x 2; each lane group
    // owns a (BM/4) x (BN/4) share of its wave's (BM/2) x (BN/2) tile: rows
    // 8q + 4(g&1) + i and columns 8p + 4(g>>1) + j, interleaved by 4 so that
    // the groups' LDS reads of one k row fall on consecutive addresses.
    double acc[QM][QN];
#pragma unroll
    for (int i = 0; i < QM; ++i)
#pragma unroll
        for (int j = 0; j < QN; ++j) acc[i][j] = 0.0;

    double ra[BM * BK / NT], rb[BN * BK / NT];
    const int nslab = int((K + BK - 1) / BK);
    load_slab<A_OC, BM>(ra, A, lda, m0, 0, M, K, t);
    load_slab<B_OC, BN>(rb, B, ldb, n0, 0, N, K, t);

    // D_g[i][j] at lane 16i + 4g + j: row 8q + 4(g&1) + i, column 8p + 4(g>>1) + j.
    const int kq = lane >> 4, grp = (lane >> 2) & 3, li = lane & 3;
    const int erow = wm * (BM / 2) + 4 * (grp & 1) + kq;
    const int ecol = wn * (BN / 2) + 4 * (grp >> 1) + li;
    // Tiles inside C (tiles up to 64 x 64) combine with C in a staged layout:
    // thread t owns row sr = t % BM of columns sc0 + SCS*k, a wave covering
    // BM consecutive rows of a column.  Those C values are loaded before the
    // K loop, so that they travel while the matrix cores run, and without
    // branches (an element outside C reads the nearest one inside; edge tiles
    // do not use them): a load behind a bounds branch is waited for at once.
    constexpr bool STAGED = BM * BN <= 64 * 64;
    constexpr int CPT = STAGED ? BM * BN / NT : 1, SCS = NT / BM;
    const int sr = t % BM, sc0 = t / BM;
    const bool inner = m0 + BM <= M && n0 + BN <= N;
    double cs[CPT];
    if (STAGED && !BETA0) {
#pragma unroll
        for (int k = 0; k < CPT; ++k)
            cs[k] = C[min(n0 + sc0 + SCS * k, N - 1) * ldc + min(m0 + sr, M - 1)];
    }
    store_slab<A_OC, BM>(As[0], ra, t);
    store_slab<B_OC, BN>(Bs[0], rb, t);
    __syncthreads();

    const int roff = wm * (BM / 2) + 4 * (grp & 1) + li;   // + 8q
    const int coff = wn * (BN / 2) + 4 * (grp >> 1) + li;  // + 8p
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) {
            load_slab<A_OC, BM>(ra, A, lda, m0, i64(s + 1) * BK, M, K, t);
            load_slab<B_OC, BN>(rb, B, ldb, n0, i64(s + 1) * BK, N, K, t);
        }
        const double* as = As[cur];
        const double* bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + kq;
            double a[QM], b[QN];
#pragma unroll
            for (int q = 0; q < QM; ++q) a[q] = as[kr * LDS_A + roff + 8 * q];
#pragma unroll
            for (int p = 0; p < QN; ++p) b[p] = bs[kr * LDS_B + coff + 8 * p];
#pragma unroll
            for (int q = 0; q < QM; ++q)
#pragma unroll
                for (int p = 0; p < QN; ++p)
                    acc[q][p] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[q], b[p], acc[q][p], 0, 0, 0);
        }
        if (s + 1 < nslab) {
            store_slab<A_OC, BM>(As[cur ^ 1], ra, t);
            store_slab<B_OC, BN>(Bs[cur ^ 1], rb, t);
        }
        __syncthreads();
    }

    if (STAGED && inner) {
        // the accumulators through LDS (the A slabs' space, free after the
        // last barrier: 2 * BK * (BM + 1) >= BN * (BM + 1) doubles), then
        // one coalesced pass over C
        static_assert(!STAGED || 2 * BK >= BN, "staging space");
        double* E = &As[0][0];
#pragma unroll
        for (int p = 0; p < QN; ++p)
#pragma unroll
            for (int q = 0; q < QM; ++q) E[(ecol + 8 * p) * LDS_A + erow + 8 * q] = acc[q][p];
        __syncthreads();
        double* cc = C + (n0 + sc0) * ldc + m0 + sr;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const double v = alpha * E[(sc0 + SCS * k) * LDS_A + sr];
            if (keep<MODE>(m0 + sr, n0 + sc0 + SCS * k))
                cc[i64(SCS * k) * ldc] = BETA0 ? v : __builtin_fma(beta, cs[k], v);
        }
        return;
    }
    // edge tiles: all loads of C (issued without branches, an
    // element outside C reading the nearest one inside and not stored)
    // before the first store
    double cv[QM][QN];
    if (!BETA0) {
#pragma unroll
        for (int p = 0; p < QN; ++p) {
            const i64 col = min(n0 + ecol + 8 * p, N - 1);
#pragma unroll
            for (int q = 0; q < QM; ++q) cv[q][p] = C[col * ldc + min(m0 + erow + 8 * q, M - 1)];
        }
    }
#pragma unroll
    for (int p = 0; p < QN; ++p) {
        const i64 col = n0 + ecol + 8 * p;
#pragma unroll
        for (int q = 0; q < QM; ++q) {
            const i64 row = m0 + erow + 8 * q;
            const double v = alpha * acc[q][p];
            if (row < M && col < N && keep<MODE>(row, col))
                C[col * ldc + row] = BETA0 ? v : __builtin_fma(beta, cv[q][p], v);
        }
    }
}

// 128 x 128 tiles on v_mfma_f64_16x16x4: each wave owns 64 x 64 of C as 4 x 4
// blocks of 16 x 16 (four accumulators per lane each).  One instruction does
// four times the work of a 4x4x4_4b one for the same two LDS reads, and the
// 16-deep slabs (66 KB of LDS, 2 x 2 slabs in flight per CU with two
// workgroups) give each barrier twice the matrix work of a 64 x 64 tile.
// Lane layout (measured on gfx950): A[i][k] at lane 16k + i, B[k][j] at lane
// 16k + j, D[i][j] in accumulator i / 4 of lane 16(i % 4) + j.
typedef double d4v __attribute__((ext_vector_type(4)));
constexpr int BK16 = 16;

// One tile; EDGE: the tile crosses an edge of C or K is not a multiple of
// the slab depth (bounds-checked loads and stores).
template <bool EDGE, bool TA, bool TB, bool BETA0, int MODE>
__device__ __forceinline__ void m16_tile(double (*As)[BK16 * 129], double (*Bs)[BK16 * 129], double* E, i64 M, i64 N, i64 K,
                                         double alpha, const double* __restrict__ A, i64 lda,
                                         const double* __restrict__ B, i64 ldb, double beta, double* __restrict__ C,
                                         i64 ldc, i64 m0, i64 n0) {
    constexpr int BM = 128, BN = 128, LDS_A = BM + PAD, LDS_B = BN + PAD;
    static_assert(LDS_A == 129 && LDS_B == 129, "LDS row length");
    constexpr bool A_OC = !TA, B_OC = TB;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    d4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = d4v{0.0, 0.0, 0.0, 0.0};

    double ra[BM * BK16 / NT], rb[BN * BK16 / NT];
    auto load = [&](i64 k0) __attribute__((always_inline)) {
        if (EDGE) {
            load_slab<A_OC, BM, BK16>(ra, A, lda, m0, k0, M, K, t);
            load_slab<B_OC, BN, BK16>(rb, B, ldb, n0, k0, N, K, t);
        } else {
            load_slab_inner<A_OC, BM, BK16>(ra, A, lda, m0, k0, t);
            load_slab_inner<B_OC, BN, BK16>(rb, B, ldb, n0, k0, t);
        }
    };
    const int nslab = int((K + BK16 - 1) / BK16);
    load(0);
    store_slab<A_OC, BM, BK16>(As[0], ra, t);
    store_slab<B_OC, BN, BK16>(Bs[0], rb, t);
    __syncthreads();

    const int lk = lane >> 4, li = lane & 15;
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) load(i64(s + 1) * BK16);
        const double* as = As[cur];
        const double* bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BK16 / 4; ++ks) {
            const int kr = ks * 4 + lk;
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = as[kr * LDS_A + wm * 64 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = bs[kr * LDS_B + wn * 64 + j * 16 + li];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nslab) {
            store_slab<A_OC, BM, BK16>(As[cur ^ 1], ra, t);
            store_slab<B_OC, BN, BK16>(Bs[cur ^ 1], rb, t);
        }
        __syncthreads();
    }

    if (!EDGE) {
        // Interior tile: C in two halves of 64 columns, each staged through
        // LDS (128 x 64 doubles, row stride 129) by the two waves that hold
        // it, then combined with C by all threads, a wave covering 64
        // consecutive rows of one column (512 contiguous bytes) and every
        // thread issuing its 32 loads before its first store: one memory
        // round trip per half.
        constexpr int LDE = 129;
        const int er = t & 127, ec0 = t >> 7;     // row in the tile, first column (then every 2nd)
        const int dn = diag_offset(m0, n0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (wn == h) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int r = 0; r < 4; ++r) E[(j * 16 + li) * LDE + wm * 64 + i * 16 + lk + 4 * r] = acc[i][j][r];
            }
            __syncthreads();
            double* cc = C + (n0 + h * 64 + ec0) * ldc + m0 + er;
            // two batches of 16 columns, the second's loads in flight while
            // the first is stored
            double cv[2][16];
            auto load16 = [&](int b) __attribute__((always_inline)) {
#pragma unroll
                for (int k = 0; k < 16; ++k) cv[b][k] = cc[i64(2 * (16 * b + k)) * ldc];
            };
            if (!BETA0) { load16(0); load16(1); }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int c = 16 * b + k;
                    const double v = alpha * E[(ec0 + 2 * c) * LDE + er];
                    const bool in = keep_local<MODE>(er, h * 64 + ec0 + 2 * c, dn);
                    if (BETA0) {
                        if (in) cc[i64(2 * c) * ldc] = v;
                    } else {
                        // outside the triangle the element gets back the value just read
                        cc[i64(2 * c) * ldc] = in ? __builtin_fma(beta, cv[b][k], v) : cv[b][k];
                    }
                }
            if (h == 0) __syncthreads();
        }
        return;
    }
    // Edge tile: one block column at a time, its C loads (issued without
    // branches; an element outside C reads the nearest one inside and is
    // not stored) before its stores.
    const int dn = diag_offset(m0, n0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const i64 col = n0 + wn * 64 + j * 16 + li;
        const i64 colc = EDGE ? min(col, N - 1) : col;
        double cv[4][4];
        if (!BETA0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const i64 row = m0 + wm * 64 + i * 16 + lk + 4 * r;
                    cv[i][r] = C[colc * ldc + (EDGE ? min(row, M - 1) : row)];
                }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 row = m0 + wm * 64 + i * 16 + lk + 4 * r;
                const double v = alpha * acc[i][j][r];
                if ((!EDGE || (row < M && col < N)) &&
                    keep_local<MODE>(wm * 64 + i * 16 + lk + 4 * r, wn * 64 + j * 16 + li, dn))
                    C[col * ldc + row] = BETA0 ? v : __builtin_fma(beta, cv[i][r], v);
            }
    }
}

template <bool TA, bool TB, bool BETA0, int MODE = PLAIN>
__global__ __launch_bounds__(NT, 2) void dgemm_m16_kernel(i64 M, i64 N, i64 K, double alpha,
                                                          const double* __restrict__ A, i64 lda,
                                                          const double* __restrict__ B, i64 ldb, double beta,
                                                          double* __restrict__ C, i64 ldc, int tiles_m, int tiles_n,
                                                          BatchArg ba = BatchArg{}) {
    // the slabs of A and B, and after the K loop the staging of C
    __shared__ __align__(16) double lds[4 * BK16 * 129];
    double (*As)[BK16 * 129] = (double (*)[BK16 * 129])lds;
    double (*Bs)[BK16 * 129] = (double (*)[BK16 * 129])(lds + 2 * BK16 * 129);
    int tm, tn;
    if (MODE == LOWER || MODE == UPPER) tri_tile_of(blockIdx.x, tiles_m, MODE == UPPER, tm, tn);
    else tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const i64 m0 = i64(tm) * 128, n0 = i64(tn) * 128;
    if (!tile_in_triangle<MODE>(m0, n0, 128, 128)) return;
    batch_operands<MODE>(ba, A, B, C);
    if (m0 + 128 <= M && n0 + 128 <= N && K % BK16 == 0)
        m16_tile<false, TA, TB, BETA0, MODE>(As, Bs, lds, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0);
    else
        m16_tile<true, TA, TB, BETA0, MODE>(As, Bs, lds, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0);
}

__global__ void dscale_kernel(i64 m, i64 n, double beta, double* __restrict__ C, i64 ldc) {
    const i64 total = m * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / m, i = idx - j * m;
        double& c = C[i + j * ldc];
        c = (beta == 0.0) ? 0.0 : c * beta;
    }
}

// C[triangle] <- beta * C[triangle] (beta == 0 writes zeros without reading)
__global__ void dscale_tri_kernel(i64 n, double beta, double* __restrict__ C, i64 ldc, int upper) {
    const i64 total = n * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / n, i = idx - j * n;
        if (upper ? i > j : i < j) continue;
        double& c = C[i + j * ldc];
        c = (beta == 0.0) ? 0.0 : c * beta;
    }
}

template <int BM, int BN, bool TA, bool TB, int MODE = PLAIN>
void launch_tiles(hipStream_t s, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda, const double* B,
                  i64 ldb, double beta, double* C, i64 ldc, unsigned batch = 1, BatchArg ba = BatchArg{}) {
    const int tiles_m = int(cdiv(m, BM)), tiles_n = int(cdiv(n, BN));
    // one triangle of a square grid of square tiles: only its tiles launch
    const bool tri = (MODE == LOWER || MODE == UPPER) && BM == BN;
    dim3 grid(tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n, batch), block(NT);
    if (beta == 0.0)
        hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, TA, TB, true, MODE>), grid, block, 0, s, m, n, k, alpha, A, lda,
                           B, ldb, beta, C, ldc, tiles_m, tiles_n, ba);
    else
        hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, TA, TB, false, MODE>), grid, block, 0, s, m, n, k, alpha, A,
                           lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, ba);
}

template <bool TA, bool TB, int MODE = PLAIN>
void launch_m16(hipStream_t s, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda, const double* B, i64 ldb,
                double beta, double* C, i64 ldc, unsigned batch = 1, BatchArg ba = BatchArg{}) {
    const int tiles_m = int(cdiv(m, 128)), tiles_n = int(cdiv(n, 128));
    const bool tri = MODE == LOWER || MODE == UPPER;   // one triangle of a square grid: only its tiles launch
    dim3 grid(tri ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n, batch), block(NT);
    if (beta == 0.0)
        hipLaunchKernelGGL((dgemm_m16_kernel<TA, TB, true, MODE>), grid, block, 0, s, m, n, k, alpha, A, lda, B, ldb,
                           beta, C, ldc, tiles_m, tiles_n, ba);
    else
        hipLaunchKernelGGL((dgemm_m16_kernel<TA, TB, false, MODE>), grid, block, 0, s, m, n, k, alpha, A, lda, B, ldb,
                           beta, C, ldc, tiles_m, tiles_n, ba);
}

// Tile shape of C per workgroup (measured on MI355X over the LU's products
// and square ones, K = 32 .. 8192): 128 x 128 on the 16x16x4 instruction
// where its grid has at least 1.75 workgroups per CU (8192^3: 64 TF/s against
// 54 for 64 x 64; 512 x 15872 x 512, 496 tiles: 170 us against 206, but
// 512 x 12000 x 512, 376 tiles: 191 us against 166); else 64 x 64 tiles on the 4x4x4_4b instruction (66 KB of
// LDS, two workgroups per CU); 32 x 32 where the 64 x 64 grid has fewer than
// about 1.5 workgroups per CU (the LU's look-ahead block, its panel updates
// and the TRSM updates: products of 32..256 rows).  A side no longer than 32
// takes a 32-wide tile.  OXIBLAS_DGEMM_TILE (m16, 64x64, 32x32) fixes the
// choice (tuning); bm = bn = 16 stands for the 16x16x4 kernel.
constexpr int NCU = 256;
void choose_tile(i64 m, i64 n, int& bm, int& bn, i64 batch = 1) {
    static int forced = -1;
    if (forced < 0) {
        forced = 0;
        if (const char* e = getenv("OXIBLAS_DGEMM_TILE")) {
            int a = 0, b = 0;
            if (sscanf(e, "%dx%d", &a, &b) == 2 && a == b && (a == 32 || a == 64)) forced = a;
            if (!strcmp(e, "m16")) forced = 16;
        }
    }
    if (forced) { bm = bn = forced; return; }
    // a batch fills the chip with its products together
    if (cdiv(m, 128) * cdiv(n, 128) * batch >= NCU * 7 / 4) { bm = bn = 16; return; }
    bm = m <= 32 ? 32 : 64;
    bn = n <= 32 ? 32 : 64;
    if (cdiv(m, bm) * cdiv(n, bn) * batch < i64(NCU) * 3 / 2) bm = bn = 32;
}

// false when no kernel takes the chosen tile shape (nothing launched)
template <bool TA, bool TB, int MODE = PLAIN>
bool launch_variant(hipStream_t s, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda, const double* B,
                    i64 ldb, double beta, double* C, i64 ldc, unsigned batch = 1, BatchArg ba = BatchArg{}) {
    int bm, bn;
    choose_tile(m, n, bm, bn, batch);
    if (bm == 16) { launch_m16<TA, TB, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba); return true; }
#define OXI_TILE(X, Y)                                                                                              \
    if (bm == X && bn == Y) {                                                                                       \
        launch_tiles<X, Y, TA, TB, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba);              \
        return true;                                                                                                \
    }
    OXI_TILE(64, 64) OXI_TILE(32, 32) OXI_TILE(64, 32) OXI_TILE(32, 64)
#undef OXI_TILE
    return false;
}

template <int MODE>
bool launch_any(hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda,
                const double* B, i64 ldb, double beta, double* C, i64 ldc, unsigned batch, BatchArg ba) {
    if (!ta && !tb) return launch_variant<false, false, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba);
    if (!ta && tb) return launch_variant<false, true, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba);
    if (ta && !tb) return launch_variant<true, false, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba);
    return launch_variant<true, true, MODE>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, ba);
}

}  // namespace

int oxi_dgemm_right64_inplace(hipStream_t s, i64 rows, const double* V, double* X, i64 ldx) {
    // X[:, 0:64] <- X[:, 0:64] * V (V 64 x 64, ld 64) in place: 32 x 64 tiles,
    // so each workgroup owns all 64 columns of its rows, and it writes them
    // only after its last read of them (the epilogue follows the K loop)
    if (rows <= 0) return OXI_OK;
    const int tiles_m = int(cdiv(rows, 32));
    hipLaunchKernelGGL((dgemm_mfma_kernel<32, 64, false, false, true>), dim3(tiles_m), dim3(NT), 0, s, rows, i64(64),
                       i64(64), 1.0, X, ldx, V, i64(64), 0.0, X, ldx, tiles_m, 1);
    return hipGetLastError() == hipSuccess ? OXI_OK : OXI_ERR_HIP;
}

int oxi_dgemm_left64_inplace(hipStream_t s, int ta, i64 ncols, const double* V, double* X, i64 ldx) {
    // X[0:64, :] <- op(V) * X[0:64, :] (V 64 x 64, ld 64) in place: 64 x 32
    // tiles, so each workgroup owns all 64 rows of its columns and writes
    // them only after its last read of them (the epilogue follows the K loop)
    if (ncols <= 0) return OXI_OK;
    const int tiles_n = int(cdiv(ncols, 32));
    if (ta)
        hipLaunchKernelGGL((dgemm_mfma_kernel<64, 32, true, false, true>), dim3(tiles_n), dim3(NT), 0, s, i64(64),
                           ncols, i64(64), 1.0, V, i64(64), X, ldx, 0.0, X, ldx, 1, tiles_n);
    else
        hipLaunchKernelGGL((dgemm_mfma_kernel<64, 32, false, false, true>), dim3(tiles_n), dim3(NT), 0, s, i64(64),
                           ncols, i64(64), 1.0, V, i64(64), X, ldx, 0.0, X, ldx, 1, tiles_n);
    return hipGetLastError() == hipSuccess ? OXI_OK : OXI_ERR_HIP;
}

bool oxi_dgemm_large_tiles(i64 m, i64 n) {
    int bm, bn;
    choose_tile(m, n, bm, bn);
    return bm == 16;
}

int oxi_launch_dscale(hipStream_t s, i64 m, i64 n, double beta, double* C, i64 ldc) {
    if (m <= 0 || n <= 0 || beta == 1.0) return OXI_OK;
    i64 total = m * n;
    int blocks = int(std::min<i64>(cdiv(total, 256), 4096));
    hipLaunchKernelGGL(dscale_kernel, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
    return hipGetLastError() == hipSuccess ? OXI_OK : OXI_ERR_HIP;
}

int oxi_launch_dgemm(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                     const double* A, i64 lda, const double* B, i64 ldb, double beta, double* C, i64 ldc) {
    if (m <= 0 || n <= 0) return OXI_OK;
    if (alpha == 0.0 || k <= 0) return oxi_launch_dscale(s, m, n, beta, C, ldc);
    if (cdiv(m, 128) * cdiv(n, 128) > (i64(1) << 31) - 1)
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm: %lld x %lld has too many tiles", (long long)m,
                             (long long)n);
    bool launched;
    if (!ta && !tb) launched = launch_variant<false, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (!ta && tb) launched = launch_variant<false, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (ta && !tb) launched = launch_variant<true, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else launched = launch_variant<true, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (!launched) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm: no kernel for the chosen tile shape");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "dgemm launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

// A batch of equal-shaped products in one launch per 65535 products
// (gemm_strided_batched!, src/blas/batched.jl:80-100): grid.y is the product.
int oxi_launch_dgemm_strided(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                             const double* A, i64 lda, i64 sa, const double* B, i64 ldb, i64 sb, double beta,
                             double* C, i64 ldc, i64 sc, i64 batch) {
    if (m <= 0 || n <= 0 || batch <= 0) return OXI_OK;
    if (alpha == 0.0 || k <= 0) {
        for (i64 z = 0; z < batch; ++z) OXI_TRY(oxi_launch_dscale(s, m, n, beta, C + z * sc, ldc));
        return OXI_OK;
    }
    if (cdiv(m, 32) * cdiv(n, 32) > (i64(1) << 31) - 1)
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm batch: %lld x %lld has too many tiles", (long long)m,
                             (long long)n);
    for (i64 z0 = 0; z0 < batch; z0 += 65535) {
        const unsigned nb = unsigned(std::min<i64>(65535, batch - z0));
        BatchArg ba{sa, sb, sc, nullptr, nullptr, nullptr};
        if (!launch_any<STRIDED>(s, ta, tb, m, n, k, alpha, A + z0 * sa, lda, B + z0 * sb, ldb, beta, C + z0 * sc,
                                 ldc, nb, ba))
            return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm batch: no kernel for the chosen tile shape");
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "dgemm batch launch: %s", hipGetErrorString(e));
    }
    return OXI_OK;
}

// A batch given by device arrays of operand pointers (gemm_batched!,
// src/blas/batched.jl:49-78), all of one shape.
int oxi_launch_dgemm_ptrs(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                          const double* const* A, i64 lda, const double* const* B, i64 ldb, double beta,
                          double* const* C, i64 ldc, i64 batch) {
    if (m <= 0 || n <= 0 || batch <= 0) return OXI_OK;
    if (cdiv(m, 32) * cdiv(n, 32) > (i64(1) << 31) - 1)
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm batch: %lld x %lld has too many tiles", (long long)m,
                             (long long)n);
    for (i64 z0 = 0; z0 < batch; z0 += 65535) {
        const unsigned nb = unsigned(std::min<i64>(65535, batch - z0));
        BatchArg ba{0, 0, 0, A + z0, B + z0, C + z0};
        if (!launch_any<PTRS>(s, ta, tb, m, n, k, alpha, nullptr, lda, nullptr, ldb, beta, nullptr, ldc, nb, ba))
            return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "dgemm batch: no kernel for the chosen tile shape");
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "dgemm batch launch: %s", hipGetErrorString(e));
    }
    return OXI_OK;
}

// C[triangle] <- alpha * op(A) op(A)^T + beta * C[triangle], the other
// triangle untouched (syrk!/herk! for Float64, src/blas/level3.jl:546-783).
// trans 0: op(A) = A (n x k); 1: op(A) = A^T (A k x n).
int oxi_launch_dsyrk(oxi_handle_t h, hipStream_t s, int upper, int trans, i64 n, i64 k, double alpha,
                     const double* A, i64 lda, double beta, double* C, i64 ldc) {
    if (n <= 0) return OXI_OK;
    if (alpha == 0.0 || k <= 0) {
        if (beta == 1.0) return OXI_OK;
        const int blocks = int(std::min<i64>(cdiv(n * n, 256), 4096));
        hipLaunchKernelGGL(dscale_tri_kernel, dim3(blocks), dim3(256), 0, s, n, beta, C, ldc, upper);
        return hipGetLastError() == hipSuccess ? OXI_OK : oxi_set_error(h, OXI_ERR_HIP, "syrk scale launch");
    }
    if (cdiv(n, 32) * cdiv(n, 32) > (i64(1) << 31) - 1)
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "syrk: n = %lld has too many tiles", (long long)n);
    const int ta = trans ? 1 : 0, tb = trans ? 0 : 1;
    const bool ok = upper ? launch_any<UPPER>(s, ta, tb, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1, BatchArg{})
                          : launch_any<LOWER>(s, ta, tb, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1, BatchArg{});
    if (!ok) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "syrk: no kernel for the chosen tile shape");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "syrk launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

namespace {
int trans_kind(char t) {
    if (t == 'N' || t == 'n') return 0;
    if (t == 'T' || t == 't' || t == 'C' || t == 'c') return 1;
    return -1;
}
bool overlap(const void* a, size_t ab, const void* b, size_t bb) {
    const char* x = (const char*)a; const char* y = (const char*)b;
    return x < y + bb && y < x + ab;
}
size_t span_bytes(i64 rows, i64 cols, i64 ld) { return rows <= 0 || cols <= 0 ? 0 : size_t((cols - 1) * ld + rows) * 8; }
}  // namespace

extern "C" int oxi_dgemm(oxi_handle_t h, char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                         int64_t ldc) {
    if (!h) return -1;
    const int ka = trans_kind(transA), kb = trans_kind(transB);
    if (ka < 0) return -2;
    if (kb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    const i64 ar = ka ? k : m, ac = ka ? m : k, br = kb ? n : k, bc = kb ? k : n;
    if (lda < std::max<i64>(1, ar)) return -9;
    if (ldb < std::max<i64>(1, br)) return -11;
    if (ldc < std::max<i64>(1, m)) return -14;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    if (alpha == 0.0 || k == 0) {
        OXI_TRY(oxi_launch_dscale(s, m, n, beta, C, ldc));
        return OXI_OK;
    }
    // Overlapping operands are copied before C is written (src/blas/gemm.jl:682-688).
    const size_t cbytes = span_bytes(m, n, ldc);
    const bool ca = overlap(C, cbytes, A, span_bytes(ar, ac, lda));
    const bool cb = overlap(C, cbytes, B, span_bytes(br, bc, ldb));
    if (ca || cb) {
        size_t need = (ca ? size_t(ar) * ac * 8 : 0) + (cb ? size_t(br) * bc * 8 : 0) + 256;
        void* ws;
        OXI_TRY(oxi_workspace(h, need, &ws));
        double* p = (double*)ws;
        if (ca) {
            OXI_HIP(h, hipMemcpy2DAsync(p, ar * 8, A, lda * 8, ar * 8, ac, hipMemcpyDeviceToDevice, s));
            A = p; lda = ar; p += size_t(ar) * ac;
            p = (double*)(((uintptr_t)p + 255) & ~uintptr_t(255));
        }
        if (cb) {
            OXI_HIP(h, hipMemcpy2DAsync(p, br * 8, B, ldb * 8, br * 8, bc, hipMemcpyDeviceToDevice, s));
            B = p; ldb = br;
        }
    }
    return oxi_launch_dgemm(h, s, ka, kb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int oxi_dgemm_strided_batched(oxi_handle_t h, char transA, char transB, int64_t m, int64_t n, int64_t k,
                                         double alpha, const double* A, int64_t lda, int64_t strideA, const double* B,
                                         int64_t ldb, int64_t strideB, double beta, double* C, int64_t ldc,
                                         int64_t strideC, int64_t batch) {
    if (!h) return -1;
    const int ka = trans_kind(transA), kb = trans_kind(transB);
    if (ka < 0) return -2;
    if (kb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    const i64 ar = ka ? k : m, ac = ka ? m : k, br = kb ? n : k, bc = kb ? k : n;
    if (lda < std::max<i64>(1, ar)) return -9;
    if (ldb < std::max<i64>(1, br)) return -12;
    if (ldc < std::max<i64>(1, m)) return -16;
    if (batch < 0) return -18;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0 || batch == 0) return OXI_OK;
    // disjoint outputs (src/blas/batched.jl:104-117): consecutive products'
    // spans of C may not meet
    const size_t cspan = span_bytes(m, n, ldc);
    if (batch > 1 && (strideC < 0 || size_t(strideC) * 8 < cspan)) return -17;
    if (!C) return -15;
    const bool active = alpha != 0.0 && k > 0;
    if (active && (!A || !B)) return !A ? -8 : -11;
    if (active && batch > 1 && (strideA < 0 || strideB < 0)) return strideA < 0 ? -10 : -13;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    if (active) {
        // an operand overlapping the outputs is copied first (batched.jl:87-88)
        const size_t call = size_t(batch - 1) * size_t(strideC) * 8 + cspan;
        const size_t aall = size_t(batch - 1) * size_t(strideA) * 8 + span_bytes(ar, ac, lda);
        const size_t ball = size_t(batch - 1) * size_t(strideB) * 8 + span_bytes(br, bc, ldb);
        const bool ca = overlap(C, call, A, aall), cb = overlap(C, call, B, ball);
        if (ca || cb) {
            const size_t need = (ca ? ((aall + 255) & ~size_t(255)) : 0) + (cb ? ((ball + 255) & ~size_t(255)) : 0) + 256;
            void* ws;
            OXI_TRY(oxi_workspace(h, need, &ws));
            char* p = (char*)ws;
            if (ca) {
                OXI_HIP(h, hipMemcpyAsync(p, A, aall, hipMemcpyDeviceToDevice, s));
                A = (const double*)p; p += (aall + 255) & ~size_t(255);
            }
            if (cb) {
                OXI_HIP(h, hipMemcpyAsync(p, B, ball, hipMemcpyDeviceToDevice, s));
                B = (const double*)p;
            }
        }
    }
    return oxi_launch_dgemm_strided(h, s, ka, kb, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                    strideC, batch);
}

extern "C" int oxi_dgemm_batched(oxi_handle_t h, char transA, char transB, int64_t m, int64_t n, int64_t k,
                                 double alpha, const double* const* A, int64_t lda, const double* const* B,
                                 int64_t ldb, double beta, double* const* C, int64_t ldc, int64_t batch) {
    if (!h) return -1;
    const int ka = trans_kind(transA), kb = trans_kind(transB);
    if (ka < 0) return -2;
    if (kb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    const i64 ar = ka ? k : m, ac = ka ? m : k, br = kb ? n : k, bc = kb ? k : n;
    if (lda < std::max<i64>(1, ar)) return -9;
    if (ldb < std::max<i64>(1, br)) return -11;
    if (ldc < std::max<i64>(1, m)) return -14;
    if (batch < 0) return -15;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0 || batch == 0) return OXI_OK;
    if (!C) return -13;
    const bool active = alpha != 0.0 && k > 0;
    if (active && (!A || !B)) return !A ? -8 : -10;
    // the pointer arrays are host arrays of device pointers; every output
    // is disjoint from the others (src/blas/batched.jl:66-68)
    const size_t cspan = span_bytes(m, n, ldc), aspan = span_bytes(ar, ac, lda), bspan = span_bytes(br, bc, ldb);
    std::vector<std::pair<uintptr_t, i64>> cs(static_cast<size_t>(batch));
    for (i64 z = 0; z < batch; ++z) {
        if (!C[z]) return -13;
        if (active && (!A[z] || !B[z])) return !A[z] ? -8 : -10;
        cs[size_t(z)] = {uintptr_t(C[z]), z};
    }
    std::sort(cs.begin(), cs.end());
    for (size_t i = 1; i < cs.size(); ++i)
        if (cs[i].first < cs[i - 1].first + cspan) return oxi_set_error(h, -13, "gemm_batched: outputs overlap");
    auto hits_c = [&](uintptr_t p, size_t bytes) {
        // any output span [c, c + cspan) meeting [p, p + bytes): the last
        // output starting before p + bytes is the only candidate (the
        // outputs are disjoint and sorted)
        auto it = std::lower_bound(cs.begin(), cs.end(), std::make_pair(p + bytes, std::numeric_limits<i64>::min()));
        if (it == cs.begin()) return false;
        --it;
        return it->first + cspan > p;
    };
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    // operands that alias an output are copied first (batched.jl:58-65)
    std::vector<const double*> pa(A ? A : nullptr, A ? A + batch : nullptr), pb(B ? B : nullptr, B ? B + batch : nullptr);
    size_t copies = 0;
    if (active)
        for (i64 z = 0; z < batch; ++z) {
            if (hits_c(uintptr_t(pa[z]), aspan)) copies += (aspan + 255) & ~size_t(255);
            if (hits_c(uintptr_t(pb[z]), bspan)) copies += (bspan + 255) & ~size_t(255);
        }
    const size_t parr = (size_t(3 * batch) * 8 + 255) & ~size_t(255);
    void* ws;
    OXI_TRY(oxi_workspace(h, parr + copies + 256, &ws));
    char* snap = (char*)ws + parr;
    if (active)
        for (i64 z = 0; z < batch; ++z) {
            if (hits_c(uintptr_t(pa[z]), aspan)) {
                OXI_HIP(h, hipMemcpyAsync(snap, pa[z], aspan, hipMemcpyDeviceToDevice, s));
                pa[z] = (const double*)snap; snap += (aspan + 255) & ~size_t(255);
            }
            if (hits_c(uintptr_t(pb[z]), bspan)) {
                OXI_HIP(h, hipMemcpyAsync(snap, pb[z], bspan, hipMemcpyDeviceToDevice, s));
                pb[z] = (const double*)snap; snap += (bspan + 255) & ~size_t(255);
            }
        }
    if (!active) {
        for (i64 z = 0; z < batch; ++z) OXI_TRY(oxi_launch_dscale(s, m, n, beta, C[z], ldc));
        return OXI_OK;
    }
    std::vector<const void*> ptrs(size_t(3 * batch));
    for (i64 z = 0; z < batch; ++z) {
        ptrs[size_t(z)] = pa[z];
        ptrs[size_t(batch + z)] = pb[z];
        ptrs[size_t(2 * batch + z)] = C[z];
    }
    const double* const* dA = (const double* const*)ws;
    OXI_HIP(h, hipMemcpyAsync(ws, ptrs.data(), ptrs.size() * 8, hipMemcpyHostToDevice, s));
    OXI_TRY(oxi_launch_dgemm_ptrs(h, s, ka, kb, m, n, k, alpha, dA, lda, dA + batch, ldb, beta,
                                  (double* const*)(dA + 2 * batch), ldc, batch));
    // the host pointer array is read by the copy above before this call returns
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

extern "C" int oxi_dsyrk(oxi_handle_t h, char uplo, char trans, int64_t n, int64_t k, double alpha, const double* A,
                         int64_t lda, double beta, double* C, int64_t ldc) {
    if (!h) return -1;
    const bool upper = uplo == 'U' || uplo == 'u';
    if (!upper && !(uplo == 'L' || uplo == 'l')) return -2;
    const int tk = trans_kind(trans);
    if (tk < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (lda < std::max<i64>(1, tk ? k : n)) return -8;
    if (ldc < std::max<i64>(1, n)) return -11;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0) return OXI_OK;
    if (!C) return -10;
    if (alpha != 0.0 && k > 0 && !A) return -7;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    // A overlapping C is copied first (src/blas/level3.jl:737-750)
    const size_t abytes = span_bytes(tk ? k : n, tk ? n : k, lda);
    if (alpha != 0.0 && k > 0 && overlap(C, span_bytes(n, n, ldc), A, abytes)) {
        void* ws;
        OXI_TRY(oxi_workspace(h, abytes + 256, &ws));
        OXI_HIP(h, hipMemcpyAsync(ws, A, abytes, hipMemcpyDeviceToDevice, s));
        A = (const double*)ws;
    }
    return oxi_launch_dsyrk(h, s, upper, tk, n, k, alpha, A, lda, beta, C, ldc);
}
