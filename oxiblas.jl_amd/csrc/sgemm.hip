// FP32 GEMM on CDNA4 matrix cores (v_mfma_f32_16x16x4f32):
// C <- alpha*op(A)*op(B) + beta*C.  Used by the FP32 working precision of the
// mixed-precision LU (the panel's recursive updates and the FP32 solves), the
// Float32 path of gemm! (src/blas/gemm.jl:625-703; GemmConfig kc=896 for
// 4-byte types, :14).
//
// A 256-thread workgroup owns a 128 x 128, 64 x 64 or 32 x 32 tile (chosen
// per product so that the grid fills the CUs), each wave a quarter of it in
// blocks of 16 x 16, BK=16 slabs staged in LDS and double-buffered.  Measured lane layout of v_mfma_f32_16x16x4f32 on
// gfx950: A[l%16][l/16] at lane l, B[l/16][l%16], D item r of lane l = row
// 4*(l/16)+r, column l%16.
#include "oxi_internal.hpp"

#include <cstdlib>

typedef float f4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 16, NT = 256;

template <bool OUTER_CONTIG, int O>
__device__ __forceinline__ void load_slab(float (&r)[O * BK / NT], const float* __restrict__ X, i64 ldx, i64 o0,
                                          i64 k0, i64 outer, i64 K, int t) {
#pragma unroll
    for (int i = 0; i < O * BK / NT; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BK; o = t / BK + i * (NT / BK); }
        const i64 go = o0 + o, gk = k0 + kk;
        r[i] = (go < outer && gk < K) ? (OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx]) : 0.0f;
    }
}

template <bool OUTER_CONTIG, int O>
__device__ __forceinline__ void store_slab(float* __restrict__ S, const float (&r)[O * BK / NT], int t) {
#pragma unroll
    for (int i = 0; i < O * BK / NT; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BK; o = t / BK + i * (NT / BK); }
        S[kk * (O + 1) + o] = r[i];
    }
}

__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
    const int total = tiles_m * tiles_n, nx = 8;
    int logical = bid;
    if (total >= nx) {
        const int xcd = bid % nx, local = bid / nx, per = total / nx, rem = total % nx;
        logical = xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
    }
    const int G = 8, group_size = G * tiles_n, group = logical / group_size, first_m = group * G;
    const int gm = min(tiles_m - first_m, G), in = logical % group_size;
    tm = first_m + in % gm;
    tn = in / gm;
}

// A BM x BN tile (BM, BN in {32, 64, 128}); the four waves split it 2 x 2,
// each wave (BM/2) x (BN/2) as QI x QJ blocks of 16 x 16.
template <int BM, int BN, bool TA, bool TB, bool BETA0>
__global__ __launch_bounds__(NT) void sgemm_mfma_kernel(i64 M, i64 N, i64 K, float alpha, const float* __restrict__ A,
                                                        i64 lda, const float* __restrict__ B, i64 ldb, float beta,
                                                        float* __restrict__ C, i64 ldc, int tiles_m, int tiles_n) {
    constexpr int LDS_A = BM + 1, LDS_B = BN + 1, QI = BM / 32, QJ = BN / 32;
    __shared__ float As[2][BK * LDS_A];
    __shared__ float Bs[2][BK * LDS_B];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const i64 m0 = i64(tm) * BM, n0 = i64(tn) * BN;
    constexpr bool A_OC = !TA, B_OC = TB;

    f4v acc[QI][QJ];
#pragma unroll
    for (int i = 0; i < QI; ++i)
#pragma unroll
        for (int j = 0; j < QJ; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    float ra[BM * BK / NT], rb[BN * BK / NT];
    const int nslab = int((K + BK - 1) / BK);
    load_slab<A_OC, BM>(ra, A, lda, m0, 0, M, K, t);
    load_slab<B_OC, BN>(rb, B, ldb, n0, 0, N, K, t);
    store_slab<A_OC, BM>(As[0], ra, t);
    store_slab<B_OC, BN>(Bs[0], rb, t);
    __syncthreads();
    const int lrow = lane >> 4, lcol = lane & 15;
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) {
            load_slab<A_OC, BM>(ra, A, lda, m0, i64(s + 1) * BK, M, K, t);
            load_slab<B_OC, BN>(rb, B, ldb, n0, i64(s + 1) * BK, N, K, t);
        }
        const float* as = As[cur];
        const float* bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + lrow;
            float a[QI], b[QJ];
#pragma unroll
            for (int i = 0; i < QI; ++i) a[i] = as[kr * LDS_A + wm * (BM / 2) + i * 16 + lcol];
#pragma unroll
            for (int j = 0; j < QJ; ++j) b[j] = bs[kr * LDS_B + wn * (BN / 2) + j * 16 + lcol];
#pragma unroll
            for (int i = 0; i < QI; ++i)
#pragma unroll
                for (int j = 0; j < QJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nslab) {
            store_slab<A_OC, BM>(As[cur ^ 1], ra, t);
            store_slab<B_OC, BN>(Bs[cur ^ 1], rb, t);
        }
        __syncthreads();
    }
    // All loads of C are issued before the first store, without branches (an
    // element outside C reads the nearest element inside it and is not
    // stored): a load behind a bounds branch is waited for before the next
    // one is issued, one memory round trip per element.
    float cv[QI][QJ][4];
    if (!BETA0) {
#pragma unroll
        for (int i = 0; i < QI; ++i)
#pragma unroll
            for (int j = 0; j < QJ; ++j) {
                const i64 col = min(n0 + wn * (BN / 2) + j * 16 + lcol, N - 1);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    cv[i][j][r] = C[col * ldc + min(m0 + wm * (BM / 2) + i * 16 + lrow * 4 + r, M - 1)];
            }
    }
#pragma unroll
    for (int i = 0; i < QI; ++i)
#pragma unroll
        for (int j = 0; j < QJ; ++j) {
            const i64 col = n0 + wn * (BN / 2) + j * 16 + lcol;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 row = m0 + wm * (BM / 2) + i * 16 + lrow * 4 + r;
                const float v = alpha * acc[i][j][r];
                if (row < M && col < N) C[col * ldc + row] = BETA0 ? v : __builtin_fmaf(beta, cv[i][j][r], v);
            }
        }
}

// Large products: 128 x 128 tiles with BK = 32 slabs (67.6 KB of LDS, two
// workgroups per CU, at most 128 registers a lane), each wave 64 x 64 as 4 x 4
// blocks of 16 x 16.  Tiles inside C with K a multiple of 32 load their slabs
// without bounds checks and stage C through LDS for coalesced column passes
// (a wave covers 64 consecutive rows of a column); edge tiles check every
// element.  (The 128 x 128 form of sgemm_mfma_kernel above needs 260
// registers, one workgroup per CU, and reads C sixteen columns at a time.)
constexpr int BKL = 32, LDL = 129;

template <bool OUTER_CONTIG, int O, bool EDGE>
__device__ __forceinline__ void load_slab_l(float (&r)[O * BKL / NT], const float* __restrict__ X, i64 ldx, i64 o0,
                                            i64 k0, i64 outer, i64 K, int t) {
#pragma unroll
    for (int i = 0; i < O * BKL / NT; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BKL; o = t / BKL + i * (NT / BKL); }
        const i64 go = o0 + o, gk = k0 + kk;
        if (EDGE) r[i] = (go < outer && gk < K) ? (OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx]) : 0.0f;
        else r[i] = OUTER_CONTIG ? X[go + gk * ldx] : X[gk + go * ldx];
    }
}
template <bool OUTER_CONTIG, int O>
__device__ __forceinline__ void store_slab_l(float* __restrict__ S, const float (&r)[O * BKL / NT], int t) {
#pragma unroll
    for (int i = 0; i < O * BKL / NT; ++i) {
        int o, kk;
        if (OUTER_CONTIG) { o = t % O; kk = t / O + i * (NT / O); }
        else { kk = t % BKL; o = t / BKL + i * (NT / BKL); }
        S[kk * LDL + o] = r[i];
    }
}

template <bool EDGE, bool TA, bool TB, bool BETA0>
__device__ __forceinline__ void sgemm_l_tile(float* lds, i64 M, i64 N, i64 K, float alpha, const float* __restrict__ A,
                                             i64 lda, const float* __restrict__ B, i64 ldb, float beta,
                                             float* __restrict__ C, i64 ldc, i64 m0, i64 n0) {
    constexpr bool A_OC = !TA, B_OC = TB;
    float* As[2] = {lds, lds + BKL * LDL};
    float* Bs[2] = {lds + 2 * BKL * LDL, lds + 3 * BKL * LDL};
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane >> 4, lcol = lane & 15;
    f4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
    float ra[128 * BKL / NT], rb[128 * BKL / NT];
    const int nslab = int((K + BKL - 1) / BKL);
    load_slab_l<A_OC, 128, EDGE>(ra, A, lda, m0, 0, M, K, t);
    load_slab_l<B_OC, 128, EDGE>(rb, B, ldb, n0, 0, N, K, t);
    store_slab_l<A_OC, 128>(As[0], ra, t);
    store_slab_l<B_OC, 128>(Bs[0], rb, t);
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) {
            load_slab_l<A_OC, 128, EDGE>(ra, A, lda, m0, i64(s + 1) * BKL, M, K, t);
            load_slab_l<B_OC, 128, EDGE>(rb, B, ldb, n0, i64(s + 1) * BKL, N, K, t);
        }
        const float* as = As[cur];
        const float* bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BKL / 4; ++ks) {
            const int kr = ks * 4 + lrow;
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = as[kr * LDL + wm * 64 + i * 16 + lcol];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = bs[kr * LDL + wn * 64 + j * 16 + lcol];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (s + 1 < nslab) {
            store_slab_l<A_OC, 128>(As[cur ^ 1], ra, t);
            store_slab_l<B_OC, 128>(Bs[cur ^ 1], rb, t);
        }
        __syncthreads();
    }
    if (!EDGE) {
        // C in two halves of 64 columns staged through LDS (128 x 64, column
        // stride 129) by the two waves that hold them; then a thread per row
        // of the tile and every other column, all its loads before its stores
        float* E = lds;
        const int er = t & 127, ec0 = t >> 7;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (wn == h) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int r = 0; r < 4; ++r) E[(j * 16 + lcol) * LDL + wm * 64 + i * 16 + 4 * lrow + r] = acc[i][j][r];
            }
            __syncthreads();
            float* cc = C + (n0 + h * 64 + ec0) * ldc + m0 + er;
            float cv[32];
            if (!BETA0) {
#pragma unroll
                for (int k = 0; k < 32; ++k) cv[k] = cc[i64(2 * k) * ldc];
            }
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float v = alpha * E[(ec0 + 2 * k) * LDL + er];
                cc[i64(2 * k) * ldc] = BETA0 ? v : __builtin_fmaf(beta, cv[k], v);
            }
            if (h == 0) __syncthreads();
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const i64 col = n0 + wn * 64 + j * 16 + lcol;
        const i64 colc = min(col, N - 1);
        float cv[4][4];
        if (!BETA0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) cv[i][r] = C[colc * ldc + min(m0 + wm * 64 + i * 16 + 4 * lrow + r, M - 1)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 row = m0 + wm * 64 + i * 16 + 4 * lrow + r;
                const float v = alpha * acc[i][j][r];
                if (row < M && col < N) C[col * ldc + row] = BETA0 ? v : __builtin_fmaf(beta, cv[i][r], v);
            }
    }
}

template <bool TA, bool TB, bool BETA0>
__global__ __launch_bounds__(NT, 2) void sgemm_l_kernel(i64 M, i64 N, i64 K, float alpha, const float* __restrict__ A,
                                                        i64 lda, const float* __restrict__ B, i64 ldb, float beta,
                                                        float* __restrict__ C, i64 ldc, int tiles_m, int tiles_n) {
    __shared__ __align__(16) float lds[4 * BKL * LDL];
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const i64 m0 = i64(tm) * 128, n0 = i64(tn) * 128;
    if (m0 + 128 <= M && n0 + 128 <= N && K % BKL == 0)
        sgemm_l_tile<false, TA, TB, BETA0>(lds, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0);
    else
        sgemm_l_tile<true, TA, TB, BETA0>(lds, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0);
}

__global__ void sscale_kernel(i64 m, i64 n, float beta, float* __restrict__ C, i64 ldc) {
    const i64 total = m * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / m, i = idx - j * m;
        float& c = C[i + j * ldc];
        c = (beta == 0.0f) ? 0.0f : c * beta;
    }
}

template <int BM, int BN, bool TA, bool TB>
void launch_tiles(hipStream_t s, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda, const float* B, i64 ldb,
                  float beta, float* C, i64 ldc) {
    const int tiles_m = int(cdiv(m, BM)), tiles_n = int(cdiv(n, BN));
    dim3 grid(tiles_m * tiles_n), block(NT);
    if (beta == 0.0f)
        hipLaunchKernelGGL((sgemm_mfma_kernel<BM, BN, TA, TB, true>), grid, block, 0, s, m, n, k, alpha, A, lda, B,
                           ldb, beta, C, ldc, tiles_m, tiles_n);
    else
        hipLaunchKernelGGL((sgemm_mfma_kernel<BM, BN, TA, TB, false>), grid, block, 0, s, m, n, k, alpha, A, lda, B,
                           ldb, beta, C, ldc, tiles_m, tiles_n);
}

// Tile shape per product, as for the FP64 GEMM: 128 x 128 where that grid has
// at least 1.75 workgroups per CU, else 64 x 64 where its grid has at least
// 1.5, else 32 x 32 (the mixed LU's panel updates, look-ahead block and TRSM
// updates are products of 32..256 rows).
template <bool TA, bool TB>
void launch_variant(hipStream_t s, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda, const float* B, i64 ldb,
                    float beta, float* C, i64 ldc) {
    constexpr i64 NCU = 256;
    if (cdiv(m, 128) * cdiv(n, 128) >= NCU * 7 / 4) {
        static const bool old_tiles = getenv("OXIBLAS_SGEMM_OLD") != nullptr;   // tuning: the first 128 x 128 form
        if (old_tiles) return launch_tiles<128, 128, TA, TB>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
        const int tiles_m = int(cdiv(m, 128)), tiles_n = int(cdiv(n, 128));
        if (beta == 0.0f)
            hipLaunchKernelGGL((sgemm_l_kernel<TA, TB, true>), dim3(tiles_m * tiles_n), dim3(NT), 0, s, m, n, k, alpha,
                               A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n);
        else
            hipLaunchKernelGGL((sgemm_l_kernel<TA, TB, false>), dim3(tiles_m * tiles_n), dim3(NT), 0, s, m, n, k,
                               alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n);
        return;
    }
    if (cdiv(m, 64) * cdiv(n, 64) >= NCU * 3 / 2) return launch_tiles<64, 64, TA, TB>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return launch_tiles<32, 32, TA, TB>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

}  // namespace

int oxi_launch_sgemm(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, float alpha, const float* A,
                     i64 lda, const float* B, i64 ldb, float beta, float* C, i64 ldc) {
    if (m <= 0 || n <= 0) return OXI_OK;
    if (alpha == 0.0f || k <= 0) {
        if (beta == 1.0f) return OXI_OK;
        const int blocks = int(std::min<i64>(cdiv(m * n, 256), 4096));
        hipLaunchKernelGGL(sscale_kernel, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
    } else if (!ta && !tb) launch_variant<false, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (!ta && tb) launch_variant<false, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (ta && !tb) launch_variant<true, false>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else launch_variant<true, true>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "sgemm launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

extern "C" int oxi_sgemm(oxi_handle_t h, char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
                         const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
                         int64_t ldc) {
    if (!h) return -1;
    auto kind = [](char t) { return (t == 'N' || t == 'n') ? 0 : (t == 'T' || t == 't' || t == 'C' || t == 'c') ? 1 : -1; };
    const int ka = kind(transA), kb = kind(transB);
    if (ka < 0) return -2;
    if (kb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    if (lda < std::max<i64>(1, ka ? k : m)) return -9;
    if (ldb < std::max<i64>(1, kb ? n : k)) return -11;
    if (ldc < std::max<i64>(1, m)) return -14;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    return oxi_launch_sgemm(h, h->stream, ka, kb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
