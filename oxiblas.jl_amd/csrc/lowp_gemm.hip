// Low-precision trailing update of the mixed-precision LU:
//   C (FP32, M x N, column-major) <- C + alpha * (sa*sb) * A_lp * B_lp
// with A_lp / B_lp in BF16, FP16 or FP8 (OCP e4m3), FP32 accumulation on the
// CDNA4 matrix cores.  This is the extension BASELINE's mixed form asks for
// beyond the reference's mixed_solve (FP32 LU + FP64 refinement,
// src/lapack/extended_lapack.jl:175-226): the O(n^3) trailing products of the
// FP32 factorization run at low precision, the panel stays FP32.
//
// Operand storage (written by the conversion kernels below from the FP32
// working matrix): A_lp row-major [M][Kp] and B_lp as [N][Kp] (k contiguous in
// both), Kp a multiple of the slab depth, zero padded.  FP8 operands carry one
// power-of-two scale per 256-row block of A and per 256-column block of B
// (exact), applied in the epilogue.
//
// Measured lane layouts on gfx950 (probe: exact small-integer products):
// v_mfma_f32_16x16x32_{bf16,f16} and v_mfma_scale_f32_16x16x128_f8f6f4 pair A
// lane l with B lane l' when l/16 == l'/16 element by element (any common
// k-assignment works; here lane group g holds k = g*8..g*8+7, resp. g*32..),
// row i = l%16 of A, column j = l%16 of B; D item r of lane l is row
// 4*(l/16)+r, column l%16.
#include "oxi_internal.hpp"

#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int BM = 128, BN = 128, NT = 256;
constexpr int SLAB_BYTES = 128;          // bytes of k per row per slab (64 bf16/f16, 128 fp8)
constexpr int ROW_STRIDE = SLAB_BYTES + 16;  // padded LDS row (conflict-free 16-byte reads)

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

// One slab: 128 rows x 128 bytes for each operand, 16 bytes per thread-load,
// 4 loads per thread per operand.
__device__ __forceinline__ void load_rows(uint4 (&r)[4], const unsigned char* __restrict__ X, i64 ldx_bytes, i64 o0,
                                          i64 outer, i64 kbyte0, int t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = t / 8 + q * 32, chunk = t % 8;
        const i64 go = o0 + row;
        r[q] = (go < outer) ? *(const uint4*)(X + go * ldx_bytes + kbyte0 + chunk * 16) : make_uint4(0, 0, 0, 0);
    }
}
__device__ __forceinline__ void store_rows(unsigned char* __restrict__ S, const uint4 (&r)[4], int t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = t / 8 + q * 32, chunk = t % 8;
        *(uint4*)(S + row * ROW_STRIDE + chunk * 16) = r[q];
    }
}

// PREC: 0 = BF16, 1 = FP16, 2 = FP8 e4m3
// vec: ldc % 4 == 0 and C 16-byte aligned.  A tile inside C then reads and
// writes each lane's four rows of a column (D items r = 0..3, consecutive
// rows) as one 16-byte streaming access: at k = nb the C read and write are
// the kernel's HBM traffic.  (Loading C before the operand loads, to overlap
// it with the products, measured no faster and spills.)
template <int PREC>
__global__ __launch_bounds__(NT, 2) void lowp_update_kernel(i64 M, i64 N, i64 Kp, float alpha,
                                                            const unsigned char* __restrict__ A,
                                                            const unsigned char* __restrict__ B,
                                                            const float* __restrict__ sa, const float* __restrict__ sb,
                                                            float* __restrict__ C, i64 ldc, int tiles_m, int tiles_n,
                                                            int vec) {
    __shared__ __attribute__((aligned(16))) unsigned char As[2][BM * ROW_STRIDE];
    __shared__ __attribute__((aligned(16))) unsigned char Bs[2][BN * ROW_STRIDE];
    constexpr int ES = PREC == 2 ? 1 : 2;  // element size
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const i64 m0 = i64(tm) * BM, n0 = i64(tn) * BN;
    const i64 ld_bytes = Kp * ES;
    const int nslab = int(ld_bytes / SLAB_BYTES);

    const int li = lane & 15, lg = lane >> 4;
    const bool full = vec && m0 + BM <= M && n0 + BN <= N;

    f4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

    uint4 ra[4], rb[4];
    load_rows(ra, A, ld_bytes, m0, M, 0, t);
    load_rows(rb, B, ld_bytes, n0, N, 0, t);
    store_rows(As[0], ra, t);
    store_rows(Bs[0], rb, t);
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
        const int cur = s & 1;
        if (s + 1 < nslab) {
            load_rows(ra, A, ld_bytes, m0, M, i64(s + 1) * SLAB_BYTES, t);
            load_rows(rb, B, ld_bytes, n0, N, i64(s + 1) * SLAB_BYTES, t);
        }
        const unsigned char* as = As[cur];
        const unsigned char* bs = Bs[cur];
        if (PREC != 2) {
            // two MFMA k-steps of 32 elements (64 bytes) per slab
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const int kb = ks * 64 + lg * 16;  // byte offset of this lane's 8 elements
                uint4 a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = *(const uint4*)(as + (wm * 64 + i * 16 + li) * ROW_STRIDE + kb);
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = *(const uint4*)(bs + (wn * 64 + j * 16 + li) * ROW_STRIDE + kb);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (PREC == 0)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                                                __builtin_bit_cast(bf16x8, b[j]),
                                                                                acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a[i]),
                                                                               __builtin_bit_cast(f16x8, b[j]),
                                                                               acc[i][j], 0, 0, 0);
                    }
            }
        } else {
            // one MFMA k-step of 128 fp8 values; this lane's 32 bytes
            const int kb = lg * 32;
            i32x8 a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const unsigned char* p = as + (wm * 64 + i * 16 + li) * ROW_STRIDE + kb;
                const uint4 lo = *(const uint4*)p, hi = *(const uint4*)(p + 16);
                a[i] = i32x8{int(lo.x), int(lo.y), int(lo.z), int(lo.w), int(hi.x), int(hi.y), int(hi.z), int(hi.w)};
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const unsigned char* p = bs + (wn * 64 + j * 16 + li) * ROW_STRIDE + kb;
                const uint4 lo = *(const uint4*)p, hi = *(const uint4*)(p + 16);
                b[j] = i32x8{int(lo.x), int(lo.y), int(lo.z), int(lo.w), int(hi.x), int(hi.y), int(hi.z), int(hi.w)};
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, 127,
                                                                                 0, 127);
        }
        if (s + 1 < nslab) {
            store_rows(As[cur ^ 1], ra, t);
            store_rows(Bs[cur ^ 1], rb, t);
        }
        __syncthreads();
    }
    // epilogue: C += alpha * sa(row block) * sb(col block) * acc
    if (full) {
        f4v cq[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                cq[i][j] = __builtin_nontemporal_load(
                    (const f4v*)(C + (n0 + wn * 64 + j * 16 + li) * ldc + m0 + wm * 64 + i * 16 + lg * 4));
        // a 256-row scale block holds the whole 64-row wave tile (m0 and
        // wm * 64 are multiples of 64), likewise for the columns
        const float rs = sa ? sa[(m0 + wm * 64) >> 8] : 1.0f;
        const float cs = alpha * (sb ? sb[(n0 + wn * 64) >> 8] : 1.0f);
        const float w = rs * cs;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                f4v v;
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = __builtin_fmaf(w, acc[i][j][r], cq[i][j][r]);
                __builtin_nontemporal_store(v, (f4v*)(C + (n0 + wn * 64 + j * 16 + li) * ldc + m0 + wm * 64 + i * 16 + lg * 4));
            }
        return;
    }
    // edge tiles and unaligned C: all loads (C and the scales) are issued
    // before the first store, without branches (an element outside C reads
    // the nearest element inside it and is not stored): a load behind a bounds
    // branch is waited for before the next one is issued, one memory round
    // trip per element.
    float cv[4][4][4], rsv[4][4], csv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const i64 col = min(n0 + wn * 64 + j * 16 + li, N - 1);
        csv[j] = alpha * (sb ? sb[col >> 8] : 1.0f);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) cv[i][j][r] = C[col * ldc + min(m0 + wm * 64 + i * 16 + lg * 4 + r, M - 1)];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) rsv[i][r] = sa ? sa[min(m0 + wm * 64 + i * 16 + lg * 4 + r, M - 1) >> 8] : 1.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const i64 col = n0 + wn * 64 + j * 16 + li;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 row = m0 + wm * 64 + i * 16 + lg * 4 + r;
                if (row < M && col < N) C[col * ldc + row] = __builtin_fmaf(csv[j] * rsv[i][r], acc[i][j][r], cv[i][j][r]);
            }
    }
}

// ---------------------------------------------------------------------------
// Conversions from the FP32 working matrix
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned char to_e4m3(float x) {
    // OCP e4m3 with round-to-nearest-even (hardware conversion on gfx950);
    // inputs are pre-scaled into [-448, 448]
    const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false);
    return (unsigned char)(packed & 0xff);
}

template <int PREC>
__device__ __forceinline__ void put(unsigned char* dst, i64 idx, float x) {
    if (PREC == 0) {
        __hip_bfloat16 v = __float2bfloat16(x);
        ((__hip_bfloat16*)dst)[idx] = v;
    } else if (PREC == 1) {
        ((__half*)dst)[idx] = __float2half_rn(x);
    } else {
        dst[idx] = to_e4m3(x);
    }
}

// A_lp[i][k] = cvt(X[i + k*ldx] / scale(i/256)) for i < M, k < K (zero up to Kp):
// the column-major block is transposed through a 32x32 LDS tile.
template <int PREC>
__global__ __launch_bounds__(256) void convert_rows_kernel(const float* __restrict__ X, i64 ldx, i64 M, i64 K, i64 Kp,
                                                           const float* __restrict__ scale,
                                                           unsigned char* __restrict__ out) {
    __shared__ float tile[32][33];
    const i64 i0 = i64(blockIdx.x) * 32, k0 = i64(blockIdx.y) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int y = ty; y < 32; y += 8) {
        const i64 i = i0 + tx, k = k0 + y;
        tile[y][tx] = (i < M && k < K) ? X[i + k * ldx] : 0.0f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const i64 i = i0 + y, k = k0 + tx;
        if (i < M && k < Kp) {
            const float inv = scale ? 1.0f / scale[i >> 8] : 1.0f;
            put<PREC>(out, i * Kp + k, tile[tx][y] * inv);
        }
    }
}

// B_lp[j][k] = cvt(X[k + j*ldx] / scale(j/256)) (k contiguous on both sides)
template <int PREC>
__global__ __launch_bounds__(256) void convert_cols_kernel(const float* __restrict__ X, i64 ldx, i64 K, i64 N, i64 Kp,
                                                           const float* __restrict__ scale,
                                                           unsigned char* __restrict__ out) {
    const i64 total = N * Kp;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / Kp, k = idx - j * Kp;
        const float inv = scale ? 1.0f / scale[j >> 8] : 1.0f;
        put<PREC>(out, idx, k < K ? X[k + j * ldx] * inv : 0.0f);
    }
}

// Power-of-two scale per 256-block along one dimension of a column-major
// block: rows (dim 0) or columns (dim 1); scale = 2^e with max/scale <= 448.
// Three steps, so that the maximum is spread over many workgroups:
// block_max_kernel (blockIdx.x = the 256-block, blockIdx.y = a slice of `cs`
// columns of it) folds |X| with fmaxf, which skips NaN, so each partial is a
// non-negative float or +inf, whose bit pattern orders like its value: an
// unsigned atomicMax on the bits, from zero, gives the block's maximum.
// scale_from_max_kernel then turns each maximum into its scale in place.
__global__ __launch_bounds__(256) void block_max_kernel(const float* __restrict__ X, i64 ldx, i64 rows, i64 cols,
                                                        int dim, int cs, unsigned* __restrict__ bits) {
    __shared__ float red[256];
    const i64 b = blockIdx.x;
    float mx = 0.0f;
    if (dim == 0) {
        const i64 r = b * 256 + threadIdx.x;
        const i64 ca = i64(blockIdx.y) * cs, ce = min(cols, ca + cs);
        if (r < rows)
            for (i64 c = ca; c < ce; ++c) mx = fmaxf(mx, fabsf(X[r + c * ldx]));
    } else {
        const i64 ca = b * 256 + i64(blockIdx.y) * cs, ce = min(min(cols, b * 256 + 256), ca + cs);
        for (i64 c = ca; c < ce; ++c)
            for (i64 r = threadIdx.x; r < rows; r += 256) mx = fmaxf(mx, fabsf(X[r + c * ldx]));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(bits + b, __float_as_uint(red[0]));
}
__global__ void scale_from_max_kernel(float* __restrict__ scale, i64 count) {
    const i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float m = scale[i];
    float s = 1.0f;
    if (m > 0.0f && m == m) {
        int e;
        frexpf(m / 448.0f, &e);      // m/448 = f * 2^e, 0.5 <= f < 1
        s = ldexpf(1.0f, e);         // m / s <= 448
    }
    scale[i] = s;
}

// ---------------------------------------------------------------------------
// Float16 gemm! (src/blas/extended_blas.jl:39-88): Float16 operands and
// output, Float32 accumulation of each element over all of K, one rounding to
// Float16 at the end: C = Float16(a * acc + b * Float32(C)) with a, b the
// Float32 values of alpha and beta rounded to Float16 (beta == 0 never reads C).
// ---------------------------------------------------------------------------
// out[o][p] (Kp-row-major) = op(X)(o, p) for o < O, p < K, zero up to Kp;
// element (o, p) of op(X) at X[o * so + p * sp].  P_CONTIG: sp == 1 (threads
// run along p), otherwise so == 1 (a 32 x 32 tile transposed through LDS).
template <bool P_CONTIG>
__global__ __launch_bounds__(256) void pack_half_kernel(const _Float16* __restrict__ X, i64 so, i64 sp, i64 O, i64 K,
                                                        i64 Kp, _Float16* __restrict__ out) {
    __shared__ _Float16 tile[32][34];
    const i64 o0 = i64(blockIdx.x) * 32, p0 = i64(blockIdx.y) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (P_CONTIG) {
        for (int y = ty; y < 32; y += 8) {
            const i64 o = o0 + y, p = p0 + tx;
            if (o < O && p < Kp) out[o * Kp + p] = p < K ? X[o * so + p] : _Float16(0);
        }
        return;
    }
    for (int y = ty; y < 32; y += 8) {
        const i64 o = o0 + tx, p = p0 + y;
        tile[y][tx] = (o < O && p < K) ? X[o + p * sp] : _Float16(0);
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const i64 o = o0 + y, p = p0 + tx;
        if (o < O && p < Kp) out[o * Kp + p] = tile[tx][y];
    }
}
__global__ void zero_f32_kernel(float* __restrict__ C, i64 total) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += i64(gridDim.x) * blockDim.x) C[i] = 0.0f;
}
// C (Float16) <- Float16(aw * acc + bw * Float32(C)); beta == 0 never reads C
__global__ void hgemm_epilogue_kernel(const float* __restrict__ acc, i64 M, i64 N, float aw, float bw, int beta0,
                                      _Float16* __restrict__ C, i64 ldc) {
    const i64 total = M * N;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / M, i = idx - j * M;
        float v = aw * acc[idx];
        if (!beta0) v += bw * float(C[i + j * ldc]);
        C[i + j * ldc] = _Float16(v);
    }
}
__global__ void hscale_kernel(i64 M, i64 N, _Float16 beta, _Float16* __restrict__ C, i64 ldc) {
    const i64 total = M * N;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / M, i = idx - j * M;
        C[i + j * ldc] = beta == _Float16(0) ? _Float16(0) : _Float16(float(beta) * float(C[i + j * ldc]));
    }
}

}  // namespace

extern "C" int oxi_hgemm(oxi_handle_t h, char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
                         const void* Av, int64_t lda, const void* Bv, int64_t ldb, float beta, void* Cv, int64_t ldc) {
    if (!h) return -1;
    const int ta = (transA == 'N' || transA == 'n') ? 0 : (transA == 'T' || transA == 't' || transA == 'C' || transA == 'c') ? 1 : -1;
    const int tb = (transB == 'N' || transB == 'n') ? 0 : (transB == 'T' || transB == 't' || transB == 'C' || transB == 'c') ? 1 : -1;
    if (ta < 0) return -2;
    if (tb < 0) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0) return -6;
    if (lda < std::max<i64>(1, ta ? k : m)) return -9;
    if (ldb < std::max<i64>(1, tb ? n : k)) return -11;
    if (ldc < std::max<i64>(1, m)) return -14;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0) return OXI_OK;
    if (!Cv) return -13;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    const _Float16* A = (const _Float16*)Av;
    const _Float16* B = (const _Float16*)Bv;
    _Float16* C = (_Float16*)Cv;
    // alpha and beta are converted to Float16 first (extended_blas.jl:43)
    const _Float16 ah = _Float16(alpha), bh = _Float16(beta);
    const unsigned eg = unsigned(std::min<i64>(cdiv(m * n, 256), 8192));
    if (ah == _Float16(0) || k == 0) {
        hipLaunchKernelGGL(hscale_kernel, dim3(eg), dim3(256), 0, s, m, n, bh, C, ldc);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? OXI_OK : oxi_set_error(h, OXI_ERR_HIP, "hgemm launch: %s", hipGetErrorString(e));
    }
    if (!A) return -8;
    if (!B) return -10;
    const i64 kslab = SLAB_BYTES / 2;
    const i64 Kp = cdiv(k, kslab) * kslab;
    const size_t abytes = (size_t(m) * Kp * 2 + 255) & ~size_t(255), bbytes = (size_t(n) * Kp * 2 + 255) & ~size_t(255);
    void* ws;
    OXI_TRY(oxi_workspace(h, abytes + bbytes + size_t(m) * size_t(n) * 4 + 256, &ws));
    _Float16* Ap = (_Float16*)ws;
    _Float16* Bp = (_Float16*)((char*)ws + abytes);
    float* acc = (float*)((char*)ws + abytes + bbytes);
    // op(A)(i, p): A[i + p*lda] ('N', i contiguous) or A[p + i*lda] ('T', p contiguous);
    // op(B)(p, j) packed as [j][p]: B[p + j*ldb] ('N') or B[j + p*ldb] ('T')
    const dim3 ga(unsigned(cdiv(m, 32)), unsigned(cdiv(Kp, 32))), gb(unsigned(cdiv(n, 32)), unsigned(cdiv(Kp, 32)));
    if (ta) hipLaunchKernelGGL(pack_half_kernel<true>, ga, dim3(256), 0, s, A, lda, i64(1), m, k, Kp, Ap);
    else hipLaunchKernelGGL(pack_half_kernel<false>, ga, dim3(256), 0, s, A, i64(1), lda, m, k, Kp, Ap);
    if (tb) hipLaunchKernelGGL(pack_half_kernel<false>, gb, dim3(256), 0, s, B, i64(1), ldb, n, k, Kp, Bp);
    else hipLaunchKernelGGL(pack_half_kernel<true>, gb, dim3(256), 0, s, B, ldb, i64(1), n, k, Kp, Bp);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(eg), dim3(256), 0, s, acc, m * n);
    const int tiles_m = int(cdiv(m, BM)), tiles_n = int(cdiv(n, BN));
    hipLaunchKernelGGL(lowp_update_kernel<1>, dim3(unsigned(tiles_m * tiles_n)), dim3(NT), 0, s, m, n, Kp, 1.0f,
                       (const unsigned char*)Ap, (const unsigned char*)Bp, (const float*)nullptr, (const float*)nullptr,
                       acc, m, tiles_m, tiles_n, int(m % 4 == 0));
    hipLaunchKernelGGL(hgemm_epilogue_kernel, dim3(eg), dim3(256), 0, s, acc, m, n, float(ah), float(bh),
                       int(bh == _Float16(0)), C, ldc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "hgemm launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

// Trailing update C (M x N, FP32, ldc) += alpha * X_A (M x K) * X_B (K x N) with
// both operands rounded to `prec` (OXI_PREC_BF16 / FP16 / FP8) from FP32
// column-major blocks.  `ws` holds the packed operands and scales.
int oxi_lowp_update(oxi_handle_t h, hipStream_t s, int prec, i64 M, i64 N, i64 K, float alpha, const float* XA,
                    i64 ldxa, const float* XB, i64 ldxb, float* C, i64 ldc, void* ws, size_t ws_bytes) {
    if (M <= 0 || N <= 0 || K <= 0) return OXI_OK;
    const int es = prec == 2 ? 1 : 2;
    const i64 kslab = SLAB_BYTES / es;
    const i64 Kp = cdiv(K, kslab) * kslab;
    const size_t abytes = size_t(M) * Kp * es, bbytes = size_t(N) * Kp * es;
    const i64 nsa = cdiv(M, 256), nsb = cdiv(N, 256);
    const size_t need = ((abytes + 255) & ~size_t(255)) + ((bbytes + 255) & ~size_t(255)) + size_t(nsa + nsb) * 4 + 256;
    if (need > ws_bytes) return oxi_set_error(h, OXI_ERR_ALLOC, "lowp update workspace too small");
    unsigned char* Ap = (unsigned char*)ws;
    unsigned char* Bp = Ap + ((abytes + 255) & ~size_t(255));
    float* sa = (float*)(Bp + ((bbytes + 255) & ~size_t(255)));
    float* sb = sa + nsa;
    if (prec == 2) {
        // sb follows sa: one fill, then the maxima of A's row blocks (slices
        // of 16 columns) and of B's column blocks (16 columns each)
        const i64 ns = nsa + nsb;
        hipLaunchKernelGGL(zero_f32_kernel, dim3(unsigned(cdiv(ns, 256))), dim3(256), 0, s, sa, ns);
        const int cs = 16;
        hipLaunchKernelGGL(block_max_kernel, dim3(unsigned(nsa), unsigned(cdiv(K, cs))), dim3(256), 0, s, XA, ldxa, M,
                           K, 0, cs, (unsigned*)sa);
        hipLaunchKernelGGL(block_max_kernel, dim3(unsigned(nsb), unsigned(256 / cs)), dim3(256), 0, s, XB, ldxb, K, N,
                           1, cs, (unsigned*)sb);
        hipLaunchKernelGGL(scale_from_max_kernel, dim3(unsigned(cdiv(ns, 256))), dim3(256), 0, s, sa, ns);
    } else {
        sa = nullptr; sb = nullptr;
    }
    dim3 gA(unsigned(cdiv(M, 32)), unsigned(cdiv(Kp, 32)));
    const unsigned gB = unsigned(std::min<i64>(cdiv(N * Kp, 256), 65535));
    if (prec == 0) {
        hipLaunchKernelGGL(convert_rows_kernel<0>, gA, dim3(256), 0, s, XA, ldxa, M, K, Kp, sa, Ap);
        hipLaunchKernelGGL(convert_cols_kernel<0>, dim3(gB), dim3(256), 0, s, XB, ldxb, K, N, Kp, sb, Bp);
    } else if (prec == 1) {
        hipLaunchKernelGGL(convert_rows_kernel<1>, gA, dim3(256), 0, s, XA, ldxa, M, K, Kp, sa, Ap);
        hipLaunchKernelGGL(convert_cols_kernel<1>, dim3(gB), dim3(256), 0, s, XB, ldxb, K, N, Kp, sb, Bp);
    } else {
        hipLaunchKernelGGL(convert_rows_kernel<2>, gA, dim3(256), 0, s, XA, ldxa, M, K, Kp, sa, Ap);
        hipLaunchKernelGGL(convert_cols_kernel<2>, dim3(gB), dim3(256), 0, s, XB, ldxb, K, N, Kp, sb, Bp);
    }
    const int tiles_m = int(cdiv(M, BM)), tiles_n = int(cdiv(N, BN));
    dim3 grid(unsigned(tiles_m * tiles_n));
    const int vec = int(ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0);
    if (prec == 0)
        hipLaunchKernelGGL(lowp_update_kernel<0>, grid, dim3(NT), 0, s, M, N, Kp, alpha, Ap, Bp, sa, sb, C, ldc, tiles_m, tiles_n, vec);
    else if (prec == 1)
        hipLaunchKernelGGL(lowp_update_kernel<1>, grid, dim3(NT), 0, s, M, N, Kp, alpha, Ap, Bp, sa, sb, C, ldc, tiles_m, tiles_n, vec);
    else
        hipLaunchKernelGGL(lowp_update_kernel<2>, grid, dim3(NT), 0, s, M, N, Kp, alpha, Ap, Bp, sa, sb, C, ldc, tiles_m, tiles_n, vec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "lowp update launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

size_t oxi_lowp_update_ws(int prec, i64 M, i64 N, i64 K) {
    const int es = prec == 2 ? 1 : 2;
    const i64 kslab = SLAB_BYTES / es;
    const i64 Kp = cdiv(K, kslab) * kslab;
    return ((size_t(M) * Kp * es + 255) & ~size_t(255)) + ((size_t(N) * Kp * es + 255) & ~size_t(255)) +
           size_t(cdiv(M, 256) + cdiv(N, 256)) * 4 + 256;
}

// C ABI: C (FP32) <- C + alpha * round(A) * round(B), the mixed-precision
// trailing product on its own (A: M x K, B: K x N, column-major FP32).
extern "C" int oxi_gemm_lowp_update(oxi_handle_t h, int prec, int64_t m, int64_t n, int64_t k, float alpha,
                                    const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc) {
    if (!h) return -1;
    if (prec < 0 || prec > 2) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (lda < std::max<i64>(1, m)) return -8;
    if (ldb < std::max<i64>(1, k)) return -10;
    if (ldc < std::max<i64>(1, m)) return -12;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0 || k == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    const size_t need = oxi_lowp_update_ws(prec, m, n, k);
    void* ws;
    OXI_TRY(oxi_workspace(h, need, &ws));
    return oxi_lowp_update(h, h->stream, prec, m, n, k, alpha, A, lda, B, ldb, C, ldc, ws, need);
}
