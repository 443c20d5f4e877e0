// Cholesky factorization A = L L^T of a symmetric positive definite FP64
// matrix and its solve: cholesky_factor / _cholesky_factor! /
// _cholesky_recursive! (src/lapack/factorizations.jl:206-286) and
// solve(F::CholeskyFactor, b) (:288-300).
//
// The recursion is the reference's (LAPACK potrf2): L11 first, then
// L21 = A21 L11^-T (right-side TRSM), A22 -= L21 L21^T (the lower-triangle
// SYRK on the matrix cores: only tiles that touch the lower triangle run),
// then L22.  Leaves of at most 64 columns are factored right-looking by one
// workgroup with the block in LDS, with the reference leaf's operations in its
// order: d = sqrt(a_cc), l_ic = a_ic / d, a_ij = fma(-l_ic, l_jc, a_ij).
// A leaf that meets a non-positive or non-finite pivot records its global
// column in a device word and stops; later leaves see the word and do
// nothing, so the first failing column is the one reported (the reference's
// PosDefException(j) / info).
#include "oxi_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace {

constexpr int CLEAF = 64;

__device__ __forceinline__ double sqrt_t(double v) { return __builtin_sqrt(v); }
__device__ __forceinline__ float sqrt_t(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename T> __device__ __forceinline__ bool pos_finite(T d);
template <> __device__ __forceinline__ bool pos_finite(double d) { return d > 0.0 && d <= 1.7976931348623157e308; }
template <> __device__ __forceinline__ bool pos_finite(float d) { return d > 0.0f && d <= 3.4028234663852886e38f; }

// A (n x n, n <= 64, lower triangle) in place; `off` is the global index of
// its first column; *info receives off + c + 1 at the first failing column c.
// One wave: lane i holds row i (a[k] = A(i, k), k <= i) in registers, and
// the column loop is unrolled, so every register index is a compile-time
// constant and no step waits at a workgroup barrier.  Per column c: the
// pivot from lane c (a readlane), its square root, every lane's multiplier
// l_ic = a_ic / r into LDS, then each lane updates its row right of c
// (a_ik = fma(-l_ic, l_kc, a_ik)) with the multipliers read back.  Elements
// right of the diagonal take the same operations on values nobody stores.
// A failing pivot (not finite and positive) stops the loop: the columns
// before it are final, the rest keep their updated values.
template <typename T>
__device__ __forceinline__ T readlane_t(T x, int l);
template <>
__device__ __forceinline__ double readlane_t(double x, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(u), l), hi = __builtin_amdgcn_readlane(unsigned(u >> 32), l);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <>
__device__ __forceinline__ float readlane_t(float x, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

template <typename T>
__global__ __launch_bounds__(64) void potrf_leaf_kernel(T* __restrict__ A, i64 lda, int n, i64 off,
                                                        i64* __restrict__ info) {
    __shared__ __align__(16) T lcol[CLEAF];   // the multipliers of the current column
    if (*info != 0) return;                   // an earlier leaf failed
    const int i = threadIdx.x;
    T a[CLEAF];
#pragma unroll
    for (int k = 0; k < CLEAF; ++k) a[k] = (i < n && k <= i) ? A[i + k * lda] : T(0);
    bool go = true;   // uniform
#pragma unroll
    for (int c = 0; c < CLEAF; ++c) {
        if (go && c < n) {
            const T d = readlane_t(a[c], c);
            if (!pos_finite(d)) {
                if (i == 0) *info = off + c + 1;
                go = false;
            } else {
                const T r = sqrt_t(d);
                const T l = i > c ? a[c] / r : T(0);
                a[c] = i == c ? r : i > c ? l : a[c];
                lcol[i] = l;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the lanes' LDS writes before the reads below
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = c + 1; k < CLEAF; ++k) a[k] = fma_t(-l, lcol[k], a[k]);
                __builtin_amdgcn_wave_barrier();   // every read of this column before the next column's writes
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CLEAF; ++k)
        if (i < n && k <= i) A[i + k * lda] = a[k];
}

// zero the strict upper triangle (the factor's storage contract)
__global__ void zero_upper_kernel(i64 n, double* __restrict__ A, i64 lda) {
    const i64 total = n * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / n, i = idx - j * n;
        if (i < j) A[i + j * lda] = 0.0;
    }
}

__global__ void set_i64_kernel(i64* p, i64 v) {
    if (threadIdx.x == 0) *p = v;
}

int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

// L21 = A21 L11^-T (X * op(T)^-1 with T = L11 lower, op = transpose) and
// A22 -= L21 L21^T: the lower-triangle SYRK in FP64; in FP32 (the working
// copy of the mixed solve, whose upper triangle is never read) the full
// product.
int chol_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 hh, const double* L11, double* A21, double* A22, i64 lda) {
    OXI_TRY(oxi_dtrsm_right(h, s, 0, 1, 0, rows, hh, L11, lda, A21, lda));
    return oxi_launch_dsyrk(h, s, 0, 0, rows, hh, -1.0, A21, lda, 1.0, A22, lda);
}
int chol_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 hh, const float* L11, float* A21, float* A22, i64 lda) {
    OXI_TRY(oxi_strsm_right(h, s, 0, 1, 0, rows, hh, L11, lda, A21, lda));
    return oxi_launch_sgemm(h, s, 0, 1, rows, rows, hh, -1.0f, A21, lda, A21, lda, 1.0f, A22, lda);
}

// _cholesky_recursive! (src/lapack/factorizations.jl:263-286) on the lower
// triangle of A (n x n); split points are multiples of the leaf width so the
// right-side TRSM can take its 64 x 64 block-inverse path.
template <typename T>
int potrf_rec(oxi_handle_t h, hipStream_t s, i64 n, T* A, i64 lda, i64 off, i64* dinfo) {
    if (n <= 0) return OXI_OK;
    if (n <= CLEAF) {
        hipLaunchKernelGGL(potrf_leaf_kernel<T>, dim3(1), dim3(64), 0, s, A, lda, int(n), off, dinfo);
        return launch_check(h, "potrf leaf");
    }
    i64 hh = n / 2;
    if (hh > CLEAF) hh = (hh / CLEAF) * CLEAF;
    OXI_TRY(potrf_rec(h, s, hh, A, lda, off, dinfo));
    OXI_TRY(chol_update(h, s, n - hh, hh, A, A + hh, A + hh + hh * lda, lda));
    return potrf_rec(h, s, n - hh, A + hh + hh * lda, lda, off + hh, dinfo);
}

// Large matrices: the same elimination in blocks of POTRF_NB columns, left
// to right (the right-looking blocked form of the recursion), with one block
// of look-ahead.  A block column is factored (its diagonal block by the
// recursion above, its rows below by the right-side TRSM) on the side
// stream, while the main stream, kept off a few CUs as for getrf, applies the
// previous block's SYRK to the rest of the trailing matrix; the next block
// column's own update comes first, so that its factorization can start.  The
// recursion's splits put a TRSM of n/2 columns at the top, one chain of
// 64-column steps; here every TRSM has POTRF_NB columns and the O(n^3) work
// is in SYRKs of rank POTRF_NB.
constexpr i64 POTRF_NB = 256;

// A[c0:, c0:c0+w] -= L[c0:, j:j+jb] L[c0:c0+w, j:j+jb]^T: the lower triangle
// of the diagonal block by SYRK, the rows below it by GEMM (FP64); the FP32
// working copy takes one product (its upper triangle is never read).
int col_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 w, i64 jb, const double* L, double* Acol, i64 lda) {
    OXI_TRY(oxi_launch_dsyrk(h, s, 0, 0, w, jb, -1.0, L, lda, 1.0, Acol, lda));
    if (rows > w) OXI_TRY(oxi_launch_dgemm(h, s, 0, 1, rows - w, w, jb, -1.0, L + w, lda, L, lda, 1.0, Acol + w, lda));
    return OXI_OK;
}
int col_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 w, i64 jb, const float* L, float* Acol, i64 lda) {
    return oxi_launch_sgemm(h, s, 0, 1, rows, w, jb, -1.0f, L, lda, L, lda, 1.0f, Acol, lda);
}
int rest_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 jb, const double* L, double* A22, i64 lda) {
    return oxi_launch_dsyrk(h, s, 0, 0, rows, jb, -1.0, L, lda, 1.0, A22, lda);
}
int rest_update(oxi_handle_t h, hipStream_t s, i64 rows, i64 jb, const float* L, float* A22, i64 lda) {
    return oxi_launch_sgemm(h, s, 0, 1, rows, rows, jb, -1.0f, L, lda, L, lda, 1.0f, A22, lda);
}
int trsm_rt(oxi_handle_t h, hipStream_t s, i64 rows, i64 w, const double* L11, double* A21, i64 lda) {
    return oxi_dtrsm_right(h, s, 0, 1, 0, rows, w, L11, lda, A21, lda);
}
int trsm_rt(oxi_handle_t h, hipStream_t s, i64 rows, i64 w, const float* L11, float* A21, i64 lda) {
    return oxi_strsm_right(h, s, 0, 1, 0, rows, w, L11, lda, A21, lda);
}

// block column c0 .. c0+w-1 (rows c0 .. n-1): diagonal block, then the rows below
template <typename T>
int panel_factor(oxi_handle_t h, hipStream_t s, i64 w, i64 rows, T* Acol, i64 lda, i64 off, i64* dinfo) {
    OXI_TRY(potrf_rec<T>(h, s, w, Acol, lda, off, dinfo));
    if (rows > w) OXI_TRY(trsm_rt(h, s, rows - w, w, Acol, Acol + w, lda));
    return OXI_OK;
}

template <typename T>
int potrf_blocked(oxi_handle_t h, hipStream_t s_caller, i64 n, T* A, i64 lda, i64* dinfo) {
    const hipStream_t ms = h->bulk ? h->bulk : s_caller, ss = h->side;
    if (ms != s_caller) {
        OXI_HIP(h, hipEventRecord(h->ev_c, s_caller));
        OXI_HIP(h, hipStreamWaitEvent(ms, h->ev_c, 0));
    }
    // 512-column blocks over the first half of large matrices (rank-512
    // SYRKs run faster while the trailing matrix is large), POTRF_NB after.
    // Measured at n = 16384: 256 throughout 43.7 ms, 384 throughout 42.5,
    // 512 throughout 42.7, this schedule 41.6-41.9 (512 then 384: 41.7-42.3).
    const i64 nb_wide = n >= 8192 ? 2 * POTRF_NB : POTRF_NB;
    auto nb_at = [&](i64 j) { return std::min<i64>(2 * j < n ? nb_wide : POTRF_NB, n - j); };
    OXI_TRY(panel_factor<T>(h, ms, nb_at(0), n, A, lda, 0, dinfo));
    bool side_pending = false;
    for (i64 j = 0, next = 0; j < n; j = next) {
        const i64 jb = nb_at(j);
        next = j + jb;
        if (side_pending) { OXI_HIP(h, hipStreamWaitEvent(ms, h->ev_b, 0)); side_pending = false; }
        if (next >= n) break;
        const i64 jn = nb_at(next);
        T* Ln = A + next + j * lda;                       // L[next:, j:j+jb]
        OXI_TRY(col_update(h, ms, n - next, jn, jb, Ln, A + next + next * lda, lda));
        OXI_HIP(h, hipEventRecord(h->ev_a, ms));
        OXI_HIP(h, hipStreamWaitEvent(ss, h->ev_a, 0));
        OXI_TRY(panel_factor<T>(h, ss, jn, n - next, A + next + next * lda, lda, next, dinfo));
        OXI_HIP(h, hipEventRecord(h->ev_b, ss));
        side_pending = true;
        const i64 r0 = next + jn;
        if (r0 < n) OXI_TRY(rest_update(h, ms, n - r0, jb, A + r0 + j * lda, A + r0 + r0 * lda, lda));
    }
    if (side_pending) OXI_HIP(h, hipStreamWaitEvent(ms, h->ev_b, 0));
    if (ms != s_caller) {
        OXI_HIP(h, hipEventRecord(h->ev_c, ms));
        OXI_HIP(h, hipStreamWaitEvent(s_caller, h->ev_c, 0));
    }
    return OXI_OK;
}

// The blocked form from 1024 columns (n = 16384: 53.0 ms against 55.3 ms for
// the recursion, both with the 256-thread leaf and the triangle-only SYRK
// grid); OXIBLAS_POTRF_RECURSIVE selects the recursion throughout (tuning).
template <typename T>
int potrf_any(oxi_handle_t h, hipStream_t s, i64 n, T* A, i64 lda, i64* dinfo) {
    static const bool recursive = getenv("OXIBLAS_POTRF_RECURSIVE") != nullptr;
    if (!recursive && n >= 4 * POTRF_NB) return potrf_blocked<T>(h, s, n, A, lda, dinfo);
    return potrf_rec<T>(h, s, n, A, lda, 0, dinfo);
}

}  // namespace

int oxi_potrf_device(oxi_handle_t h, hipStream_t s, i64 n, double* A, i64 lda, i64* dinfo) {
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(64), 0, s, dinfo, i64(0));
    OXI_TRY(launch_check(h, "potrf info"));
    return potrf_any<double>(h, s, n, A, lda, dinfo);
}
int oxi_potrf_device(oxi_handle_t h, hipStream_t s, i64 n, float* A, i64 lda, i64* dinfo) {
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(64), 0, s, dinfo, i64(0));
    OXI_TRY(launch_check(h, "potrf info"));
    return potrf_any<float>(h, s, n, A, lda, dinfo);
}

extern "C" int oxi_dpotrf(oxi_handle_t h, char uplo, int64_t n, double* A, int64_t lda, int64_t* info) {
    if (!h) return -1;
    if (!(uplo == 'L' || uplo == 'l')) return -2;
    if (n < 0) return -3;
    if (lda < std::max<i64>(1, n)) return -5;
    if (!info) return -6;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = 0;
    if (n == 0) return OXI_OK;
    if (!A) return -4;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    OXI_TRY(oxi_potrf_device(h, s, n, A, lda, h->dev_info));
    i64 v = 0;
    OXI_HIP(h, hipMemcpyAsync(&v, h->dev_info, sizeof(i64), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    *info = v;
    if (v == 0) {
        // the strict upper triangle of the factor is zero (factorizations.jl:256-260)
        hipLaunchKernelGGL(zero_upper_kernel, dim3(unsigned(std::min<i64>(cdiv(n * n, 256), 8192))), dim3(256), 0, s,
                           n, A, lda);
        OXI_TRY(launch_check(h, "potrf zero upper"));
        OXI_HIP(h, hipStreamSynchronize(s));
    }
    return OXI_OK;
}

extern "C" int oxi_dpotrs(oxi_handle_t h, char uplo, int64_t n, int64_t nrhs, const double* L, int64_t ldl,
                          double* B, int64_t ldb) {
    if (!h) return -1;
    if (!(uplo == 'L' || uplo == 'l')) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldl < std::max<i64>(1, n)) return -6;
    if (ldb < std::max<i64>(1, n)) return -8;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0 || nrhs == 0) return OXI_OK;
    if (!L) return -5;
    if (!B) return -7;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    i64 z = 0;
    OXI_TRY(oxi_first_zero_diag(h, s, L, ldl, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "potrs: L(%lld,%lld) is zero", (long long)z, (long long)z);
    // L Y = B, then L^T X = Y (factorizations.jl:288-300): up to
    // OXI_CHAIN_MAX_RHS right-hand sides through the chained solve (its
    // workspace prepared for this factor), more through the left trsm
    if (oxi_trsm_chain_fits(n, nrhs, ldb)) {
        void* ws;
        OXI_TRY(oxi_workspace(h, oxi_trsm_chain_ws_bytes(n), &ws));
        OXI_TRY(oxi_trsm_chain_begin(h, s, ws));
        OXI_TRY(oxi_trsm_chain_prepare(h, s, 2, n, L, ldl, ws));
        OXI_TRY(oxi_trsm_chain(h, s, 2, 0, n, nrhs, L, ldl, B, ldb, ws));
        OXI_TRY(oxi_trsm_chain(h, s, 2, 1, n, nrhs, L, ldl, B, ldb, ws));
        return oxi_trsm_chain_check(h, s, ws);
    }
    OXI_TRY(oxi_dtrsm_left(h, s, 0, 0, 0, n, nrhs, L, ldl, B, ldb));
    OXI_TRY(oxi_dtrsm_left(h, s, 0, 1, 0, n, nrhs, L, ldl, B, ldb));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}
