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
// 256 threads: thread i + 64 q takes row i and, in the update of column c,
// the columns j = c + 1 + q, c + 5 + q, ... up to i.  Per column: the pivot
// (read by every thread), the column scaled by its square root, a barrier,
// the rank-1 update of the trailing lower triangle, a barrier.  The square
// roots go to the diagonal at the end (the column step reads only the
// diagonal element of its own column).
template <typename T>
__global__ __launch_bounds__(256) void potrf_leaf_kernel(T* __restrict__ A, i64 lda, int n, i64 off,
                                                         i64* __restrict__ info) {
    __shared__ T S[CLEAF * (CLEAF + 1)];   // S[j * 65 + i] = A(i, j)
    __shared__ T R[CLEAF];                 // the square roots of the pivots
    if (*info != 0) return;                 // an earlier leaf failed
    const int t = threadIdx.x, i = t & 63, q = t >> 6;
    for (int j = q; j < n; j += 4)
        if (i < n && i >= j) S[j * (CLEAF + 1) + i] = A[i + j * lda];
    __syncthreads();
    int done = n;
    for (int c = 0; c < n; ++c) {
        const T d = S[c * (CLEAF + 1) + c];
        if (!pos_finite(d)) {                  // not finite and positive: stop at this column (uniform)
            if (t == 0) *info = off + c + 1;
            done = c;
            break;
        }
        const T r = sqrt_t(d);
        if (q == 0) {
            if (i == c) R[c] = r;
            else if (i > c && i < n) S[c * (CLEAF + 1) + i] = S[c * (CLEAF + 1) + i] / r;
        }
        __syncthreads();
        if (i > c && i < n) {
            // all loads of the step before its stores (the element lists of
            // the threads are disjoint, the column read lies left of them)
            const T lic = S[c * (CLEAF + 1) + i];
            T lj[CLEAF / 4], aij[CLEAF / 4];
#pragma unroll
            for (int k = 0; k < CLEAF / 4; ++k) {
                const int j = 4 * k + q;
                lj[k] = S[c * (CLEAF + 1) + j];
                aij[k] = S[j * (CLEAF + 1) + i];
            }
#pragma unroll
            for (int k = 0; k < CLEAF / 4; ++k) {
                const int j = 4 * k + q;
                if (j > c && j <= i) S[j * (CLEAF + 1) + i] = fma_t(-lic, lj[k], aij[k]);
            }
        }
        __syncthreads();
    }
    // the columns before a failing one are final; the rest keep their
    // updated values (the reference's leaf returns with the same contents)
    if (q == 0 && i < done) S[i * (CLEAF + 1) + i] = R[i];
    __syncthreads();
    for (int j = q; j < n; j += 4)
        if (i < n && i >= j) A[i + j * lda] = S[j * (CLEAF + 1) + i];
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
        hipLaunchKernelGGL(potrf_leaf_kernel<T>, dim3(1), dim3(256), 0, s, A, lda, int(n), off, dinfo);
        return launch_check(h, "potrf leaf");
    }
    i64 hh = n / 2;
    if (hh > CLEAF) hh = (hh / CLEAF) * CLEAF;
    OXI_TRY(potrf_rec(h, s, hh, A, lda, off, dinfo));
    OXI_TRY(chol_update(h, s, n - hh, hh, A, A + hh, A + hh + hh * lda, lda));
    return potrf_rec(h, s, n - hh, A + hh + hh * lda, lda, off + hh, dinfo);
}

// Large matrices: the same elimination in blocks of POTRF_NB columns,
// left to right (the right-looking blocked form of the recursion: each
// block's diagonal part by the recursion above, its rows below by the
// right-side TRSM, then one SYRK of the whole trailing matrix).  The
// recursion's own splits put a TRSM of n/2 columns at the top, one chain of
// 64-column steps; here every TRSM has POTRF_NB columns and the O(n^3) work
// is in SYRKs of rank POTRF_NB over the trailing matrix.
constexpr i64 POTRF_NB = 256;

template <typename T>
int potrf_blocked(oxi_handle_t h, hipStream_t s, i64 n, T* A, i64 lda, i64* dinfo) {
    for (i64 j = 0; j < n; j += POTRF_NB) {
        const i64 jb = std::min(POTRF_NB, n - j);
        T* Ajj = A + j + j * lda;
        OXI_TRY(potrf_rec<T>(h, s, jb, Ajj, lda, j, dinfo));
        if (j + jb < n) OXI_TRY(chol_update(h, s, n - j - jb, jb, Ajj, Ajj + jb, Ajj + jb + jb * lda, lda));
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
    // L Y = B, then L^T X = Y (factorizations.jl:288-300)
    OXI_TRY(oxi_dtrsm_left(h, s, 0, 0, 0, n, nrhs, L, ldl, B, ldb));
    OXI_TRY(oxi_dtrsm_left(h, s, 0, 1, 0, n, nrhs, L, ldl, B, ldb));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}
