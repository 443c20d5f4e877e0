// What is built on the LU factors and the residual: the inverse
// (native_inverse, src/lapack/dense_utilities.jl:14-36), log|det| and its
// sign (logabsdet_factor, :177-187; LinearAlgebra.logabsdet(F::LU),
// src/api/linalg_api.jl:185-191), the componentwise backward error
// (backward_error, dense_utilities.jl:214-228), the Hager-Higham estimate of
// ||A^-1||_1 (inverse_norm_estimate, src/lapack/expert_refinement.jl:45-60),
// refinement with given factors (refine_solve, :89-104) and the equilibrated
// expert solve (expert_solve, src/lapack/advanced_factorizations.jl:591-645).
//
// Every matrix pass is a device kernel; the host only runs the drivers'
// loops and reads the handful of scalars their decisions need.  Sums over a
// row or column run in a fixed order (fixed chunks, fixed reduction trees),
// so the results are deterministic; their rounding differs from the
// reference's sequential sums by the usual reassociation bound.
#include "oxi_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

typedef unsigned long long u64;
constexpr int TPB = 256;
constexpr i64 KC = 256;   // columns per partial-sum chunk of the residual

__device__ __forceinline__ u64 dbits(double v) { return __builtin_bit_cast(u64, v); }

int launch_ok(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

__global__ void fill_u64_kernel(u64* p, int n, u64 v) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// ---- residual and componentwise backward error ------------------------------
// Partial sums of column chunk blockIdx.y: P = sum_k a_ik x_kr, Q = sum_k
// |a_ik| |x_kr|, for right-hand side blockIdx.z; rows i over blockIdx.x.
__global__ __launch_bounds__(TPB) void resid_partial_kernel(i64 m, i64 n, const double* __restrict__ A, i64 lda,
                                                            const double* __restrict__ X, i64 ldx,
                                                            double* __restrict__ P, double* __restrict__ Q) {
    __shared__ double xs[KC];
    const int t = threadIdx.x;
    const i64 i = i64(blockIdx.x) * TPB + t;
    const i64 r = blockIdx.z;
    const i64 k0 = i64(blockIdx.y) * KC, k1 = min(n, k0 + KC);
    const int len = int(k1 - k0);
    for (int k = t; k < len; k += TPB) xs[k] = X[(k0 + k) + r * ldx];
    __syncthreads();
    if (i >= m) return;
    double v = 0, d = 0;
    const double* a = A + i + k0 * lda;
#pragma unroll 8
    for (int k = 0; k < len; ++k) {
        const double av = a[k * lda];
        v = __builtin_fma(av, xs[k], v);
        d = __builtin_fma(fabs(av), fabs(xs[k]), d);
    }
    const i64 o = (i64(blockIdx.y) + i64(gridDim.y) * r) * m + i;
    P[o] = v;
    Q[o] = d;
}

// R = B - A X and the largest |r_i| / (|A||x| + |b|)_i (0/0 = 0, r/0 = Inf,
// non-finite denominators Inf; NaN wins, as max in the reference); stat[0]
// the error and stat[1] the sum of squares of R per block (fixed order:
// block partials to `sq`, summed by sumsq_final_kernel).
__global__ __launch_bounds__(TPB) void resid_final_kernel(i64 m, i64 nrhs, int nchunk, const double* __restrict__ P,
                                                          const double* __restrict__ Q, const double* __restrict__ B,
                                                          i64 ldb, double* __restrict__ R, i64 ldr, u64* err,
                                                          double* __restrict__ sq) {
    __shared__ double red[TPB];
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    const i64 r = blockIdx.y;
    double rel = 0, s2 = 0;
    if (i < m) {
        double s = 0, q = 0;
        for (int c = 0; c < nchunk; ++c) {
            const i64 o = (i64(c) + i64(nchunk) * r) * m + i;
            s += P[o];
            q += Q[o];
        }
        const double b = B[i + r * ldb];
        const double value = b - s, den = fabs(b) + q;
        R[i + r * ldr] = value;
        rel = !(den <= 1.7976931348623157e308) ? INFINITY : den == 0 ? (value == 0 ? 0.0 : INFINITY) : fabs(value) / den;
        s2 = value * value;
    }
    u64 rb = dbits(rel);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 other = __shfl_xor(rb, o);
        rb = other > rb ? other : rb;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(err, rb);
    red[threadIdx.x] = s2;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sq[blockIdx.x + i64(gridDim.x) * r] = red[0];
}

// one block: out[0] = sum of v[0..len) in a fixed order
__global__ __launch_bounds__(TPB) void sum_final_kernel(const double* __restrict__ v, i64 len, double* out) {
    __shared__ double red[TPB];
    double s = 0;
    for (i64 i = threadIdx.x; i < len; i += TPB) s += v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// colsum[j] = sum_i |A[i,j]| (one block per column, fixed order)
__global__ __launch_bounds__(TPB) void col_abs_sum_kernel(i64 m, const double* __restrict__ A, i64 lda,
                                                          double* __restrict__ colsum) {
    __shared__ double red[TPB];
    const i64 j = blockIdx.x;
    double s = 0;
    for (i64 i = threadIdx.x; i < m; i += TPB) s += fabs(A[i + j * lda]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) colsum[j] = red[0];
}

// one block: out = max of v[0..len) (NaN wins)
__global__ __launch_bounds__(TPB) void max_final_kernel(const double* __restrict__ v, i64 len, double* out) {
    __shared__ u64 red[TPB];
    u64 b = 0;
    for (i64 i = threadIdx.x; i < len; i += TPB) {
        const u64 x = dbits(fabs(v[i]));
        b = x > b ? x : b;
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x + w] > red[threadIdx.x] ? red[threadIdx.x + w] : red[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = __builtin_bit_cast(double, red[0]);
}

// ---- log|det| ---------------------------------------------------------------
// One block: sum of log|u_ii| (fixed order), the count of negative and zero
// pivots, and the count of interchanges (ipiv[i] != i+1).
__global__ __launch_bounds__(TPB) void logdet_kernel(i64 n, const double* __restrict__ LU, i64 lda,
                                                     const i64* __restrict__ ipiv, double* out) {
    __shared__ double red[TPB];
    __shared__ int cnt[3][TPB];
    double s = 0;
    int neg = 0, zero = 0, swaps = 0;
    for (i64 i = threadIdx.x; i < n; i += TPB) {
        const double u = LU[i + i * lda];
        s += log(fabs(u));
        neg += u < 0;
        zero += u == 0 || u != u;
        swaps += ipiv[i] != i + 1;
    }
    red[threadIdx.x] = s;
    cnt[0][threadIdx.x] = neg; cnt[1][threadIdx.x] = zero; cnt[2][threadIdx.x] = swaps;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[threadIdx.x] += red[threadIdx.x + w];
            for (int c = 0; c < 3; ++c) cnt[c][threadIdx.x] += cnt[c][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = double(cnt[0][0]);
        out[2] = double(cnt[1][0]);
        out[3] = double(cnt[2][0]);
    }
}

// ---- inverse --------------------------------------------------------------------
__global__ void identity_kernel(i64 n, double* __restrict__ X, i64 ldx) {
    const i64 total = n * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 j = idx / n, i = idx - j * n;
        X[i + j * ldx] = i == j ? 1.0 : 0.0;
    }
}
// out[:, perm[k]-1] = X[:, k] (native_inverse's column permutation)
__global__ void permute_cols_kernel(i64 n, const double* __restrict__ X, i64 ldx, const i64* __restrict__ perm,
                                    double* __restrict__ out, i64 ldo) {
    const i64 total = n * n;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 k = idx / n, i = idx - k * n;
        out[i + (perm[k] - 1) * ldo] = X[i + k * ldx];
    }
}

// ---- expert solve: scaling --------------------------------------------------------
// rowmax[i] = max_j |A[i,j]| over column chunk blockIdx.y, combined by atomicMax
// on the bit patterns (non-negative doubles order as integers)
__global__ __launch_bounds__(TPB) void row_absmax_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                                         u64* __restrict__ rowmax) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    const i64 k0 = i64(blockIdx.y) * KC, k1 = min(n, k0 + KC);
    u64 b = 0;
    for (i64 k = k0; k < k1; ++k) {
        const u64 x = dbits(fabs(A[i + k * lda]));
        b = x > b ? x : b;
    }
    atomicMax(&rowmax[i], b);
}
// S = A ./ rows (row i by its maximum), then column scales cols[j] = max_i
// |S[i,j]| (1 when zero), and S[:, j] ./= cols[j]: one block per column
__global__ __launch_bounds__(TPB) void scale_cols_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                                         const double* __restrict__ rows, double* __restrict__ S,
                                                         i64 lds, double* __restrict__ cols) {
    __shared__ u64 red[TPB];
    const i64 j = blockIdx.x;
    u64 b = 0;
    for (i64 i = threadIdx.x; i < n; i += TPB) {
        const double v = A[i + j * lda] / rows[i];
        S[i + j * lds] = v;
        const u64 x = dbits(fabs(v));
        b = x > b ? x : b;
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x + w] > red[threadIdx.x] ? red[threadIdx.x + w] : red[threadIdx.x];
        __syncthreads();
    }
    const double cmax = __builtin_bit_cast(double, red[0]);
    const double c = cmax == 0 ? 1.0 : cmax;
    if (threadIdx.x == 0) cols[j] = c;
    for (i64 i = threadIdx.x; i < n; i += TPB) S[i + j * lds] = S[i + j * lds] / c;
}
// Y[i, r] = X[i, r] / d[i] (rows) ; with add: X[i, r] += Y[i, r] / d[i]
__global__ void div_rows_kernel(i64 n, i64 nrhs, const double* __restrict__ X, i64 ldx, const double* __restrict__ d,
                                double* __restrict__ Y, i64 ldy) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 r = idx / n, i = idx - r * n;
        Y[i + r * ldy] = X[i + r * ldx] / d[i];
    }
}
__global__ void add_div_rows_kernel(i64 n, i64 nrhs, const double* __restrict__ C, i64 ldc,
                                    const double* __restrict__ d, double* __restrict__ X, i64 ldx) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 r = idx / n, i = idx - r * n;
        X[i + r * ldx] += d ? C[i + r * ldc] / d[i] : C[i + r * ldc];
    }
}
__global__ void copy_kernel(i64 n, i64 nrhs, const double* __restrict__ X, i64 ldx, double* __restrict__ Y, i64 ldy) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 r = idx / n, i = idx - r * n;
        Y[i + r * ldy] = X[i + r * ldx];
    }
}

unsigned egrid(i64 total) { return unsigned(std::max<i64>(1, std::min<i64>(cdiv(total, TPB), 8192))); }

// ---------------------------------------------------------------------------
// Host-side building blocks on the handle's second workspace.
// ---------------------------------------------------------------------------
struct Resid {
    i64 m = 0, n = 0, nrhs = 0;
    int nchunk = 0;
    double *P = nullptr, *Q = nullptr, *sq = nullptr, *scal = nullptr;
    u64* err = nullptr;
    static size_t bytes(i64 m, i64 n, i64 nrhs) {
        const size_t nch = size_t(cdiv(std::max<i64>(n, 1), KC));
        const size_t part = (nch * size_t(nrhs) * size_t(m) * 8 + 255) & ~size_t(255);
        const size_t sq = (size_t(cdiv(std::max<i64>(m, 1), TPB)) * size_t(nrhs) * 8 + 255) & ~size_t(255);
        return 2 * part + sq + 256 + 256;
    }
    void place(char* p, i64 m_, i64 n_, i64 nrhs_) {
        m = m_; n = n_; nrhs = nrhs_;
        nchunk = int(cdiv(std::max<i64>(n, 1), KC));
        const size_t part = (size_t(nchunk) * size_t(nrhs) * size_t(m) * 8 + 255) & ~size_t(255);
        P = (double*)p; Q = (double*)(p + part);
        sq = (double*)(p + 2 * part);
        const size_t sqb = (size_t(cdiv(std::max<i64>(m, 1), TPB)) * size_t(nrhs) * 8 + 255) & ~size_t(255);
        err = (u64*)(p + 2 * part + sqb);
        scal = (double*)(p + 2 * part + sqb + 256);
    }
};

// R = B - A X, error and ||R||_F to host (synchronises)
int residual(oxi_handle_t h, hipStream_t s, Resid& w, const double* A, i64 lda, const double* X, i64 ldx,
             const double* B, i64 ldb, double* R, i64 ldr, double* err, double* rnorm) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, w.err, 1, u64(0));
    if (w.n > 0) {
        hipLaunchKernelGGL(resid_partial_kernel, dim3(unsigned(cdiv(w.m, TPB)), unsigned(w.nchunk), unsigned(w.nrhs)),
                           dim3(TPB), 0, s, w.m, w.n, A, lda, X, ldx, w.P, w.Q);
    } else {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, s, (u64*)w.P, int(w.m * w.nrhs), u64(0));
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, s, (u64*)w.Q, int(w.m * w.nrhs), u64(0));
    }
    OXI_TRY(launch_ok(h, "residual"));
    const unsigned gx = unsigned(cdiv(w.m, TPB));
    hipLaunchKernelGGL(resid_final_kernel, dim3(gx, unsigned(w.nrhs)), dim3(TPB), 0, s, w.m, w.nrhs,
                       std::max(w.nchunk, 1), w.P, w.Q, B, ldb, R, ldr, w.err, w.sq);
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(TPB), 0, s, w.sq, i64(gx) * w.nrhs, w.scal);
    OXI_TRY(launch_ok(h, "residual"));
    u64 e = 0;
    double ss = 0;
    OXI_HIP(h, hipMemcpyAsync(&e, w.err, 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipMemcpyAsync(&ss, w.scal, 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    std::memcpy(err, &e, 8);
    if (rnorm) *rnorm = std::sqrt(ss);
    return OXI_OK;
}

// ||X||_1 (largest column sum of |x|) of an m x ncols matrix, to host
int norm1(oxi_handle_t h, hipStream_t s, i64 m, i64 ncols, const double* X, i64 ldx, double* colsum, double* scal,
          double* out) {
    if (m == 0 || ncols == 0) { *out = 0; return OXI_OK; }
    hipLaunchKernelGGL(col_abs_sum_kernel, dim3(unsigned(ncols)), dim3(TPB), 0, s, m, X, ldx, colsum);
    hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(TPB), 0, s, colsum, ncols, scal);
    OXI_TRY(launch_ok(h, "norm1"));
    OXI_HIP(h, hipMemcpyAsync(out, scal, 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

// Solves with stored factors: the permutation on the device, the trsv
// workspaces of both solve kinds prepared on first use.
struct Solver {
    oxi_handle_t h;
    hipStream_t s;
    i64 n;
    const double* LU;
    i64 lda;
    i64* dperm;
    double* Y;           // n x nmax
    void* ctrl[2];
    bool ready[2] = {false, false};
    static size_t bytes(i64 n, i64 nmax) {
        return ((size_t(n) * 8 + 255) & ~size_t(255)) + ((size_t(n) * size_t(nmax) * 8 + 255) & ~size_t(255)) +
               2 * ((oxi_trsv_ws_bytes(n, 8) + 255) & ~size_t(255));
    }
    int place(char* p, i64 nmax, const i64* ipiv) {
        dperm = (i64*)p;
        p += (size_t(n) * 8 + 255) & ~size_t(255);
        Y = (double*)p;
        p += (size_t(n) * size_t(nmax) * 8 + 255) & ~size_t(255);
        ctrl[0] = p;
        ctrl[1] = p + ((oxi_trsv_ws_bytes(n, 8) + 255) & ~size_t(255));
        std::vector<i64> perm;
        OXI_TRY(oxi_perm_from_ipiv(h, s, ipiv, n, perm));
        OXI_HIP(h, hipMemcpyAsync(dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        return OXI_OK;
    }
    // B <- op(A)^-1 B in place (tk 0: A, 1: A^T)
    int solve(int tk, i64 nrhs, double* B, i64 ldb) {
        void* c = nullptr;
        if (nrhs <= OXI_TRSV_MAX_RHS) {
            c = ctrl[tk];
            if (!ready[tk]) {
                OXI_TRY(oxi_trsv_begin(h, s, c));
                OXI_TRY(oxi_trsv_prepare(h, s, tk, n, LU, lda, (double*)nullptr, c));
                ready[tk] = true;
            }
        }
        return oxi_getrs_perm_f64(h, s, tk, n, nrhs, LU, lda, dperm, B, ldb, Y, c);
    }
    int check() {
        for (int k = 0; k < 2; ++k)
            if (ready[k]) OXI_TRY(oxi_trsv_check(h, s, ctrl[k]));
        return OXI_OK;
    }
};

// Hager-Higham estimate of ||A^-1||_1 (inverse_norm_estimate,
// src/lapack/expert_refinement.jl:45-60); x, y, z: device vectors of n.
int inverse_norm(Solver& sv, i64 maxiter, double* x, double* colsum, double* scal, double* est) {
    const i64 n = sv.n;
    oxi_handle_t h = sv.h;
    hipStream_t s = sv.s;
    if (n == 0) { *est = 0; return OXI_OK; }
    const size_t nn = static_cast<size_t>(n);
    std::vector<double> hx(nn, 1.0 / double(n)), hy(nn), hz(nn);
    double estimate = 0;
    i64 previous = -1;
    for (i64 it = 0; it < maxiter; ++it) {
        OXI_HIP(h, hipMemcpyAsync(x, hx.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_TRY(sv.solve(0, 1, x, n));
        OXI_HIP(h, hipMemcpyAsync(hy.data(), x, size_t(n) * 8, hipMemcpyDeviceToHost, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        double next = 0;
        for (i64 i = 0; i < n; ++i) next += std::fabs(hy[size_t(i)]);
        estimate = std::max(estimate, next);
        // signs, then z = A^-T signs (the adjoint solve of an LU factor)
        for (i64 i = 0; i < n; ++i) hz[size_t(i)] = hy[size_t(i)] == 0 ? 1.0 : (hy[size_t(i)] > 0 ? 1.0 : -1.0);
        OXI_HIP(h, hipMemcpyAsync(x, hz.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_TRY(sv.solve(1, 1, x, n));
        OXI_HIP(h, hipMemcpyAsync(hz.data(), x, size_t(n) * 8, hipMemcpyDeviceToHost, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        i64 index = 0;
        double zmax = -1, dot = 0;
        for (i64 i = 0; i < n; ++i) {
            const double a = std::fabs(hz[size_t(i)]);
            if (a > zmax) { zmax = a; index = i; }   // argmax: the first largest
            dot += hz[size_t(i)] * hx[size_t(i)];
        }
        if (zmax <= dot) break;
        if (index == previous) break;
        std::fill(hx.begin(), hx.end(), 0.0);
        hx[size_t(index)] = 1.0;
        previous = index;
    }
    for (i64 i = 0; i < n; ++i)
        hx[size_t(i)] = ((i % 2 == 0) ? 1.0 : -1.0) * (1.0 + double(i) / double(std::max<i64>(1, n - 1)));
    OXI_HIP(h, hipMemcpyAsync(x, hx.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
    OXI_TRY(sv.solve(0, 1, x, n));
    OXI_HIP(h, hipMemcpyAsync(hy.data(), x, size_t(n) * 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    double alt = 0;
    for (i64 i = 0; i < n; ++i) alt += std::fabs(hy[size_t(i)]);
    *est = std::max(estimate, 2.0 * alt / (3.0 * double(n)));
    (void)colsum; (void)scal;
    return OXI_OK;
}

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int oxi_dlogabsdet(oxi_handle_t h, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv,
                              double* logabs, double* sign) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (lda < std::max<i64>(1, n)) return -4;
    if (!logabs) return -6;
    if (!sign) return -7;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0) { *logabs = 0; *sign = 1; return OXI_OK; }
    if (!LU) return -3;
    if (!ipiv) return -5;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    void* ws;
    OXI_TRY(oxi_workspace(h, 256, &ws));
    hipLaunchKernelGGL(logdet_kernel, dim3(1), dim3(TPB), 0, s, n, LU, lda, ipiv, (double*)ws);
    OXI_TRY(launch_ok(h, "logabsdet"));
    double out[4];
    OXI_HIP(h, hipMemcpyAsync(out, ws, sizeof(out), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    // sum of log|u_ii| (-Inf with a zero pivot) and the sign: the product of
    // the pivots' signs (0 with a zero pivot), negated per interchange
    *logabs = out[0];
    const i64 neg = i64(out[1]), zero = i64(out[2]), swaps = i64(out[3]);
    *sign = zero ? 0.0 : (((neg + swaps) % 2) ? -1.0 : 1.0);
    return OXI_OK;
}

extern "C" int oxi_dgetri(oxi_handle_t h, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv, double* Ainv,
                          int64_t ldainv) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (lda < std::max<i64>(1, n)) return -4;
    if (ldainv < std::max<i64>(1, n)) return -7;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0) return OXI_OK;
    if (!LU) return -3;
    if (!ipiv) return -5;
    if (!Ainv) return -6;
    {
        const char* a = (const char*)LU; const char* b = (const char*)Ainv;
        const size_t ab = size_t((n - 1) * lda + n) * 8, bb = size_t((n - 1) * ldainv + n) * 8;
        if (a < b + bb && b < a + ab) return oxi_set_error(h, -6, "getri: the inverse overlaps the factors");
    }
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    i64 z = 0;
    OXI_TRY(oxi_first_zero_diag(h, s, LU, lda, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "getri: U(%lld,%lld) is zero", (long long)z, (long long)z);
    std::vector<i64> perm;
    OXI_TRY(oxi_perm_from_ipiv(h, s, ipiv, n, perm));
    void* ws;
    OXI_TRY(oxi_workspace2(h, al(size_t(n) * size_t(n) * 8) + al(size_t(n) * 8), &ws));
    double* X = (double*)ws;
    i64* dperm = (i64*)((char*)ws + al(size_t(n) * size_t(n) * 8));
    OXI_HIP(h, hipMemcpyAsync(dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
    // inv(A) = inv(U) inv(L) P (dense_utilities.jl:14-36): X = U^-1 (left
    // solve of the identity), X <- X L^-1 (right solve, unit lower), then
    // the columns in the order of the permutation
    hipLaunchKernelGGL(identity_kernel, dim3(egrid(n * n)), dim3(TPB), 0, s, n, X, n);
    OXI_TRY(launch_ok(h, "getri"));
    OXI_TRY(oxi_dtrsm_left(h, s, 1, 0, 0, n, n, LU, lda, X, n));
    OXI_TRY(oxi_dtrsm_right(h, s, 0, 0, 1, n, n, LU, lda, X, n));
    hipLaunchKernelGGL(permute_cols_kernel, dim3(egrid(n * n)), dim3(TPB), 0, s, n, X, n, dperm, Ainv, ldainv);
    OXI_TRY(launch_ok(h, "getri"));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

extern "C" int oxi_dbackward_error(oxi_handle_t h, int64_t m, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                   const double* X, int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                                   double* error) {
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < std::max<i64>(1, m)) return -6;
    if (ldx < std::max<i64>(1, n)) return -8;
    if (ldb < std::max<i64>(1, m)) return -10;
    if (R && ldr < std::max<i64>(1, m)) return -12;
    if (!error) return -13;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *error = 0;
    if (m == 0 || nrhs == 0) return OXI_OK;
    if (!A && n > 0) return -5;
    if (!X && n > 0) return -7;
    if (!B) return -9;
    if (nrhs > 65535) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "backward_error: more than 65535 columns");
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    const size_t rb = R ? 0 : al(size_t(m) * size_t(nrhs) * 8);
    void* ws;
    OXI_TRY(oxi_workspace2(h, Resid::bytes(m, n, nrhs) + rb, &ws));
    Resid w;
    w.place((char*)ws + rb, m, n, nrhs);
    double* Rout = R ? R : (double*)ws;
    return residual(h, s, w, A, lda, X, ldx, B, ldb, Rout, R ? ldr : m, error, nullptr);
}

extern "C" int oxi_dinverse_norm_estimate(oxi_handle_t h, int64_t n, const double* LU, int64_t lda,
                                          const int64_t* ipiv, int32_t maxiter, double* estimate) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (lda < std::max<i64>(1, n)) return -4;
    if (maxiter < 1) return -6;
    if (!estimate) return -7;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *estimate = 0;
    if (n == 0) return OXI_OK;
    if (!LU) return -3;
    if (!ipiv) return -5;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    i64 z = 0;
    OXI_TRY(oxi_first_zero_diag(h, s, LU, lda, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "inverse_norm_estimate: U(%lld,%lld) is zero", (long long)z,
                                (long long)z);
    void* ws;
    OXI_TRY(oxi_workspace2(h, Solver::bytes(n, 1) + al(size_t(n) * 8) + 256, &ws));
    Solver sv{h, s, n, LU, lda};
    OXI_TRY(sv.place((char*)ws, 1, ipiv));
    double* x = (double*)((char*)ws + Solver::bytes(n, 1));
    OXI_TRY(inverse_norm(sv, maxiter, x, nullptr, nullptr, estimate));
    return sv.check();
}

extern "C" int oxi_drefine_solve(oxi_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                 const double* LU, int64_t ldlu, const int64_t* ipiv, const double* B, int64_t ldb,
                                 double* X, int64_t ldx, int32_t maxiter, double rtol, oxi_refine_info* info) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (lda < std::max<i64>(1, n)) return -5;
    if (ldlu < std::max<i64>(1, n)) return -7;
    if (ldb < std::max<i64>(1, n)) return -10;
    if (ldx < std::max<i64>(1, n)) return -12;
    if (maxiter < 0) return -13;
    if (!(rtol >= 0)) return -14;
    if (!info) return -15;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = oxi_refine_info{};
    info->converged = 1;
    info->rcond_estimate = 1.0;
    if (n == 0 || nrhs == 0) return OXI_OK;
    if (!A) return -4;
    if (!LU) return -6;
    if (!ipiv) return -8;
    if (!B) return -9;
    if (!X) return -11;
    if (nrhs > 65535) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "refine_solve: more than 65535 columns");
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    i64 z = 0;
    OXI_TRY(oxi_first_zero_diag(h, s, LU, ldlu, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "refine_solve: U(%lld,%lld) is zero", (long long)z, (long long)z);
    const size_t sb = al(Solver::bytes(n, std::max<i64>(nrhs, 1))), rb = al(Resid::bytes(n, n, nrhs)),
                 nr = al(size_t(n) * size_t(nrhs) * 8), vb = al(size_t(std::max(n, nrhs)) * 8);
    void* ws;
    OXI_TRY(oxi_workspace2(h, sb + rb + nr + vb + vb + 256, &ws));
    char* p = (char*)ws;
    Solver sv{h, s, n, LU, ldlu};
    OXI_TRY(sv.place(p, std::max<i64>(nrhs, 1), ipiv));
    Resid w;
    w.place(p + sb, n, n, nrhs);
    double* R = (double*)(p + sb + rb);
    double* colsum = (double*)(p + sb + rb + nr);
    double* x = (double*)(p + sb + rb + nr + vb);
    double* scal = (double*)(p + sb + rb + nr + 2 * vb);
    // X = solve(F, B), then X .+= solve(F, residual) until the componentwise
    // error is <= rtol (expert_refinement.jl:89-104)
    hipLaunchKernelGGL(copy_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, B, ldb, X, ldx);
    OXI_TRY(launch_ok(h, "refine"));
    OXI_TRY(sv.solve(0, nrhs, X, ldx));
    double err = INFINITY, rnorm = 0;
    i64 iterations = 0;
    for (i64 it = 0; it <= maxiter; ++it) {
        OXI_TRY(residual(h, s, w, A, lda, X, ldx, B, ldb, R, n, &err, &rnorm));
        iterations = it;
        if (err <= rtol || it == maxiter) break;
        OXI_TRY(sv.solve(0, nrhs, R, n));
        hipLaunchKernelGGL(add_div_rows_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, R, n,
                           (const double*)nullptr, X, ldx);
        OXI_TRY(launch_ok(h, "refine"));
    }
    // condition estimates (not certified, as the reference says)
    double inv_norm = 0, anorm = 0, xnorm = 0, rnorm1 = 0;
    OXI_TRY(inverse_norm(sv, 8, x, colsum, scal, &inv_norm));
    OXI_TRY(norm1(h, s, n, n, A, lda, (double*)w.P, scal, &anorm));   // P has room for n column sums
    OXI_TRY(norm1(h, s, n, nrhs, X, ldx, colsum, scal, &xnorm));
    OXI_TRY(norm1(h, s, n, nrhs, R, n, colsum, scal, &rnorm1));
    OXI_TRY(sv.check());
    const double condition = anorm * inv_norm;
    const double delta = xnorm == 0 ? (rnorm == 0 ? 0.0 : INFINITY) : inv_norm * rnorm1 / xnorm;
    info->iterations = iterations;
    info->converged = err <= rtol;
    info->backward_error = err;
    info->residual_norm = rnorm;
    info->rcond_estimate = condition == 0 ? 1.0 : std::min(1.0, 1.0 / condition);
    info->forward_error_estimate = delta < 1 ? delta / (1 - delta) : INFINITY;
    return OXI_OK;
}

extern "C" int oxi_dexpert_solve(oxi_handle_t h, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                                 const double* B, int64_t ldb, double* X, int64_t ldx, int32_t maxiter, double rtol,
                                 int32_t equilibrate, double* LU, int64_t ldlu, int64_t* ipiv, double* rows,
                                 double* cols, oxi_refine_info* info) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (lda < std::max<i64>(1, n)) return -5;
    if (ldb < std::max<i64>(1, n)) return -7;
    if (ldx < std::max<i64>(1, n)) return -9;
    if (maxiter < 0) return -10;
    if (!(rtol >= 0)) return -11;
    if (ldlu < std::max<i64>(1, n)) return -14;
    if (!info) return -18;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = oxi_refine_info{};
    info->converged = 1;
    info->rcond_estimate = 1.0;
    if (n == 0) return OXI_OK;
    if (!A) return -4;
    if (!B && nrhs) return -6;
    if (!X && nrhs) return -8;
    if (!LU) return -13;
    if (!ipiv) return -15;
    if (!rows) return -16;
    if (!cols) return -17;
    if (nrhs > 65535) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "expert_solve: more than 65535 columns");
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    // 1. scaling (advanced_factorizations.jl:597-614): rows by their largest
    // |entry| (a zero row is SingularException(i)), then columns
    if (equilibrate) {
        void* ws;
        OXI_TRY(oxi_workspace2(h, al(size_t(n) * 8), &ws));
        u64* rmax = (u64*)ws;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, s, rmax, int(std::min<i64>(n, 1 << 30)), u64(0));
        if (n > (i64(1) << 30)) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "expert_solve: n too large");
        hipLaunchKernelGGL(row_absmax_kernel, dim3(unsigned(cdiv(n, TPB)), unsigned(cdiv(n, KC))), dim3(TPB), 0, s, n,
                           A, lda, rmax);
        OXI_TRY(launch_ok(h, "expert scaling"));
        std::vector<double> hr(static_cast<size_t>(n));
        OXI_HIP(h, hipMemcpyAsync(hr.data(), rmax, size_t(n) * 8, hipMemcpyDeviceToHost, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        for (i64 i = 0; i < n; ++i)
            if (hr[size_t(i)] == 0)
                return oxi_set_error(h, OXI_ERR_SINGULAR, "expert_solve: zero row U(%lld,%lld)", (long long)(i + 1),
                                     (long long)(i + 1));
        OXI_HIP(h, hipMemcpyAsync(rows, hr.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(scale_cols_kernel, dim3(unsigned(n)), dim3(TPB), 0, s, n, A, lda, rows, LU, ldlu, cols);
        OXI_TRY(launch_ok(h, "expert scaling"));
    } else {
        std::vector<double> ones(static_cast<size_t>(n), 1.0);
        OXI_HIP(h, hipMemcpyAsync(rows, ones.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_HIP(h, hipMemcpyAsync(cols, ones.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_HIP(h, hipMemcpy2DAsync(LU, size_t(ldlu) * 8, A, size_t(lda) * 8, size_t(n) * 8, size_t(n),
                                    hipMemcpyDeviceToDevice, s));
    }
    // 2. the LU of the scaled matrix (lu_factor with partial pivoting)
    i64 finfo = 0;
    OXI_TRY(oxi_dgetrf(h, n, n, LU, ldlu, ipiv, 0, &finfo));
    if (finfo != 0)
        return oxi_set_error(h, OXI_ERR_SINGULAR, "expert_solve: U(%lld,%lld) is zero", (long long)finfo,
                             (long long)finfo);
    if (nrhs == 0) return OXI_OK;
    const size_t sb = al(Solver::bytes(n, nrhs)), rb = al(Resid::bytes(n, n, nrhs)), nr = al(size_t(n) * size_t(nrhs) * 8);
    void* ws;
    OXI_TRY(oxi_workspace2(h, sb + rb + 2 * nr + 256, &ws));
    char* p = (char*)ws;
    Solver sv{h, s, n, LU, ldlu};
    OXI_TRY(sv.place(p, nrhs, ipiv));
    Resid w;
    w.place(p + sb, n, n, nrhs);
    double* R = (double*)(p + sb + rb);
    double* T = (double*)(p + sb + rb + nr);
    // 3. X = solve(F, B ./ rows) ./ cols, refinement on the original system
    hipLaunchKernelGGL(div_rows_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, B, ldb, rows, T, n);
    OXI_TRY(launch_ok(h, "expert"));
    OXI_TRY(sv.solve(0, nrhs, T, n));
    hipLaunchKernelGGL(div_rows_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, T, n, cols, X, ldx);
    OXI_TRY(launch_ok(h, "expert"));
    double err = INFINITY, rnorm = 0;
    i64 iterations = 0;
    for (i64 it = 0; it <= maxiter; ++it) {
        OXI_TRY(residual(h, s, w, A, lda, X, ldx, B, ldb, R, n, &err, &rnorm));
        iterations = it;
        if (err <= rtol || it == maxiter) break;
        hipLaunchKernelGGL(div_rows_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, R, n, rows, T, n);
        OXI_TRY(launch_ok(h, "expert"));
        OXI_TRY(sv.solve(0, nrhs, T, n));
        hipLaunchKernelGGL(add_div_rows_kernel, dim3(egrid(n * nrhs)), dim3(TPB), 0, s, n, nrhs, T, n, cols, X, ldx);
        OXI_TRY(launch_ok(h, "expert"));
    }
    OXI_TRY(sv.check());
    info->iterations = iterations;
    info->converged = err <= rtol;
    info->backward_error = err;
    info->residual_norm = rnorm;
    info->rcond_estimate = 0;            // not computed by expert_solve
    info->forward_error_estimate = 0;
    return OXI_OK;
}
