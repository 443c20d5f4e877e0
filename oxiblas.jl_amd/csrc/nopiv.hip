// Low-precision LU without pivoting after a random butterfly transform (RBT):
// the communication-free alternative to the partially pivoted FP32 factor of
// the mixed-precision solve (mixed_solve, src/lapack/extended_lapack.jl:114-239).
//
// Partial pivoting needs one device-wide pivot search per column; the FP32
// factor of mixed_solve is then bounded by that chain exactly like the FP64
// one.  A two-sided random butterfly transform A' = U^T A V (Parker 1995;
// Baboulin, Dongarra, Herrmann, Tomov 2013) makes pivoting unnecessary with
// high probability, so the factor runs as a blocked LU with no row
// interchanges: one workgroup factors each 128 x 128 diagonal block in LDS,
// the panel solves are triangular solves with its triangles, and the
// trailing update is the FP32 or BF16/FP16/FP8 GEMM.
// The FP64 refinement of the mixed solve then restores full accuracy:
// A^-1 r = V A'^-1 U^T r, two O(n) butterfly applications around each solve.
//
// Butterflies: level 0 pairs (i, i + n/2); level 1 pairs (i, i + n/4) inside
// each half.  A butterfly B = (1/sqrt2) [R0 R1; R0 -R1] with random diagonal
// R0, R1 (entries exp(u/10), u uniform in [-1/2, 1/2], from a fixed seed) acts
// on a pair (x_i, x_{i+d}) as
//   B^T form:  y_i = c R[i] (x_i + x_{i+d}),   y_{i+d} = c R[i+d] (x_i - x_{i+d})
//   B form:    y_i = c (R[i] x_i + R[i+d] x_{i+d}),  y_{i+d} = c (R[i] x_i - R[i+d] x_{i+d})
// with R[i] = R0 entry, R[i+d] = R1 entry, c = 1/sqrt2.  U = diag(B1a, B1b) B0
// and V alike (independent draws); U^T A applies level 1 then level 0 in B^T
// form on rows, A V the same on columns (which is also the B^T form), V z
// applies level 0 then level 1 in B form.  n must be a multiple of 4.
#include "oxi_internal.hpp"

#include <cmath>
#include <random>
#include <vector>

namespace {

constexpr int NBD = 128;   // diagonal block of the factorization
constexpr int DT = 1024;   // threads of the diagonal-block kernel

// ---------------------------------------------------------------------------
// Butterfly transforms
// ---------------------------------------------------------------------------

// One level on both sides of the n x n matrix A (column-major): the 2 x 2
// groups {i, i+d} x {j, j+d} with i mod 2d < d, j mod 2d < d; rows get the
// B^T form with RU, columns the B^T form with RV.
__global__ void rbt_matrix_level_kernel(float* __restrict__ A, i64 lda, i64 n, i64 d, const float* __restrict__ RU,
                                        const float* __restrict__ RV) {
    const i64 half = n / 2;   // groups per dimension
    const i64 gi = i64(blockIdx.x) * blockDim.x + threadIdx.x;   // row group (fast: coalesced)
    const i64 gj = blockIdx.y + i64(blockIdx.z) * 65535;           // column group
    if (gi >= half || gj >= half) return;
    const i64 i = (gi / d) * 2 * d + gi % d, j = (gj / d) * 2 * d + gj % d;
    const float c = 0.70710678118654752f;
    const float a00 = A[i + j * lda], a10 = A[i + d + j * lda];
    const float a01 = A[i + (j + d) * lda], a11 = A[i + d + (j + d) * lda];
    // rows: B^T form
    const float r0 = c * RU[i], r1 = c * RU[i + d];
    const float b00 = r0 * (a00 + a10), b10 = r1 * (a00 - a10);
    const float b01 = r0 * (a01 + a11), b11 = r1 * (a01 - a11);
    // columns: B^T form
    const float s0 = c * RV[j], s1 = c * RV[j + d];
    A[i + j * lda] = s0 * (b00 + b01);
    A[i + (j + d) * lda] = s1 * (b00 - b01);
    A[i + d + j * lda] = s0 * (b10 + b11);
    A[i + d + (j + d) * lda] = s1 * (b10 - b11);
}

// One level on vectors (n x ncols, column-major, ld): tform = B^T form, else B form.
template <typename TV>
__global__ void rbt_vec_level_kernel(TV* __restrict__ X, i64 ldx, i64 n, i64 ncols, i64 d, const TV* __restrict__ R,
                                     int tform) {
    const i64 g = i64(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 col = blockIdx.y;
    if (g >= n / 2 || col >= ncols) return;
    const i64 i = (g / d) * 2 * d + g % d;
    TV* x = X + col * ldx;
    const TV c = TV(0.70710678118654752440);
    const TV x0 = x[i], x1 = x[i + d];
    if (tform) {
        x[i] = c * R[i] * (x0 + x1);
        x[i + d] = c * R[i + d] * (x0 - x1);
    } else {
        const TV y0 = R[i] * x0, y1 = R[i + d] * x1;
        x[i] = c * (y0 + y1);
        x[i + d] = c * (y0 - y1);
    }
}

// ---------------------------------------------------------------------------
// Diagonal block: LU without pivoting in LDS.  info: first zero pivot
// (global, 1-based) if none before.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(DT) void nopiv_diag_kernel(float* __restrict__ A, i64 lda, i64 k0, int nb, i64* info) {
    __shared__ float B[NBD][NBD + 1];   // B[i][c] = A(k0 + i, k0 + c)
    const int t = threadIdx.x;
    for (int e = t; e < NBD * NBD; e += DT) {
        const int i = e % NBD, c = e / NBD;
        B[i][c] = (i < nb && c < nb) ? A[(k0 + i) + (k0 + c) * lda] : (i == c ? 1.0f : 0.0f);
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const float d = B[j][j];
        if (d == 0.0f) {
            if (t == 0 && *info == 0) *info = k0 + j + 1;
            __syncthreads();
            continue;
        }
        for (int i = j + 1 + t; i < nb; i += DT) B[i][j] = B[i][j] / d;
        __syncthreads();
        const int rem = nb - 1 - j;
        for (int e = t; e < rem * rem; e += DT) {
            const int i = j + 1 + e % rem, c = j + 1 + e / rem;
            B[i][c] = __builtin_fmaf(-B[i][j], B[j][c], B[i][c]);
        }
        __syncthreads();
    }
    for (int e = t; e < nb * nb; e += DT) {
        const int i = e % nb, c = e / nb;
        A[(k0 + i) + (k0 + c) * lda] = B[i][c];
    }
}

int check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

}  // namespace

// Random butterflies of one solve: R[level][side][n] (side 0 = U, 1 = V),
// FP32 and FP64 copies; deterministic for a given n.
int oxi_rbt_make(oxi_handle_t h, hipStream_t s, i64 n, float* R32, double* R64) {
    std::mt19937_64 gen(0x0B1A5ull ^ uint64_t(n));
    std::uniform_real_distribution<double> u(-0.5, 0.5);
    std::vector<double> r(size_t(4) * n);
    for (auto& v : r) v = std::exp(u(gen) / 10.0);
    std::vector<float> rf(r.begin(), r.end());
    std::vector<double> rd(rf.begin(), rf.end());   // the same (FP32-exact) values in both copies
    OXI_HIP(h, hipMemcpyAsync(R32, rf.data(), rf.size() * 4, hipMemcpyHostToDevice, s));
    OXI_HIP(h, hipMemcpyAsync(R64, rd.data(), rd.size() * 8, hipMemcpyHostToDevice, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

// A <- U^T A V (n x n FP32, n % 4 == 0)
int oxi_rbt_matrix(oxi_handle_t h, hipStream_t s, i64 n, float* A, i64 lda, const float* R32) {
    if (n % 4) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "rbt: n = %lld is not a multiple of 4", (long long)n);
    const i64 half = n / 2;
    const dim3 g(unsigned(cdiv(half, 256)), unsigned(std::min<i64>(half, 65535)), unsigned(cdiv(half, 65535)));
    // level 1 (d = n/4) first, then level 0 (d = n/2)
    hipLaunchKernelGGL(rbt_matrix_level_kernel, g, dim3(256), 0, s, A, lda, n, n / 4, R32 + 2 * n, R32 + 3 * n);
    hipLaunchKernelGGL(rbt_matrix_level_kernel, g, dim3(256), 0, s, A, lda, n, n / 2, R32, R32 + n);
    return check(h, "rbt matrix");
}

// X <- U^T X (ut = 1) or X <- V X (ut = 0), X n x ncols
template <typename TV>
int rbt_vec_t(oxi_handle_t h, hipStream_t s, int ut, i64 n, i64 ncols, TV* X, i64 ldx, const TV* R) {
    const dim3 g(unsigned(cdiv(n / 2, 256)), unsigned(ncols));
    if (ut) {
        hipLaunchKernelGGL(rbt_vec_level_kernel<TV>, g, dim3(256), 0, s, X, ldx, n, ncols, n / 4, R + 2 * n, 1);
        hipLaunchKernelGGL(rbt_vec_level_kernel<TV>, g, dim3(256), 0, s, X, ldx, n, ncols, n / 2, R, 1);
    } else {
        hipLaunchKernelGGL(rbt_vec_level_kernel<TV>, g, dim3(256), 0, s, X, ldx, n, ncols, n / 2, R + n, 0);
        hipLaunchKernelGGL(rbt_vec_level_kernel<TV>, g, dim3(256), 0, s, X, ldx, n, ncols, n / 4, R + 3 * n, 0);
    }
    return check(h, "rbt vector");
}
int oxi_rbt_vec(oxi_handle_t h, hipStream_t s, int ut, i64 n, i64 ncols, float* X, i64 ldx, const float* R32) {
    return rbt_vec_t<float>(h, s, ut, n, ncols, X, ldx, R32);
}
int oxi_rbt_vec(oxi_handle_t h, hipStream_t s, int ut, i64 n, i64 ncols, double* X, i64 ldx, const double* R64) {
    return rbt_vec_t<double>(h, s, ut, n, ncols, X, ldx, R64);
}

size_t oxi_getrf_nopiv_ws(int prec, i64 n) {
    return (prec >= 0 ? oxi_lowp_update_ws(prec, n, n, NBD) : 0) + 256;
}

// LU without pivoting of the n x n FP32 matrix A (column-major), trailing
// updates in prec (-1 FP32, else BF16/FP16/FP8).  info: first zero pivot
// (1-based) or 0; the factorization goes on past a zero pivot (the caller
// treats it as a failure).
int oxi_getrf_nopiv32(oxi_handle_t h, hipStream_t s, int prec, i64 n, float* A, i64 lda, i64* dinfo, void* ws,
                      size_t ws_bytes) {
    if (ws_bytes < oxi_getrf_nopiv_ws(prec, n)) return oxi_set_error(h, OXI_ERR_ALLOC, "getrf_nopiv: workspace too small");
    for (i64 k = 0; k < n; k += NBD) {
        const int nb = int(std::min<i64>(NBD, n - k));
        const i64 r = n - k - nb;   // rows / columns after the block
        hipLaunchKernelGGL(nopiv_diag_kernel, dim3(1), dim3(DT), 0, s, A, lda, k, nb, dinfo);
        OXI_TRY(check(h, "nopiv diag"));
        if (r <= 0) break;
        float* A11 = A + k + k * lda;
        float* A21 = A + (k + nb) + k * lda;
        float* A12 = A + k + (k + nb) * lda;
        float* A22 = A + (k + nb) + (k + nb) * lda;
        // L21 = A21 U11^-1, U12 = L11^-1 A12 (substitution, in place)
        OXI_TRY(oxi_strsm_right(h, s, 1, 0, 0, r, nb, A11, lda, A21, lda));
        OXI_TRY(oxi_strsm_left(h, s, 0, 0, 1, nb, r, A11, lda, A12, lda));
        // A22 -= L21 U12
        if (prec < 0)
            OXI_TRY(oxi_launch_sgemm(h, s, 0, 0, r, r, nb, -1.0f, A21, lda, A12, lda, 1.0f, A22, lda));
        else
            OXI_TRY(oxi_lowp_update(h, s, prec, r, r, nb, -1.0f, A21, lda, A12, lda, A22, lda, ws, ws_bytes));
    }
    return OXI_OK;
}
