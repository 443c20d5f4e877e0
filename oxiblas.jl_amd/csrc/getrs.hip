// LU solves (ldiv!(F, B), ldiv!(transpose(F), B): src/api/linalg_api.jl:168-184)
// with the factors of oxi_dgetrf.
#include "oxi_internal.hpp"

#include <vector>
#include <algorithm>

namespace {
// Row gather / scatter of a column-major X (n x nrhs) with a 1-based permutation.
__global__ void permute_rows_kernel(const double* __restrict__ X, i64 ldx, double* __restrict__ Y, i64 ldy, i64 n,
                                    i64 nrhs, const i64* __restrict__ perm, int inverse) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 col = idx / n, i = idx - col * n;
        const i64 pi = perm[i] - 1;
        if (!inverse) Y[i + col * ldy] = X[pi + col * ldx];   // Y = X[p, :]
        else Y[pi + col * ldy] = X[i + col * ldx];            // Y[p, :] = X
    }
}


int trans_kind(char t) {
    if (t == 'N' || t == 'n') return 0;
    if (t == 'T' || t == 't' || t == 'C' || t == 'c') return 1;
    return -1;
}
int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}
}  // namespace

extern "C" int oxi_dgetrs(oxi_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                          const int64_t* ipiv, double* B, int64_t ldb) {
    if (!h) return -1;
    const int tk = trans_kind(trans);
    if (tk < 0) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < std::max<i64>(1, n)) return -6;
    if (ldb < std::max<i64>(1, n)) return -9;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0 || nrhs == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    i64 z = 0;
    OXI_TRY(oxi_first_zero_diag(h, s, LU, lda, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "getrs: U(%lld,%lld) is zero", (long long)z, (long long)z);
    // permutation from the interchanges (src/lapack/factorizations.jl:72-75)
    std::vector<i64> piv(n), perm(n);
    OXI_HIP(h, hipMemcpyAsync(piv.data(), ipiv, size_t(n) * 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    for (i64 i = 0; i < n; ++i) perm[i] = i + 1;
    for (i64 j = 0; j < n; ++j) {
        const i64 p = piv[j] - 1;
        if (p < 0 || p >= n) return oxi_set_error(h, -7, "getrs: ipiv[%lld] = %lld out of range", (long long)j, (long long)piv[j]);
        std::swap(perm[j], perm[p]);
    }
    const size_t xbytes = size_t(n) * size_t(nrhs) * 8;
    void* ws;
    OXI_TRY(oxi_workspace(h, xbytes + size_t(n) * 8 + 512, &ws));
    double* Y = (double*)ws;
    i64* dperm = (i64*)((char*)ws + ((xbytes + 255) & ~size_t(255)));
    OXI_HIP(h, hipMemcpyAsync(dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
    const unsigned pg = unsigned(std::min<i64>(cdiv(n * nrhs, 256), 8192));
    if (!tk) {
        // B <- B[p, :] ; L \ ; U \   (src/api/linalg_api.jl:168-174)
        hipLaunchKernelGGL(permute_rows_kernel, dim3(pg), dim3(256), 0, s, B, ldb, Y, n, n, nrhs, dperm, 0);
        OXI_TRY(launch_check(h, "permute"));
        OXI_HIP(h, hipMemcpy2DAsync(B, ldb * 8, Y, n * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, s));
        OXI_TRY(oxi_dtrsm_left(h, s, 0, 0, 1, n, nrhs, LU, lda, B, ldb));
        OXI_TRY(oxi_dtrsm_left(h, s, 1, 0, 0, n, nrhs, LU, lda, B, ldb));
    } else {
        // U^T \ ; L^T \ ; inverse interchanges   (src/api/linalg_api.jl:176-184)
        OXI_TRY(oxi_dtrsm_left(h, s, 1, 1, 0, n, nrhs, LU, lda, B, ldb));
        OXI_TRY(oxi_dtrsm_left(h, s, 0, 1, 1, n, nrhs, LU, lda, B, ldb));
        hipLaunchKernelGGL(permute_rows_kernel, dim3(pg), dim3(256), 0, s, B, ldb, Y, n, n, nrhs, dperm, 1);
        OXI_TRY(launch_check(h, "permute"));
        OXI_HIP(h, hipMemcpy2DAsync(B, ldb * 8, Y, n * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, s));
    }
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}
