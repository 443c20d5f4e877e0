// LU solves (ldiv!(F, B), ldiv!(transpose(F), B): src/api/linalg_api.jl:168-184)
// with the factors of oxi_dgetrf (FP64) or oxi_sgetrf_lowp (FP32).
#include "oxi_internal.hpp"

#include <vector>
#include <algorithm>

namespace {
// Row gather / scatter of a column-major X (n x nrhs) with a 1-based permutation.
template <typename T>
__global__ void permute_rows_kernel(const T* __restrict__ X, i64 ldx, T* __restrict__ Y, i64 ldy, i64 n, i64 nrhs,
                                    const i64* __restrict__ perm, int inverse) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 col = idx / n, i = idx - col * n;
        const i64 pi = perm[i] - 1;
        if (!inverse) Y[i + col * ldy] = X[pi + col * ldx];   // Y = X[p, :]
        else Y[pi + col * ldy] = X[i + col * ldx];            // Y[p, :] = X
    }
}

template <typename T>
__global__ void copy_cols_kernel(const T* __restrict__ Y, i64 ldy, T* __restrict__ B, i64 ldb, i64 n, i64 nrhs) {
    const i64 total = n * nrhs;
    for (i64 idx = i64(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += i64(gridDim.x) * blockDim.x) {
        const i64 col = idx / n, i = idx - col * n;
        B[i + col * ldb] = Y[i + col * ldy];
    }
}

__global__ void diag_zero_f32_kernel(const float* __restrict__ T, i64 ldt, i64 n, i64* __restrict__ out) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x)
        if (T[i + i * ldt] == 0.0f) atomicMin((unsigned long long*)out, (unsigned long long)(i + 1));
}
__global__ void set_i64_kernel(i64* p, i64 v) { *p = v; }

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

int trsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs, const double* A,
              i64 lda, double* B, i64 ldb) {
    return oxi_dtrsm_left(h, s, upper, trans, unit, m, nrhs, A, lda, B, ldb);
}
int trsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs, const float* A,
              i64 lda, float* B, i64 ldb) {
    return oxi_strsm_left(h, s, upper, trans, unit, m, nrhs, A, lda, B, ldb);
}

int chain_triangle(oxi_handle_t h, hipStream_t s, int tk, int second, i64 n, i64 nrhs, const double* LU, i64 lda,
                   double* B, i64 ldb, void* mws) {
    return oxi_trsm_chain(h, s, tk, second, n, nrhs, LU, lda, B, ldb, mws);
}
int chain_triangle(oxi_handle_t h, hipStream_t, int, int, i64, i64, const float*, i64, float*, i64, void*) {
    return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "trsm chain: FP64 only");
}

// Whether a solve takes the chained multi-RHS triangles (their workspace is
// then prepared for the factors).
bool use_chain(i64 n, i64 nrhs, i64 ldb) { return nrhs > OXI_TRSV_MAX_RHS && oxi_trsm_chain_fits(n, nrhs, ldb); }

// Two triangles of one solve: per column with the single-RHS chain kernel
// when there are few right-hand sides (the reference's vector branch uses
// trsv!, src/api/linalg_api.jl:168-184), with the multi-RHS chain up to
// OXI_CHAIN_MAX_RHS (FP64; mws its workspace, prepared for the kind tk),
// else the blocked trsm.
template <typename T>
int triangles(oxi_handle_t h, hipStream_t s, int tk, i64 n, i64 nrhs, const T* LU, i64 lda, T* B, i64 ldb,
              void* ctrl, void* mws, hipEvent_t second_ready) {
    if (ctrl && nrhs <= OXI_TRSV_MAX_RHS) {
        for (i64 j = 0; j < nrhs; ++j) {
            OXI_TRY(oxi_trsv(h, s, tk ? 2 : 0, n, LU, lda, B + j * ldb, ctrl));
            if (second_ready && j == 0) OXI_HIP(h, hipStreamWaitEvent(s, second_ready, 0));
            OXI_TRY(oxi_trsv(h, s, tk ? 3 : 1, n, LU, lda, B + j * ldb, ctrl));
        }
        return OXI_OK;
    }
    if (mws && use_chain(n, nrhs, ldb)) {
        OXI_TRY(chain_triangle(h, s, tk, 0, n, nrhs, LU, lda, B, ldb, mws));
        OXI_TRY(chain_triangle(h, s, tk, 1, n, nrhs, LU, lda, B, ldb, mws));
        return OXI_OK;
    }
    if (!tk) {
        OXI_TRY(trsm_left(h, s, 0, 0, 1, n, nrhs, LU, lda, B, ldb));
        OXI_TRY(trsm_left(h, s, 1, 0, 0, n, nrhs, LU, lda, B, ldb));
    } else {
        OXI_TRY(trsm_left(h, s, 1, 1, 0, n, nrhs, LU, lda, B, ldb));
        OXI_TRY(trsm_left(h, s, 0, 1, 1, n, nrhs, LU, lda, B, ldb));
    }
    return OXI_OK;
}

// The solve proper, with a device permutation and a workspace Y (n x nrhs);
// no synchronisation.
template <typename T>
int getrs_perm(oxi_handle_t h, hipStream_t s, int tk, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* dperm, T* B,
               i64 ldb, T* Y, void* ctrl, void* mws = nullptr, hipEvent_t second_ready = nullptr) {
    const unsigned pg = unsigned(std::min<i64>(cdiv(n * nrhs, 256), 8192));
    if (!tk) {
        // B <- B[p, :] ; L \ ; U \   (src/api/linalg_api.jl:168-174)
        hipLaunchKernelGGL(permute_rows_kernel<T>, dim3(pg), dim3(256), 0, s, B, ldb, Y, n, n, nrhs, dperm, 0);
        hipLaunchKernelGGL(copy_cols_kernel<T>, dim3(pg), dim3(256), 0, s, Y, n, B, ldb, n, nrhs);
        OXI_TRY(launch_check(h, "permute"));
        OXI_TRY(triangles<T>(h, s, 0, n, nrhs, LU, lda, B, ldb, ctrl, mws, second_ready));
    } else {
        // U^T \ ; L^T \ ; inverse interchanges   (src/api/linalg_api.jl:176-184)
        OXI_TRY(triangles<T>(h, s, 1, n, nrhs, LU, lda, B, ldb, ctrl, mws, second_ready));
        hipLaunchKernelGGL(permute_rows_kernel<T>, dim3(pg), dim3(256), 0, s, B, ldb, Y, n, n, nrhs, dperm, 1);
        hipLaunchKernelGGL(copy_cols_kernel<T>, dim3(pg), dim3(256), 0, s, Y, n, B, ldb, n, nrhs);
        OXI_TRY(launch_check(h, "permute"));
    }
    return OXI_OK;
}

int first_zero_diag(oxi_handle_t h, hipStream_t s, const double* T, i64 ldt, i64 n, i64* z) {
    return oxi_first_zero_diag(h, s, T, ldt, n, z);
}
int first_zero_diag(oxi_handle_t h, hipStream_t s, const float* T, i64 ldt, i64 n, i64* z) {
    void* ws;
    OXI_TRY(oxi_workspace(h, 256, &ws));
    i64* d = (i64*)ws;
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, d, i64(0x7fffffffffffffffLL));
    hipLaunchKernelGGL(diag_zero_f32_kernel, dim3(unsigned(std::min<i64>(cdiv(n, 256), 1024))), dim3(256), 0, s, T,
                       ldt, n, d);
    OXI_TRY(launch_check(h, "diag check"));
    i64 v = 0;
    OXI_HIP(h, hipMemcpyAsync(&v, d, sizeof(i64), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    *z = (v == i64(0x7fffffffffffffffLL)) ? 0 : v;
    return OXI_OK;
}

template <typename T>
int getrs_api(oxi_handle_t h, char trans, i64 n, i64 nrhs, const T* LU, i64 lda, const i64* ipiv, T* B, i64 ldb) {
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
    OXI_TRY(first_zero_diag(h, s, LU, lda, n, &z));
    if (z) return oxi_set_error(h, OXI_ERR_SINGULAR, "getrs: U(%lld,%lld) is zero", (long long)z, (long long)z);
    std::vector<i64> perm;
    OXI_TRY(oxi_perm_from_ipiv(h, s, ipiv, n, perm));
    const size_t xbytes = size_t(n) * size_t(nrhs) * sizeof(T);
    void* ws;
    const size_t xpad = (xbytes + 255) & ~size_t(255), ppad = (size_t(n) * 8 + 255) & ~size_t(255);
    const bool chain = nrhs <= OXI_TRSV_MAX_RHS;
    const bool mchain = sizeof(T) == 8 && use_chain(n, nrhs, ldb);
    OXI_TRY(oxi_workspace(h, xpad + ppad + (chain ? oxi_trsv_ws_bytes(n, sizeof(T)) : mchain ? oxi_trsm_chain_ws_bytes(n) : 256),
                          &ws));
    T* Y = (T*)ws;
    i64* dperm = (i64*)((char*)ws + xpad);
    void* ctrl = (char*)ws + xpad + ppad;
    OXI_HIP(h, hipMemcpyAsync(dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
    hipEvent_t second_ready = nullptr;
    if (chain) {
        OXI_TRY(oxi_trsv_begin(h, s, ctrl));
        if constexpr (sizeof(T) == 8) {
            OXI_TRY(oxi_trsv_prepare_split(h, s, tk, n, (const double*)LU, lda, ctrl));
            second_ready = h->ev_b;
        } else {
            OXI_TRY(oxi_trsv_prepare(h, s, tk, n, LU, lda, (T*)nullptr, ctrl));
        }
    }
    if (mchain) {
        OXI_TRY(oxi_trsm_chain_begin(h, s, ctrl));
        OXI_TRY(oxi_trsm_chain_prepare(h, s, tk, n, (const double*)LU, lda, ctrl));
    }
    OXI_TRY(getrs_perm<T>(h, s, tk, n, nrhs, LU, lda, dperm, B, ldb, Y, chain ? ctrl : nullptr, mchain ? ctrl : nullptr,
                          second_ready));
    if (chain) return oxi_trsv_check(h, s, ctrl);    // the chain kernel ran: its status
    if (mchain) return oxi_trsm_chain_check(h, s, ctrl);
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

}  // namespace

// Permutation p (1-based) from the interchanges, applied in order
// (src/lapack/factorizations.jl:72-75); synchronises the stream.
int oxi_perm_from_ipiv(oxi_handle_t h, hipStream_t s, const i64* ipiv, i64 n, std::vector<i64>& perm) {
    std::vector<i64> piv(n);
    OXI_HIP(h, hipMemcpyAsync(piv.data(), ipiv, size_t(n) * 8, hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    perm.resize(n);
    for (i64 i = 0; i < n; ++i) perm[i] = i + 1;
    for (i64 j = 0; j < n; ++j) {
        const i64 p = piv[j] - 1;
        if (p < 0 || p >= n)
            return oxi_set_error(h, -7, "ipiv[%lld] = %lld out of range", (long long)j, (long long)piv[j]);
        std::swap(perm[j], perm[p]);
    }
    return OXI_OK;
}

int oxi_getrs_perm_f64(oxi_handle_t h, hipStream_t s, int trans, i64 n, i64 nrhs, const double* LU, i64 lda,
                       const i64* dperm, double* B, i64 ldb, double* Y, void* ctrl) {
    return getrs_perm<double>(h, s, trans, n, nrhs, LU, lda, dperm, B, ldb, Y, ctrl);
}
int oxi_getrs_perm_f32(oxi_handle_t h, hipStream_t s, int trans, i64 n, i64 nrhs, const float* LU, i64 lda,
                       const i64* dperm, float* B, i64 ldb, float* Y, void* ctrl) {
    return getrs_perm<float>(h, s, trans, n, nrhs, LU, lda, dperm, B, ldb, Y, ctrl);
}

extern "C" int oxi_dgetrs(oxi_handle_t h, char trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                          const int64_t* ipiv, double* B, int64_t ldb) {
    return getrs_api<double>(h, trans, n, nrhs, LU, lda, ipiv, B, ldb);
}

extern "C" int oxi_sgetrs(oxi_handle_t h, char trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
                          const int64_t* ipiv, float* B, int64_t ldb) {
    return getrs_api<float>(h, trans, n, nrhs, LU, lda, ipiv, B, ldb);
}

// ---------------------------------------------------------------------------
// Solve plans: what a solve with given factors prepares -- the permutation on
// the device and the inverses of the 128 x 128 diagonal blocks of both
// triangles for both solve kinds -- made once per factor and reused by every
// later solve with those factors.
struct oxi_getrs_plan_s {
    int device = 0;
    i64 n = 0;
    const double* LU = nullptr;
    i64 lda = 0;
    i64* dperm = nullptr;
    void* ctrl[2] = {nullptr, nullptr};
    bool prepared[2] = {false, false};   // the inverses of kind 'N' (0) / 'T' (1), made on first use
    void* mws[2] = {nullptr, nullptr};   // the multi-RHS chain's workspace per kind, made and prepared on first use
    i64 zero_pivot = 0;       // first zero on U's diagonal (1-based), 0 if none
};

extern "C" int oxi_dgetrs_plan_create(oxi_handle_t h, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv,
                                      oxi_getrs_plan_t* plan) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (lda < std::max<i64>(1, n)) return -4;
    if (!plan) return -6;
    *plan = nullptr;
    if (n > 0 && !LU) return -3;
    if (n > 0 && !ipiv) return -5;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    auto* p = new oxi_getrs_plan_s;
    p->device = h->device; p->n = n; p->LU = LU; p->lda = lda;
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(s);
        if (p->dperm) (void)hipFree(p->dperm);
        for (auto c : p->ctrl) if (c) (void)hipFree(c);
        delete p;
        return rc;
    };
    if (n == 0) { *plan = p; return OXI_OK; }
    int rc = first_zero_diag(h, s, LU, lda, n, &p->zero_pivot);
    if (rc != OXI_OK) return fail(rc);
    std::vector<i64> perm;
    if ((rc = oxi_perm_from_ipiv(h, s, ipiv, n, perm)) != OXI_OK) return fail(rc);
    const size_t cb = oxi_trsv_ws_bytes(n, sizeof(double));
    if (hipMalloc(&p->dperm, size_t(n) * 8) != hipSuccess || hipMalloc(&p->ctrl[0], cb) != hipSuccess ||
        hipMalloc(&p->ctrl[1], cb) != hipSuccess)
        return fail(oxi_set_error(h, OXI_ERR_ALLOC, "getrs plan: allocation failed"));
    if (hipMemcpyAsync(p->dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(oxi_set_error(h, OXI_ERR_HIP, "getrs plan: copy failed"));
    if (hipStreamSynchronize(s) != hipSuccess) return fail(oxi_set_error(h, OXI_ERR_HIP, "getrs plan: sync failed"));
    *plan = p;
    return OXI_OK;
}

extern "C" int oxi_dgetrs_plan(oxi_handle_t h, oxi_getrs_plan_t plan, char trans, int64_t nrhs, double* B,
                               int64_t ldb) {
    if (!h) return -1;
    if (!plan) return -2;
    const int tk = trans_kind(trans);
    if (tk < 0) return -3;
    if (nrhs < 0) return -4;
    const i64 n = plan->n;
    if (ldb < std::max<i64>(1, n)) return -6;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0 || nrhs == 0) return OXI_OK;
    if (!B) return -5;
    if (plan->device != h->device) return oxi_set_error(h, -2, "getrs plan: made on another device");
    if (plan->zero_pivot)
        return oxi_set_error(h, OXI_ERR_SINGULAR, "getrs: U(%lld,%lld) is zero", (long long)plan->zero_pivot,
                             (long long)plan->zero_pivot);
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    void* ws;
    OXI_TRY(oxi_workspace(h, size_t(n) * size_t(nrhs) * 8 + 256, &ws));
    const bool chain = nrhs <= OXI_TRSV_MAX_RHS;
    void* ctrl = chain ? plan->ctrl[tk] : nullptr;
    if (chain) {
        OXI_TRY(oxi_trsv_begin(h, s, ctrl));
        if (!plan->prepared[tk]) {
            OXI_TRY(oxi_trsv_prepare(h, s, tk, n, plan->LU, plan->lda, (double*)nullptr, ctrl));
            plan->prepared[tk] = true;
        }
    }
    const bool mchain = use_chain(n, nrhs, ldb);
    if (mchain) {
        if (!plan->mws[tk]) {
            void* m = nullptr;
            if (hipMalloc(&m, oxi_trsm_chain_ws_bytes(n)) != hipSuccess)
                return oxi_set_error(h, OXI_ERR_ALLOC, "getrs plan: allocation failed");
            if (oxi_trsm_chain_prepare(h, s, tk, n, plan->LU, plan->lda, m) != OXI_OK) {
                (void)hipStreamSynchronize(s);
                (void)hipFree(m);
                return OXI_ERR_HIP;
            }
            plan->mws[tk] = m;
        }
        OXI_TRY(oxi_trsm_chain_begin(h, s, plan->mws[tk]));
    }
    OXI_TRY(getrs_perm<double>(h, s, tk, n, nrhs, plan->LU, plan->lda, plan->dperm, B, ldb, (double*)ws, ctrl,
                               mchain ? plan->mws[tk] : nullptr));
    if (chain) return oxi_trsv_check(h, s, ctrl);
    if (mchain) return oxi_trsm_chain_check(h, s, plan->mws[tk]);
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}

extern "C" int oxi_dgetrs_plan_destroy(oxi_handle_t h, oxi_getrs_plan_t plan) {
    if (!h) return -1;
    if (!plan) return OXI_OK;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    (void)hipSetDevice(plan->device);
    (void)hipStreamSynchronize(h->stream);
    if (plan->dperm) (void)hipFree(plan->dperm);
    for (auto c : plan->ctrl) if (c) (void)hipFree(c);
    for (auto m : plan->mws) if (m) (void)hipFree(m);
    delete plan;
    return OXI_OK;
}

// Pullback of X = A \ B with the factors of oxi_dgetrf: Bbar = A^-T Xbar
// (ldiv!(transpose(F), .), src/api/linalg_api.jl:176-184) and
// Abar = -Bbar X^T (a rank-nrhs FP64 MFMA GEMM).
extern "C" int oxi_dgesv_adjoint(oxi_handle_t h, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                                 const int64_t* ipiv, const double* X, int64_t ldx, const double* Xbar,
                                 int64_t ldxbar, double* Bbar, int64_t ldbbar, double* Abar, int64_t ldabar) {
    if (!h) return -1;
    if (n < 0) return -2;
    if (nrhs < 0) return -3;
    if (lda < std::max<i64>(1, n)) return -5;
    if (ldx < std::max<i64>(1, n)) return -8;
    if (ldxbar < std::max<i64>(1, n)) return -10;
    if (ldbbar < std::max<i64>(1, n)) return -12;
    if (Abar && ldabar < std::max<i64>(1, n)) return -14;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (n == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    if (nrhs == 0) {
        if (Abar) OXI_TRY(oxi_launch_dscale(s, n, n, 0.0, Abar, ldabar));
        OXI_HIP(h, hipStreamSynchronize(s));
        return OXI_OK;
    }
    if (!LU) return -4;
    if (!ipiv) return -6;
    if (!X) return -7;
    if (!Xbar) return -9;
    if (!Bbar) return -11;
    // Bbar <- Xbar, then the transposed solve in place
    OXI_HIP(h, hipMemcpy2DAsync(Bbar, size_t(ldbbar) * 8, Xbar, size_t(ldxbar) * 8, size_t(n) * 8, size_t(nrhs),
                                hipMemcpyDeviceToDevice, s));
    OXI_TRY(getrs_api<double>(h, 'T', n, nrhs, LU, lda, ipiv, Bbar, ldbbar));
    if (Abar) {
        OXI_TRY(oxi_launch_dgemm(h, s, 0, 1, n, n, nrhs, -1.0, Bbar, ldbbar, X, ldx, 0.0, Abar, ldabar));
        OXI_HIP(h, hipStreamSynchronize(s));
    }
    return OXI_OK;
}
