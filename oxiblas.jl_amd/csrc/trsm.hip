// Triangular solves (trsm!, src/blas/level3.jl:309-544) on gfx950: leaves of
// at most 64 rows/columns solved in registers or with one wave per right-hand
// side, and the reference's recursive halving with FP64 MFMA GEMM updates in
// between (_trsm_blocks_left! / _trsm_blocks_right!, level3.jl:338-376).
#include "oxi_internal.hpp"

#include <algorithm>

namespace {
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr int TRL = 64;

// Right side: X (rows x n) <- X * op(T)^-1, n <= 64.  Element order matches
// the reference's right leaf (src/blas/level3.jl:319-333):
// op(T) upper: x_c = (b_c - sum_{p<c} x_p U[p][c]) / U[c][c], c ascending;
// op(T) lower: x_c = (b_c - sum_{p>c} x_p L[p][c]) / L[c][c], c descending.
// One wave per row, lane q = column q: lane q keeps column q of op(T) in
// registers, and the solve is a column sweep (step c: lane c finishes x_c,
// the wave reads it with readlane, every later lane applies one FMA), so each
// element sees exactly the operations above, in that order.  A wave takes
// `rpw` consecutive rows in turn.
__device__ __forceinline__ double rlane(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(u), l);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(u >> 32), l);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ float rlane(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), l));
}

template <typename TY, bool LOWER, bool UNIT, bool TRANS, int NW>
__global__ __launch_bounds__(256) void trsm_right_leaf(TY* __restrict__ X, i64 ldx, i64 rows, int rpw,
                                                       const TY* __restrict__ T, i64 ldt) {
    const int lane = threadIdx.x & 63;
    const i64 wave = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    const i64 r0 = wave * rpw;
    if (r0 >= rows) return;
    const int q = lane < NW ? lane : NW - 1;          // lanes beyond the width mirror the last column
    TY u[NW];                                         // u[c] = op(T)[c][q]
#pragma unroll
    for (int c = 0; c < NW; ++c) u[c] = TRANS ? T[q + i64(c) * ldt] : T[c + i64(q) * ldt];
    const i64 r1 = min(rows, r0 + rpw);
    for (i64 r = r0; r < r1; ++r) {
        TY x = X[r + i64(q) * ldx];
        if (!LOWER) {
#pragma unroll
            for (int c = 0; c < NW; ++c) {
                if (!UNIT && lane == c) x = x / u[c];
                const TY xc = rlane(x, c);
                if (lane > c) x = fma_t(-xc, u[c], x);
            }
        } else {
#pragma unroll
            for (int c = NW - 1; c >= 0; --c) {
                if (!UNIT && lane == c) x = x / u[c];
                const TY xc = rlane(x, c);
                if (lane < c) x = fma_t(-xc, u[c], x);
            }
        }
        if (lane < NW) X[r + i64(q) * ldx] = x;
    }
}

// Any width n <= 64: the width-64 layout with the steps past n skipped.
template <typename TY, bool LOWER, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void trsm_right_leaf_any(TY* __restrict__ X, i64 ldx, i64 rows, int rpw, int n,
                                                           const TY* __restrict__ T, i64 ldt) {
    const int lane = threadIdx.x & 63;
    const i64 wave = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    const i64 r0 = wave * rpw;
    if (r0 >= rows) return;
    const int q = lane < n ? lane : n - 1;
    TY u[TRL];
#pragma unroll
    for (int c = 0; c < TRL; ++c) u[c] = c < n ? (TRANS ? T[q + i64(c) * ldt] : T[c + i64(q) * ldt]) : TY(0);
    const i64 r1 = min(rows, r0 + rpw);
    for (i64 r = r0; r < r1; ++r) {
        TY x = X[r + i64(q) * ldx];
        if (!LOWER) {
#pragma unroll
            for (int c = 0; c < TRL; ++c) {
                if (c < n) {
                    if (!UNIT && lane == c) x = x / u[c];
                    const TY xc = rlane(x, c);
                    if (lane > c) x = fma_t(-xc, u[c], x);
                }
            }
        } else {
#pragma unroll
            for (int c = TRL - 1; c >= 0; --c) {
                if (c < n) {
                    if (!UNIT && lane == c) x = x / u[c];
                    const TY xc = rlane(x, c);
                    if (lane < c) x = fma_t(-xc, u[c], x);
                }
            }
        }
        if (lane < n) X[r + i64(q) * ldx] = x;
    }
}

template <typename TY, bool LOWER, bool UNIT, bool TRANS>
void launch_right_leaf_w(hipStream_t s, i64 rows, int n, const TY* T, i64 ldt, TY* X, i64 ldx) {
    // about 1024 waves in flight; consecutive rows per wave share cache lines
    const int rpw = int(std::max<i64>(1, cdiv(rows, 1024)));
    const i64 waves = cdiv(rows, rpw);
    const dim3 grid(unsigned(cdiv(waves, 4))), block(256);
    if (n == 64) hipLaunchKernelGGL((trsm_right_leaf<TY, LOWER, UNIT, TRANS, 64>), grid, block, 0, s, X, ldx, rows, rpw, T, ldt);
    else if (n == 32) hipLaunchKernelGGL((trsm_right_leaf<TY, LOWER, UNIT, TRANS, 32>), grid, block, 0, s, X, ldx, rows, rpw, T, ldt);
    else if (n == 16) hipLaunchKernelGGL((trsm_right_leaf<TY, LOWER, UNIT, TRANS, 16>), grid, block, 0, s, X, ldx, rows, rpw, T, ldt);
    else if (n == 8) hipLaunchKernelGGL((trsm_right_leaf<TY, LOWER, UNIT, TRANS, 8>), grid, block, 0, s, X, ldx, rows, rpw, T, ldt);
    else hipLaunchKernelGGL((trsm_right_leaf_any<TY, LOWER, UNIT, TRANS>), grid, block, 0, s, X, ldx, rows, rpw, n, T, ldt);
}

template <typename TY, bool LOWER>
void launch_right_leaf(hipStream_t s, int unit, int trans, i64 rows, int n, const TY* T, i64 ldt, TY* X, i64 ldx) {
    if (unit) {
        if (trans) launch_right_leaf_w<TY, LOWER, true, true>(s, rows, n, T, ldt, X, ldx);
        else launch_right_leaf_w<TY, LOWER, true, false>(s, rows, n, T, ldt, X, ldx);
    } else {
        if (trans) launch_right_leaf_w<TY, LOWER, false, true>(s, rows, n, T, ldt, X, ldx);
        else launch_right_leaf_w<TY, LOWER, false, false>(s, rows, n, T, ldt, X, ldx);
    }
}

// Inverses of the 64 x 64 diagonal blocks of an upper op(T): block b's
// V = U_bb^-1 goes to inv + 4096 b, column-major (ld 64, zeros below the
// diagonal); with strip > 0, to its place in a column-major strip of ld
// `ldv` whose rows z*strip .. (z+1)*strip - 1 hold the strip x strip block z
// (inv + 64 b + ((64 b) % strip) * ldv).  Thread j computes column j right-looking, with the column's
// 64 partial sums in registers: for k = j down to 0, V[k][j] = acc_k /
// U[k][k], then acc_i -= U[i][k] V[k][j] for every i < k.  The k steps form
// the only dependent chain (64 divisions); the i updates of a step are
// independent.  Lanes with k > j take no update at step k (the terms the
// substitution V[i][j] = (delta_ij - sum_{k=i+1..j} U[i][k] V[k][j]) /
// U[i][i] does not have), so only the order of each sum differs from it.
template <bool UNIT, bool TRANS>
__global__ __launch_bounds__(64) void trsm_inv64_kernel(const double* __restrict__ T, i64 ldt,
                                                        double* __restrict__ inv, i64 ldv = 64, int strip = 0) {
    __shared__ double Ut[64][65];   // Ut[c][r] = op(T)[r][c]: columns contiguous
    __shared__ double V[64][65];
    const int tj = threadIdx.x;
    const i64 o = i64(blockIdx.x) * 64;
    // T's element (o + tj, o + q), the thread index along the contiguous
    // direction, 64 loads per thread in flight; it is op(T)[tj][q], or
    // op(T)[q][tj] when TRANS
    double r[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) r[q] = T[(o + tj) + (o + q) * ldt];
#pragma unroll
    for (int q = 0; q < 64; ++q) {
        if (TRANS) Ut[tj][q] = r[q];   // element (row q, column tj) of op(T)
        else Ut[q][tj] = r[q];         // element (row tj, column q)
    }
    __syncthreads();
    const int j = tj;
    double acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = i == j ? 1.0 : 0.0;
#pragma unroll
    for (int k = 63; k >= 0; --k) {
        const double x = UNIT ? acc[k] : acc[k] / Ut[k][k];
        V[k][j] = k > j ? 0.0 : x;
        if (k <= j) {
#pragma unroll
            for (int i = 0; i < k; ++i) acc[i] = __builtin_fma(-Ut[k][i], x, acc[i]);
        }
    }
    __syncthreads();
    // column-major store, the thread index along rows
    const i64 d = 64 * i64(blockIdx.x);
    double* out = strip > 0 ? inv + d + (d % strip) * ldv : inv + 64 * d;
    const i64 ld = strip > 0 ? ldv : 64;
#pragma unroll 8
    for (int c = 0; c < 64; ++c) out[tj + ld * c] = V[tj][c];
}

// Left side: X (n x nrhs) <- op(T)^-1 X, one wave per right-hand side, lane =
// row.  LOWER refers to op(T).  Per element: s = b_i; s = fma(-T[i][p], x_p, s)
// in order of p; s / T[i][i] when not unit (src/blas/level3.jl:309-318,
// src/blas/triangular_microkernels.jl:31-59).
template <typename TY, bool LOWER, bool UNIT>
__global__ __launch_bounds__(256) void trsm_left_leaf(TY* __restrict__ X, i64 ldx, i64 nrhs, int n,
                                                      const TY* __restrict__ T, i64 ldt, int trans) {
    __shared__ TY S[TRL * (TRL + 1)];  // S[p*(TRL+1) + i] = op(T)[i][p]
    for (int e = threadIdx.x; e < TRL * TRL; e += blockDim.x) {
        const int pp = e / TRL, i = e % TRL;
        TY val = TY(0);
        if (pp < n && i < n) val = trans ? T[pp + i64(i) * ldt] : T[i + i64(pp) * ldt];
        S[pp * (TRL + 1) + i] = val;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const i64 col = i64(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (col >= nrhs) return;
    TY* xc = X + col * ldx;
    TY x = lane < n ? xc[lane] : TY(0);
    if (LOWER) {
        for (int pp = 0; pp < n; ++pp) {
            if (!UNIT && lane == pp) x = x / S[pp * (TRL + 1) + pp];
            const TY xp = __shfl(x, pp);
            if (lane > pp) x = fma_t(-S[pp * (TRL + 1) + lane], xp, x);
        }
    } else {
        for (int pp = n - 1; pp >= 0; --pp) {
            if (!UNIT && lane == pp) x = x / S[pp * (TRL + 1) + pp];
            const TY xp = __shfl(x, pp);
            if (lane < pp) x = fma_t(-S[pp * (TRL + 1) + lane], xp, x);
        }
    }
    if (lane < n) xc[lane] = x;
}

// Diagonal check for the non-unit solves: first zero on the diagonal of T (1-based) or 0.
__global__ void diag_zero_kernel(const double* __restrict__ T, i64 ldt, i64 n, i64* __restrict__ out) {
    for (i64 i = i64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += i64(gridDim.x) * blockDim.x)
        if (T[i + i * ldt] == 0.0) atomicMin((unsigned long long*)out, (unsigned long long)(i + 1));
}
__global__ void set_i64_kernel(i64* p, i64 v) { *p = v; }


int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}
}  // namespace

template <typename TY>
int gemm_t(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, TY alpha, const TY* A, i64 lda,
           const TY* B, i64 ldb, TY beta, TY* C, i64 ldc);
template <>
int gemm_t<double>(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha, const double* A,
                   i64 lda, const double* B, i64 ldb, double beta, double* C, i64 ldc) {
    return oxi_launch_dgemm(h, s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
template <>
int gemm_t<float>(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, float alpha, const float* A,
                  i64 lda, const float* B, i64 ldb, float beta, float* C, i64 ldc) {
    return oxi_launch_sgemm(h, s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Right-side triangular solve X <- X * op(T)^-1 where op(T) is upper
// triangular (T upper and trans == 0, or T lower and trans == 1), recursive
// halving with GEMM updates (src/blas/level3.jl:358-376).
// inv: when not null (FP64 only), the inverses of op(T)'s 64 x 64 diagonal
// blocks (n a multiple of 64), and each 64-column leaf is the in-place
// product X_b <- X_b V_b on the matrix cores instead of a column sweep.
template <typename TY>
static int trsm_right_upper(oxi_handle_t h, hipStream_t s, int trans, int unit, i64 rows, i64 n, const TY* T,
                            i64 ldt, TY* X, i64 ldx, const double* inv = nullptr) {
    if (rows <= 0 || n <= 0) return OXI_OK;
    if (n <= TRL) {
        if constexpr (sizeof(TY) == 8) {
            if (inv) {
                if (oxi_dgemm_right64_inplace(s, rows, inv, (double*)X, ldx) != OXI_OK)
                    return oxi_set_error(h, OXI_ERR_HIP, "trsm in-place block product launch failed");
                return OXI_OK;
            }
        }
        launch_right_leaf<TY, false>(s, unit, trans, rows, int(n), T, ldt, X, ldx);
        return launch_check(h, "trsm_right_leaf");
    }
    i64 hh = n / 2;
    if (hh > TRL) hh = (hh / TRL) * TRL;
    // X1 = X[:, :hh], X2 = X[:, hh:]; op(T) = [U11 U12; 0 U22]
    OXI_TRY(trsm_right_upper(h, s, trans, unit, rows, hh, T, ldt, X, ldx, inv));
    // X2 -= X1 * U12, U12 = op(T)[0:hh, hh:n]
    const TY* U12 = trans ? T + hh : T + hh * ldt;  // op(T)[0][hh]: trans -> T[hh][0]
    OXI_TRY(gemm_t<TY>(h, s, 0, trans, rows, n - hh, hh, TY(-1), X, ldx, U12, ldt, TY(1), X + hh * ldx, ldx));
    return trsm_right_upper(h, s, trans, unit, rows, n - hh, T + hh + hh * ldt, ldt, X + hh * ldx, ldx,
                            inv ? inv + (hh / TRL) * 4096 : nullptr);
}

// Tall right-side solves (FP64, rows >= 1024, n a multiple of 64): the
// diagonal blocks' inverses first, then the recursion with in-place products
// for the leaves (a leaf's column sweep is a chain of 64 dependent steps per
// row; the product runs on the matrix cores).
static const double* right_upper_inverses(oxi_handle_t h, hipStream_t s, int trans, int unit, i64 rows, i64 n,
                                          const double* T, i64 ldt) {
    if (rows < 1024 || n < TRL || n % TRL || n / TRL > TRSM_INV_BLOCKS || !h->trsm_inv) return nullptr;
    const dim3 grid(unsigned(n / TRL)), block(64);
    if (unit) {
        if (trans) hipLaunchKernelGGL((trsm_inv64_kernel<true, true>), grid, block, 0, s, T, ldt, h->trsm_inv);
        else hipLaunchKernelGGL((trsm_inv64_kernel<true, false>), grid, block, 0, s, T, ldt, h->trsm_inv);
    } else {
        if (trans) hipLaunchKernelGGL((trsm_inv64_kernel<false, true>), grid, block, 0, s, T, ldt, h->trsm_inv);
        else hipLaunchKernelGGL((trsm_inv64_kernel<false, false>), grid, block, 0, s, T, ldt, h->trsm_inv);
    }
    return hipGetLastError() == hipSuccess ? h->trsm_inv : nullptr;
}

// Right-side solve with op(T) lower: X <- X * op(T)^-1 (backward over columns).
template <typename TY>
static int trsm_right_lower(oxi_handle_t h, hipStream_t s, int trans, int unit, i64 rows, i64 n, const TY* T,
                            i64 ldt, TY* X, i64 ldx);

// Left solves with many right-hand sides (FP64, m a multiple of 64 and at
// most 64 * TRSM_INV_BLOCKS rows): the inverses of op(A)'s 64 x 64 diagonal
// blocks first, so that every leaf of the recursion is one in-place product
// on the matrix cores instead of a column sweep per right-hand side.  For an
// upper op(A) the kernel gives V_b = op(A)_bb^-1; for a lower op(A) it gives
// V_b = (op(A)_bb^T)^-1 and the leaf multiplies by V_b^T.
static const double* left_inverses(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs,
                                   const double* A, i64 lda) {
    if (m < 2 * TRL || m % TRL || m / TRL > TRSM_INV_BLOCKS || nrhs < 16 || !h->trsm_inv) return nullptr;
    const bool lower_op = (!upper) != (trans != 0);
    const bool ktrans = lower_op ? !trans : (trans != 0);
    const dim3 grid(unsigned(m / TRL)), block(64);
    if (unit) {
        if (ktrans) hipLaunchKernelGGL((trsm_inv64_kernel<true, true>), grid, block, 0, s, A, lda, h->trsm_inv);
        else hipLaunchKernelGGL((trsm_inv64_kernel<true, false>), grid, block, 0, s, A, lda, h->trsm_inv);
    } else {
        if (ktrans) hipLaunchKernelGGL((trsm_inv64_kernel<false, true>), grid, block, 0, s, A, lda, h->trsm_inv);
        else hipLaunchKernelGGL((trsm_inv64_kernel<false, false>), grid, block, 0, s, A, lda, h->trsm_inv);
    }
    return hipGetLastError() == hipSuccess ? h->trsm_inv : nullptr;
}

// Left solves with a moderate number of right-hand sides on large triangles
// (FP64, m a multiple of TRB and at least 2 TRB, 16 .. TRB_MAX_RHS
// right-hand sides).  The recursion's updates are skinny (m/2 x nrhs x m/2 at
// the top, and every level costs about as much, one product after the
// other, each on a few workgroups); here op(A) is swept in blocks of TRB
// rows: X_b = op(A)_bb^-1 B_b, one product with the block's explicit
// inverse, then B_rest -= op(A)[rest, b] X_b, whose rows spread over the
// chip.  The TRB x TRB inverses come from the 64 x 64 ones by doubling: for
// U = [U11 U12; 0 U22], U^-1 = [V11, -V11 U12 V22; 0, V22], two batched
// products per doubling and position.  As in the 64 x 64 leaves, each
// block's solve is a product with its inverse rather than a substitution.
constexpr int TRB = 256, TRB_MAX_RHS = 512;

static bool left_blocked_applies(i64 m, i64 nrhs) {
    return m >= 2 * TRB && m % TRB == 0 && nrhs >= 16 && nrhs <= TRB_MAX_RHS;
}

static int trsm_left_blocked(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs,
                             const double* A, i64 lda, double* B, i64 ldb) {
    const bool lower_op = (!upper) != (trans != 0);   // op(A) lower?
    // the upper matrix U' whose inverse the kernels build: U' = op(A) for an
    // upper op(A), op(A)^T for a lower one; U'(i, j) = ktrans ? A(j, i) : A(i, j)
    const bool ktrans = lower_op ? !trans : (trans != 0);
    // S: m x TRB strip (ld m), rows z*TRB.. hold block z's inverse; T: the
    // doubling's temporaries; X: the solution until the final copy
    const size_t sS = size_t(m) * TRB, sT = size_t(m) * TRB / 2, sX = size_t(m) * size_t(nrhs);
    void* ws;
    OXI_TRY(oxi_workspace3(h, (sS + sT + sX) * sizeof(double), &ws));
    double* S = (double*)ws;
    double* T = S + sS;
    double* X = T + sT;
    if (oxi_launch_dscale(s, m, TRB, 0.0, S, m) != OXI_OK) return oxi_set_error(h, OXI_ERR_HIP, "trsm strip clear");
    const dim3 grid(unsigned(m / TRL)), block(64);
    if (unit) {
        if (ktrans) hipLaunchKernelGGL((trsm_inv64_kernel<true, true>), grid, block, 0, s, A, lda, S, m, TRB);
        else hipLaunchKernelGGL((trsm_inv64_kernel<true, false>), grid, block, 0, s, A, lda, S, m, TRB);
    } else {
        if (ktrans) hipLaunchKernelGGL((trsm_inv64_kernel<false, true>), grid, block, 0, s, A, lda, S, m, TRB);
        else hipLaunchKernelGGL((trsm_inv64_kernel<false, false>), grid, block, 0, s, A, lda, S, m, TRB);
    }
    OXI_TRY(launch_check(h, "trsm block inverses"));
    const i64 nz = m / TRB;
    for (i64 w = TRL; w < TRB; w *= 2) {
        for (i64 l = 0; l < TRB; l += 2 * w) {
            // in every block z: p = local rows/columns [l, l+w), q = [l+w, l+2w)
            const double* Upq = ktrans ? A + (l + w) + l * lda : A + l + (l + w) * lda;
            const double* Vp = S + l + l * m;
            const double* Vq = S + (l + w) + (l + w) * m;
            double* Xpq = S + l + (l + w) * m;
            OXI_TRY(oxi_launch_dgemm_strided(h, s, ktrans ? 1 : 0, 0, w, w, w, 1.0, Upq, lda, TRB * (lda + 1), Vq, m,
                                             TRB, 0.0, T, w, w * w, nz));
            OXI_TRY(oxi_launch_dgemm_strided(h, s, 0, 0, w, w, w, -1.0, Vp, m, TRB, T, w, w * w, 0.0, Xpq, m, TRB,
                                             nz));
        }
    }
    if (lower_op) {
        // op(A)_bb^-1 = (U'_bb^-1)^T
        for (i64 b0 = 0; b0 < m; b0 += TRB) {
            OXI_TRY(oxi_launch_dgemm(h, s, 1, 0, TRB, nrhs, TRB, 1.0, S + b0, m, B + b0, ldb, 0.0, X + b0, m));
            const i64 r0 = b0 + TRB;
            if (r0 < m) {
                const double* Ar = trans ? A + b0 + r0 * lda : A + r0 + b0 * lda;   // op(A)[r0][b0]
                OXI_TRY(oxi_launch_dgemm(h, s, trans ? 1 : 0, 0, m - r0, nrhs, TRB, -1.0, Ar, lda, X + b0, m, 1.0,
                                         B + r0, ldb));
            }
        }
    } else {
        for (i64 b0 = m - TRB; b0 >= 0; b0 -= TRB) {
            OXI_TRY(oxi_launch_dgemm(h, s, 0, 0, TRB, nrhs, TRB, 1.0, S + b0, m, B + b0, ldb, 0.0, X + b0, m));
            if (b0 > 0) {
                const double* Ar = trans ? A + b0 : A + b0 * lda;                   // op(A)[0][b0]
                OXI_TRY(oxi_launch_dgemm(h, s, trans ? 1 : 0, 0, b0, nrhs, TRB, -1.0, Ar, lda, X + b0, m, 1.0, B,
                                         ldb));
            }
        }
    }
    OXI_HIP(h, hipMemcpy2DAsync(B, size_t(ldb) * 8, X, size_t(m) * 8, size_t(m) * 8, size_t(nrhs),
                                hipMemcpyDeviceToDevice, s));
    return OXI_OK;
}

template <typename TY>
static int trsm_left_t(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs, const TY* A,
                       i64 lda, TY* B, i64 ldb, const double* inv = nullptr) {
    if (m <= 0 || nrhs <= 0) return OXI_OK;
    const bool lower_op = (!upper) != (trans != 0);  // op(A) lower?
    if constexpr (sizeof(TY) == 8) {
        if (inv && m == TRL) {
            if (oxi_dgemm_left64_inplace(s, lower_op ? 1 : 0, nrhs, inv, (double*)B, ldb) != OXI_OK)
                return oxi_set_error(h, OXI_ERR_HIP, "trsm in-place block product launch failed");
            return OXI_OK;
        }
    }
    if (m <= TRL) {
        dim3 grid(unsigned(cdiv(nrhs, 4)));
        if (lower_op) {
            if (unit) hipLaunchKernelGGL((trsm_left_leaf<TY, true, true>), grid, dim3(256), 0, s, B, ldb, nrhs, int(m), A, lda, trans);
            else hipLaunchKernelGGL((trsm_left_leaf<TY, true, false>), grid, dim3(256), 0, s, B, ldb, nrhs, int(m), A, lda, trans);
        } else {
            if (unit) hipLaunchKernelGGL((trsm_left_leaf<TY, false, true>), grid, dim3(256), 0, s, B, ldb, nrhs, int(m), A, lda, trans);
            else hipLaunchKernelGGL((trsm_left_leaf<TY, false, false>), grid, dim3(256), 0, s, B, ldb, nrhs, int(m), A, lda, trans);
        }
        return launch_check(h, "trsm_left_leaf");
    }
    i64 hh = m / 2;
    if (hh > TRL) hh = (hh / TRL) * TRL;
    // op(A) = [A11 A12; A21 A22] in op coordinates; op(A)[i][j] = trans ? A[j][i] : A[i][j]
    const TY* A11 = A;
    const TY* A22 = A + hh + hh * lda;
    const double* inv2 = inv ? inv + (hh / TRL) * 4096 : nullptr;   // the blocks of A22
    if (lower_op) {
        OXI_TRY(trsm_left_t(h, s, upper, trans, unit, hh, nrhs, A11, lda, B, ldb, inv));
        // B2 -= op(A)[hh:, :hh] * B1 ; op(A)[hh][0] = trans ? A[0][hh] : A[hh][0]
        const TY* A21 = trans ? A + hh * lda : A + hh;
        OXI_TRY(gemm_t<TY>(h, s, trans, 0, m - hh, nrhs, hh, TY(-1), A21, lda, B, ldb, TY(1), B + hh, ldb));
        return trsm_left_t(h, s, upper, trans, unit, m - hh, nrhs, A22, lda, B + hh, ldb, inv2);
    }
    OXI_TRY(trsm_left_t(h, s, upper, trans, unit, m - hh, nrhs, A22, lda, B + hh, ldb, inv2));
    // B1 -= op(A)[:hh, hh:] * B2 ; op(A)[0][hh] = trans ? A[hh][0] : A[0][hh]
    const TY* A12 = trans ? A + hh : A + hh * lda;
    OXI_TRY(gemm_t<TY>(h, s, trans, 0, hh, nrhs, m - hh, TY(-1), A12, lda, B + hh, ldb, TY(1), B, ldb));
    return trsm_left_t(h, s, upper, trans, unit, hh, nrhs, A11, lda, B, ldb, inv);
}

// Right-side solve with op(T) lower: X <- X * op(T)^-1.  op(T) = [L11 0; L21 L22]:
// X2 = B2 * L22^-1 first, then B1 -= X2 * L21, then X1 = B1 * L11^-1.
template <typename TY>
static int trsm_right_lower(oxi_handle_t h, hipStream_t s, int trans, int unit, i64 rows, i64 n, const TY* T,
                            i64 ldt, TY* X, i64 ldx) {
    if (rows <= 0 || n <= 0) return OXI_OK;
    if (n <= TRL) {
        launch_right_leaf<TY, true>(s, unit, trans, rows, int(n), T, ldt, X, ldx);
        return launch_check(h, "trsm_right_leaf");
    }
    i64 hh = n / 2;
    if (hh > TRL) hh = (hh / TRL) * TRL;
    OXI_TRY(trsm_right_lower(h, s, trans, unit, rows, n - hh, T + hh + hh * ldt, ldt, X + hh * ldx, ldx));
    const TY* L21 = trans ? T + hh * ldt : T + hh;  // op(T)[hh][0]
    OXI_TRY(gemm_t<TY>(h, s, 0, trans, rows, hh, n - hh, TY(-1), X + hh * ldx, ldx, L21, ldt, TY(1), X, ldx));
    return trsm_right_lower(h, s, trans, unit, rows, hh, T, ldt, X, ldx);
}

template <typename TY>
static int trsm_right_t(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n, const TY* A,
                        i64 lda, TY* B, i64 ldb) {
    const bool upper_op = (upper != 0) != (trans != 0);
    if (upper_op) {
        const double* inv = nullptr;
        if constexpr (sizeof(TY) == 8) inv = right_upper_inverses(h, s, trans, unit, m, n, (const double*)A, lda);
        return trsm_right_upper(h, s, trans, unit, m, n, A, lda, B, ldb, inv);
    }
    return trsm_right_lower(h, s, trans, unit, m, n, A, lda, B, ldb);
}


int oxi_dtrsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs, const double* A,
                   i64 lda, double* B, i64 ldb) {
    static const bool blocked = !getenv("OXIBLAS_TRSM_RECURSIVE");   // tuning: the recursion throughout
    if (blocked && left_blocked_applies(m, nrhs))
        return trsm_left_blocked(h, s, upper, trans, unit, m, nrhs, A, lda, B, ldb);
    const double* inv = left_inverses(h, s, upper, trans, unit, m, nrhs, A, lda);
    return trsm_left_t<double>(h, s, upper, trans, unit, m, nrhs, A, lda, B, ldb, inv);
}
int oxi_dtrsm_right(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n, const double* A,
                    i64 lda, double* B, i64 ldb) {
    return trsm_right_t<double>(h, s, upper, trans, unit, m, n, A, lda, B, ldb);
}
int oxi_strsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs, const float* A,
                   i64 lda, float* B, i64 ldb) {
    return trsm_left_t<float>(h, s, upper, trans, unit, m, nrhs, A, lda, B, ldb);
}
int oxi_strsm_right(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n, const float* A,
                    i64 lda, float* B, i64 ldb) {
    return trsm_right_t<float>(h, s, upper, trans, unit, m, n, A, lda, B, ldb);
}

static int trans_kind(char t) {
    if (t == 'N' || t == 'n') return 0;
    if (t == 'T' || t == 't' || t == 'C' || t == 'c') return 1;
    return -1;
}

// First zero on the diagonal of an n x n triangle (1-based), synchronising.
int oxi_first_zero_diag(oxi_handle_t h, hipStream_t s, const double* T, i64 ldt, i64 n, i64* where) {
    void* ws;
    OXI_TRY(oxi_workspace(h, 256, &ws));
    i64* d = (i64*)ws;
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(1), 0, s, d, i64(0x7fffffffffffffffLL));
    hipLaunchKernelGGL(diag_zero_kernel, dim3(unsigned(std::min<i64>(cdiv(n, 256), 1024))), dim3(256), 0, s, T, ldt, n, d);
    OXI_TRY(launch_check(h, "diag check"));
    i64 v = 0;
    OXI_HIP(h, hipMemcpyAsync(&v, d, sizeof(i64), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    *where = (v == i64(0x7fffffffffffffffLL)) ? 0 : v;
    return OXI_OK;
}

extern "C" int oxi_dtrsm(oxi_handle_t h, char side, char uplo, char trans, char diag, int64_t m, int64_t n,
                         double alpha, const double* A, int64_t lda, double* B, int64_t ldb, int64_t* singular_index) {
    if (!h) return -1;
    const bool left = side == 'L' || side == 'l';
    if (!left && !(side == 'R' || side == 'r')) return -2;
    const bool upper = uplo == 'U' || uplo == 'u';
    if (!upper && !(uplo == 'L' || uplo == 'l')) return -3;
    const int tk = trans_kind(trans);
    if (tk < 0) return -4;
    const bool unit = diag == 'U' || diag == 'u';
    if (!unit && !(diag == 'N' || diag == 'n')) return -5;
    if (m < 0) return -6;
    if (n < 0) return -7;
    const i64 na = left ? m : n;
    if (lda < std::max<i64>(1, na)) return -10;
    if (ldb < std::max<i64>(1, m)) return -12;
    if (singular_index) *singular_index = 0;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 || n == 0) return OXI_OK;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    if (alpha == 0.0) return oxi_launch_dscale(s, m, n, 0.0, B, ldb);
    if (!unit) {
        i64 z = 0;
        OXI_TRY(oxi_first_zero_diag(h, s, A, lda, na, &z));
        if (z) {
            if (singular_index) *singular_index = z;
            return oxi_set_error(h, OXI_ERR_SINGULAR, "trsm: zero diagonal at %lld", (long long)z);
        }
    }
    if (alpha != 1.0) OXI_TRY(oxi_launch_dscale(s, m, n, alpha, B, ldb));
    if (left) return oxi_dtrsm_left(h, s, upper, tk, unit, m, n, A, lda, B, ldb);
    return oxi_dtrsm_right(h, s, upper, tk, unit, m, n, A, lda, B, ldb);
}

