// Mixed-precision LU solve with FP64 refinement on the GPU:
// mixed_solve(A, B)  (src/lapack/factorizations.jl:419-427 ->
// src/lapack/extended_lapack.jl:175-226).
//
//   1. radix equilibration E = Dr^-1 A Dc^-1 with power-of-two scales
//      (equilibrate(A; radix=true), src/lapack/dense_utilities.jl:191-212),
//      fused with the finite check of A (:182) and the row sums of |A|;
//   2. E rounded to FP32 and factored by the FP32 LU whose trailing products
//      run in BF16 / FP16 / FP8 / FP32 (oxi_getrf32); a singular or
//      overflowing low factor falls back to the FP64 LU of E (:195-211);
//   3. the correction step of the reference (:114-131): scale the residual
//      by the row scales, a binary magnitude, low-precision solve, scale
//      back by the column scales;
//   4. residual and componentwise backward error in FP64
//      (_refinement_residual!, src/lapack/expert_refinement.jl:67-81),
//      deterministic (fixed-order partial sums);
//   5. classical iterative refinement (the reference's loop), or GMRES-IR:
//      left-preconditioned GMRES in FP64 on A d = r, the preconditioner being
//      the low factors applied in FP64 arithmetic (Carson & Higham, SIAM J.
//      Sci. Comput. 40(2), 2018), for condition numbers where the classical
//      iteration stalls.
// The residual, corrections and X stay FP64; A is never modified.
#include "oxi_internal.hpp"

#include <cfloat>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>

namespace {

typedef unsigned long long u64;
constexpr int TPB = 256;

__device__ __forceinline__ u64 dbits(double v) { return __builtin_bit_cast(u64, v); }
__device__ __forceinline__ double bitsd(u64 v) { return __builtin_bit_cast(double, v); }

// round(Int, log2(v)) for finite v > 0, with log2 correctly rounded to double
// and ties to even (Julia's round).  v = m 2^ep with m in [1, 2): the answer
// is ep or ep + 1 by m against sqrt(2), except when log2(v) rounds to exactly
// ep + 1/2 (then the even one).
__device__ int round_log2(double v) {
    int e;
    const double f = frexp(v, &e);
    const int ep = e - 1;
    const double m = 2.0 * f;
    const double S_HI = 1.4142135623730951, S_LO = -0x1.bdd3413b26456p-54;   // sqrt(2) = S_HI + S_LO
    const double dl = ((m - S_HI) - S_LO) * 1.0201394465967895;              // log2(m) - 1/2, first order
    const double a = fabs(double(ep) + 0.5);
    int ea;
    frexp(a, &ea);                                                            // a in [2^(ea-1), 2^ea)
    double half = ldexp(1.0, ea - 54);                                        // half an ulp of a
    if (a == ldexp(1.0, ea - 1) && (ep >= 0 ? dl < 0 : dl > 0)) half *= 0.5;  // finer spacing below a power of two
    if (fabs(dl) < half) return (ep & 1) ? ep + 1 : ep;
    return dl > 0 ? ep + 1 : ep;
}
// ldexp(1, clamp(round(Int, log2(v)), exponent(floatmin), exponent(floatmax)))
__device__ double radix_scale(double v) {
    const int e = min(1023, max(-1022, round_log2(v)));
    return ldexp(1.0, e);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Atomic max of a non-negative double (bit pattern order = value order; NaN
// patterns sort above +inf).  One atomic per wave.
__device__ __forceinline__ void wave_atomic_max(u64* dst, double v) {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) atomicMax(dst, dbits(v));
}

__global__ void fill_u64_kernel(u64* p, int count) {
    if (int(threadIdx.x) < count) p[threadIdx.x] = 0;
}

// ---- equilibration (src/lapack/dense_utilities.jl:191-212) ------------------
// Per-row partial max |a| and sum |a| over a column chunk; flags[0] on a
// non-finite entry.
__global__ void row_stats_kernel(const double* __restrict__ A, i64 lda, i64 m, i64 n, i64 kc,
                                 double* __restrict__ pmax, double* __restrict__ psum, unsigned* __restrict__ flags) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= m) return;
    const i64 k0 = i64(blockIdx.y) * kc, k1 = min(n, k0 + kc);
    double mx = 0, s = 0;
    bool bad = false;
#pragma unroll 8
    for (i64 k = k0; k < k1; ++k) {
        const double a = fabs(A[i + k * lda]);
        bad |= !(a <= DBL_MAX);
        mx = fmax(mx, a);
        s += a;
    }
    if (bad) atomicOr(&flags[0], 1u);
    pmax[i64(blockIdx.y) * m + i] = mx;
    psum[i64(blockIdx.y) * m + i] = s;
}

__device__ __forceinline__ double scale_of(double v, int radix) { return radix ? radix_scale(v) : v; }

__global__ void row_scale_kernel(i64 m, int nchunk, const double* __restrict__ pmax, const double* __restrict__ psum,
                                 int radix, double* __restrict__ rows, double* __restrict__ rinv,
                                 double* __restrict__ rowsum, u64* anorm) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    double s = 0;
    if (i < m) {
        double mx = 0;
        for (int c = 0; c < nchunk; ++c) {
            mx = fmax(mx, pmax[i64(c) * m + i]);
            s += psum[i64(c) * m + i];
        }
        const double r = mx > 0 ? scale_of(mx, radix) : 1.0;
        rows[i] = r;
        if (rinv) rinv[i] = 1.0 / r;      // exact for the power-of-two scales of radix mode
        if (rowsum) rowsum[i] = s;
    }
    if (anorm) wave_atomic_max(anorm, s);
}

// column max of |A ./ rows| -> cols
__global__ void col_scale_kernel(const double* __restrict__ A, i64 lda, i64 m, i64 n, const double* __restrict__ rows,
                                 int radix, double* __restrict__ cols, double* __restrict__ cinv) {
    __shared__ double part[TPB / 64];
    for (i64 j = blockIdx.x; j < n; j += gridDim.x) {
        double mx = 0;
        for (i64 i = threadIdx.x; i < m; i += TPB) mx = fmax(mx, fabs(A[i + j * lda] / rows[i]));
        mx = wave_max(mx);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < TPB / 64; ++w) mx = fmax(mx, part[w]);
            const double c = mx > 0 ? scale_of(mx, radix) : 1.0;
            cols[j] = c;
            if (cinv) cinv[j] = 1.0 / c;
        }
        __syncthreads();
    }
}

// E = (A ./ rows) ./ cols rounded to TO (two divisions, as the reference's two
// passes); flags[1] on a non-finite result.  E may be A itself.
template <typename TO>
__global__ void equil_convert_kernel(const double* A, i64 lda, i64 m, i64 n, const double* __restrict__ rows,
                                     const double* __restrict__ cols, TO* E, i64 lde, unsigned* __restrict__ flags) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= m) return;
    const double ri = rows[i];
    bool bad = false;
    for (i64 j = blockIdx.y; j < n; j += gridDim.y) {
        const TO v = TO((A[i + j * lda] / ri) / cols[j]);
        bad |= !(fabs(double(v)) <= DBL_MAX);
        E[i + j * lde] = v;
    }
    if (bad) atomicOr(&flags[1], 1u);
}

// the largest row maximum (one global scale, the Cholesky form of the mixed
// solve, src/lapack/extended_lapack.jl:187-189)
__global__ void global_max_kernel(const double* __restrict__ pmax, i64 total, u64* out) {
    double mx = 0;
    for (i64 i = i64(blockIdx.x) * TPB + threadIdx.x; i < total; i += i64(gridDim.x) * TPB) mx = fmax(mx, pmax[i]);
    wave_atomic_max(out, mx);
}
// rows = scale (1 for a zero matrix), cols = 1 and E (TO) = A ./ scale
template <typename TO>
__global__ void global_scale_kernel(const double* A, i64 lda, i64 n, const u64* scalebits, double* __restrict__ rows,
                                    double* __restrict__ rinv, double* __restrict__ cols, double* __restrict__ cinv,
                                    TO* E, i64 lde, unsigned* __restrict__ flags) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    double sc = bitsd(*scalebits);
    if (sc == 0) sc = 1.0;
    if (blockIdx.y == 0) {
        rows[i] = sc; rinv[i] = 1.0 / sc; cols[i] = 1.0; cinv[i] = 1.0;
    }
    if (!E) return;
    bool bad = false;
    for (i64 j = blockIdx.y; j < n; j += gridDim.y) {
        const TO v = TO(A[i + j * lda] / sc);
        bad |= !(fabs(double(v)) <= DBL_MAX);
        E[i + j * lde] = v;
    }
    if (bad) atomicOr(&flags[1], 1u);
}

__global__ void finite_kernel(const double* __restrict__ B, i64 ldb, i64 n, i64 nrhs, unsigned* flags) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    bool bad = false;
    for (i64 r = blockIdx.y; r < nrhs; r += gridDim.y) bad |= !(fabs(B[i + r * ldb]) <= DBL_MAX);
    if (bad) atomicOr(&flags[0], 1u);
}

// ---- correction step (src/lapack/extended_lapack.jl:114-131) ---------------
__global__ void corr_scale_kernel(const double* __restrict__ R, i64 ldr, i64 n, i64 nrhs,
                                  const double* __restrict__ rows, double* __restrict__ S, u64* mag,
                                  unsigned* flags) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    double a = 0;
    if (i < n) {
        for (i64 r = blockIdx.y; r < nrhs; r += gridDim.y) {
            const double v = R[i + r * ldr] / rows[i];
            S[i + r * n] = v;
            a = fmax(a, fabs(v));
            if (!(fabs(v) <= DBL_MAX)) atomicOr(&flags[2], 1u);
        }
    }
    wave_atomic_max(mag, a);
}

template <typename TL>
__global__ void corr_low_kernel(const double* __restrict__ S, i64 n, i64 nrhs, const u64* mag,
                                TL* __restrict__ Y, double* magout, unsigned* flags) {
    const double mg = bitsd(*mag);
    double magnitude = 0;
    if (mg > 0) {
        int e;
        frexp(mg, &e);
        magnitude = ldexp(1.0, e - 1);   // ldexp(1, exponent(mg))
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *magout = magnitude;
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    for (i64 r = blockIdx.y; r < nrhs; r += gridDim.y) {
        const TL y = magnitude > 0 ? TL(S[i + r * n] / magnitude) : TL(0);
        if (!(fabs(double(y)) <= DBL_MAX) || (sizeof(TL) == 4 && !(fabs(double(y)) <= double(FLT_MAX))))
            atomicOr(&flags[3], 1u);
        Y[i + r * n] = y;
    }
}

template <typename TL>
__global__ void corr_back_kernel(const TL* __restrict__ Y, i64 n, i64 nrhs, const double* magout,
                                 const double* __restrict__ cols, double* __restrict__ X, i64 ldx, int init) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    const double mg = *magout;
    for (i64 r = blockIdx.y; r < nrhs; r += gridDim.y) {
        const double c = mg > 0 ? (double(Y[i + r * n]) * mg) / cols[i] : 0.0;
        X[i + r * ldx] = init ? c : X[i + r * ldx] + c;
    }
}

// ---- residual / backward error / matrix-vector product ---------------------
// Partial sums over a column chunk: P = sum a*x, Q = sum |a||x| (G right-hand
// sides per block); max |x| and the finiteness of X recorded on the way.
template <int G, bool DEN>
__global__ __launch_bounds__(TPB) void res_partial_kernel(const double* __restrict__ A, i64 lda, i64 n,
                                                          const double* __restrict__ X, i64 ldx, i64 nrhs, i64 kc,
                                                          double* __restrict__ P, double* __restrict__ Q, u64* xmax,
                                                          unsigned* flags) {
    __shared__ double xs[G][TPB];
    const int t = threadIdx.x;
    const i64 i = i64(blockIdx.x) * TPB + t;
    const i64 k0 = i64(blockIdx.y) * kc, k1 = min(n, k0 + kc);
    const i64 r0 = i64(blockIdx.z) * G;
    const int g = int(min(i64(G), nrhs - r0));
    double v[G], d[G];
#pragma unroll
    for (int r = 0; r < G; ++r) v[r] = d[r] = 0;
    double xm = 0;
    for (i64 kk = k0; kk < k1; kk += TPB) {
        const int len = int(min(i64(TPB), k1 - kk));
        __syncthreads();
#pragma unroll
        for (int r = 0; r < G; ++r) {
            double xv = 0;
            if (t < len && r < g) xv = X[(kk + t) + (r0 + r) * ldx];
            xs[r][t] = xv;
            xm = fmax(xm, fabs(xv));
            if (!(fabs(xv) <= DBL_MAX)) atomicOr(&flags[4], 1u);
        }
        __syncthreads();
        if (i < n) {
#pragma unroll 8
            for (int k = 0; k < len; ++k) {
                const double a = A[i + (kk + k) * lda];
#pragma unroll
                for (int r = 0; r < G; ++r) {
                    v[r] = __builtin_fma(a, xs[r][k], v[r]);
                    if (DEN) d[r] = __builtin_fma(fabs(a), fabs(xs[r][k]), d[r]);
                }
            }
        }
    }
    if (xmax && blockIdx.x == 0) wave_atomic_max(xmax, xm);
    if (i < n)
#pragma unroll
        for (int r = 0; r < G; ++r)
            if (r < g) {
                const i64 o = (i64(blockIdx.y) * nrhs + r0 + r) * n + i;
                P[o] = v[r];
                if (DEN) Q[o] = d[r];
            }
}

// R = B - A X, componentwise backward error max_i |r_i| / (|A||x| + |b|)_i
// with the reference's conventions for zero and non-finite denominators;
// stats[0] = berr, [1] = max |r|, [2] = max |b| (as ordered bit patterns).
__global__ void res_final_kernel(i64 n, i64 nrhs, int nchunk, const double* __restrict__ P,
                                 const double* __restrict__ Q, const double* __restrict__ B, i64 ldb,
                                 double* __restrict__ R, i64 ldr, u64* stats) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    double rel = 0, rm = 0, bm = 0;
    if (i < n) {
        for (i64 r = blockIdx.y; r < nrhs; r += gridDim.y) {
            const double b = B[i + r * ldb];
            double s = 0, q = 0;
            for (int c = 0; c < nchunk; ++c) {
                const i64 o = (i64(c) * nrhs + r) * n + i;
                s += P[o];
                q += Q[o];
            }
            const double value = b - s, den = fabs(b) + q;
            R[i + r * ldr] = value;
            const double rl = !(den <= DBL_MAX) ? INFINITY
                              : den == 0       ? (value == 0 ? 0.0 : INFINITY)
                                               : fabs(value) / den;
            rel = (rl != rl || rl > rel) ? rl : rel;
            rm = fmax(rm, fabs(value));
            bm = fmax(bm, fabs(b));
        }
    }
    // NaN must survive the reduction: compare bit patterns
    u64 rb = dbits(rel);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 other = __shfl_xor(rb, o);
        rb = other > rb ? other : rb;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&stats[0], rb);
    wave_atomic_max(&stats[1], rm);
    wave_atomic_max(&stats[2], bm);
}

// y = A x from the partial sums (fixed order)
__global__ void mv_final_kernel(i64 n, int nchunk, const double* __restrict__ P, double* __restrict__ y) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int c = 0; c < nchunk; ++c) s += P[i64(c) * n + i];
    y[i] = s;
}

// ---- GMRES building blocks ---------------------------------------------------
// y = P Dr^-1 v (row interchanges of the factor after the row scaling)
__global__ void prec_in_kernel(const double* __restrict__ v, i64 n, const i64* __restrict__ perm,
                               const double* __restrict__ rinv, double* __restrict__ y) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    const i64 p = perm[i] - 1;
    y[i] = v[p] * rinv[p];
}
__global__ void prec_out_kernel(double* __restrict__ y, i64 n, const double* __restrict__ cinv) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i < n) y[i] = y[i] * cinv[i];
}
// part[c * nseg + seg] = sum over the segment of V[:, c] .* w
__global__ void dots_kernel(const double* __restrict__ V, i64 ldv, const double* __restrict__ w, i64 n, i64 seglen,
                            int nseg, double* __restrict__ part) {
    __shared__ double red[TPB / 64];
    const int seg = blockIdx.x, c = blockIdx.y;
    const i64 a = i64(seg) * seglen, b = min(n, a + seglen);
    double s = 0;
    for (i64 i = a + threadIdx.x; i < b; i += TPB) s = __builtin_fma(V[i + c * ldv], w[i], s);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int q = 1; q < TPB / 64; ++q) t += red[q];
        part[i64(c) * nseg + seg] = t;
    }
}
__global__ void sum_parts_kernel(const double* __restrict__ part, int nseg, int ncols, double* __restrict__ out) {
    const int c = threadIdx.x;
    if (c >= ncols) return;
    double s = 0;
    for (int q = 0; q < nseg; ++q) s += part[i64(c) * nseg + q];
    out[c] = s;
}
// w += sign * V[:, 0:ncols] h
__global__ void vupdate_kernel(const double* __restrict__ V, i64 ldv, int ncols, const double* __restrict__ h,
                               double sign, double* __restrict__ w, i64 n) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int c = 0; c < ncols; ++c) s = __builtin_fma(V[i + c * ldv], h[c], s);
    w[i] = __builtin_fma(sign, s, w[i]);
}
__global__ void vdiv_kernel(double* __restrict__ w, i64 n, double d) {
    const i64 i = i64(blockIdx.x) * TPB + threadIdx.x;
    if (i < n) w[i] = w[i] / d;
}

// ---- driver ------------------------------------------------------------------
constexpr size_t ALIGN = 256;
size_t pad(size_t b) { return (b + ALIGN - 1) & ~(ALIGN - 1); }

struct Mixed {
    oxi_handle_t h;
    hipStream_t s;
    i64 n, nrhs;
    const double* A;
    i64 lda;
    const double* B;
    i64 ldb;
    double* X;
    i64 ldx;
    // device state
    float* A32;
    double* E64 = nullptr;     // FP64 factor after a fallback (own allocation)
    i64* ipiv;
    i64* dperm;
    double *rows, *cols, *rinv, *cinv, *rowsum;
    double *R, *S, *Yd, *Ywd;
    float *Yf, *Ywf;
    double *P, *Q;
    i64 kc;
    int nchunk;
    u64* stats;        // [0] berr [1] max|r| [2] max|b| [3] max|x| [4] ||A||_inf [5] magnitude max
    unsigned* flags;   // [0] non-finite input [1] matrix conversion [2],[3] RHS conversion [4] non-finite X
    double* magout;
    void* ctrl;        // trsv workspace (control words and diagonal-block inverses)
    int prepared = -1; // what the inverses in ctrl were made for (see prepare())
    // random butterfly transform (pivot == 1): A' = U^T A_eq V factored without pivoting
    bool rbt = false;
    float* R32 = nullptr;
    double* R64 = nullptr;
    // GMRES
    int m = 0;
    double *V = nullptr, *tmp = nullptr, *part = nullptr, *hdev = nullptr;
    int nseg = 1;
    i64 seglen = 1;
    bool fallback = false;
    bool chol = false;         // Cholesky factors (factorization=:cholesky): L, then L^T solves
    double anorm = 0;

    dim3 vgrid(i64 cols) const {
        return dim3(unsigned(cdiv(n, TPB)), unsigned(std::max<i64>(1, std::min<i64>(cols, 65535))));
    }
};

int launch_ok(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

int read_stats(Mixed& c, u64* st, unsigned* fl) {
    OXI_HIP(c.h, hipMemcpyAsync(st, c.stats, 8 * sizeof(u64), hipMemcpyDeviceToHost, c.s));
    OXI_HIP(c.h, hipMemcpyAsync(fl, c.flags, 8 * sizeof(unsigned), hipMemcpyDeviceToHost, c.s));
    OXI_HIP(c.h, hipStreamSynchronize(c.s));
    return OXI_OK;
}
double from_bits(u64 b) { double d; std::memcpy(&d, &b, 8); return d; }

// Residual chunking: partial sums of at most ~256 MiB.
void plan_residual(Mixed& c) {
    const i64 groups = c.nrhs;   // partial entries per chunk = n * nrhs
    i64 maxchunks = std::max<i64>(1, (i64(256) << 20) / (16 * c.n * groups));
    i64 kc = std::max<i64>(256, cdiv(c.n, std::min<i64>(maxchunks, 512)));
    kc = cdiv(kc, 64) * 64;
    c.kc = kc;
    c.nchunk = int(cdiv(c.n, kc));
}

// P/Q partials of A * X(:, cols) for `ncols` columns starting at X
int matprod_partials(Mixed& c, const double* Xp, i64 ldxp, i64 ncols, bool den, u64* xmax) {
    if (ncols == 1) {
        dim3 grid(unsigned(cdiv(c.n, TPB)), unsigned(c.nchunk), 1);
        if (den)
            hipLaunchKernelGGL((res_partial_kernel<1, true>), grid, dim3(TPB), 0, c.s, c.A, c.lda, c.n, Xp, ldxp,
                               ncols, c.kc, c.P, c.Q, xmax, c.flags);
        else
            hipLaunchKernelGGL((res_partial_kernel<1, false>), grid, dim3(TPB), 0, c.s, c.A, c.lda, c.n, Xp, ldxp,
                               ncols, c.kc, c.P, c.Q, xmax, c.flags);
    } else {
        dim3 grid(unsigned(cdiv(c.n, TPB)), unsigned(c.nchunk), unsigned(cdiv(ncols, 4)));
        hipLaunchKernelGGL((res_partial_kernel<4, true>), grid, dim3(TPB), 0, c.s, c.A, c.lda, c.n, Xp, ldxp, ncols,
                           c.kc, c.P, c.Q, xmax, c.flags);
    }
    return launch_ok(c.h, "residual");
}

// R = B - A X; returns the componentwise and normwise backward errors and
// whether X is finite.
int residual(Mixed& c, double* berr, double* nberr, bool* xfinite) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, c.stats, 4);
    OXI_TRY(matprod_partials(c, c.X, c.ldx, c.nrhs, true, &c.stats[3]));
    hipLaunchKernelGGL(res_final_kernel, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.n, c.nrhs, c.nchunk, c.P, c.Q, c.B,
                       c.ldb, c.R, c.n, c.stats);
    OXI_TRY(launch_ok(c.h, "residual"));
    u64 st[8];
    unsigned fl[8];
    OXI_TRY(read_stats(c, st, fl));
    if (fl[2] | fl[3]) return oxi_set_error(c.h, OXI_ERR_OVERFLOW, "mixed_solve: low precision RHS conversion failed");
    *berr = from_bits(st[0]);
    const double rmax = from_bits(st[1]), bmax = from_bits(st[2]), xmax = from_bits(st[3]);
    const double den = c.anorm * xmax + bmax;
    *nberr = den > 0 ? rmax / den : (rmax == 0 ? 0.0 : INFINITY);
    *xfinite = fl[4] == 0;
    return OXI_OK;
}

// X (init) or X + correction, from the residual Rin (n x nrhs, ld ldr)
// The diagonal-block inverses of the single-RHS triangular solves, made once
// per kind: 0 FP32 factors and FP32 vectors (IR correction), 1 FP32 factors
// and FP64 vectors (GMRES preconditioner), 2 FP64 factors (after a fallback).
int prepare(Mixed& c, int kind) {
    if (c.prepared == kind) return OXI_OK;
    if (kind == 0) OXI_TRY(oxi_trsv_prepare(c.h, c.s, 0, c.n, (const float*)c.A32, c.n, (float*)nullptr, c.ctrl));
    else if (kind == 1) OXI_TRY(oxi_trsv_prepare(c.h, c.s, 0, c.n, (const float*)c.A32, c.n, (double*)nullptr, c.ctrl));
    else OXI_TRY(oxi_trsv_prepare(c.h, c.s, 0, c.n, (const double*)c.E64, c.n, (double*)nullptr, c.ctrl));
    c.prepared = kind;
    return OXI_OK;
}

// The correction with Cholesky factors (solve(F::CholeskyFactor, .),
// src/lapack/factorizations.jl:288-300): L then L^T, in FP32 on the low
// factor or in FP64 after a fallback.
int chol_correction(Mixed& c, const double* Rin, i64 ldr, bool init) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, &c.stats[5], 1);
    hipLaunchKernelGGL(corr_scale_kernel, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, Rin, ldr, c.n, c.nrhs, c.rows, c.S,
                       &c.stats[5], c.flags);
    if (!c.fallback) {
        hipLaunchKernelGGL(corr_low_kernel<float>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.S, c.n, c.nrhs, &c.stats[5],
                           c.Yf, c.magout, c.flags);
        OXI_TRY(launch_ok(c.h, "correction"));
        OXI_TRY(oxi_strsm_left(c.h, c.s, 0, 0, 0, c.n, c.nrhs, c.A32, c.n, c.Yf, c.n));
        OXI_TRY(oxi_strsm_left(c.h, c.s, 0, 1, 0, c.n, c.nrhs, c.A32, c.n, c.Yf, c.n));
        hipLaunchKernelGGL(corr_back_kernel<float>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.Yf, c.n, c.nrhs, c.magout,
                           c.cols, c.X, c.ldx, int(init));
    } else {
        hipLaunchKernelGGL(corr_low_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.S, c.n, c.nrhs,
                           &c.stats[5], c.Yd, c.magout, c.flags);
        OXI_TRY(launch_ok(c.h, "correction"));
        OXI_TRY(oxi_dtrsm_left(c.h, c.s, 0, 0, 0, c.n, c.nrhs, c.E64, c.n, c.Yd, c.n));
        OXI_TRY(oxi_dtrsm_left(c.h, c.s, 0, 1, 0, c.n, c.nrhs, c.E64, c.n, c.Yd, c.n));
        hipLaunchKernelGGL(corr_back_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.Yd, c.n, c.nrhs, c.magout,
                           c.cols, c.X, c.ldx, int(init));
    }
    return launch_ok(c.h, "correction");
}

int correction(Mixed& c, const double* Rin, i64 ldr, bool init) {
    // after the butterflies (no pivoting) the few-RHS correction solves run in
    // FP64 arithmetic with the FP32 factors: the triangles of a factor without
    // pivoting are less well conditioned, and FP32 block-inverse products
    // would slow the contraction down
    const bool wide = c.rbt && c.nrhs <= OXI_TRSV_MAX_RHS;
    if (c.chol) return chol_correction(c, Rin, ldr, init);
    if (c.nrhs <= OXI_TRSV_MAX_RHS) OXI_TRY(prepare(c, c.fallback ? 2 : wide ? 1 : 0));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, &c.stats[5], 1);
    hipLaunchKernelGGL(corr_scale_kernel, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, Rin, ldr, c.n, c.nrhs, c.rows, c.S,
                       &c.stats[5], c.flags);
    if (wide) {
        hipLaunchKernelGGL(corr_low_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.S, c.n, c.nrhs,
                           &c.stats[5], c.Yd, c.magout, c.flags);
        OXI_TRY(launch_ok(c.h, "correction"));
        OXI_TRY(oxi_rbt_vec(c.h, c.s, 1, c.n, c.nrhs, c.Yd, c.n, c.R64));
        for (i64 j = 0; j < c.nrhs; ++j) {
            OXI_TRY(oxi_trsv(c.h, c.s, 0, c.n, (const float*)c.A32, c.n, c.Yd + j * c.n, c.ctrl));
            OXI_TRY(oxi_trsv(c.h, c.s, 1, c.n, (const float*)c.A32, c.n, c.Yd + j * c.n, c.ctrl));
        }
        OXI_TRY(oxi_rbt_vec(c.h, c.s, 0, c.n, c.nrhs, c.Yd, c.n, c.R64));
        hipLaunchKernelGGL(corr_back_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.Yd, c.n, c.nrhs, c.magout,
                           c.cols, c.X, c.ldx, int(init));
    } else if (!c.fallback) {
        hipLaunchKernelGGL(corr_low_kernel<float>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.S, c.n, c.nrhs, &c.stats[5],
                           c.Yf, c.magout, c.flags);
        OXI_TRY(launch_ok(c.h, "correction"));
        if (c.rbt) OXI_TRY(oxi_rbt_vec(c.h, c.s, 1, c.n, c.nrhs, c.Yf, c.n, c.R32));
        OXI_TRY(oxi_getrs_perm_f32(c.h, c.s, 0, c.n, c.nrhs, c.A32, c.n, c.dperm, c.Yf, c.n, c.Ywf, c.ctrl));
        if (c.rbt) OXI_TRY(oxi_rbt_vec(c.h, c.s, 0, c.n, c.nrhs, c.Yf, c.n, c.R32));
        hipLaunchKernelGGL(corr_back_kernel<float>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.Yf, c.n, c.nrhs, c.magout,
                           c.cols, c.X, c.ldx, int(init));
    } else {
        hipLaunchKernelGGL(corr_low_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.S, c.n, c.nrhs,
                           &c.stats[5], c.Yd, c.magout, c.flags);
        OXI_TRY(launch_ok(c.h, "correction"));
        OXI_TRY(oxi_getrs_perm_f64(c.h, c.s, 0, c.n, c.nrhs, c.E64, c.n, c.dperm, c.Yd, c.n, c.Ywd, c.ctrl));
        hipLaunchKernelGGL(corr_back_kernel<double>, c.vgrid(c.nrhs), dim3(TPB), 0, c.s, c.Yd, c.n, c.nrhs, c.magout,
                           c.cols, c.X, c.ldx, int(init));
    }
    return launch_ok(c.h, "correction");
}

// out = M^-1 v with M^-1 = Dc^-1 U^-1 L^-1 P Dr^-1 (low factors, FP64 arithmetic)
int precond(Mixed& c, const double* v, double* out) {
    OXI_TRY(prepare(c, c.fallback ? 2 : 1));
    const unsigned g = unsigned(cdiv(c.n, TPB));
    hipLaunchKernelGGL(prec_in_kernel, dim3(g), dim3(TPB), 0, c.s, v, c.n, c.dperm, c.rinv, out);
    OXI_TRY(launch_ok(c.h, "precond"));
    if (!c.fallback) {
        if (c.rbt) OXI_TRY(oxi_rbt_vec(c.h, c.s, 1, c.n, 1, out, c.n, c.R64));
        OXI_TRY(oxi_trsv(c.h, c.s, 0, c.n, (const float*)c.A32, c.n, out, c.ctrl));
        OXI_TRY(oxi_trsv(c.h, c.s, 1, c.n, (const float*)c.A32, c.n, out, c.ctrl));
        if (c.rbt) OXI_TRY(oxi_rbt_vec(c.h, c.s, 0, c.n, 1, out, c.n, c.R64));
    } else {
        OXI_TRY(oxi_trsv(c.h, c.s, 0, c.n, (const double*)c.E64, c.n, out, c.ctrl));
        OXI_TRY(oxi_trsv(c.h, c.s, 1, c.n, (const double*)c.E64, c.n, out, c.ctrl));
    }
    hipLaunchKernelGGL(prec_out_kernel, dim3(g), dim3(TPB), 0, c.s, out, c.n, c.cinv);
    return launch_ok(c.h, "precond");
}

// hdev[0:ncols] = V[:, 0:ncols]^T w
int dots(Mixed& c, const double* V, int ncols, const double* w, double* hdev) {
    hipLaunchKernelGGL(dots_kernel, dim3(unsigned(c.nseg), unsigned(ncols)), dim3(TPB), 0, c.s, V, c.n, w, c.n,
                       c.seglen, c.nseg, c.part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64 * unsigned(cdiv(ncols, 64))), 0, c.s, c.part, c.nseg,
                       ncols, hdev);
    return launch_ok(c.h, "dots");
}

// One GMRES cycle on M^-1 A d = M^-1 r for column j of the residual; X(:, j) += d.
int gmres_correction(Mixed& c, i64 j, double tau, i64* inner) {
    const i64 n = c.n;
    const int m = c.m;
    const unsigned g = unsigned(cdiv(n, TPB));
    double* V = c.V;
    double* h1 = c.hdev;            // m+1
    double* h2 = c.hdev + (m + 1);  // m+1
    double* hn = c.hdev + 2 * (m + 1);
    double* ydev = c.hdev + 3 * (m + 1);
    OXI_TRY(precond(c, c.R + j * n, V));
    OXI_TRY(dots(c, V, 1, V, hn));
    double beta2 = 0;
    OXI_HIP(c.h, hipMemcpyAsync(&beta2, hn, 8, hipMemcpyDeviceToHost, c.s));
    OXI_HIP(c.h, hipStreamSynchronize(c.s));
    const double beta = std::sqrt(beta2);
    if (!(beta > 0) || !std::isfinite(beta)) return OXI_OK;
    hipLaunchKernelGGL(vdiv_kernel, dim3(g), dim3(TPB), 0, c.s, V, n, beta);
    std::vector<double> H(size_t(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1, 0.0), hh(2 * (m + 1) + 1);
    gv[0] = beta;
    int kend = 0;
    for (int k = 0; k < m; ++k) {
        double* w = V + i64(k + 1) * n;
        // w = M^-1 A v_k
        OXI_TRY(matprod_partials(c, V + i64(k) * n, n, 1, false, nullptr));
        hipLaunchKernelGGL(mv_final_kernel, dim3(g), dim3(TPB), 0, c.s, n, c.nchunk, c.P, c.tmp);
        OXI_TRY(precond(c, c.tmp, w));
        // classical Gram-Schmidt, twice
        OXI_TRY(dots(c, V, k + 1, w, h1));
        hipLaunchKernelGGL(vupdate_kernel, dim3(g), dim3(TPB), 0, c.s, V, n, k + 1, h1, -1.0, w, n);
        OXI_TRY(dots(c, V, k + 1, w, h2));
        hipLaunchKernelGGL(vupdate_kernel, dim3(g), dim3(TPB), 0, c.s, V, n, k + 1, h2, -1.0, w, n);
        OXI_TRY(dots(c, w, 1, w, hn));
        OXI_HIP(c.h, hipMemcpyAsync(hh.data(), c.hdev, sizeof(double) * (2 * (m + 1) + 1), hipMemcpyDeviceToHost,
                                    c.s));
        OXI_HIP(c.h, hipStreamSynchronize(c.s));
        ++*inner;
        double* col = &H[size_t(k) * (m + 1)];
        for (int i = 0; i <= k; ++i) col[i] = hh[i] + hh[(m + 1) + i];
        const double hk1 = std::sqrt(hh[2 * (m + 1)]);
        col[k + 1] = hk1;
        for (int i = 0; i < k; ++i) {
            const double t0 = cs[i] * col[i] + sn[i] * col[i + 1];
            col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
            col[i] = t0;
        }
        const double r = std::hypot(col[k], col[k + 1]);
        cs[k] = r > 0 ? col[k] / r : 1.0;
        sn[k] = r > 0 ? col[k + 1] / r : 0.0;
        col[k] = r;
        col[k + 1] = 0;
        gv[k + 1] = -sn[k] * gv[k];
        gv[k] = cs[k] * gv[k];
        kend = k + 1;
        if (!(std::fabs(gv[k + 1]) > tau * beta) || !(hk1 > 0) || !std::isfinite(hk1)) break;
        hipLaunchKernelGGL(vdiv_kernel, dim3(g), dim3(TPB), 0, c.s, w, n, hk1);
    }
    // H y = g (upper triangular, kend x kend); X(:, j) += V y
    std::vector<double> y(kend);
    for (int i = kend - 1; i >= 0; --i) {
        double s = gv[i];
        for (int q = i + 1; q < kend; ++q) s -= H[size_t(q) * (m + 1) + i] * y[q];
        y[i] = H[size_t(i) * (m + 1) + i] != 0 ? s / H[size_t(i) * (m + 1) + i] : 0.0;
    }
    OXI_HIP(c.h, hipMemcpyAsync(ydev, y.data(), sizeof(double) * kend, hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(vupdate_kernel, dim3(g), dim3(TPB), 0, c.s, V, n, kend, ydev, 1.0, c.X + j * c.ldx, n);
    OXI_TRY(launch_ok(c.h, "gmres update"));
    return OXI_OK;
}

// Equilibration of A (m x n) on the stream: scales, optional reciprocals and
// row sums of |A|, ||A||_inf into *anorm; E (TO) = (A ./ rows) ./ cols when E
// is given.  pmax/psum: cdiv(n, EQ_KC) * m doubles each.  No synchronisation.
constexpr i64 EQ_KC = 256;
template <typename TO>
int equilibrate_dev(oxi_handle_t h, hipStream_t s, i64 m, i64 n, const double* A, i64 lda, int radix, double* rows,
                    double* cols, double* rinv, double* cinv, double* rowsum, double* pmax, double* psum,
                    unsigned* flags, u64* anorm, TO* E, i64 lde) {
    const int nchunk = int(cdiv(n, EQ_KC));
    hipLaunchKernelGGL(row_stats_kernel, dim3(unsigned(cdiv(m, TPB)), unsigned(nchunk)), dim3(TPB), 0, s, A, lda, m,
                       n, EQ_KC, pmax, psum, flags);
    hipLaunchKernelGGL(row_scale_kernel, dim3(unsigned(cdiv(m, TPB))), dim3(TPB), 0, s, m, nchunk, pmax, psum, radix,
                       rows, rinv, rowsum, anorm);
    hipLaunchKernelGGL(col_scale_kernel, dim3(unsigned(std::min<i64>(n, 65535))), dim3(TPB), 0, s, A, lda, m, n,
                       rows, radix, cols, cinv);
    if (E)
        hipLaunchKernelGGL(equil_convert_kernel<TO>,
                           dim3(unsigned(cdiv(m, TPB)), unsigned(std::max<i64>(1, std::min<i64>(n, 65535)))),
                           dim3(TPB), 0, s, A, lda, m, n, rows, cols, E, lde, flags);
    return launch_ok(h, "equilibrate");
}

struct EventPair {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    ~EventPair() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

}  // namespace

extern "C" int oxi_dgesv_mixed(oxi_handle_t h, int prec, int solver, int64_t n, int64_t nrhs, const double* A,
                               int64_t lda, const double* B, int64_t ldb, double* X, int64_t ldx, int32_t maxiter,
                               double rtol, int32_t gmres_restart, int32_t pivot, oxi_ir_info* info) {
    if (!h) return -1;
    if (prec < OXI_PREC_BF16 || prec > OXI_PREC_FP32) return -2;
    if (solver < 0 || solver > 2) return -3;
    if (n < 0) return -4;
    if (nrhs < 0) return -5;
    if (lda < std::max<i64>(1, n)) return -7;
    if (ldb < std::max<i64>(1, n)) return -9;
    if (ldx < std::max<i64>(1, n)) return -11;
    if (maxiter < 0) return -12;
    if (!std::isfinite(rtol) || rtol < 0) return -13;
    if (gmres_restart < 0 || gmres_restart > 1000) return -14;
    if (pivot < 0 || pivot > 2) return -15;
    if (!info) return -16;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = oxi_ir_info{};
    info->converged = 1;
    if (n == 0 || nrhs == 0) return OXI_OK;
    if (!A) return -6;
    if (!B) return -8;
    if (!X) return -10;
    (void)hipSetDevice(h->device);

    Mixed c;
    c.h = h; c.s = h->stream; c.n = n; c.nrhs = nrhs;
    c.A = A; c.lda = lda; c.B = B; c.ldb = ldb; c.X = X; c.ldx = ldx;
    plan_residual(c);
    const bool want_gmres = solver != 0;
    c.m = want_gmres ? (gmres_restart > 0 ? gmres_restart : 30) : 0;
    c.nseg = int(std::min<i64>(64, cdiv(n, TPB)));
    c.seglen = cdiv(n, c.nseg);
    const int nrowchunk = int(cdiv(n, EQ_KC));

    // workspace layout
    const size_t nn = size_t(n) * size_t(n), nr = size_t(n) * size_t(nrhs);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += pad(bytes); return o; };
    const size_t oA32 = take(nn * 4), oipiv = take(size_t(n) * 8), operm = take(size_t(n) * 8);
    const size_t orows = take(size_t(n) * 8), ocols = take(size_t(n) * 8), orinv = take(size_t(n) * 8),
                 ocinv = take(size_t(n) * 8), orowsum = take(size_t(n) * 8);
    const size_t oR = take(nr * 8), oS = take(nr * 8), oYd = take(nr * 8), oYwd = take(nr * 8), oYf = take(nr * 4),
                 oYwf = take(nr * 4);
    const size_t partial = std::max(size_t(c.nchunk) * nr, size_t(nrowchunk) * size_t(n));
    const size_t oP = take(partial * 8), oQ = take(partial * 8);
    const size_t ostats = take(8 * 8), oflags = take(8 * 4), omag = take(8),
                 octrl = take(oxi_trsv_ws_bytes(n, sizeof(double)));
    const size_t opart = take(size_t(c.nseg) * size_t(c.m + 1) * 8), ohdev = take(size_t(4 * (c.m + 1) + 1) * 8);
    // the random butterfly transform applies to the FP32 factor and needs
    // n % 4 == 0; otherwise the pivoted factor runs (factors rounded below
    // FP32 without pivoting lose too much to unbounded multipliers)
    const bool want_rbt = pivot == 1 && n % 4 == 0 && prec == OXI_PREC_FP32;
    const int lprec = prec == OXI_PREC_FP32 ? -1 : prec;
    const size_t nopiv_bytes = want_rbt ? oxi_getrf_nopiv_ws(lprec, n) : 0;
    size_t oR32 = 0, oR64 = 0, onopiv = 0, odinfo = 0;
    if (want_rbt) {
        oR32 = take(size_t(4) * n * 4);
        oR64 = take(size_t(4) * n * 8);
        odinfo = take(8);
        onopiv = take(nopiv_bytes);
    }
    size_t oV = 0, otmp = 0;
    if (want_gmres) {
        oV = take(size_t(n) * size_t(c.m + 1) * 8);
        otmp = take(size_t(n) * 8);
    }
    void* ws;
    OXI_TRY(oxi_workspace2(h, off, &ws));
    char* base = static_cast<char*>(ws);
    c.A32 = (float*)(base + oA32);
    c.ipiv = (i64*)(base + oipiv);
    c.dperm = (i64*)(base + operm);
    c.rows = (double*)(base + orows); c.cols = (double*)(base + ocols);
    c.rinv = (double*)(base + orinv); c.cinv = (double*)(base + ocinv); c.rowsum = (double*)(base + orowsum);
    c.R = (double*)(base + oR); c.S = (double*)(base + oS); c.Yd = (double*)(base + oYd); c.Ywd = (double*)(base + oYwd);
    c.Yf = (float*)(base + oYf); c.Ywf = (float*)(base + oYwf);
    c.P = (double*)(base + oP); c.Q = (double*)(base + oQ);
    c.stats = (u64*)(base + ostats); c.flags = (unsigned*)(base + oflags); c.magout = (double*)(base + omag);
    c.ctrl = base + octrl;
    c.part = (double*)(base + opart); c.hdev = (double*)(base + ohdev);
    if (want_gmres) { c.V = (double*)(base + oV); c.tmp = (double*)(base + otmp); }
    if (want_rbt) { c.R32 = (float*)(base + oR32); c.R64 = (double*)(base + oR64); }

    EventPair ev;
    for (auto& e : ev.e) OXI_HIP(h, hipEventCreate(&e));
    OXI_HIP(h, hipEventRecord(ev.e[0], c.s));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, c.stats, 8);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, (u64*)c.flags, 4);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, (u64*)c.ctrl, 2);   // chain-kernel status

    // 1. finite check + equilibration
    hipLaunchKernelGGL(finite_kernel, c.vgrid(nrhs), dim3(TPB), 0, c.s, B, ldb, n, nrhs, c.flags);
    OXI_TRY(equilibrate_dev<float>(h, c.s, n, n, A, lda, 1, c.rows, c.cols, c.rinv, c.cinv, c.rowsum, c.P, c.Q,
                                   c.flags, &c.stats[4], c.A32, n));
    {
        u64 st[8];
        unsigned fl[8];
        OXI_TRY(read_stats(c, st, fl));
        if (fl[0]) return oxi_set_error(h, OXI_ERR_NOT_FINITE, "mixed_solve: finite mixed input required");
        c.anorm = from_bits(st[4]);
        // 2. low-precision factor (random butterflies and no pivoting, or
        // partial pivoting), FP64 fallback
        i64 finfo = 0;
        if (!fl[1] && want_rbt) {
            i64* dinfo = (i64*)(base + odinfo);
            OXI_TRY(oxi_rbt_make(h, c.s, n, c.R32, c.R64));
            OXI_TRY(oxi_rbt_matrix(h, c.s, n, c.A32, n, c.R32));
            hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, (u64*)dinfo, 1);
            OXI_TRY(launch_ok(h, "rbt"));
            OXI_TRY(oxi_getrf_nopiv32(h, c.s, lprec, n, c.A32, n, dinfo, base + onopiv, nopiv_bytes));
            i64 zinfo = 0;
            OXI_HIP(h, hipMemcpyAsync(&zinfo, dinfo, 8, hipMemcpyDeviceToHost, c.s));
            OXI_HIP(h, hipStreamSynchronize(c.s));
            if (zinfo == 0) {
                c.rbt = true;
                std::vector<i64> ident(n);
                for (i64 i = 0; i < n; ++i) ident[i] = i + 1;
                OXI_HIP(h, hipMemcpyAsync(c.ipiv, ident.data(), size_t(n) * 8, hipMemcpyHostToDevice, c.s));
                OXI_HIP(h, hipStreamSynchronize(c.s));
            } else {
                // a zero pivot after the transform: the pivoted factor of the original matrix
                hipLaunchKernelGGL(equil_convert_kernel<float>,
                                   dim3(unsigned(cdiv(n, TPB)), unsigned(std::min<i64>(n, 65535))), dim3(TPB), 0, c.s,
                                   A, lda, n, n, c.rows, c.cols, c.A32, n, c.flags);
                OXI_TRY(launch_ok(h, "equilibrate"));
            }
        }
        if (!fl[1] && !c.rbt) OXI_TRY(oxi_getrf32(h, n, n, c.A32, n, c.ipiv, 0, &finfo, lprec, pivot == 2));
        if (fl[1] || finfo != 0) {
            c.fallback = true;
            if (hipMalloc(&c.E64, nn * 8) != hipSuccess) {
                c.E64 = nullptr;
                return oxi_set_error(h, OXI_ERR_ALLOC, "mixed_solve: FP64 fallback factor allocation failed");
            }
        }
    }
    struct FreeE64 { double*& p; ~FreeE64() { if (p) { (void)hipDeviceSynchronize(); (void)hipFree(p); } } } free_e64{c.E64};
    if (c.fallback) {
        hipLaunchKernelGGL(equil_convert_kernel<double>, dim3(unsigned(cdiv(n, TPB)), unsigned(std::min<i64>(n, 65535))),
                           dim3(TPB), 0, c.s, A, lda, n, n, c.rows, c.cols, c.E64, n, c.flags);
        OXI_TRY(launch_ok(h, "equilibrate"));
        i64 finfo = 0;
        OXI_TRY(oxi_dgetrf(h, n, n, c.E64, n, c.ipiv, 0, &finfo));
        if (finfo != 0)
            return oxi_set_error(h, OXI_ERR_SINGULAR, "mixed_solve: singular matrix (U(%lld,%lld) is zero)",
                                 (long long)finfo, (long long)finfo);
    }
    {
        std::vector<i64> perm;
        OXI_TRY(oxi_perm_from_ipiv(h, c.s, c.ipiv, n, perm));
        OXI_HIP(h, hipMemcpyAsync(c.dperm, perm.data(), size_t(n) * 8, hipMemcpyHostToDevice, c.s));
    }
    OXI_HIP(h, hipEventRecord(ev.e[1], c.s));

    // 3-5. first solve and refinement
    OXI_TRY(correction(c, B, ldb, true));
    double err = INFINITY, nerr = INFINITY, prev = INFINITY;
    bool xfinite = true, used_gmres = false;
    i64 iterations = 0, inner = 0;
    int mode = solver == 1 ? 1 : 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        OXI_TRY(residual(c, &err, &nerr, &xfinite));
        iterations = iteration;
        if (err <= rtol || iteration == maxiter || !std::isfinite(err)) break;
        if (solver == 2 && mode == 0 && iteration > 0 && !(err <= 0.5 * prev)) mode = 1;   // refinement stalls
        prev = err;
        if (mode == 0) {
            OXI_TRY(correction(c, c.R, n, false));
        } else {
            used_gmres = true;
            // a GMRES cycle reduces the preconditioned residual by tau: by
            // 1e-8, or by a hundredth of what the backward error still has
            // to lose when that asks for less (at most 1e-2)
            const double tau = std::max(1e-8, std::min(1e-2, 0.01 * rtol / err));
            for (i64 j = 0; j < nrhs; ++j) OXI_TRY(gmres_correction(c, j, tau, &inner));
        }
    }
    OXI_HIP(h, hipEventRecord(ev.e[2], c.s));
    // nrm2 of the last residual (the reference's `residual` field)
    double rss = 0;
    for (i64 j = 0; j < nrhs; ++j) {
        OXI_TRY(dots(c, c.R + j * n, 1, c.R + j * n, c.hdev));
        double v = 0;
        OXI_HIP(h, hipMemcpyAsync(&v, c.hdev, 8, hipMemcpyDeviceToHost, c.s));
        OXI_HIP(h, hipStreamSynchronize(c.s));
        rss += v;
    }
    OXI_TRY(oxi_trsv_check(h, c.s, c.ctrl));
    OXI_HIP(h, hipEventSynchronize(ev.e[2]));
    float f_ms = 0, r_ms = 0;
    (void)hipEventElapsedTime(&f_ms, ev.e[0], ev.e[1]);
    (void)hipEventElapsedTime(&r_ms, ev.e[1], ev.e[2]);

    const bool finite = xfinite && std::isfinite(err);
    info->iterations = iterations;
    info->inner_iterations = inner;
    info->converged = err <= rtol;
    info->fallback = c.fallback;
    info->reason = !finite ? 2 : c.fallback ? 3 : err <= rtol ? 0 : 1;
    info->solver = used_gmres ? 1 : 0;
    info->backward_error = err;
    info->normwise_backward_error = nerr;
    info->residual_norm = std::sqrt(rss);
    info->factor_ms = f_ms;
    info->refine_ms = r_ms;
    return OXI_OK;
}

// mixed_solve(A, B; factorization=:cholesky) (src/lapack/extended_lapack.jl:175-226
// with _mixed_factor's :cholesky, :161): A scaled by its largest |entry|,
// the FP32 Cholesky of the lower triangle, FP64 fallback on
// PosDefException, the reference's correction step with L and L^T solves,
// classical refinement.
extern "C" int oxi_dposv_mixed(oxi_handle_t h, int prec, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                               const double* B, int64_t ldb, double* X, int64_t ldx, int32_t maxiter, double rtol,
                               oxi_ir_info* info) {
    if (!h) return -1;
    if (prec != OXI_PREC_FP32) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (lda < std::max<i64>(1, n)) return -6;
    if (ldb < std::max<i64>(1, n)) return -8;
    if (ldx < std::max<i64>(1, n)) return -10;
    if (maxiter < 0) return -11;
    if (!std::isfinite(rtol) || rtol < 0) return -12;
    if (!info) return -13;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = oxi_ir_info{};
    info->converged = 1;
    if (n == 0 || nrhs == 0) return OXI_OK;
    if (!A) return -5;
    if (!B) return -7;
    if (!X) return -9;
    (void)hipSetDevice(h->device);

    Mixed c;
    c.h = h; c.s = h->stream; c.n = n; c.nrhs = nrhs; c.chol = true;
    c.A = A; c.lda = lda; c.B = B; c.ldb = ldb; c.X = X; c.ldx = ldx;
    plan_residual(c);
    const int nrowchunk = int(cdiv(n, EQ_KC));
    const size_t nn = size_t(n) * size_t(n), nr = size_t(n) * size_t(nrhs);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += pad(bytes); return o; };
    const size_t oA32 = take(nn * 4);
    const size_t orows = take(size_t(n) * 8), ocols = take(size_t(n) * 8), orinv = take(size_t(n) * 8),
                 ocinv = take(size_t(n) * 8), orowsum = take(size_t(n) * 8);
    const size_t oR = take(nr * 8), oS = take(nr * 8), oYd = take(nr * 8), oYf = take(nr * 4);
    const size_t partial = std::max(size_t(c.nchunk) * nr, size_t(nrowchunk) * size_t(n));
    const size_t oP = take(partial * 8), oQ = take(partial * 8);
    const size_t ostats = take(8 * 8), oflags = take(8 * 4), omag = take(8), oscale = take(8), odinfo = take(8);
    void* ws;
    OXI_TRY(oxi_workspace2(h, off, &ws));
    char* base = static_cast<char*>(ws);
    c.A32 = (float*)(base + oA32);
    c.rows = (double*)(base + orows); c.cols = (double*)(base + ocols);
    c.rinv = (double*)(base + orinv); c.cinv = (double*)(base + ocinv); c.rowsum = (double*)(base + orowsum);
    c.R = (double*)(base + oR); c.S = (double*)(base + oS); c.Yd = (double*)(base + oYd); c.Yf = (float*)(base + oYf);
    c.P = (double*)(base + oP); c.Q = (double*)(base + oQ);
    c.stats = (u64*)(base + ostats); c.flags = (unsigned*)(base + oflags); c.magout = (double*)(base + omag);
    u64* scalebits = (u64*)(base + oscale);
    i64* dinfo = (i64*)(base + odinfo);

    EventPair ev;
    for (auto& e : ev.e) OXI_HIP(h, hipEventCreate(&e));
    OXI_HIP(h, hipEventRecord(ev.e[0], c.s));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, c.stats, 8);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, (u64*)c.flags, 4);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, c.s, scalebits, 1);
    // 1. finite check, ||A||_inf, the global scale and the FP32 copy
    hipLaunchKernelGGL(finite_kernel, c.vgrid(nrhs), dim3(TPB), 0, c.s, B, ldb, n, nrhs, c.flags);
    hipLaunchKernelGGL(row_stats_kernel, dim3(unsigned(cdiv(n, TPB)), unsigned(nrowchunk)), dim3(TPB), 0, c.s, A, lda,
                       n, n, EQ_KC, c.P, c.Q, c.flags);
    hipLaunchKernelGGL(row_scale_kernel, dim3(unsigned(cdiv(n, TPB))), dim3(TPB), 0, c.s, n, nrowchunk, c.P, c.Q, 0,
                       c.rows, c.rinv, c.rowsum, &c.stats[4]);
    hipLaunchKernelGGL(global_max_kernel, dim3(unsigned(std::min<i64>(cdiv(i64(nrowchunk) * n, TPB), 4096))), dim3(TPB),
                       0, c.s, c.P, i64(nrowchunk) * n, scalebits);
    hipLaunchKernelGGL(global_scale_kernel<float>,
                       dim3(unsigned(cdiv(n, TPB)), unsigned(std::min<i64>(n, 65535))), dim3(TPB), 0, c.s, A, lda, n,
                       scalebits, c.rows, c.rinv, c.cols, c.cinv, c.A32, n, c.flags);
    OXI_TRY(launch_ok(h, "mixed cholesky scaling"));
    {
        u64 st[8];
        unsigned fl[8];
        OXI_TRY(read_stats(c, st, fl));
        if (fl[0]) return oxi_set_error(h, OXI_ERR_NOT_FINITE, "mixed_solve: finite mixed input required");
        c.anorm = from_bits(st[4]);
        // 2. the FP32 Cholesky; FP64 on PosDefException or overflow
        i64 finfo = 0;
        if (!fl[1]) {
            OXI_TRY(oxi_potrf_device(h, c.s, n, c.A32, n, dinfo));
            OXI_HIP(h, hipMemcpyAsync(&finfo, dinfo, 8, hipMemcpyDeviceToHost, c.s));
            OXI_HIP(h, hipStreamSynchronize(c.s));
        }
        if (fl[1] || finfo != 0) {
            c.fallback = true;
            if (hipMalloc(&c.E64, nn * 8) != hipSuccess) {
                c.E64 = nullptr;
                return oxi_set_error(h, OXI_ERR_ALLOC, "mixed_solve: FP64 fallback factor allocation failed");
            }
        }
    }
    struct FreeE64 { double*& p; ~FreeE64() { if (p) { (void)hipDeviceSynchronize(); (void)hipFree(p); } } } free_e64{c.E64};
    if (c.fallback) {
        hipLaunchKernelGGL(global_scale_kernel<double>,
                           dim3(unsigned(cdiv(n, TPB)), unsigned(std::min<i64>(n, 65535))), dim3(TPB), 0, c.s, A, lda,
                           n, scalebits, c.rows, c.rinv, c.cols, c.cinv, c.E64, n, c.flags);
        OXI_TRY(launch_ok(h, "mixed cholesky scaling"));
        OXI_TRY(oxi_potrf_device(h, c.s, n, c.E64, n, dinfo));
        i64 finfo = 0;
        OXI_HIP(h, hipMemcpyAsync(&finfo, dinfo, 8, hipMemcpyDeviceToHost, c.s));
        OXI_HIP(h, hipStreamSynchronize(c.s));
        if (finfo != 0)
            return oxi_set_error(h, OXI_ERR_NOT_POSDEF, "mixed_solve: not positive definite (A(%lld,%lld))",
                                 (long long)finfo, (long long)finfo);
    }
    OXI_HIP(h, hipEventRecord(ev.e[1], c.s));
    // 3-5. first solve and classical refinement (the reference's loop)
    OXI_TRY(correction(c, B, ldb, true));
    double err = INFINITY, nerr = INFINITY;
    bool xfinite = true;
    i64 iterations = 0;
    for (int iteration = 0; iteration <= maxiter; ++iteration) {
        OXI_TRY(residual(c, &err, &nerr, &xfinite));
        iterations = iteration;
        if (err <= rtol || iteration == maxiter || !std::isfinite(err)) break;
        OXI_TRY(correction(c, c.R, n, false));
    }
    OXI_HIP(h, hipEventRecord(ev.e[2], c.s));
    double rss = 0;
    {
        // nrm2 of the last residual: the partial-sum machinery of the residual, by column
        std::vector<double> hR(nr);
        OXI_HIP(h, hipMemcpyAsync(hR.data(), c.R, nr * 8, hipMemcpyDeviceToHost, c.s));
        OXI_HIP(h, hipStreamSynchronize(c.s));
        for (double v : hR) rss += v * v;
    }
    OXI_HIP(h, hipEventSynchronize(ev.e[2]));
    float f_ms = 0, r_ms = 0;
    (void)hipEventElapsedTime(&f_ms, ev.e[0], ev.e[1]);
    (void)hipEventElapsedTime(&r_ms, ev.e[1], ev.e[2]);
    const bool finite = xfinite && std::isfinite(err);
    info->iterations = iterations;
    info->inner_iterations = 0;
    info->converged = err <= rtol;
    info->fallback = c.fallback;
    info->reason = !finite ? 2 : c.fallback ? 3 : err <= rtol ? 0 : 1;
    info->solver = 0;
    info->backward_error = err;
    info->normwise_backward_error = nerr;
    info->residual_norm = std::sqrt(rss);
    info->factor_ms = f_ms;
    info->refine_ms = r_ms;
    return OXI_OK;
}

extern "C" int oxi_dequilibrate(oxi_handle_t h, int64_t m, int64_t n, const double* A, int64_t lda, double* E,
                                int64_t lde, double* rows, double* cols, int radix) {
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (lda < std::max<i64>(1, m)) return -5;
    if (lde < std::max<i64>(1, m)) return -7;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (m == 0 && n == 0) return OXI_OK;
    if (m > 0 && !rows) return -8;
    if (n > 0 && !cols) return -9;
    if (m > 0 && n > 0 && (!A || !E)) return !A ? -4 : -6;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    if (m == 0 || n == 0) {
        // no entries: every scale is one
        std::vector<double> ones(size_t(std::max<i64>(m, n)), 1.0);
        if (m > 0) OXI_HIP(h, hipMemcpyAsync(rows, ones.data(), size_t(m) * 8, hipMemcpyHostToDevice, s));
        if (n > 0) OXI_HIP(h, hipMemcpyAsync(cols, ones.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        return OXI_OK;
    }
    const size_t part = size_t(cdiv(n, EQ_KC)) * size_t(m) * 8;
    void* ws;
    OXI_TRY(oxi_workspace(h, 2 * pad(part) + 256, &ws));
    double* pmax = (double*)ws;
    double* psum = (double*)((char*)ws + pad(part));
    unsigned* flags = (unsigned*)((char*)ws + 2 * pad(part));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, (u64*)flags, 4);
    // the scales first, then check, then write E (E may be A)
    OXI_TRY(equilibrate_dev<double>(h, s, m, n, A, lda, radix ? 1 : 0, rows, cols, nullptr, nullptr, nullptr, pmax,
                                    psum, flags, nullptr, nullptr, 0));
    unsigned fl[8];
    OXI_HIP(h, hipMemcpyAsync(fl, flags, sizeof(fl), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    if (fl[0]) return oxi_set_error(h, OXI_ERR_NOT_FINITE, "equilibrate: non-finite entry");
    hipLaunchKernelGGL(equil_convert_kernel<double>,
                       dim3(unsigned(cdiv(m, TPB)), unsigned(std::min<i64>(n, 65535))), dim3(TPB), 0, s, A, lda, m, n,
                       rows, cols, E, lde, flags);
    OXI_TRY(launch_ok(h, "equilibrate"));
    OXI_HIP(h, hipStreamSynchronize(s));
    return OXI_OK;
}
