// Triangular solve with one right-hand side against LU factors:
// trsv!(x, A; uplo, trans, unit)  (src/blas/level2.jl trsv!, the vector branch
// of the LU solve, src/api/linalg_api.jl:168-184).
//
// The n rows are cut into blocks of RB=128 rows in solve order; block b needs
// x of every earlier block.
//
// Preparation (oxi_trsv_prepare, once per factor and triangle): the inverse
// of every RB x RB diagonal block, computed by parallel column sweeps, in the
// arithmetic type.  A solve then does no sequential sweep: each block's x is
// a matrix-vector product with its inverse.
//
// The solve (one launch per triangle): workgroups take block numbers from an
// atomic ticket, so a workgroup only ever waits on blocks owned by
// workgroups that are already running (no dependence on dispatch order, no
// deadlock).  A workgroup
//   1. stages the inverse of its diagonal block in LDS,
//   2. streams the off-diagonal part of its block row, chunk by chunk, as the
//      earlier blocks publish their x (one agent-scope `ready` counter that
//      only grows: block b stores b+1 after its x is written),
//   3. reduces the partial sums into r = b_blk - sum and forms x = inv * r,
//   4. writes its x and then bumps `ready`.
// The matrix is read once (HBM-bound: n^2/2 elements per triangle).
//
// The last chunk is taken out of the dependence chain: with G_b = inv_b *
// T_{b,b-1} (prepared with the inverses), a block computes
// y = inv_b (b_blk - sum over chunks < b-1) while block b-1 is still being
// solved, and only x = y - G_b x_{b-1} waits for block b-1.  The chain per
// block is then one publish, one x fetch and one 128 x 128 product from
// registers.  Using the block inverse instead of substitution inside the block
// changes the rounding (not the result beyond it); the solve stays
// deterministic.
//
// The matrix type TM and the arithmetic/vector type TV may differ: FP32
// factors solved in FP64 arithmetic is the preconditioner of GMRES-IR.
// Solve order and memory access: `which` 0 = L (unit) forward, 1 = U backward,
// 2 = U^T forward, 3 = L^T (unit) backward; backward solves use logical index
// i' = n-1-i, so the same code runs both directions with coalesced access.
#include "oxi_internal.hpp"

#include <cstdlib>

namespace {

constexpr int RB = 128, NT = 512;
constexpr int XB = 8;                          // published chunks whose x is fetched in one round
constexpr unsigned SPIN_LIMIT = 1u << 20;

struct TrsvCtrl { unsigned ticket, ready, err, pad; };

template <typename TM> struct Vec;
template <> struct Vec<double> { typedef double type __attribute__((ext_vector_type(2))); static constexpr int N = 2; };
template <> struct Vec<float> { typedef float type __attribute__((ext_vector_type(4))); static constexpr int N = 4; };

__device__ __forceinline__ double bcast(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(u), l);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(u >> 32), l);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
__device__ __forceinline__ float bcast(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), l));
}
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// x is exchanged between workgroups with device-scope (agent) relaxed stores
// and loads, which are coherent across XCDs; the publishing wave waits for
// its stores to be acknowledged before it bumps `ready`.  (An agent-scope
// release/acquire pair would write back and invalidate the L2 instead.)
template <typename V>
__device__ __forceinline__ void put_agent(V* p, V v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V>
__device__ __forceinline__ V get_agent(const V* p) {
    return __hip_atomic_load(const_cast<V*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per launch: the ticket and the ready counter start at zero.  The error word
// is sticky across the launches of one solve (a timed-out launch leaves x
// unwritten, and a later launch must not hide that); it is cleared once per
// solve by trsv_status_clear and read at the end by oxi_trsv_check.
__global__ void trsv_ctrl_reset(TrsvCtrl* c) {
    if (threadIdx.x == 0) { c->ticket = 0; c->ready = 0; }
}
__global__ void trsv_status_clear(TrsvCtrl* c) {
    if (threadIdx.x == 0) { c->ticket = 0; c->ready = 0; c->err = 0; c->pad = 0; }
}

// The diagonal block of block b, D[k * (RB + 1) + i] = T(i, k) in logical
// (solve-order) indices, into LDS.
template <typename TM, bool TRANS, bool BWD>
__device__ __forceinline__ void load_diag(const TM* __restrict__ M, i64 ld, i64 p0, int rows, TM* D, int t) {
    for (int e = t; e < RB * RB; e += NT) {
        const int pr = e % RB, pc = e / RB;
        if (pr < rows && pc < rows) {
            const TM v = M[(p0 + pr) + (p0 + pc) * ld];
            const int lr = BWD ? rows - 1 - pr : pr, lc = BWD ? rows - 1 - pc : pc;
            const int i = TRANS ? lc : lr, k = TRANS ? lr : lc;
            D[k * (RB + 1) + i] = v;
        }
    }
}

// Inverse of every diagonal block (logical indices, lower triangular):
// inv[b][k * RB + i] = Dinv(i, k), zero above the diagonal and outside the
// block.  Wave w solves T X = I for columns 16w .. 16w+15, four at a time, by
// column sweep (lane i owns rows i and i+64; x_k broadcast by readlane).
template <typename TM, typename TV, bool TRANS, bool UNIT, bool BWD>
__global__ __launch_bounds__(NT) void trsv_dinv_kernel(const TM* __restrict__ M, i64 ld, i64 n, TV* __restrict__ inv) {
    __shared__ TM D[RB * (RB + 1)];
    __shared__ TV rinv[RB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int b = blockIdx.x;
    const i64 lo = i64(b) * RB;
    const int rows = int(min(i64(RB), n - lo));
    const i64 p0 = BWD ? n - lo - rows : lo;
    load_diag<TM, TRANS, BWD>(M, ld, p0, rows, D, t);
    __syncthreads();
    if (!UNIT)
        for (int k = t; k < rows; k += NT) rinv[k] = TV(1) / TV(D[k * (RB + 1) + k]);
    __syncthreads();
    TV* out = inv + size_t(b) * RB * RB;
    constexpr int CPW = RB / (NT / 64);   // columns per wave
    for (int c0 = CPW * w; c0 < CPW * (w + 1); c0 += 4) {
        TV s[4][2];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            s[r][0] = lane == c0 + r ? TV(1) : TV(0);
            s[r][1] = lane + 64 == c0 + r ? TV(1) : TV(0);
        }
        for (int k = c0; k < rows; ++k) {
            const TV d0 = TV(D[k * (RB + 1) + lane]), d1 = TV(D[k * (RB + 1) + 64 + lane]);
            const TV rk = UNIT ? TV(1) : rinv[k];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const TV sk = k < 64 ? s[r][0] : s[r][1];
                const TV xk = bcast(UNIT ? sk : sk * rk, k & 63);
                if (lane > k) s[r][0] = fma_t(-d0, xk, s[r][0]);
                if (lane + 64 > k) s[r][1] = fma_t(-d1, xk, s[r][1]);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = c0 + r;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = lane + 64 * h;
                TV v = TV(0);
                if (c < rows && i < rows) v = UNIT ? s[r][h] : s[r][h] * rinv[i];
                out[c * RB + i] = v;
            }
        }
    }
}

// G_b = inv_b * T_{b,b-1} for b >= 1, G[q * RB + i] over logical rows i of
// block b and physical columns q of block b-1 (the order the chain kernel
// holds x_{b-1} in), as the chain kernel's last chunk would use them.  The
// block T_{b,b-1} goes to LDS in the working type; the inverse is read from
// memory.  Block 0 has no G.
template <typename TM, typename TV, bool TRANS, bool BWD>
__global__ __launch_bounds__(NT) void trsv_g_kernel(const TM* __restrict__ M, i64 ld, i64 n, const TV* __restrict__ inv,
                                                    TV* __restrict__ G) {
    __shared__ TV Tp[RB * (RB + 1)];   // Tp[q * (RB + 1) + pr]: physical row pr of block b, column q of block b-1
    const int t = threadIdx.x;
    const int b = blockIdx.x + 1;
    const i64 lo = i64(b) * RB;
    const int rows = int(min(i64(RB), n - lo));
    const i64 p0 = BWD ? n - lo - rows : lo;
    const i64 q0 = BWD ? n - i64(b) * RB : i64(b - 1) * RB;
    for (int e = t; e < RB * RB; e += NT) {
        const int pr = e % RB, q = e / RB;
        TV v = TV(0);
        if (pr < rows) v = TV(TRANS ? M[(q0 + q) + (p0 + pr) * ld] : M[(p0 + pr) + (q0 + q) * ld]);
        Tp[q * (RB + 1) + pr] = v;
    }
    __syncthreads();
    const TV* di = inv + size_t(b) * RB * RB;
    TV* g = G + size_t(b) * RB * RB;
    const int i = t % RB, qg = t / RB;
    constexpr int QPT = RB / (NT / RB);
    TV acc[QPT];
#pragma unroll
    for (int k = 0; k < QPT; ++k) acc[k] = TV(0);
    for (int j = 0; j < rows; ++j) {
        const TV d = di[j * RB + i];
        const int pr = BWD ? rows - 1 - j : j;
#pragma unroll
        for (int k = 0; k < QPT; ++k) acc[k] = fma_t(d, Tp[(qg * QPT + k) * (RB + 1) + pr], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < QPT; ++k) g[(qg * QPT + k) * RB + i] = acc[k];
}

// Two device-scope loads issued back to back and waited for once (relaxed
// atomic loads written in C++ are issued one at a time, each waited for).
__device__ __forceinline__ void load2_agent(const double* p0, const double* p1, double& a, double& b) {
    asm volatile("global_load_dwordx2 %0, %2, off sc1\n\tglobal_load_dwordx2 %1, %3, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=&v"(a), "=&v"(b) : "v"(p0), "v"(p1) : "memory");
}
__device__ __forceinline__ void load2_agent(const float* p0, const float* p1, float& a, float& b) {
    asm volatile("global_load_dword %0, %2, off sc1\n\tglobal_load_dword %1, %3, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=&v"(a), "=&v"(b) : "v"(p0), "v"(p1) : "memory");
}

template <typename TM, typename TV, bool TRANS, bool UNIT, bool BWD, bool VECOK>
__global__ __launch_bounds__(NT) void trsv_chain_kernel(const TM* __restrict__ M, i64 ld, i64 n, TV* x,
                                                        TrsvCtrl* ctrl, const TV* __restrict__ inv,
                                                        const TV* __restrict__ G) {
    constexpr int VEC = VECOK ? Vec<TM>::N : 1;
    constexpr int LPC = RB / VEC;              // no-trans: lanes per column
    constexpr int CG = NT / LPC;               // no-trans: column groups
    constexpr int NRED = TRANS ? 1 : CG;
    constexpr int LPR = 8;                     // trans: lanes per row
    constexpr int RPT = RB / (NT / LPR);       // trans: rows per thread
    constexpr int ITS = RB / (LPR * VEC);      // trans: iterations per chunk
    constexpr int KP = NT / RB;                // product: threads per row
    static_assert(XB * RB == 2 * NT, "two x loads per thread per round");
    __shared__ TV DI[RB * RB];                 // DI[k * RB + i] = inverse(i, k), logical indices
    __shared__ TV red[NRED > KP ? NRED : KP][RB];
    __shared__ TV rv[RB];
    __shared__ TV xs[XB * RB];                 // x of the chunks being applied
    __shared__ int s_b, s_ready;

    const int t = threadIdx.x;
    if (t == 0) s_b = int(atomicAdd(&ctrl->ticket, 1u));
    __syncthreads();
    const int b = s_b;
    const i64 lo = i64(b) * RB;
    const int rows = int(min(i64(RB), n - lo));
    const i64 p0 = BWD ? n - lo - rows : lo;   // first physical row of the block

    // 1. the inverse of the diagonal block -> LDS
    {
        const TV* src = inv + size_t(b) * RB * RB;
        for (int e = t; e < RB * RB; e += NT) DI[e] = src[e];
    }

    // with G: this thread's part of G_b (row i, columns part, part+4, ...),
    // loaded before any wait
    constexpr int GPT = RB / (NT / RB);
    TV gr[GPT];
    if (G && b > 0) {
        const TV* g = G + size_t(b) * RB * RB;
        const int i = t % RB, part = t / RB;
#pragma unroll
        for (int k = 0; k < GPT; ++k) gr[k] = g[(part + 4 * k) * RB + i];
    }
    const int cend = G ? b - 1 : b;            // chunks summed before the final step
    // this block's right-hand side, loaded before any wait (its rows are
    // written only by this workgroup, at the end)
    TV xb0 = TV(0);
    if (t < RB && t < rows) xb0 = x[p0 + (BWD ? rows - 1 - t : t)];

    // the last chunk summed before the final step (no transpose, vector
    // loads): its part of the matrix loaded before any wait, so that no HBM
    // round trip lies between that chunk's x and this block's final step
    constexpr int NQ = RB / CG;
    typedef typename Vec<TM>::type TMV;
    TMV pre[(!TRANS && VECOK) ? NQ : 1];
    const bool use_pre = !TRANS && VECOK && cend >= 1 && (t % LPC) * VEC < rows;
    if constexpr (!TRANS && VECOK) {
        if (use_pre) {
            const i64 ql = BWD ? n - i64(cend) * RB : i64(cend - 1) * RB;
            const int pr = (t % LPC) * VEC, cg = t / LPC;
#pragma unroll
            for (int i = 0; i < NQ; ++i)
                pre[i] = *reinterpret_cast<const TMV*>(M + (p0 + pr) + (ql + cg + i * CG) * ld);
        }
    }

    // 2. off-diagonal block row, chunk by chunk as earlier blocks publish
    TV acc[TRANS ? RPT : VEC];
#pragma unroll
    for (int r = 0; r < (TRANS ? RPT : VEC); ++r) acc[r] = TV(0);
    int avail = 0;
    for (int c = 0; c < cend;) {
        if (c >= avail) {
            if (t == 0) {
                unsigned v, spins = 0;
                while ((v = __hip_atomic_load(&ctrl->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) <=
                       unsigned(c)) {
                    if (__hip_atomic_load(&ctrl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                        ++spins > SPIN_LIMIT) {
                        __hip_atomic_store(&ctrl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                s_ready = int(v);
            }
            __syncthreads();
            avail = s_ready;
            __syncthreads();
            if (avail <= c) return;            // timed out: uniform exit
        }
        // x of up to XB published chunks into LDS with one round of
        // device-scope loads (two per thread, issued together);
        // xs[k * RB + e] = element e (physical order) of chunk c + k
        const int nbat = min(min(avail, cend) - c, XB);
        {
            const int e0 = t, e1 = t + NT;
            const int k0 = e0 / RB, k1 = e1 / RB;
            const i64 q0 = BWD ? n - i64(c + k0 + 1) * RB : i64(c + k0) * RB;
            const i64 q1 = BWD ? n - i64(c + k1 + 1) * RB : i64(c + k1) * RB;
            // entries beyond the batch read the first entry again (in bounds) and are not used
            const TV* a0 = k0 < nbat ? &x[q0 + (e0 - k0 * RB)] : &x[BWD ? n - i64(c + 1) * RB : i64(c) * RB];
            const TV* a1 = k1 < nbat ? &x[q1 + (e1 - k1 * RB)] : a0;
            TV v0, v1;
            load2_agent(a0, a1, v0, v1);
            if (k0 < nbat) xs[e0] = v0;
            if (k1 < nbat) xs[e1] = v1;
        }
        __syncthreads();
        for (int k = 0; k < nbat; ++k) {
            const i64 q0 = BWD ? n - i64(c + k + 1) * RB : i64(c + k) * RB;   // chunk = full block
            const TV* xc = xs + k * RB;
            if constexpr (!TRANS && VECOK) {
                if (use_pre && c + k == cend - 1) {
                    const int cg = t / LPC;
#pragma unroll
                    for (int i = 0; i < NQ; ++i) {
                        const TV xv = xc[cg + i * CG];
#pragma unroll
                        for (int v = 0; v < VEC; ++v) acc[v] = fma_t(TV(pre[i][v]), xv, acc[v]);
                    }
                    continue;
                }
            }
            if constexpr (!TRANS) {
                // T = M: block rows (contiguous) x chunk columns
                const int pr = (t % LPC) * VEC, cg = t / LPC;
                if (pr < rows) {
#pragma unroll 4
                    for (int q = cg; q < RB; q += CG) {
                        const TV xv = xc[q];
                        const TM* src = M + (p0 + pr) + (q0 + q) * ld;
                        if constexpr (VECOK) {
                            const typename Vec<TM>::type m = *reinterpret_cast<const typename Vec<TM>::type*>(src);
#pragma unroll
                            for (int v = 0; v < VEC; ++v) acc[v] = fma_t(TV(m[v]), xv, acc[v]);
                        } else {
                            acc[0] = fma_t(TV(src[0]), xv, acc[0]);
                        }
                    }
                }
            } else {
                // T = M^T: block rows are columns of M, chunk indices are rows of M (contiguous)
                const int u = t % LPR, rg = t / LPR;
#pragma unroll 2
                for (int it = 0; it < ITS; ++it) {
                    const int jj = (it * LPR + u) * VEC;
                    TV xv[VEC];
#pragma unroll
                    for (int v = 0; v < VEC; ++v) xv[v] = xc[jj + v];
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        const int pr = rg + r * (NT / LPR);
                        if (pr < rows) {
                            const TM* src = M + (q0 + jj) + (p0 + pr) * ld;
                            if constexpr (VECOK) {
                                const typename Vec<TM>::type m = *reinterpret_cast<const typename Vec<TM>::type*>(src);
#pragma unroll
                                for (int v = 0; v < VEC; ++v) acc[r] = fma_t(TV(m[v]), xv[v], acc[r]);
                            } else {
                                acc[r] = fma_t(TV(src[0]), xv[0], acc[r]);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();                        // xs is refilled next
        c += nbat;
    }

    // 3. partial sums -> LDS (fixed order: deterministic), r = b_blk - sum
    if constexpr (!TRANS) {
        const int pr = (t % LPC) * VEC, cg = t / LPC;
#pragma unroll
        for (int v = 0; v < VEC; ++v) red[cg][pr + v] = acc[v];
    } else {
        const int u = t % LPR, rg = t / LPR;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            TV a = acc[r];
#pragma unroll
            for (int o = 1; o < LPR; o <<= 1) a += __shfl_xor(a, o);
            if (u == 0) red[0][rg + r * (NT / LPR)] = a;
        }
    }
    __syncthreads();
    if (t < RB) {
        const int i = t;   // logical
        TV r = TV(0);
        if (i < rows) {
            const int pr = BWD ? rows - 1 - i : i;
            TV sum = red[0][pr];
#pragma unroll
            for (int g = 1; g < NRED; ++g) sum += red[g][pr];
            r = xb0 - sum;
        }
        rv[i] = r;
    }
    __syncthreads();

    // 4. x = inverse * r: KP threads per row over interleaved columns, then wave 0 combines
    {
        const int i = t % RB, part = t / RB;
        TV a = TV(0);
        for (int k = part; k <= i; k += KP) a = fma_t(DI[k * RB + i], rv[k], a);
        red[part][i] = a;
    }
    __syncthreads();
    if (G && b > 0) {
        // y = inverse * (b - sum over chunks < b-1) -> rv; then x_{b-1}
        if (t < RB) {
            TV yi = red[0][t];
#pragma unroll
            for (int g = 1; g < KP; ++g) yi += red[g][t];
            rv[t] = yi;
        }
        if (t == 0) {
            unsigned v, spins = 0;
            while ((v = __hip_atomic_load(&ctrl->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < unsigned(b)) {
                if (__hip_atomic_load(&ctrl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > SPIN_LIMIT) {
                    __hip_atomic_store(&ctrl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    v = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            s_ready = int(v);
        }
        __syncthreads();
        if (s_ready < b) return;               // timed out: uniform exit
        const i64 q0 = BWD ? n - i64(b) * RB : i64(b - 1) * RB;
        if (t < RB) xs[t] = get_agent(&x[q0 + t]);
        __syncthreads();
        {
            const int i = t % RB, part = t / RB;
            TV a = TV(0);
#pragma unroll
            for (int k = 0; k < GPT; ++k) a = fma_t(gr[k], xs[part + 4 * k], a);
            red[part][i] = a;
        }
        __syncthreads();
        if (t >= 64) return;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = t + 64 * h;
            if (i < rows) {
                TV s2 = red[0][i];
#pragma unroll
                for (int g = 1; g < KP; ++g) s2 += red[g][i];
                put_agent(&x[p0 + (BWD ? rows - 1 - i : i)], rv[i] - s2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);   // x of this block acknowledged before `ready` moves
        if (t == 0) __hip_atomic_store(&ctrl->ready, unsigned(b + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (t >= 64) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int i = t + 64 * h;
        if (i < rows) {
            TV xi = red[0][i];
#pragma unroll
            for (int g = 1; g < KP; ++g) xi += red[g][i];
            put_agent(&x[p0 + (BWD ? rows - 1 - i : i)], xi);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);   // x of this block acknowledged before `ready` moves
    if (t == 0) __hip_atomic_store(&ctrl->ready, unsigned(b + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename TM, typename TV, bool TRANS, bool UNIT, bool BWD>
void launch_dir(hipStream_t s, i64 n, const TM* M, i64 ld, TV* x, TrsvCtrl* c, const TV* inv, const TV* G,
                bool vecok) {
    const unsigned grid = unsigned(cdiv(n, RB));
    if (vecok)
        hipLaunchKernelGGL((trsv_chain_kernel<TM, TV, TRANS, UNIT, BWD, true>), dim3(grid), dim3(NT), 0, s, M, ld, n,
                           x, c, inv, G);
    else
        hipLaunchKernelGGL((trsv_chain_kernel<TM, TV, TRANS, UNIT, BWD, false>), dim3(grid), dim3(NT), 0, s, M, ld,
                           n, x, c, inv, G);
}

// Workspace: the control words, then the block inverses of the two triangles
// of a solve (slot 0: the first triangle, slot 1: the second), then their
// G blocks (slots 2 and 3).
size_t slot_bytes(i64 n, size_t tv) { return (size_t(cdiv(n, RB)) * RB * RB * tv + 255) & ~size_t(255); }
template <typename TV>
TV* slot_of(void* ws, i64 n, int slot) {
    return (TV*)((char*)ws + 256 + size_t(slot) * slot_bytes(n, sizeof(TV)));
}

template <typename TM, typename TV>
int prepare_t(oxi_handle_t h, hipStream_t s, int tk, i64 n, const TM* M, i64 ld, void* ws, bool split = false) {
    if (n <= 0) return OXI_OK;
    const unsigned grid = unsigned(cdiv(n, RB));
    // split: the second triangle's part on the side stream (it starts after
    // the work already on s; its end is recorded in ev_b), so that it runs
    // beside the first triangle's solve
    hipStream_t s2 = s;
    if (split) {
        OXI_HIP(h, hipEventRecord(h->ev_a, s));
        OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
        s2 = h->side;
    }
    TV* i0 = slot_of<TV>(ws, n, 0);
    TV* i1 = slot_of<TV>(ws, n, 1);
    TV* g0 = slot_of<TV>(ws, n, 2);
    TV* g1 = slot_of<TV>(ws, n, 3);
    if (!tk) {
        hipLaunchKernelGGL((trsv_dinv_kernel<TM, TV, false, true, false>), dim3(grid), dim3(NT), 0, s, M, ld, n, i0);
        hipLaunchKernelGGL((trsv_dinv_kernel<TM, TV, false, false, true>), dim3(grid), dim3(NT), 0, s2, M, ld, n, i1);
    } else {
        hipLaunchKernelGGL((trsv_dinv_kernel<TM, TV, true, false, false>), dim3(grid), dim3(NT), 0, s, M, ld, n, i0);
        hipLaunchKernelGGL((trsv_dinv_kernel<TM, TV, true, true, true>), dim3(grid), dim3(NT), 0, s2, M, ld, n, i1);
    }
    if (grid > 1) {
        if (!tk) {
            hipLaunchKernelGGL((trsv_g_kernel<TM, TV, false, false>), dim3(grid - 1), dim3(NT), 0, s, M, ld, n, i0, g0);
            hipLaunchKernelGGL((trsv_g_kernel<TM, TV, false, true>), dim3(grid - 1), dim3(NT), 0, s2, M, ld, n, i1, g1);
        } else {
            hipLaunchKernelGGL((trsv_g_kernel<TM, TV, true, false>), dim3(grid - 1), dim3(NT), 0, s, M, ld, n, i0, g0);
            hipLaunchKernelGGL((trsv_g_kernel<TM, TV, true, true>), dim3(grid - 1), dim3(NT), 0, s2, M, ld, n, i1, g1);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "trsv prepare launch: %s", hipGetErrorString(e));
    if (split) OXI_HIP(h, hipEventRecord(h->ev_b, s2));
    return OXI_OK;
}

template <typename TM, typename TV>
int trsv_t(oxi_handle_t h, hipStream_t s, int which, i64 n, const TM* M, i64 ld, TV* x, void* ws) {
    if (n <= 0) return OXI_OK;
    if (n > i64(RB) * 0x7fffffff) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "trsv: n too large");
    TrsvCtrl* c = static_cast<TrsvCtrl*>(ws);
    constexpr int VEC = Vec<TM>::N;
    const bool vecok = (reinterpret_cast<uintptr_t>(M) % 16 == 0) && (ld % VEC == 0) && (n % VEC == 0);
    const int slot = (which == 0 || which == 2) ? 0 : 1;
    const TV* inv = slot_of<TV>(ws, n, slot);
    static const bool plain = getenv("OXIBLAS_TRSV_PLAIN") != nullptr;   // tuning: every chunk on the chain
    const TV* G = plain ? nullptr : slot_of<TV>(ws, n, 2 + slot);
    hipLaunchKernelGGL(trsv_ctrl_reset, dim3(1), dim3(64), 0, s, c);
    switch (which) {
        case 0: launch_dir<TM, TV, false, true, false>(s, n, M, ld, x, c, inv, G, vecok); break;
        case 1: launch_dir<TM, TV, false, false, true>(s, n, M, ld, x, c, inv, G, vecok); break;
        case 2: launch_dir<TM, TV, true, false, false>(s, n, M, ld, x, c, inv, G, vecok); break;
        case 3: launch_dir<TM, TV, true, true, true>(s, n, M, ld, x, c, inv, G, vecok); break;
        default: return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "trsv: bad triangle %d", which);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "trsv launch: %s", hipGetErrorString(e));
    return OXI_OK;
}

}  // namespace

size_t oxi_trsv_ws_bytes(i64 n, size_t tv) { return 256 + 4 * slot_bytes(std::max<i64>(n, 1), tv); }

int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* M, i64 ld, double*, void* ws) {
    return prepare_t<double, double>(h, s, tk, n, M, ld, ws);
}
int oxi_trsv_prepare_split(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* M, i64 ld, void* ws) {
    return prepare_t<double, double>(h, s, tk, n, M, ld, ws, true);
}
int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const float* M, i64 ld, float*, void* ws) {
    return prepare_t<float, float>(h, s, tk, n, M, ld, ws);
}
int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const float* M, i64 ld, double*, void* ws) {
    return prepare_t<float, double>(h, s, tk, n, M, ld, ws);
}

int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const double* M, i64 ld, double* x, void* ctrl) {
    return trsv_t<double, double>(h, s, which, n, M, ld, x, ctrl);
}
int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const float* M, i64 ld, float* x, void* ctrl) {
    return trsv_t<float, float>(h, s, which, n, M, ld, x, ctrl);
}
int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const float* M, i64 ld, double* x, void* ctrl) {
    return trsv_t<float, double>(h, s, which, n, M, ld, x, ctrl);
}

int oxi_trsv_begin(oxi_handle_t h, hipStream_t s, void* ctrl) {
    hipLaunchKernelGGL(trsv_status_clear, dim3(1), dim3(64), 0, s, static_cast<TrsvCtrl*>(ctrl));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "trsv status clear: %s", hipGetErrorString(e));
    return OXI_OK;
}

int oxi_trsv_check(oxi_handle_t h, hipStream_t s, const void* ctrl) {
    unsigned err = 0;
    OXI_HIP(h, hipMemcpyAsync(&err, &static_cast<const TrsvCtrl*>(ctrl)->err, sizeof(unsigned),
                              hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    if (err) return oxi_set_error(h, OXI_ERR_TIMEOUT, "trsv: block chain timed out");
    return OXI_OK;
}
