// Panel leaf kernel template (see panel.hip for the algorithm); included by
// panel.hip (FP64) and panel32.hip (FP32, the working precision of the
// mixed-precision LU).
#pragma once
#include "oxi_internal.hpp"

#include <utility>
#include <type_traits>

namespace {

// Pivot key with the reference's strict '>' scan semantics: NaN in the
// diagonal position wins (nothing compares greater), NaN elsewhere never wins.
__device__ __forceinline__ double pivot_key(double v, bool diag) {
    double a = __builtin_fabs(v);
    if (a != a) return diag ? __builtin_huge_val() : -1.0;
    return a;
}
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ bool better(double k1, long long r1, double k2, long long r2) {
    return k1 > k2 || (k1 == k2 && r1 < r2);
}

// Cross-workgroup exchange without agent-scope fences.  A release/acquire
// pair at agent scope writes back and invalidates the whole L2 of the XCD,
// which costs microseconds per pivot column while the trailing GEMM keeps
// the L2 full of dirty lines.  Instead, the exchanged data itself uses
// device-scope (agent) relaxed stores and loads, which are coherent across
// XCDs without touching other lines; each wave waits for its stores to be
// acknowledged (s_waitcnt 0) before the workgroup barrier that precedes the
// counter increment, and readers load the data only after observing the
// counter.
__device__ __forceinline__ void put_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double get_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Up to this many workgroups, every workgroup reads all published records
// (key, row and the candidate's 32 values) in one pass after the barrier, so
// the winner's row needs no second round trip.
constexpr int LEAF_MERGE_P = 64;
constexpr int REC_W = 3 + OXI_LEAF_W;   // doubles of a record that are read

template <typename T, int LTT, int RPT>
__global__ __launch_bounds__(LTT) void lu_leaf_kernel(LeafArgs a) {
    constexpr int LT = LTT;
    __shared__ T u_sh[OXI_LEAF_W], dg_sh[OXI_LEAF_W];
    __shared__ double rec_sh[LEAF_MERGE_P * REC_W + OXI_LEAF_W];
    __shared__ double wkey[LT / 64];
    __shared__ long long wrow[LT / 64];
    __shared__ long long win_row;
    __shared__ int win_wg, dead;
    __shared__ double lds_rec[OXI_PUB_STRIDE];  // P == 1 exchange
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int p = blockIdx.x;
    const i64 R = i64(LT) * RPT;
    const int w = a.w;
    const bool global_x = a.P > 1;
    // a barrier of an earlier leaf timed out: the factorization is already
    // reported as failed, leave at once
    if (__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    T v[RPT][OXI_LEAF_W];
    T* Wm = (T*)a.Wm;
    i64 lr[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        lr[s] = i64(p) * R + i64(s) * LT + t;
        const bool ok = lr[s] < a.mrows;
        const T* src = Wm + a.c0 + (a.c0 + lr[s]) * a.ldw;
#pragma unroll
        for (int c = 0; c < OXI_LEAF_W; ++c) v[s][c] = (ok && c < w) ? src[c] : T(0);
    }
    if (t == 0) dead = 0;
    __syncthreads();

    // one elimination step per leaf column; the column index is a
    // compile-time constant so that v[][] stays in registers
    auto step = [&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if (j < w && !dead) {
        // ---- local candidate
        double bk = -2.0; long long br = (1ll << 62);
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (lr[s] < a.mrows && lr[s] >= j) {
                double k = pivot_key(double(v[s][j]), lr[s] == j);
                if (better(k, lr[s], bk, br)) { bk = k; br = lr[s]; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            double ok = __shfl_xor(bk, off);
            long long orr = __shfl_xor(br, off);
            if (better(ok, orr, bk, br)) { bk = ok; br = orr; }
        }
        if (lane == 0) { wkey[wave] = bk; wrow[wave] = br; }
        __syncthreads();
        if (t == 0) {
            double k = wkey[0]; long long r = wrow[0];
            for (int q = 1; q < LT / 64; ++q)
                if (better(wkey[q], wrow[q], k, r)) { k = wkey[q]; r = wrow[q]; }
            wkey[0] = k; wrow[0] = r;
        }
        __syncthreads();
        const double lkey = wkey[0];
        const long long lrow = wrow[0];
        // ---- publish: record + candidate row; WG 0 also publishes diag row j
        double* rec = global_x ? a.pub + (size_t(j) * a.P + p) * OXI_PUB_STRIDE : lds_rec;
        double* drow = a.pub + (size_t(w) * a.P + j) * OXI_PUB_STRIDE;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (lr[s] == lrow && lrow < a.mrows) {
                if (global_x) {
                    put_agent(&rec[0], lkey); put_agent(&rec[1], __longlong_as_double(lrow));
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) put_agent(&rec[3 + c], double(v[s][c]));
                } else {
                    rec[0] = lkey; rec[1] = __longlong_as_double(lrow);
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) rec[3 + c] = double(v[s][c]);
                }
            }
            if (p == 0 && lr[s] == j) {
                if (global_x) {
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) put_agent(&drow[c], double(v[s][c]));
                } else {
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) dg_sh[c] = v[s][c];
                }
            }
        }
        if (lrow >= a.mrows && t == 0) {  // no candidate rows in this WG
            if (global_x) {
                put_agent(&rec[0], -2.0); put_agent(&rec[1], __longlong_as_double((long long)(1ll << 62)));
            } else {
                rec[0] = -2.0; rec[1] = __longlong_as_double((long long)(1ll << 62));
            }
        }
        if (global_x) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);      // this wave's record stores are acknowledged
        }
        __syncthreads();
        // ---- device-wide exchange
        if (global_x) {
            if (t == 0) {
                const unsigned target = a.base + unsigned(a.P) * unsigned(j + 1);
                __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                long spins = 0;
                while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    if (++spins > (1L << 24)) { dead = 1; __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            __syncthreads();
            if (!dead && a.P <= LEAF_MERGE_P) {
                // all records and the diagonal row into LDS in one pass
                const int nrec = a.P * REC_W;
                for (int e = t; e < nrec + OXI_LEAF_W; e += LT) {
                    if (e < nrec) {
                        const int q = e / REC_W, f = e - q * REC_W;
                        rec_sh[e] = get_agent(&a.pub[(size_t(j) * a.P + q) * OXI_PUB_STRIDE + f]);
                    } else {
                        rec_sh[e] = get_agent(&drow[e - nrec]);
                    }
                }
                __syncthreads();
                if (wave == 0) {
                    double k = -3.0; long long r = (1ll << 62); int wg = 0;
                    for (int q = lane; q < a.P; q += 64) {
                        const double kq = rec_sh[q * REC_W];
                        const long long rr = __double_as_longlong(rec_sh[q * REC_W + 1]);
                        if (better(kq, rr, k, r)) { k = kq; r = rr; wg = q; }
                    }
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) {
                        double ok = __shfl_xor(k, off);
                        long long orr = __shfl_xor(r, off);
                        int ow = __shfl_xor(wg, off);
                        if (better(ok, orr, k, r)) { k = ok; r = orr; wg = ow; }
                    }
                    if (lane == 0) win_row = r;
                    if (lane < OXI_LEAF_W) {
                        u_sh[lane] = T(rec_sh[wg * REC_W + 3 + lane]);
                        dg_sh[lane] = T(rec_sh[nrec + lane]);
                    }
                }
            } else if (!dead) {
                // winner over all P records (wave 0)
                if (wave == 0) {
                    double k = -3.0; long long r = (1ll << 62); int wg = 0;
                    for (int q = lane; q < a.P; q += 64) {
                        const double* rq = a.pub + (size_t(j) * a.P + q) * OXI_PUB_STRIDE;
                        double kq = get_agent(&rq[0]); long long rr = __double_as_longlong(get_agent(&rq[1]));
                        if (better(kq, rr, k, r)) { k = kq; r = rr; wg = q; }
                    }
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) {
                        double ok = __shfl_xor(k, off);
                        long long orr = __shfl_xor(r, off);
                        int ow = __shfl_xor(wg, off);
                        if (better(ok, orr, k, r)) { k = ok; r = orr; wg = ow; }
                    }
                    if (lane == 0) { win_row = r; win_wg = wg; }
                }
                __syncthreads();
                const double* rw = a.pub + (size_t(j) * a.P + win_wg) * OXI_PUB_STRIDE;
                if (t < OXI_LEAF_W) { u_sh[t] = T(get_agent(&rw[3 + t])); dg_sh[t] = T(get_agent(&drow[t])); }
            }
            __syncthreads();
        } else {
            if (t == 0) win_row = lrow;
            if (t < OXI_LEAF_W) u_sh[t] = T(lds_rec[3 + t]);
            __syncthreads();
        }
        if (!dead) {
            const long long piv = win_row;
            if (p == 0 && t == 0) a.ipiv[a.c0 + j] = a.c0 + piv + 1;
            // ---- swap rows j and piv across the leaf
            if (piv != j) {
#pragma unroll
                for (int s = 0; s < RPT; ++s) {
                    if (lr[s] == j) {
#pragma unroll
                        for (int c = 0; c < OXI_LEAF_W; ++c) v[s][c] = u_sh[c];
                    } else if (lr[s] == piv) {
#pragma unroll
                        for (int c = 0; c < OXI_LEAF_W; ++c) v[s][c] = dg_sh[c];
                    }
                }
            }
            const T d = u_sh[j];
            if (d == T(0)) {
                if (p == 0 && t == 0 && *a.info == 0) *a.info = a.c0 + j + 1;
            } else {
                // ---- scale and rank-1 update of the rows below the pivot
#pragma unroll
                for (int s = 0; s < RPT; ++s) {
                    if (lr[s] > j && lr[s] < a.mrows) {
                        const T l = v[s][j] / d;
                        v[s][j] = l;
#pragma unroll
                        for (int c = j + 1; c < OXI_LEAF_W; ++c)
                            v[s][c] = fma_t(-l, u_sh[c], v[s][c]);
                    }
                }
            }
        }
        __syncthreads();
        }
    };
    [&]<int... J>(std::integer_sequence<int, J...>) __attribute__((always_inline)) {
        (step(std::integral_constant<int, J>{}), ...);
    }(std::make_integer_sequence<int, OXI_LEAF_W>{});
    // ---- write back
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        if (lr[s] < a.mrows) {
            T* dst = Wm + a.c0 + (a.c0 + lr[s]) * a.ldw;
#pragma unroll
            for (int c = 0; c < OXI_LEAF_W; ++c) if (c < w) dst[c] = v[s][c];
        }
    }
}

}  // namespace
