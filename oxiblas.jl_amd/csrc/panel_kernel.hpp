// Panel leaf kernel template (the algorithm is described in panel.hip);
// included by panel.hip (FP64) and panel32.hip (FP32, the working precision
// of the mixed-precision LU).
#pragma once
#include "oxi_internal.hpp"

#include <utility>
#include <type_traits>

namespace {

typedef unsigned long long u64;

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

// ---------------------------------------------------------------------------
// Cross-workgroup exchange of one pivot column: self-validating records.
//
// Every workgroup publishes one record per column: its best candidate (key,
// row) and that row's w values, plus a tag (getrf call and global column) and
// a checksum over all other fields keyed by the tag.  Every workgroup then
// loads all records and accepts a record only when its checksum matches what
// the fields and the expected tag give; it reloads the others until all are
// valid.  Nothing orders the writer's stores: a record that is still partly
// old (an earlier column or call) fails its checksum and is simply read
// again, so one round trip suffices once the last record has landed.  The
// data goes through device-scope (agent) relaxed stores and loads, which are
// coherent across XCDs without the L2 write-back / invalidate that an
// agent-scope release / acquire pair costs.
// ---------------------------------------------------------------------------
constexpr int REC_N = 4 + OXI_LEAF_W;           // key, row, tag, 32 values, checksum
constexpr int REC_CHK = REC_N - 1;
constexpr int REC_SLOTS = OXI_LEAF_MAX_P + 1;   // records per column: P candidates + the diagonal row

__device__ __forceinline__ void put_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double get_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 bits(double v) { return __builtin_bit_cast(u64, v); }
__device__ __forceinline__ double from_bits(u64 v) { return __builtin_bit_cast(double, v); }
// field f's share of the checksum (fields 0 .. REC_N-2)
__device__ __forceinline__ u64 rec_term(u64 b, u64 tag, int f) {
    return (b ^ (tag * 0x9E3779B97F4A7C15ull + u64(f + 1) * 0xD1B54A32D192ED03ull)) * 0x94D049BB133111EBull;
}

// Publish a record: fields 0..REC_N-2 from vals(f) (field 2 is the tag), then
// the checksum.
template <typename F>
__device__ __forceinline__ void publish(double* rec, u64 tag, F vals) {
    u64 h = 0;
#pragma unroll
    for (int f = 0; f < REC_CHK; ++f) {
        const double v = f == 2 ? from_bits(tag) : vals(f);
        h ^= rec_term(bits(v), tag, f);
        put_agent(&rec[f], v);
    }
    put_agent(&rec[REC_CHK], from_bits(h));
}

template <typename T, int LTT, int RPT>
__global__ __launch_bounds__(LTT) void lu_leaf_kernel(LeafArgs a) {
    constexpr int LT = LTT, NWV = LT / 64;
    constexpr int RPW = 16;                           // records one wave validates per pass (4 lanes each)
    constexpr int PASS = NWV * RPW;
    constexpr int NPASS_MAX = (REC_SLOTS + PASS - 1) / PASS;
    __shared__ double wkey[2][NWV];
    __shared__ long long wrow[2][NWV];
    __shared__ double rec_sh[2][REC_SLOTS * REC_N];   // by column parity
    __shared__ int dead;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int p = blockIdx.x;
    const i64 R = i64(LT) * RPT;
    const int w = a.w;
    const bool global_x = a.P > 1;
    // a wait of an earlier leaf timed out: the factorization is already
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

    // One elimination step per leaf column.  The column index is a
    // compile-time constant so that v[][] stays in registers and each step
    // touches only the columns right of its pivot.
    auto step = [&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        if (j >= w || dead) return;
        const int par = j & 1;
        // ---- local candidate: waves, then every thread over the waves
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
        if (lane == 0) { wkey[par][wave] = bk; wrow[par][wave] = br; }
        __syncthreads();
        double lkey = wkey[par][0];
        long long lrow = wrow[par][0];
#pragma unroll
        for (int q = 1; q < NWV; ++q)
            if (better(wkey[par][q], wrow[par][q], lkey, lrow)) { lkey = wkey[par][q]; lrow = wrow[par][q]; }

        // ---- exchange: the winner's row (up) and row j (dp), as doubles in LDS
        double* rs = rec_sh[par];
        const double* up;
        const double* dp;
        long long piv;
        if (!global_x) {
            // one workgroup: the candidate is the winner
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                if (lr[s] == lrow && lrow < a.mrows)
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) rs[3 + c] = double(v[s][c]);
                if (lr[s] == j)
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) rs[REC_N + 3 + c] = double(v[s][c]);
            }
            __syncthreads();
            piv = lrow;
            up = rs + 3;
            dp = rs + REC_N + 3;
        } else {
            const u64 tag = a.tag0 + u64(a.c0 + j);
            double* rec = a.pub + (size_t(j) * a.P + p) * OXI_PUB_STRIDE;
            double* drec = a.pub + (size_t(w) * a.P + j) * OXI_PUB_STRIDE;
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                if (lr[s] == lrow && lrow < a.mrows)
                    publish(rec, tag, [&](int f) {
                        return f == 0 ? lkey : f == 1 ? __longlong_as_double(lrow) : double(v[s][f - 3]);
                    });
                if (p == 0 && lr[s] == j)
                    publish(drec, tag, [&](int f) { return f < 3 ? 0.0 : double(v[s][f - 3]); });
            }
            if (lrow >= a.mrows && t == 0)   // no candidate rows in this workgroup
                publish(rec, tag, [&](int f) {
                    return f == 0 ? -2.0 : f == 1 ? __longlong_as_double((long long)(1ll << 62)) : 0.0;
                });
            // gather: record q (q < P candidates, q == P row j) by lanes 4i..4i+3 of
            // one wave; lane sub holds fields sub, sub+4, ..., sub+32
            const int P1 = a.P + 1, npass = (P1 + PASS - 1) / PASS;
            const int sub = lane & 3;
            bool okp[NPASS_MAX];
#pragma unroll
            for (int ps = 0; ps < NPASS_MAX; ++ps) okp[ps] = false;
            unsigned spins = 0;
            while (true) {
                bool all_ok = true;
#pragma unroll
                for (int ps = 0; ps < NPASS_MAX; ++ps) {
                    const int q = ps * PASS + wave * RPW + (lane >> 2);
                    if (ps < npass && q < P1 && !okp[ps]) {
                        const double* src = q < a.P ? a.pub + (size_t(j) * a.P + q) * OXI_PUB_STRIDE : drec;
                        double fv[REC_N / 4];
                        u64 hx = 0;
#pragma unroll
                        for (int i = 0; i < REC_N / 4; ++i) {
                            const int f = sub + 4 * i;
                            fv[i] = get_agent(&src[f]);
                            if (f < REC_CHK) hx ^= rec_term(bits(fv[i]), tag, f);
                        }
                        hx ^= __shfl_xor(hx, 1);
                        hx ^= __shfl_xor(hx, 2);
                        // the checksum sits in lane sub 3 (field 35), the tag in sub 2 (field 2)
                        const u64 chk = bits(__shfl(fv[REC_N / 4 - 1], (lane & ~3) | 3));
                        const u64 tg = bits(__shfl(fv[0], (lane & ~3) | 2));
                        if (hx == chk && tg == tag) {
                            okp[ps] = true;
#pragma unroll
                            for (int i = 0; i < REC_N / 4; ++i) rs[q * REC_N + sub + 4 * i] = fv[i];
                        }
                    }
                    if (ps < npass && q < P1 && !okp[ps]) all_ok = false;
                }
                if (__syncthreads_and(all_ok)) break;
                if (++spins > (1u << 22)) {
                    if (t == 0) __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    dead = 1;   // every thread writes the same value: uniform exit
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (spins > (1u << 22)) return;
            // the winner over the P candidates, in every wave
            double k = -3.0; long long r = (1ll << 62); int wq = 0;
            for (int q = lane; q < a.P; q += 64) {
                const double kq = rs[q * REC_N];
                const long long rq = __double_as_longlong(rs[q * REC_N + 1]);
                if (better(kq, rq, k, r)) { k = kq; r = rq; wq = q; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ok = __shfl_xor(k, off);
                const long long orr = __shfl_xor(r, off);
                const int ow = __shfl_xor(wq, off);
                if (better(ok, orr, k, r)) { k = ok; r = orr; wq = ow; }
            }
            piv = r;
            up = rs + wq * REC_N + 3;
            dp = rs + a.P * REC_N + 3;
        }

        // ---- interchange, scale, rank-1 update
        if (p == 0 && t == 0) a.ipiv[a.c0 + j] = a.c0 + piv + 1;
        if (piv != j) {
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                if (lr[s] == j) {
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) v[s][c] = T(up[c]);
                } else if (lr[s] == piv) {
#pragma unroll
                    for (int c = 0; c < OXI_LEAF_W; ++c) v[s][c] = T(dp[c]);
                }
            }
        }
        const T d = T(up[j]);
        if (d == T(0)) {
            if (p == 0 && t == 0 && *a.info == 0) *a.info = a.c0 + j + 1;
        } else {
            T ucol[OXI_LEAF_W];
#pragma unroll
            for (int c = j + 1; c < OXI_LEAF_W; ++c) ucol[c] = T(up[c]);
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
                if (lr[s] > j && lr[s] < a.mrows) {
                    const T l = v[s][j] / d;
                    v[s][j] = l;
#pragma unroll
                    for (int c = j + 1; c < OXI_LEAF_W; ++c) v[s][c] = fma_t(-l, ucol[c], v[s][c]);
                }
            }
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
