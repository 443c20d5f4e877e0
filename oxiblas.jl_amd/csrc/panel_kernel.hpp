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
static_assert(REC_N == 36, "load9_agent reads 9 fields per lane");

__device__ __forceinline__ void put_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double get_agent(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 bits(double v) { return __builtin_bit_cast(u64, v); }
__device__ __forceinline__ double from_bits(u64 v) { return __builtin_bit_cast(double, v); }
// Nine device-scope loads 32 bytes apart (one lane's fields of a record),
// issued back to back and waited for once.  Relaxed atomic loads written in
// C++ are issued one at a time, each waited for before the next.
__device__ __forceinline__ void load9_agent(const double* p, double (&f)[9]) {
    asm volatile(
        "global_load_dwordx2 %0, %9, off sc1\n\t"
        "global_load_dwordx2 %1, %9, off offset:32 sc1\n\t"
        "global_load_dwordx2 %2, %9, off offset:64 sc1\n\t"
        "global_load_dwordx2 %3, %9, off offset:96 sc1\n\t"
        "global_load_dwordx2 %4, %9, off offset:128 sc1\n\t"
        "global_load_dwordx2 %5, %9, off offset:160 sc1\n\t"
        "global_load_dwordx2 %6, %9, off offset:192 sc1\n\t"
        "global_load_dwordx2 %7, %9, off offset:224 sc1\n\t"
        "global_load_dwordx2 %8, %9, off offset:256 sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(f[0]), "=&v"(f[1]), "=&v"(f[2]), "=&v"(f[3]), "=&v"(f[4]), "=&v"(f[5]), "=&v"(f[6]), "=&v"(f[7]),
          "=&v"(f[8])
        : "v"(p)
        : "memory");
}
// field f's share of the checksum (fields 0 .. REC_N-2)
__device__ __forceinline__ u64 rec_term(u64 b, u64 tag, int f) {
    return (b ^ (tag * 0x9E3779B97F4A7C15ull + u64(f + 1) * 0xD1B54A32D192ED03ull)) * 0x94D049BB133111EBull;
}

// ---------------------------------------------------------------------------
// Cross-lane helpers (DPP: data moves inside the VALU, no LDS round trip).
// ---------------------------------------------------------------------------
template <int CTRL, int ROWM = 0xf, int BANKM = 0xf>
__device__ __forceinline__ u64 dpp64(u64 x) {
    // lanes whose source is outside the row (or masked off) keep their own value
    const int lo = __builtin_amdgcn_update_dpp(int(unsigned(x)), int(unsigned(x)), CTRL, ROWM, BANKM, false);
    const int hi = __builtin_amdgcn_update_dpp(int(unsigned(x >> 32)), int(unsigned(x >> 32)), CTRL, ROWM, BANKM, false);
    return (u64(unsigned(hi)) << 32) | unsigned(lo);
}
template <int CTRL, int ROWM = 0xf, int BANKM = 0xf>
__device__ __forceinline__ double dppd(double x) { return from_bits(dpp64<CTRL, ROWM, BANKM>(bits(x))); }
__device__ __forceinline__ u64 readlane64(u64 x, int l) {
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(x), l), hi = __builtin_amdgcn_readlane(unsigned(x >> 32), l);
    return (u64(hi) << 32) | lo;
}
// maximum over the 64 lanes of a wave (all lanes active, no NaN), in every lane:
// inclusive prefix maxima inside each row of 16, then across rows to lane 63
__device__ __forceinline__ double wave_max(double x) {
    x = __builtin_fmax(x, dppd<0x111>(x));         // row_shr:1
    x = __builtin_fmax(x, dppd<0x112>(x));         // row_shr:2
    x = __builtin_fmax(x, dppd<0x114>(x));         // row_shr:4
    x = __builtin_fmax(x, dppd<0x118>(x));         // row_shr:8
    x = __builtin_fmax(x, dppd<0x142, 0xa>(x));    // row_bcast:15 into rows 1, 3
    x = __builtin_fmax(x, dppd<0x143, 0xc>(x));    // row_bcast:31 into rows 2, 3
    return from_bits(readlane64(bits(x), 63));
}
// XOR over the 16 lanes of each row, in every lane of the row
__device__ __forceinline__ u64 row_xor16(u64 h) {
    h ^= dpp64<0x128>(h);   // row_ror:8
    h ^= dpp64<0x124>(h);   // row_ror:4
    h ^= dpp64<0x122>(h);   // row_ror:2
    h ^= dpp64<0x121>(h);   // row_ror:1
    return h;
}
// XOR over each quad of lanes, and a quad member's value, in every lane of the quad
__device__ __forceinline__ u64 quad_xor(u64 h) {
    h ^= dpp64<0xB1>(h);    // quad_perm [1,0,3,2]
    h ^= dpp64<0x4E>(h);    // quad_perm [2,3,0,1]
    return h;
}
template <int SEL>
__device__ __forceinline__ u64 quad_bcast(u64 x) { return dpp64<SEL * 0x55>(x); }   // quad_perm [SEL x 4]
__device__ __forceinline__ int first_lane(u64 m) { return int(__builtin_ctzll(m)); }

// Publish one record by the 16 lanes of a row of a wave: lane L stores fields
// L, L + 16 and L + 32 (the checksum is field 35 = lane 3's third) with
// device-scope stores; the checksum is reduced across the row.  Fields: key
// k0, row k1, the tag, then the w values row[0 .. 31] (LDS) when has, else
// zeros.  Branch-free: every lane reads its three row values (clamped into
// the row) and selects.
__device__ __forceinline__ void publish16(double* rec, u64 tag, int L, const double* row, bool has, double k0,
                                          double k1) {
    const double r0 = row[L < 3 ? 0 : L - 3], r1 = row[L + 13], r2 = row[L < 3 ? L + 29 : OXI_LEAF_W - 1];
    const double f0 = L == 0 ? k0 : L == 1 ? k1 : L == 2 ? from_bits(tag) : has ? r0 : 0.0;
    const double f1 = has ? r1 : 0.0;
    const double f2 = has ? r2 : 0.0;
    u64 h = rec_term(bits(f0), tag, L) ^ rec_term(bits(f1), tag, L + 16);
    if (L < 3) h ^= rec_term(bits(f2), tag, L + 32);
    h = row_xor16(h);
    put_agent(&rec[L], f0);
    put_agent(&rec[L + 16], f1);
    if (L < 4) put_agent(&rec[L + 32], L < 3 ? f2 : from_bits(h));
}

// ---------------------------------------------------------------------------
// The leaf kernel.  Thread t of workgroup p owns the row at position
// p*R + s*LT + t (s < RPT) in registers.  The registers of a row rotate by
// one per column: at column j register c holds column j + c, so the pivot
// column is always register 0 and the rank-1 update is the same 31 FMAs at
// every step (v[c] <- fma(-l, u[c+1], v[c+1])).  The column loop is rolled
// and every register index is a compile-time constant.
//
// The exchange of column j+1 is overlapped with the update of column j: once
// the multipliers of column j are known, every row computes only its next
// column (register 1 after the update), the candidates for column j+1 are
// chosen from those values, the candidate rows (and position j+1's row) are
// updated on their own and published, and only then do all rows take the
// full update while the records travel.
//
// Values leave the registers as they become final: the multiplier of column
// j goes to memory at step j (at the row's position at that time), and the
// pivot row at position j goes to memory at step j (later interchanges only
// involve positions below j).  Interchanges of later steps are applied to
// the stored multipliers afterwards by leaf_swaps_kernel, as getrf2 applies
// the right half's interchanges to the left half.
// ---------------------------------------------------------------------------
template <typename T, int LTT, int RPT>
__global__ __launch_bounds__(LTT) void lu_leaf_kernel(LeafArgs a) {
    constexpr int LT = LTT, NWV = LT / 64, W = OXI_LEAF_W;
    constexpr int RPW = 16;                           // records one wave validates per pass (4 lanes each)
    constexpr int PASS = NWV * RPW;
    constexpr int NPASS_MAX = (REC_SLOTS + PASS - 1) / PASS;
    constexpr unsigned SPIN_LIMIT = 1u << 22;
    __shared__ double wkey[2][NWV];
    __shared__ long long wrow[2][NWV];
    // LDS copies of records keep field f at offset f + 1 of a 40-double slot,
    // so that the row values (fields 3..34) start 16-byte aligned
    constexpr int RS_N = 40;
    __shared__ __align__(16) double wvals[2][NWV + 1][W];        // each wave's best row; [NWV]: position j
    __shared__ __align__(16) double rec_sh[2][REC_SLOTS * RS_N];  // gathered records, by column parity
    __shared__ __align__(16) double raw[NWV + 1][W];              // candidate rows before the pending update
    __shared__ double raw_l[NWV + 1];                            // and their multipliers
    __shared__ long long lrow_sh[2];                  // P == 1: the winner's position
    __shared__ int lw_sh[2];                          // P == 1: the winner's wave
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int p = blockIdx.x;
    const i64 R = i64(LT) * RPT;
    const int w = a.w;
    const bool global_x = a.P > 1;
    // polling slots of the record exchange (see gather); vmask is first
    // used after the barrier of the first select_publish
    __shared__ u64 vmask[2][2];   // validated records of the column, by column parity
    if (t < 4) vmask[t >> 1][t & 1] = 0;
    const int P1 = a.P + 1;
    const u64 full0 = P1 >= 64 ? ~0ull : (1ull << P1) - 1;
    const u64 full1 = P1 <= 64 ? 0ull : P1 >= 128 ? ~0ull : (1ull << (P1 - 64)) - 1;
    const int slot = wave * RPW + (lane >> 2);
    int qs[NPASS_MAX];   // records of this lane's slot; P1 = none
#pragma unroll
    for (int ps = 0; ps < NPASS_MAX; ++ps) qs[ps] = min(ps * PASS + slot, P1);

    T v[RPT][W];
    T* Wm = (T*)a.Wm;
    i64 lr[RPT];
    T* wrow_of[RPT];   // this position's row of W (leaf column 0)
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        lr[s] = i64(p) * R + i64(s) * LT + t;
        const bool ok = lr[s] < a.mrows;
        wrow_of[s] = Wm + a.c0 + (a.c0 + lr[s]) * a.ldw;
#pragma unroll
        for (int c = 0; c < W; ++c) v[s][c] = (ok && c < w) ? wrow_of[s][c] : T(0);
    }
    // a wait of an earlier leaf timed out: the factorization is already
    // reported as failed, leave at once.  The flag is read with the rows (one
    // memory round trip for both; the rows lie inside the panel either way)
    // and nothing has been stored before the test.
    const int failed = __hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the rows have arrived before the loop: otherwise the first use inside a
    // conditional block leaves a wait on all vector memory operations (the
    // stores of the previous column included) at the top of every column
    __builtin_amdgcn_s_waitcnt(0);
    if (failed != 0) return;

    // ---- candidates of column jn and their publication.  x[s]: row s's
    // value in column jn.  pend: the registers still await the update of
    // column jn - 1 (multipliers lm, pivot row upp, upd false for a zero
    // pivot); the candidate rows are then updated here, one column per lane.
    auto select_publish = [&](int jn, const T (&x)[RPT], bool pend, const T (&lm)[RPT], const double* upp,
                              bool upd) __attribute__((always_inline)) {
        const int parn = jn & 1;
        double bk = -2.0; long long br = (1ll << 62); int bs = -1;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (lr[s] < a.mrows && lr[s] >= jn) {
                const double k = pivot_key(double(x[s]), lr[s] == jn);
                if (better(k, lr[s], bk, br)) { bk = k; br = lr[s]; bs = s; }
            }
        }
        const double wk = wave_max(bk);
        // rows of slot s increase with the lane, and every row of slot s is
        // below every row of slot s + 1
        int wl = 0, ws = -1;
#pragma unroll
        for (int s = RPT - 1; s >= 0; --s) {
            const u64 m = __ballot(bk == wk && bs == s);
            if (m) { wl = first_lane(m); ws = s; }
        }
        const long long wr = (long long)readlane64(u64(br), wl);
        // the wave's candidate and position jn's row: as they are (pend false)
        // or through raw[] and one updated column per lane
        double (*dst)[W] = pend ? raw : wvals[parn];
        bool diag_here = false;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            const bool cand = lane == wl && ws == s, diag = lr[s] == jn;
            diag_here |= diag;
            if (cand)
#pragma unroll
                for (int c = 0; c < W; ++c) dst[wave][c] = double(v[s][c]);
            if (diag)
#pragma unroll
                for (int c = 0; c < W; ++c) dst[NWV][c] = double(v[s][c]);
            if (pend && cand) raw_l[wave] = double(lm[s]);
            if (pend && diag) raw_l[NWV] = double(lm[s]);
        }
        if (pend) {
            diag_here = __ballot(diag_here) != 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the lanes' LDS writes before the reads below
            __builtin_amdgcn_wave_barrier();
            if (lane < W) {
                auto updated = [&](int slot) __attribute__((always_inline)) -> double {
                    if (lane + 1 >= W) return 0.0;
                    const T r = T(raw[slot][lane + 1]);
                    return double(upd ? fma_t(-T(raw_l[slot]), T(upp[lane + 1]), r) : r);
                };
                if (ws >= 0) wvals[parn][wave][lane] = updated(wave);
                if (diag_here) wvals[parn][NWV][lane] = updated(NWV);
            }
        }
        if (lane == 0) { wkey[parn][wave] = wk; wrow[parn][wave] = ws >= 0 ? wr : (1ll << 62); }
        __syncthreads();
        double lkey = wkey[parn][0];
        long long lrow = wrow[parn][0];
        int lw = 0;
#pragma unroll
        for (int q = 1; q < NWV; ++q)
            if (better(wkey[parn][q], wrow[parn][q], lkey, lrow)) { lkey = wkey[parn][q]; lrow = wrow[parn][q]; lw = q; }
        if (!global_x) {
            if (t == 0) { lrow_sh[parn] = lrow; lw_sh[parn] = lw; }
            return;
        }
        const u64 tag = a.tag0 + u64(a.c0 + jn);
        double* rec = a.pub + (size_t(jn) * a.P + p) * OXI_PUB_STRIDE;
        double* drec = a.pub + (size_t(w) * a.P + jn) * OXI_PUB_STRIDE;
        if (wave == 0 && lane < 16) {
            const bool has = lrow < a.mrows;
            publish16(rec, tag, lane, wvals[parn][lw], has, has ? lkey : -2.0,
                      __longlong_as_double(has ? lrow : (long long)(1ll << 62)));
        }
        if (p == 0 && wave == NWV - 1 && lane < 16)
            publish16(drec, tag, lane, wvals[parn][NWV], true, 0.0, 0.0);
    };

    // ---- the records of column jn into rec_sh (P > 1), then a barrier, so
    // that no wave overwrites the buffers of column jn - 1 while another
    // still reads them.  Returns false on timeout.
    //
    // Record q (q < P candidates, q == P position jn) is polled by the 4 lanes
    // of slot q % PASS (lane sub holds fields sub, sub+4, ..., sub+32).  The
    // waves poll on their own, without a barrier per round, and a record
    // whose tag is not yet the column's is passed over without hashing.  A
    // validated record goes to rec_sh and its bit to vmask (LDS); every wave
    // stops once all bits are set.  (Polling each record from several waves
    // at staggered times was slower: a wave cannot leave a poll in flight,
    // so the barrier then waits for the late copies' round trips.)  vmask of
    // the other parity is cleared after the barrier: its readers (column
    // jn - 1) passed the barrier of select_publish before this one.
    auto gather = [&](int jn) __attribute__((always_inline)) -> bool {
        if (!global_x) { __syncthreads(); return true; }
        const int parn = jn & 1;
        const u64 tag = a.tag0 + u64(a.c0 + jn);
        const double* drec = a.pub + (size_t(w) * a.P + jn) * OXI_PUB_STRIDE;
        double* rs = rec_sh[parn];
        u64* vm = vmask[parn];
        const int sub = lane & 3;
        unsigned spins = 0;
        while (true) {
            const u64 m0 = __hip_atomic_load(&vm[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const u64 m1 = __hip_atomic_load(&vm[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (__builtin_amdgcn_readfirstlane(int(m0 == full0 && m1 == full1))) break;
#pragma unroll
            for (int ps = 0; ps < NPASS_MAX; ++ps) {
                const int q = qs[ps];
                const bool have = q >= P1 || (((q < 64 ? m0 >> q : m1 >> (q - 64)) & 1) != 0);
                if (!have) {
                    const double* src = q < a.P ? a.pub + (size_t(jn) * a.P + q) * OXI_PUB_STRIDE : drec;
                    double fv[REC_N / 4];
                    load9_agent(src + sub, fv);
                    // the tag sits in quad lane 2 (field 2): a record of an
                    // earlier column or call is passed over without hashing
                    const u64 tg = quad_bcast<2>(bits(fv[0]));
                    if (tg == tag) {
                        u64 hx = 0;
#pragma unroll
                        for (int i = 0; i < REC_N / 4; ++i) {
                            const int f = sub + 4 * i;
                            if (f < REC_CHK) hx ^= rec_term(bits(fv[i]), tag, f);
                        }
                        hx = quad_xor(hx);
                        // the checksum sits in quad lane 3 (field 35)
                        const u64 chk = quad_bcast<3>(bits(fv[REC_N / 4 - 1]));
                        if (hx == chk) {
#pragma unroll
                            for (int i = 0; i < REC_N / 4; ++i) rs[q * RS_N + 1 + sub + 4 * i] = fv[i];
                            if (sub == 0)
                                __hip_atomic_fetch_or(&vm[q >> 6], 1ull << (q & 63), __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
            }
            if (++spins > SPIN_LIMIT) break;   // uniform per wave
            __builtin_amdgcn_s_sleep(1);
        }
        __syncthreads();
        const bool ok = __hip_atomic_load(&vm[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == full0 &&
                        __hip_atomic_load(&vm[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == full1;
        if (!ok) {   // the same in every thread: no wave writes vmask after the barrier
            if (t == 0) __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        if (t < 2) vmask[parn ^ 1][t] = 0;
        return true;
    };

    // column 0: the registers as loaded
    {
        T x0[RPT];
#pragma unroll
        for (int s = 0; s < RPT; ++s) x0[s] = v[s][0];
        select_publish(0, x0, false, x0, nullptr, false);
        if (!gather(0)) return;
    }

#pragma unroll 1
    for (int j0 = 0; j0 < w; ++j0) {
        const int j = __builtin_amdgcn_readfirstlane(j0);
        const int par = j & 1;
        // ---- the winner of column j: the pivot row (up) and position j's
        // row (dp), as doubles in LDS, in the rotated register order
        const double* up;
        const double* dp;
        long long piv;
        if (!global_x) {
            piv = lrow_sh[par];
            up = wvals[par][lw_sh[par]];
            dp = wvals[par][NWV];
        } else {
            // the winner over the P candidates, in every wave; record q holds
            // rows of workgroup q only, so ties go to the lowest q
            const double* rs = rec_sh[par];
            double k = -3.0; long long r = (1ll << 62); int wq = -1;
            for (int q = lane; q < a.P; q += 64) {
                const double kq = rs[q * RS_N + 1];
                const long long rq = __double_as_longlong(rs[q * RS_N + 2]);
                if (better(kq, rq, k, r)) { k = kq; r = rq; wq = q; }
            }
            const double gk = wave_max(k);
            const u64 m0 = __ballot(k == gk && wq >= 0 && wq < 64);
            const int gl = m0 ? first_lane(m0) : first_lane(__ballot(k == gk && wq >= 64));
            piv = (long long)readlane64(u64(r), gl);
            const int gq = __builtin_amdgcn_readlane(wq, gl);
            up = (const double*)__builtin_assume_aligned(rs + gq * RS_N + 4, 16);
            dp = (const double*)__builtin_assume_aligned(rs + a.P * RS_N + 4, 16);
        }

        // ---- interchange (the registers: columns >= j), store position j,
        // scale, store the multipliers.  Only position piv takes a row
        // (position j's, dp): position j's registers are not read again
        // (candidates, multipliers and later interchanges all lie below j),
        // its final row is the pivot row, stored from up.
        if (p == 0 && t == 0) a.ipiv[a.c0 + j] = a.c0 + piv + 1;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (piv != j && lr[s] == piv) {
#pragma unroll
                for (int c = 0; c < W; ++c) v[s][c] = T(dp[c]);
            }
        }
        // position j now holds the pivot row, final in columns j .. w-1
        if (p == 0 && wave == 0 && lane < w - j) Wm[a.c0 + j + (a.c0 + j) * a.ldw + lane] = T(up[lane]);
        const T d = T(up[0]);
        const bool upd = d != T(0);
        if (!upd && p == 0 && t == 0 && *a.info == 0) *a.info = a.c0 + j + 1;
        T l[RPT];
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            l[s] = upd ? v[s][0] / d : v[s][0];   // a zero pivot leaves the column as it is
            if (lr[s] > j && lr[s] < a.mrows) wrow_of[s][j] = l[s];
        }

        // ---- column j+1: its values, candidates and records first
        const bool more = j + 1 < w;
        if (more) {
            T x[RPT];
            const T u1 = T(up[1]);
#pragma unroll
            for (int s = 0; s < RPT; ++s) x[s] = upd ? fma_t(-l[s], u1, v[s][1]) : v[s][1];
            select_publish(j + 1, x, true, l, up, upd);
        }

        // ---- the rank-1 update with the rotation, while the records travel
        if (upd) {
            T u[W];
#pragma unroll
            for (int c = 1; c < W; ++c) u[c] = T(up[c]);
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
#pragma unroll
                for (int c = 0; c + 1 < W; ++c) v[s][c] = fma_t(-l[s], u[c + 1], v[s][c + 1]);
                v[s][W - 1] = T(0);
            }
        } else {
#pragma unroll
            for (int s = 0; s < RPT; ++s) {
#pragma unroll
                for (int c = 0; c + 1 < W; ++c) v[s][c] = v[s][c + 1];
                v[s][W - 1] = T(0);
            }
        }
        if (more && !gather(j + 1)) return;
    }
}

// The slots of the interchanges of steps k_lo .. w-1 of a leaf at c0 (one
// wave): lane i < n (n = w - k_lo) is step k_lo + i's own position; lane
// n + i is the pivot position p_i of step k_lo + i when it lies below the
// leaf's w rows and i is its first occurrence.  pos = -1 for unused lanes.
// The interchanges are then replayed on slot ids: afterwards lane s holds
// the slot whose original content ends at slot s.  Every lane follows its
// own slot backwards through the transpositions (step i exchanges slot i
// with slot sb_i, held by lane i), so the steps carry no cross-lane
// dependence: one readlane and two selects each.
// pi_all: lane i < w holds step i's pivot position relative to c0 (-1 beyond).
__device__ __forceinline__ void leaf_slots(i64 pi_all, int w, int k_lo, int lane, i64& pos, int& src) {
    const int n = w - k_lo;
    const i64 pi_sh = (i64)__shfl((long long)pi_all, min(k_lo + lane, 63));
    const i64 pi = lane < n ? pi_sh : -1;
    // unrolled to the largest leaf without a bound test (the readlanes of
    // different steps are independent and interleave): lanes i >= n hold
    // pi = -1, which matches only lanes >= n, whose first is not used
    int first = lane;
#pragma unroll
    for (int i = 0; i < OXI_LEAF_W; ++i) {
        const i64 q = (i64)readlane64(u64(pi), i);
        if (i < first && q == pi) first = i;
    }
    pos = lane < n ? k_lo + lane : -1;
    const int src_i = lane - n;   // differs per lane: a shuffle, not readlane
    const int src_l = (src_i >= 0 && src_i < n) ? src_i : 0;
    const i64 pos_out = (i64)__shfl((long long)pi, src_l);
    const int first_out = __shfl(first, src_l);
    if (src_i >= 0 && src_i < n && pos_out >= w && first_out == src_i) pos = pos_out;
    // the slot that step i exchanges with slot i (lane i < n)
    const int sbv = lane < n ? (pi < w ? int(pi - k_lo) : n + first) : lane;
    // the content that ends at slot `lane`: the transpositions of steps
    // n-1 .. 0 applied to it in turn (steps i >= n exchange slot i with
    // itself: sbv = lane there)
    int x = lane;
#pragma unroll
    for (int i = OXI_LEAF_W - 1; i >= 0; --i) {
        const int sb = __builtin_amdgcn_readlane(sbv, i);
        x = x == i ? sb : x == sb ? i : x;
    }
    src = x;
}

// A leaf's interchanges applied where the leaf did not apply them, blocks of
// two waves:
//  * its own multiplier columns c < w - 1 (one block each, the first wave):
//    column c takes the interchanges of steps c+1 .. w-1 (the multipliers of
//    column c were stored at the positions of step c);
//  * the other columns of the enclosing panel, [x_lo, c0) and [c0+w, x_hi),
//    64 per block (a lane per column): all w interchanges, half of the slots
//    by each wave.
// Applying a leaf's interchanges to the rest of its panel at once gives the
// same columns as the getrf2 recursion, which applies them level by level:
// a left subtree's multipliers are not read again once its right sibling
// has started, and the right sibling's columns receive them before any
// update reads them.
template <typename T>
__global__ __launch_bounds__(128) void leaf_swaps_kernel(T* Wm, i64 ldw, i64 c0, int w, const i64* ipiv,
                                                         const int* flags, i64 x_lo, i64 x_hi, int first_block = 0) {
    const int lane = threadIdx.x & 63;
    const int half = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    // the pivots are loaded with the flag, not after its test, and without a
    // branch (lanes >= w read entry c0): one memory round trip for both
    // (ipiv lies in its own allocation, always readable)
    const i64 pv = ipiv[c0 + (lane < w ? lane : 0)];
    if (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const i64 pi_all = lane < w ? pv - 1 - c0 : -1;
    const int b = int(blockIdx.x) + first_block;   // first_block = w - 1: the leaf's own columns are final
    if (b < w - 1) {
        if (half != 0) return;
        const int c = b;
        i64 pos; int src;
        leaf_slots(pi_all, w, c + 1, lane, pos, src);
        const i64 psrc = (i64)__shfl((long long)pos, src);
        T v = T(0);
        if (pos >= 0 && psrc != pos) v = Wm[(c0 + psrc) * ldw + c0 + c];
        if (pos >= 0 && psrc != pos) Wm[(c0 + pos) * ldw + c0 + c] = v;
        return;
    }
    const i64 fb = b - (w - 1);
    const i64 tl = (c0 - x_lo + 63) / 64;
    i64 x;
    bool ok;
    if (fb < tl) { x = x_lo + fb * 64 + lane; ok = x < c0; }
    else { x = c0 + w + (fb - tl) * 64 + lane; ok = x < x_hi; }
    i64 pos; int src;
    leaf_slots(pi_all, w, 0, lane, pos, src);
    // slot s moves when its source differs (at most 2w <= 64 slots).  Both
    // waves compute all slots; wave h moves slots 32h .. 32h+31.  All loads
    // of the block have returned before any store (a wait and a barrier) and
    // are issued without branches: a slot that does not move reads its own
    // position, a lane outside the columns reads column c0 (both in bounds),
    // and only moving slots are stored.  Lane s computes slot s's row
    // offsets once; the loops only read them back (two readlanes and an add
    // per slot and access).
    const i64 psrc = (i64)__shfl((long long)pos, src);
    const bool moves = pos >= 0 && src != lane;
    const u64 mv = __ballot(moves);
    const i64 offs = (c0 + (moves ? psrc : 0)) * ldw;
    const i64 offd = (c0 + (moves ? pos : 0)) * ldw;
    const i64 xr = ok ? x : c0;
    const int s0 = 32 * half;
    T v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = Wm[(i64)readlane64(u64(offs), s0 + k) + xr];
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; ++k)
        if ((mv >> (s0 + k)) & 1) {
            const i64 to = (i64)readlane64(u64(offd), s0 + k);
            if (ok) Wm[to + x] = v[k];
        }
}

}  // namespace
