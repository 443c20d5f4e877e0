// Tournament-pivoting panel leaf for the FP32 factor of the mixed-precision
// solve.  No counterpart in the reference: its mixed_solve factors the Float32
// copy with partial pivoting (src/lapack/extended_lapack.jl:183,195-198); the
// mixed solve's contract is the refinement's convergence and accuracy, not
// the factor's pivots, and partial pivoting stays the default.
//
// The partial-pivoting leaf (panel_kernel.hpp) exchanges one record per
// column between its P workgroups: w device-scope round trips per leaf, each
// behind a chain of workgroup barriers.  Here the w pivots of a leaf are
// chosen by a tournament (communication-avoiding LU: Grigori, Demmel and
// Xiang, "CALU: a communication optimal LU factorization algorithm", SIAM J.
// Matrix Anal. Appl. 32(4), 2011) whose every round is partial pivoting
// inside one wave, with lane reductions and no barrier per column:
//  1. each wave eliminates its own rows (in registers), only to pick its w
//     candidate rows;
//  2. wave 0 loads the original rows of its workgroup's four candidate sets
//     and picks w of them the same way;
//  3. a tree over the workgroups, four sets to one per level: the leader of
//     each set publishes its positions (one tagged, checksummed record), every
//     workgroup of a group gathers the group's records, loads those original
//     rows and picks w of them; the last level keeps its pivot rows (the
//     leaf's rows of U and the top of its L) and every workgroup ends with the
//     same pivots;
//  4. after every workgroup's "loaded" flag, each one eliminates its own
//     original rows with the pivot rows (per row exactly the operations of
//     partial pivoting with these pivots: l_c = a_c / u_cc, a_k = fma(-l_c,
//     u_ck, a_k), a zero pivot leaves the column and updates nothing) and
//     stores them at their final positions; workgroup 0 stores the pivot rows,
//     ipiv and info.
// Ties in a pivot search go to the row of lowest position; NaN never wins.
//
#include "panel_kernel.hpp"

#include <climits>

namespace {

constexpr int TW = OXI_LEAF_W;           // 32
constexpr int TLT = 256, TNWV = TLT / 64;
constexpr int TP_LEVELS = 3;             // tree levels over workgroups: 32 -> 8 -> 2 -> 1
constexpr int TP_REC = 36;               // record fields: tag, count, 32 positions, spare, checksum
constexpr unsigned TP_SPIN_LIMIT = 1u << 22;
static_assert(OXI_TP_MAX_P <= 64, "three levels of four-way merges");

__device__ __forceinline__ float tp_key(float v) {
    const float a = __builtin_fabsf(v);
    return a != a ? -1.0f : a;
}
__device__ __forceinline__ bool tp_better(float k1, int r1, float k2, int r2) {
    return k1 > k2 || (k1 == k2 && r1 < r2);
}
__device__ __forceinline__ u64 wave_xor64(u64 x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x ^= (u64)__shfl_xor((long long)x, o);
    return x;
}
template <int CTRL, int ROWM = 0xf>
__device__ __forceinline__ float dppf(float x) {
    // lanes whose source is outside the row (or masked off) keep their own value
    const int i = __builtin_bit_cast(int, x);
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(i, i, CTRL, ROWM, 0xf, false));
}
// maximum over the 64 lanes of a wave (all lanes active, no NaN), in every
// lane: inclusive prefix maxima inside each row of 16, then across rows
__device__ __forceinline__ float wave_maxf(float x) {
    x = __builtin_fmaxf(x, dppf<0x111>(x));         // row_shr:1
    x = __builtin_fmaxf(x, dppf<0x112>(x));         // row_shr:2
    x = __builtin_fmaxf(x, dppf<0x114>(x));         // row_shr:4
    x = __builtin_fmaxf(x, dppf<0x118>(x));         // row_shr:8
    x = __builtin_fmaxf(x, dppf<0x142, 0xa>(x));    // row_bcast:15 into rows 1, 3
    x = __builtin_fmaxf(x, dppf<0x143, 0xc>(x));    // row_bcast:31 into rows 2, 3
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}

// Partial pivoting inside one wave over the rows it holds (NR per lane:
// values a in column order, positions id, INT_MAX for no row or an
// eliminated one) for columns 0 .. w-1.  Lane 0 writes step c's pivot
// position to pv[c] (INT_MAX once no row is left).  The pivot row reaches
// every lane through LDS: its lane writes it to bc (W floats of this wave),
// or with KEEP to prow[c] (the leaf's pivot row c as partial pivoting leaves
// it: multipliers of the earlier steps, then U from column c on).  The
// column loop is unrolled: every register index is a compile-time constant
// and a step touches only the columns right of its own.  Every row takes the
// update (rows out of the search hold values nobody reads again).
template <int NR, bool KEEP>
__device__ __forceinline__ void wave_gepp(float (&a)[NR][TW], int (&id)[NR], int w, int* pv, float* bc,
                                          float (*prow)[TW]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < TW; ++c) {
        if (c < w) {
            float bk = -2.0f;
            int bid = INT_MAX, bs = 0;
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                const float k = id[s] != INT_MAX ? tp_key(a[s][c]) : -2.0f;
                if (tp_better(k, id[s], bk, bid)) { bk = k; bid = id[s]; bs = s; }
            }
            const float wk = wave_maxf(bk);
            if (wk == -2.0f) {   // no row left (uniform)
                if (lane == 0)
                    for (int k = c; k < w; ++k) pv[k] = INT_MAX;
                return;
            }
            u64 m = __ballot(bk == wk);
            int ol = first_lane(m);
            int oid = __builtin_amdgcn_readlane(bid, ol);
            m &= m - 1;
            while (m) {   // equal keys in several lanes (rare): the lowest position
                const int l2 = first_lane(m);
                const int id2 = __builtin_amdgcn_readlane(bid, l2);
                if (id2 < oid) { oid = id2; ol = l2; }
                m &= m - 1;
            }
            if (lane == 0) pv[c] = oid;
            // the pivot row in every lane
            float* dst = (float*)__builtin_assume_aligned(KEEP ? prow[c] : bc, 16);
            if (lane == ol) {
                // value selects: a pointer chosen between the rows' arrays
                // would put them in scratch memory
#pragma unroll
                for (int k = 0; k < TW; ++k) {
                    float x = a[0][k];
#pragma unroll
                    for (int s = 1; s < NR; ++s) x = bs == s ? a[s][k] : x;
                    dst[k] = x;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the lane's LDS writes before the reads below
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int s = 0; s < NR; ++s)
                if (id[s] == oid) id[s] = INT_MAX;
            const float d = dst[c];
            if (d != 0.0f) {   // a zero pivot leaves the column and updates nothing (uniform)
                float l[NR];
#pragma unroll
                for (int s = 0; s < NR; ++s) l[s] = a[s][c] / d;
#pragma unroll
                for (int k = c + 1; k < TW; ++k) {
                    const float uk = dst[k];
#pragma unroll
                    for (int s = 0; s < NR; ++s) a[s][k] = __builtin_fmaf(-l[s], uk, a[s][k]);
                }
#pragma unroll
                for (int s = 0; s < NR; ++s) a[s][c] = l[s];
            }
        }
        __builtin_amdgcn_sched_barrier(0);   // keep each step's loads inside it (register pressure)
    }
}

// The original rows at leaf positions ids (two per lane) into registers, in
// column order.
__device__ __forceinline__ void load_rows(const LeafArgs& a, const int (&ids)[2], float (&v)[2][TW]) {
    const float* Wm = (const float*)a.Wm;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const bool ok = ids[s] != INT_MAX;
        const float* row = Wm + a.c0 + (a.c0 + i64(ok ? ids[s] : 0)) * a.ldw;
#pragma unroll
        for (int c = 0; c < TW; ++c) v[s][c] = (ok && c < a.w) ? row[c] : 0.0f;
    }
}

template <int RPT>
__global__ __launch_bounds__(TLT) __attribute__((amdgpu_waves_per_eu(1, 1))) void lu_leaf_tp_kernel(LeafArgs a) {
    constexpr i64 R = i64(TLT) * RPT;
    __shared__ int pv[TNWV][TW];                      // each wave's candidates
    __shared__ int gset[TW];                          // this workgroup's set at the current level
    __shared__ int uni[4 * TW];                       // a merge's input positions
    __shared__ int pid[TW];                           // the leaf's pivots
    __shared__ __align__(16) float prow[TW][TW];      // the pivot rows
    __shared__ __align__(16) float bc[TNWV][TW];      // each wave's pivot row of the step
    __shared__ int ipiv_rel[TW];
    __shared__ int map_from[2 * TW], map_to[2 * TW];
    __shared__ int bad;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int p = blockIdx.x;
    const int w = a.w;
    float* Wm = (float*)a.Wm;
    const u64 tag_base = a.tag0 + 8 * u64(a.c0);      // level l's records: tag_base + l; the flags: + 7
    const u64 tag_flag = tag_base + 7;
    double* flg = a.pub + TP_LEVELS * OXI_TP_MAX_P * TP_REC;   // [P]
    // tree levels over the workgroups: 4^levels >= P
    const int nlev = a.P <= 1 ? 0 : a.P <= 4 ? 1 : a.P <= 16 ? 2 : 3;

    // ---- own rows: original values (kept to the end)
    float o[RPT][TW];
    int lr[RPT];
    bool live[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        lr[s] = int(i64(p) * R + i64(s) * TLT + t);
        live[s] = lr[s] < a.mrows;
        const float* row = Wm + a.c0 + (a.c0 + i64(live[s] ? lr[s] : 0)) * a.ldw;
#pragma unroll
        for (int c = 0; c < TW; ++c) o[s][c] = (live[s] && c < w) ? row[c] : 0.0f;
    }
    const int failed = __hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    if (failed != 0) return;
    if (t == 0) bad = 0;

    // ---- 1. each wave's candidates
    {
        float v[RPT][TW];
        int id[RPT];
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            id[s] = live[s] ? lr[s] : INT_MAX;
#pragma unroll
            for (int c = 0; c < TW; ++c) v[s][c] = o[s][c];
        }
        wave_gepp<RPT, false>(v, id, w, pv[wave], bc[wave], nullptr);
    }
    __syncthreads();

    // ---- 2. and 3. the workgroup's set, then the tree (wave 0)
    if (wave == 0) {
        int ids[2];
        float v[2][TW];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int k = lane + 64 * s;
            ids[s] = (k % TW) < w ? pv[k / TW][k % TW] : INT_MAX;
        }
        load_rows(a, ids, v);
        if (nlev == 0) wave_gepp<2, true>(v, ids, w, pid, bc[0], prow);
        else wave_gepp<2, false>(v, ids, w, gset, bc[0], nullptr);
        for (int lev = 1; lev <= nlev; ++lev) {
            const int sh = 2 * (lev - 1);
            const int si = p >> sh;                             // this workgroup's set at level lev - 1
            const int nsets = (a.P + (1 << sh) - 1) >> sh;
            const int g0 = si & ~3, g1 = min(g0 + 4, nsets);    // the group's sets
            const u64 tag = tag_base + u64(lev);
            double* recs = a.pub + (lev - 1) * OXI_TP_MAX_P * TP_REC;
            if ((p & ((1 << sh) - 1)) == 0) {                  // the set's leader publishes it
                const int pos = lane >= 2 && lane < 2 + TW && lane - 2 < w ? gset[lane - 2] : INT_MAX;
                double f = pos == INT_MAX ? -1.0 : double(pos);
                if (lane == 0) f = from_bits(tag);
                else if (lane == 1) f = double(__builtin_popcountll(__ballot(pos != INT_MAX)));
                else if (lane >= 2 + TW) f = 0.0;
                const u64 h = wave_xor64(lane < TP_REC - 1 ? rec_term(bits(f), tag, lane) : 0ull);
                if (lane == TP_REC - 1) f = from_bits(h);
                if (lane < TP_REC) put_agent(&recs[size_t(si) * TP_REC + lane], f);
            }
            // the group's sets: this workgroup's own from LDS, the others' records
            if (lane < TW) uni[(si - g0) * TW + lane] = lane < w ? gset[lane] : INT_MAX;
            for (int q = g1 - g0; q < 4; ++q)
                if (lane < TW) uni[q * TW + lane] = INT_MAX;
            unsigned have = 1u << (si - g0), spins = 0;
            const unsigned want = (1u << (g1 - g0)) - 1;
            while (have != want) {
                for (int j = g0; j < g1; ++j) {
                    if ((have >> (j - g0)) & 1) continue;
                    const double f = lane < TP_REC ? get_agent(&recs[size_t(j) * TP_REC + lane]) : 0.0;
                    if (readlane64(bits(f), 0) != tag) continue;
                    const u64 hx = wave_xor64(lane < TP_REC - 1 ? rec_term(bits(f), tag, lane) : 0ull);
                    if (hx != readlane64(bits(f), TP_REC - 1)) continue;
                    if (lane >= 2 && lane < 2 + TW) uni[(j - g0) * TW + lane - 2] = f < 0.0 ? INT_MAX : int(f);
                    have |= 1u << (j - g0);
                }
                if (have == want) break;
                if (++spins > TP_SPIN_LIMIT) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (have != want) {   // timed out (uniform in the wave)
                if (lane == 0) {
                    bad = 1;
                    __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                break;
            }
#pragma unroll
            for (int s = 0; s < 2; ++s) ids[s] = uni[lane + 64 * s];
            load_rows(a, ids, v);
            if (lev == nlev) wave_gepp<2, true>(v, ids, w, pid, bc[0], prow);
            else wave_gepp<2, false>(v, ids, w, gset, bc[0], nullptr);
        }
        // every row this workgroup reads is in registers: raise the flag
        __builtin_amdgcn_s_waitcnt(0);
        if (lane == 0) put_agent(&flg[p], from_bits(tag_flag));
    }
    __syncthreads();
    if (bad) return;   // uniform: written before the barrier

    // ---- final positions: the interchanges of the w pivots replayed on the
    // positions they touch (wave 0; lane i < w: position i, lane w + k:
    // pivot k's position when it lies below the leaf's w rows)
    if (wave == 0) {
        int pos = -1, content = -1;
        if (lane < w) { pos = lane; content = lane; }
        else if (lane < 2 * w && lane - w < w) {
            const int pk = pid[lane - w];
            if (pk != INT_MAX && pk >= w) { pos = pk; content = pk; }
        }
        for (int c = 0; c < w; ++c) {
            const int target = pid[c];
            const u64 m = __ballot(pos >= 0 && content == target);
            const int L = m ? first_lane(m) : c;
            const int ppos = __builtin_amdgcn_readlane(pos, L);
            if (lane == 0) ipiv_rel[c] = ppos;
            const int cc = __builtin_amdgcn_readlane(content, c), cl = __builtin_amdgcn_readlane(content, L);
            if (lane == c) content = cl;
            if (lane == L) content = cc;
        }
        if (lane < 2 * TW) { map_from[lane] = pos >= 0 ? content : -1; map_to[lane] = pos; }
    }
    // the other workgroups' flags: nobody stores before every workgroup has
    // read what it needs of the leaf's original rows
    {
        unsigned spins = 0;
        while (true) {
            int ok = 1;
            if (t < a.P) ok = bits(get_agent(&flg[t])) == tag_flag;
            ok = __syncthreads_and(ok);
            if (ok) break;
            if (++spins > TP_SPIN_LIMIT) {
                if (t == 0) __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }

    // ---- 4. own rows: eliminated with the pivot rows, stored at their final
    // positions (pivot rows are stored by workgroup 0 from prow)
    {
        bool mine[RPT];
        int dest[RPT];
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            bool piv = false;
            for (int c = 0; c < w; ++c) piv |= pid[c] == lr[s];
            mine[s] = live[s] && !piv;
            dest[s] = lr[s];
            if (mine[s] && lr[s] < w) {
                for (int k = 0; k < 2 * TW; ++k)
                    if (map_from[k] == lr[s]) dest[s] = map_to[k];
            }
        }
        // the columns in order, unrolled (compile-time indices, no ring):
        // l_c = a_c / u_cc, a_k = fma(-l_c, u_ck, a_k) for k > c; a zero
        // pivot leaves the column.  Rows that are not stored take the same
        // operations on values nobody reads.
#pragma unroll
        for (int c = 0; c < TW; ++c) {
            const float d = c < w ? prow[c][c] : 0.0f;
            if (d != 0.0f) {   // past the leaf, or a zero pivot: the column stays (uniform)
                float l[RPT];
#pragma unroll
                for (int s = 0; s < RPT; ++s) l[s] = o[s][c] / d;
#pragma unroll
                for (int k = c + 1; k < TW; ++k) {
                    const float uck = prow[c][k];   // U row c, column k
#pragma unroll
                    for (int s = 0; s < RPT; ++s) o[s][k] = __builtin_fmaf(-l[s], uck, o[s][k]);
                }
#pragma unroll
                for (int s = 0; s < RPT; ++s) o[s][c] = l[s];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (!mine[s]) continue;
            float* out = Wm + a.c0 + (a.c0 + i64(dest[s])) * a.ldw;
#pragma unroll
            for (int c = 0; c < TW; ++c)
                if (c < w) out[c] = o[s][c];
        }
    }
    if (p == 0) {
        for (int k = t; k < w * TW; k += TLT) {
            const int r = k / TW, c = k % TW;
            if (c < w) Wm[a.c0 + c + (a.c0 + i64(r)) * a.ldw] = prow[r][c];
        }
        if (t < w) a.ipiv[a.c0 + t] = a.c0 + ipiv_rel[t] + 1;
        if (t == 0) {
            for (int c = 0; c < w; ++c)
                if (prow[c][c] == 0.0f) {
                    if (*a.info == 0) *a.info = a.c0 + c + 1;
                    break;
                }
        }
    }
}

}  // namespace

int oxi_launch_lu_leaf_tp(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind) {
    if (a.P > OXI_TP_MAX_P) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "tournament leaf: %d workgroups", a.P);
    if (kind == 0) hipLaunchKernelGGL((lu_leaf_tp_kernel<1>), dim3(unsigned(a.P)), dim3(TLT), 0, s, a);
    else hipLaunchKernelGGL((lu_leaf_tp_kernel<2>), dim3(unsigned(a.P)), dim3(TLT), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "lu_leaf_tp launch: %s", hipGetErrorString(e));
    // the leaf's interchanges on the rest of its panel (its own columns are final)
    const unsigned nblk = unsigned(cdiv(a.c0 - a.xlo, 64) + cdiv(a.xhi - a.c0 - a.w, 64));
    if (nblk > 0) {
        hipLaunchKernelGGL(leaf_swaps_kernel<float>, dim3(nblk), dim3(128), 0, s, (float*)a.Wm, a.ldw, a.c0, a.w,
                           (const i64*)a.ipiv, (const int*)a.flags, a.xlo, a.xhi, std::max(a.w - 1, 0));
        e = hipGetLastError();
        if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "leaf_swaps launch: %s", hipGetErrorString(e));
    }
    return OXI_OK;
}
