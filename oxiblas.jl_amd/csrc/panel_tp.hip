// Tournament-pivoting panel leaf for the FP32 factor of the mixed-precision
// solve.  No counterpart in the reference: its mixed_solve factors the Float32
// copy with partial pivoting (src/lapack/extended_lapack.jl:183,195-198); the
// mixed solve's contract is the refinement's convergence and accuracy, not
// the factor's pivots, and partial pivoting stays the default.
//
// The partial-pivoting leaf (panel_kernel.hpp) exchanges one record per
// column between its P workgroups: w device-scope round trips per leaf.  Here
// the w pivots of a leaf are chosen by a tournament (communication-avoiding
// LU: Grigori, Demmel and Xiang, "CALU: a communication optimal LU
// factorization algorithm", SIAM J. Matrix Anal. Appl. 32(4), 2011), with a
// flat reduction tree of one level and two exchanges per leaf:
//  1. each workgroup runs Gaussian elimination with partial pivoting on its
//     own rows (in registers), only to pick its w candidate rows;
//  2. the workgroups publish their candidates' positions (one tagged,
//     checksummed record each), and every workgroup gathers all records;
//  3. every workgroup loads the candidates' original rows (at most 32 P <=
//     1024) and runs the same elimination on them: its pivots, in order, are
//     the leaf's pivots, and the eliminated pivot rows are the leaf's rows of
//     U and the top of its L.  Then it raises a tagged "loaded" flag;
//  4. after every workgroup's flag, each one eliminates its own original rows
//     with the pivot rows (per row exactly the operations of partial pivoting
//     with these pivots: l_c = a_c / u_cc, a_k = fma(-l_c, u_ck, a_k), a zero
//     pivot leaves the column and updates nothing) and stores them at their
//     final positions; workgroup 0 stores the pivot rows, ipiv and info.
// Ties in a pivot search go to the row of lowest position; NaN never wins.
//
// Rows sit in registers as a ring: at step c register r holds column c + r
// and registers W - c .. W - 1 the multipliers of steps 0 .. c - 1, so every
// register index is a compile-time constant in the rolled column loop; after
// W steps a row's registers are its W multipliers in order.
#include "panel_kernel.hpp"

#include <climits>

namespace {

constexpr int TW = OXI_LEAF_W;           // 32
constexpr int TLT = 256, TNWV = TLT / 64;
constexpr int TP_CR = OXI_TP_MAX_P * TW / TLT;   // candidate rows per thread (4)
constexpr int TP_REC = 36;               // record fields: tag, count, 32 positions, spare, checksum
constexpr unsigned TP_SPIN_LIMIT = 1u << 22;
static_assert(OXI_TP_MAX_P * TW % TLT == 0, "candidate slots per thread");

struct TpSel {
    float key;
    int id;
    int pad[2];
    float ring[TW];
};

__device__ __forceinline__ float tp_key(float v) {
    const float a = __builtin_fabsf(v);
    return a != a ? -1.0f : a;
}
__device__ __forceinline__ bool tp_better(float k1, int r1, float k2, int r2) {
    return k1 > k2 || (k1 == k2 && r1 < r2);
}
__device__ __forceinline__ int wave_min_int(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    return x;
}
__device__ __forceinline__ u64 wave_xor64(u64 x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x ^= (u64)__shfl_xor((long long)x, o);
    return x;
}

// One step of elimination on the rings of a thread's rows: the pivot value
// d (column c of the pivot row) and u[r] = column c + 1 + r of the pivot row
// (r < W - 1 - c).  l = v[0] / d; registers 1 .. W-1-c (columns c+1 .. W-1)
// get fma(-l, u, v), the ring rotates by one and l enters at the end.  A zero
// pivot leaves the column: l = v[0], nothing updated.
template <int NR>
__device__ __forceinline__ void ring_step(float (&v)[NR][TW], const bool (&live)[NR], int c, float d,
                                          const float (&u)[TW - 1]) {
    const bool upd = d != 0.0f;
#pragma unroll
    for (int s = 0; s < NR; ++s) {
        if (!live[s]) continue;
        const float l = upd ? v[s][0] / d : v[s][0];
#pragma unroll
        for (int r = 0; r + 1 < TW; ++r)
            v[s][r] = (upd && r < TW - 1 - c) ? __builtin_fmaf(-l, u[r], v[s][r + 1]) : v[s][r + 1];
        v[s][TW - 1] = l;
    }
}

// Gaussian elimination with partial pivoting over the rows the workgroup's
// threads hold (NR each, positions id, live = still eligible), for columns
// 0 .. w-1.  pid[c] receives step c's pivot position (INT_MAX when no row is
// left).  With `prow`, step c's pivot ring (the full row as partial pivoting
// leaves it: multipliers left of c, U from c on) goes to prow[c][0..W-1] in
// column order, U's rotated row to urot[c] and the pivot to ucc[c].
template <int NR, bool KEEP>
__device__ __forceinline__ void gepp_rounds(float (&v)[NR][TW], const int (&id)[NR], bool (&live)[NR], int w,
                                            TpSel (*sel)[TNWV], int* pid, float (*prow)[TW], float (*urot)[TW],
                                            float* ucc) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll 1
    for (int c0 = 0; c0 < w; ++c0) {
        const int c = __builtin_amdgcn_readfirstlane(c0);
        const int par = c & 1;
        float bk = -2.0f;
        int bid = INT_MAX, bs = -1;
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            if (live[s]) {
                const float k = tp_key(v[s][0]);
                if (tp_better(k, id[s], bk, bid)) { bk = k; bid = id[s]; bs = s; }
            }
        }
        const float wk = float(wave_max(double(bk)));
        const int wid = wave_min_int(bk == wk ? bid : INT_MAX);
        const u64 own = __ballot(bs >= 0 && bk == wk && bid == wid);
        if (own) {
            const int ol = first_lane(own);
            if (lane == ol) {
#pragma unroll
                for (int s = 0; s < NR; ++s)
                    if (s == bs)
#pragma unroll
                        for (int r = 0; r < TW; ++r) sel[par][wave].ring[r] = v[s][r];
            }
        }
        if (lane == 0) { sel[par][wave].key = own ? wk : -3.0f; sel[par][wave].id = own ? wid : INT_MAX; }
        __syncthreads();
        float gk = sel[par][0].key;
        int gid = sel[par][0].id, gw = 0;
#pragma unroll
        for (int q = 1; q < TNWV; ++q)
            if (tp_better(sel[par][q].key, sel[par][q].id, gk, gid)) { gk = sel[par][q].key; gid = sel[par][q].id; gw = q; }
        if (t == 0) pid[c] = gid;
        if (gid == INT_MAX) continue;   // no row left (uniform)
        const float* pr = sel[par][gw].ring;
        if (KEEP && t < TW) {
            prow[c][(c + t) & (TW - 1)] = pr[t];
            if (t + 1 < TW) urot[c][t] = pr[t + 1];
            if (t == 0) ucc[c] = pr[0];
        }
#pragma unroll
        for (int s = 0; s < NR; ++s)
            if (id[s] == gid) live[s] = false;
        float u[TW - 1];
#pragma unroll
        for (int r = 0; r + 1 < TW; ++r) u[r] = pr[r + 1];
        ring_step<NR>(v, live, c, pr[0], u);
    }
}

template <int RPT>
__global__ __launch_bounds__(TLT) void lu_leaf_tp_kernel(LeafArgs a) {
    constexpr i64 R = i64(TLT) * RPT;
    __shared__ TpSel sel[2][TNWV];
    __shared__ int pid_loc[TW], pid[TW];
    __shared__ int cand[OXI_TP_MAX_P * TW];
    __shared__ float prow[TW][TW], urot[TW][TW];
    __shared__ float ucc[TW];
    __shared__ int ipiv_rel[TW];
    __shared__ int map_from[2 * TW], map_to[2 * TW];
    __shared__ int nvalid;
    __shared__ int bad;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int p = blockIdx.x;
    const int w = a.w;
    float* Wm = (float*)a.Wm;
    const u64 tag_rec = a.tag0 + 2 * u64(a.c0);
    const u64 tag_flag = tag_rec + 1;
    double* recs = a.pub;                                  // [P][TP_REC]
    double* flg = a.pub + OXI_TP_MAX_P * TP_REC;           // [P]

    // ---- own rows: original values (kept to the end) and a working ring
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

    // ---- 1. local candidates
    {
        float v[RPT][TW];
        bool lv[RPT];
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            lv[s] = live[s];
#pragma unroll
            for (int c = 0; c < TW; ++c) v[s][c] = o[s][c];
        }
        gepp_rounds<RPT, false>(v, lr, lv, w, sel, pid_loc, nullptr, nullptr, nullptr);
    }
    __syncthreads();

    // ---- 2. publish them, gather everybody's (wave 0 writes; lanes = fields)
    if (wave == 0) {
        int cnt = 0;
        for (int c = 0; c < w; ++c) cnt += pid_loc[c] != INT_MAX;
        double f = 0.0;
        if (lane == 0) f = from_bits(tag_rec);
        else if (lane == 1) f = double(cnt);
        else if (lane < 2 + TW) f = (lane - 2 < w) ? double(pid_loc[lane - 2] == INT_MAX ? -1 : pid_loc[lane - 2]) : -1.0;
        const u64 h = wave_xor64(lane < TP_REC - 1 ? rec_term(bits(f), tag_rec, lane) : 0ull);
        if (lane == TP_REC - 1) f = from_bits(h);
        if (lane < TP_REC) put_agent(&recs[size_t(p) * TP_REC + lane], f);
    }
    // every wave gathers records q = wave, wave + 4, ...; a round is a pass
    // over the missing ones, then a barrier and a count
    if (t == 0) { nvalid = 0; bad = 0; }
    __syncthreads();
    {
        u64 have = 0;   // bit k: this wave's k-th record is in (wave-uniform)
        unsigned spins = 0;
        while (true) {
            for (int q = wave, k = 0; q < a.P; q += TNWV, ++k) {
                if ((have >> k) & 1) continue;
                const double f = lane < TP_REC ? get_agent(&recs[size_t(q) * TP_REC + lane]) : 0.0;
                const u64 tg = readlane64(bits(f), 0);
                if (tg != tag_rec) continue;
                const u64 hx = wave_xor64(lane < TP_REC - 1 ? rec_term(bits(f), tag_rec, lane) : 0ull);
                if (hx != readlane64(bits(f), TP_REC - 1)) continue;
                if (lane >= 2 && lane < 2 + TW) cand[q * TW + lane - 2] = int(f);
                have |= 1ull << k;
                if (lane == 0) atomicAdd(&nvalid, 1);
            }
            __syncthreads();
            const int nv = nvalid;
            __syncthreads();
            if (nv >= a.P) break;
            if (++spins > TP_SPIN_LIMIT) {
                if (t == 0) __hip_atomic_store(a.flags, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;   // uniform: every thread read the same count
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }

    // ---- 3. the tournament: the candidates' original rows, the elimination
    {
        float cr[TP_CR][TW];
        int cid[TP_CR];
        bool clive[TP_CR];
#pragma unroll
        for (int i = 0; i < TP_CR; ++i) {
            const int k = t + i * TLT;
            const int pos = k < a.P * TW ? cand[k] : -1;
            clive[i] = pos >= 0;
            cid[i] = clive[i] ? pos : INT_MAX;
            const float* row = Wm + a.c0 + (a.c0 + i64(clive[i] ? pos : 0)) * a.ldw;
#pragma unroll
            for (int c = 0; c < TW; ++c) cr[i][c] = (clive[i] && c < w) ? row[c] : 0.0f;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        // every row this workgroup reads is in registers: raise the flag
        if (t == 0) put_agent(&flg[p], from_bits(tag_flag));
        gepp_rounds<TP_CR, true>(cr, cid, clive, w, sel, pid, prow, urot, ucc);
    }
    __syncthreads();

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
#pragma unroll 1
        for (int c0 = 0; c0 < TW; ++c0) {
            const int c = __builtin_amdgcn_readfirstlane(c0);
            const float d = c < w ? ucc[c] : 0.0f;
            float u[TW - 1];
#pragma unroll
            for (int r = 0; r + 1 < TW; ++r) u[r] = c < w ? urot[c][r] : 0.0f;
            ring_step<RPT>(o, mine, c, d, u);
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
                if (ucc[c] == 0.0f) {
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
