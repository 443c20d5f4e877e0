// Blocked LU with partial pivoting (getrf), row interchanges (laswp) and
// the driver of the LU path on gfx950.
//
// Replaces _lu_factor! / _lu_panel! / _lu_pivot / _lu_swap!
// (src/lapack/factorizations.jl:30-158) and trsm! (src/blas/level3.jl:309-544).
//
// Working layout.  getrf runs on W = A^T (A's rows are W's columns), held in
// place for square matrices and in the handle workspace otherwise.  In that
// layout a row interchange of A moves two contiguous vectors (coalesced) instead
// of two strided rows, one row of a panel is a contiguous segment, and the
// trailing update A22 -= L21*U12 is the NN product W22 -= W12'*W21'.  The
// factors are transposed back into A (column-major, LAPACK layout) at the end.
//
// Algorithm (the reference's, in W's coordinates):
//   for each panel of nb columns:
//     getrf2 recursion on the panel (halving, leaves of <= 32 columns)
//     interchanges applied to the columns left and right of the panel
//     U12 <- L11^-1 A12           (right-side solve with L11^T in W)
//     A22 <- A22 - L21*U12        (FP64 MFMA GEMM)
// The leaf eliminates column by column exactly as _lu_panel!'s leaf: the pivot
// is the first row of largest |value| (strict '>' semantics, NaN never beats
// a number except in the diagonal position), rows are swapped across the leaf,
// the column is divided by the pivot and the leaf's remaining columns get the
// rank-1 update fma(-l, u, a); a zero pivot records info and skips the column.
//
// The leaf runs as one persistent kernel over P workgroups (P co-resident,
// P <= number of CUs) that each own a contiguous block of rows in registers;
// per column every workgroup publishes its candidate (key, row and the whole
// row) as a self-validating record and reads all the others until they are
// valid (bounded wait with a timeout flag so every wave exits), so all pick
// the same winner; with P == 1 the exchange stays in LDS (panel.hip).  A
// small kernel after each leaf applies its interchanges to its multipliers
// and to the rest of the panel.
#include "oxi_internal.hpp"

#include <vector>
#include <algorithm>
#include <cstdlib>

namespace {

// ---------------------------------------------------------------------------
// Transposes
// ---------------------------------------------------------------------------
constexpr int TT = 32;  // transpose tile

// out (n x m, ldo) <- in^T, in is m x n (ldi)
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T* __restrict__ in, i64 ldi, T* __restrict__ out,
                                                        i64 ldo, i64 m, i64 n) {
    __shared__ T tile[TT][TT + 1];
    const i64 i0 = i64(blockIdx.x) * TT, j0 = i64(blockIdx.y) * TT;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8
    for (int y = ty; y < TT; y += 8) {
        i64 i = i0 + tx, j = j0 + y;
        if (i < m && j < n) tile[y][tx] = in[i + j * ldi];
    }
    __syncthreads();
    for (int y = ty; y < TT; y += 8) {
        i64 j = j0 + tx, i = i0 + y;  // out[j + i*ldo] = in[i + j*ldi]
        if (i < m && j < n) out[j + i * ldo] = tile[tx][y];
    }
}

// In-place transpose of the leading n x n block: tile (bi, bj) with bi < bj is
// swapped with tile (bj, bi); diagonal tiles are transposed in LDS.
template <typename T>
__global__ __launch_bounds__(256) void transpose_square_kernel(T* __restrict__ A, i64 lda, i64 n, int tiles,
                                                               i64 id0) {
    __shared__ T t1[TT][TT + 1], t2[TT][TT + 1];
    // linear block id -> (bi, bj) with bi <= bj: row bi of the upper triangle
    // starts at S(bi) = bi*tiles - bi*(bi-1)/2; the root of S(b) = id, then
    // exact integer corrections
    const i64 id = id0 + blockIdx.x, tl = tiles;
    auto start = [&](i64 b) { return b * tl - b * (b - 1) / 2; };
    const double q = double(2 * tl + 1);
    i64 bi = i64((q - sqrt(q * q - 8.0 * double(id))) * 0.5);
    if (bi < 0) bi = 0;
    if (bi > tl - 1) bi = tl - 1;
    while (bi > 0 && start(bi) > id) --bi;
    while (bi + 1 < tl && start(bi + 1) <= id) ++bi;
    const int bj = int(bi + (id - start(bi)));
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const i64 r0 = bi * TT, c0 = i64(bj) * TT;
    if (c0 + TT <= n) {
        // both tiles inside A (bi <= bj): every load issued before any wait
        T a1[TT / 8], a2[TT / 8];
#pragma unroll
        for (int k = 0; k < TT / 8; ++k) {
            const int y = ty + 8 * k;
            a1[k] = A[(r0 + tx) + (c0 + y) * lda];
            a2[k] = A[(c0 + tx) + (r0 + y) * lda];
        }
#pragma unroll
        for (int k = 0; k < TT / 8; ++k) {
            t1[ty + 8 * k][tx] = a1[k];
            t2[ty + 8 * k][tx] = a2[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TT / 8; ++k) {
            const int y = ty + 8 * k;
            A[(c0 + tx) + (r0 + y) * lda] = t1[tx][y];
            if (bi != bj) A[(r0 + tx) + (c0 + y) * lda] = t2[tx][y];
        }
        return;
    }
    for (int y = ty; y < TT; y += 8) {
        i64 r = r0 + tx, c = c0 + y;
        if (r < n && c < n) t1[y][tx] = A[r + c * lda];       // tile (bi, bj): t1[col][row]
        i64 r2 = c0 + tx, c2 = r0 + y;
        if (bi != bj && r2 < n && c2 < n) t2[y][tx] = A[r2 + c2 * lda];  // tile (bj, bi)
    }
    __syncthreads();
    for (int y = ty; y < TT; y += 8) {
        // new tile (bj, bi) element (c0+tx, r0+y) = old A[r0+y][c0+tx] = t1[tx][y]
        i64 r = c0 + tx, c = r0 + y;
        if (r < n && c < n) A[r + c * lda] = t1[tx][y];
        if (bi != bj) {
            i64 r2 = r0 + tx, c2 = c0 + y;  // new tile (bi, bj) element = old A[c0+y][r0+tx] = t2[tx][y]
            if (r2 < n && c2 < n) A[r2 + c2 * lda] = t2[tx][y];
        }
    }
}

// ---------------------------------------------------------------------------
// Row interchanges.
// perm_pairs_kernel: one workgroup turns the interchanges k1..k2 (0-based
// steps of ipiv, 1-based rows) into (dst, src) pairs of the net permutation,
// new[dst] = old[src].
//
// LU pivots satisfy p_s >= k1+s, and then the net permutation has a closed
// form that needs no sequential replay: position k1+s is final after step s
// and holds what position p_s held just before step s.  A position q >= k1+s
// changes before step s only at steps u < s with p_u = q, where it receives
// val(u), the content of position k1+u just before step u.  So with
//   prev(t)  = the latest u < t with p_u = k1+t,
//   same(s)  = the latest u < s with p_u = p_s,
//   val(t)   = prev(t) >= 0 ? val(prev(t)) : k1+t     (pointer jumping),
// position k1+s ends with same(s) >= 0 ? val(same(s)) : p_s, and a pivot row
// q beyond k2 ends with val(u) for the last step u with p_u = q.  Other
// interchange lists (and the reverse order) are replayed in order by one
// thread in LDS.
// ---------------------------------------------------------------------------
constexpr int PERM_MAX_STEPS = 512;
constexpr int HASH = 2048;

__global__ __launch_bounds__(256) void perm_pairs_kernel(const i64* __restrict__ ipiv, i64 poff, i64 k1, int cnt,
                                                         int inverse, int* __restrict__ pairs /*[2*2*cnt+1]*/) {
    __shared__ int pv[PERM_MAX_STEPS];     // 0-based pivot row of each step
    __shared__ int prv[PERM_MAX_STEPS];    // prev(t)
    __shared__ int sm[PERM_MAX_STEPS];     // same(s)
    __shared__ int val[PERM_MAX_STEPS], nxt[PERM_MAX_STEPS];
    __shared__ int islast[PERM_MAX_STEPS];
    __shared__ int slot_row[2 * PERM_MAX_STEPS], slot_src[2 * PERM_MAX_STEPS];
    __shared__ int hkey[HASH], hval[HASH];
    __shared__ int np, ordered, unresolved;
    const int t = threadIdx.x, nt = blockDim.x;
    const int K1 = int(k1);
    if (t == 0) { np = 0; ordered = 1; unresolved = 0; }
    for (int i = t; i < cnt; i += nt) pv[i] = int(ipiv[k1 + i] - 1 - poff);
    __syncthreads();
    for (int i = t; i < cnt; i += nt) {
        if (pv[i] < K1 + i) ordered = 0;
        prv[i] = -1; islast[i] = 1;
    }
    __syncthreads();

    if (!inverse && ordered) {
        // prev(t): every u with p_u = k1+t has u <= t
        for (int u = t; u < cnt; u += nt) {
            const int w = pv[u] - K1;
            if (w > u && w < cnt) atomicMax(&prv[w], u);
        }
        // same(s): previous occurrence of the same pivot row (broadcast reads)
        for (int s = t; s < cnt; s += nt) {
            const int ps = pv[s];
            int b = -1;
            for (int u = 0; u < s; ++u) b = (pv[u] == ps) ? u : b;
            sm[s] = b;
        }
        __syncthreads();
        for (int i = t; i < cnt; i += nt) {
            val[i] = prv[i] < 0 ? K1 + i : -1;
            nxt[i] = prv[i];
            if (sm[i] >= 0) islast[sm[i]] = 0;
        }
        // val by pointer jumping (chains point to earlier steps)
        for (int round = 0; round < 12; ++round) {
            __syncthreads();
            if (t == 0) unresolved = 0;
            int nv[PERM_MAX_STEPS / 256], nn[PERM_MAX_STEPS / 256];
            for (int k = 0, i = t; i < cnt; i += nt, ++k) {
                nv[k] = val[i]; nn[k] = nxt[i];
                if (nv[k] < 0) {
                    const int r = nn[k];
                    if (val[r] >= 0) { nv[k] = val[r]; nn[k] = -1; }
                    else nn[k] = nxt[r];
                }
            }
            __syncthreads();
            for (int k = 0, i = t; i < cnt; i += nt, ++k) {
                val[i] = nv[k]; nxt[i] = nn[k];
                if (nv[k] < 0) unresolved = 1;
            }
            __syncthreads();
            if (!unresolved) break;
        }
    }
    if (!inverse && ordered && !unresolved) {
        // pairs: the panel positions, then pivot rows beyond k2
        for (int s = t; s < cnt; s += nt) {
            const int src = sm[s] >= 0 ? val[sm[s]] : pv[s];
            if (src != K1 + s) {
                const int q = atomicAdd(&np, 1);
                pairs[1 + 2 * q] = K1 + s; pairs[2 + 2 * q] = src;
            }
            if (pv[s] >= K1 + cnt && islast[s] && val[s] != pv[s]) {
                const int q = atomicAdd(&np, 1);
                pairs[1 + 2 * q] = pv[s]; pairs[2 + 2 * q] = val[s];
            }
        }
        __syncthreads();
        if (t == 0) pairs[0] = np;
        return;
    }

    // general replay, also the safety net should a chain stay unresolved:
    // rows k1..k2 occupy slots 0..cnt-1; other rows get slots
    // through an LDS hash table
    for (int i = t; i < HASH; i += nt) hkey[i] = -1;
    for (int i = t; i < cnt; i += nt) { slot_row[i] = K1 + i; slot_src[i] = K1 + i; }
    __syncthreads();
    if (t == 0) {
        int ns = cnt;
        auto find = [&](int row) -> int {
            if (row >= K1 && row < K1 + cnt) return row - K1;
            unsigned hsh = (unsigned(row) * 2654435761u) & (HASH - 1);
            while (true) {
                if (hkey[hsh] == row) return hval[hsh];
                if (hkey[hsh] == -1) {
                    hkey[hsh] = row; hval[hsh] = ns;
                    slot_row[ns] = row; slot_src[ns] = row;
                    return ns++;
                }
                hsh = (hsh + 1) & (HASH - 1);
            }
        };
        for (int s = 0; s < cnt; ++s) {
            const int step = inverse ? cnt - 1 - s : s;
            const int j = K1 + step;
            const int p = pv[step];
            if (p == j) continue;
            const int a = find(j), b = find(p);
            const int tmp = slot_src[a]; slot_src[a] = slot_src[b]; slot_src[b] = tmp;
        }
        int n2 = 0;
        for (int q = 0; q < ns; ++q)
            if (slot_src[q] != slot_row[q]) { pairs[1 + 2 * n2] = slot_row[q]; pairs[2 + 2 * n2] = slot_src[q]; ++n2; }
        pairs[0] = n2;
    }
}

// Apply pairs to columns of W: W[c][dst] <- W[c][src] for every c in
// [c0, c1) (W column = A row, W row = A column).  A workgroup stages 16 W rows
// x all pairs through LDS so the reads complete before any write.  The pair
// list is copied to LDS first, and each thread issues its loads eight at a
// time without branches (a padding slot reloads the last pair): a pair index
// read from global memory, or a load behind a bounds branch, would cost one
// memory round trip per pair.
constexpr int PAIR_ROWS = 16;
template <typename T>
__global__ __launch_bounds__(256) void apply_pairs_w_kernel(T* __restrict__ Wm, i64 ldw, i64 c0, i64 c1,
                                                            const int* __restrict__ pairs, i64 roff) {
    // [npairs][PAIR_ROWS] values, then the pair list (dst, src) per pair
    extern __shared__ __attribute__((aligned(16))) unsigned char buf_raw[];
    T* buf = (T*)buf_raw;
    const int np = pairs[0];
    if (np == 0) return;
    int* pl = (int*)(buf_raw + size_t(np) * PAIR_ROWS * sizeof(T));
    const int t = threadIdx.x;
    for (int i = t; i < 2 * np; i += blockDim.x) pl[i] = pairs[1 + i];
    __syncthreads();
    const i64 cbase = c0 + i64(blockIdx.x) * PAIR_ROWS;
    const int r = t % PAIR_ROWS, q0 = t / PAIR_ROWS, qs = blockDim.x / PAIR_ROWS;
    const i64 c = cbase + r;
    const bool ok = c < c1;
    const i64 cr = ok ? c : c0;                      // a column inside the range to read
    constexpr int U = 8;
    for (int q = q0; q < np; q += U * qs) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int qq = min(q + u * qs, np - 1);
            v[u] = Wm[cr + (i64(pl[2 * qq + 1]) + roff) * ldw];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int qq = q + u * qs;
            if (qq < np) buf[qq * PAIR_ROWS + r] = v[u];
        }
    }
    __syncthreads();
    if (ok)
        for (int q = q0; q < np; q += qs) Wm[c + (i64(pl[2 * q]) + roff) * ldw] = buf[q * PAIR_ROWS + r];
}

// Column-major (LAPACK) form: rows of A (n x ncols, lda) permuted by pairs.
__global__ __launch_bounds__(256) void apply_pairs_cm_kernel(double* __restrict__ A, i64 lda, i64 ncols,
                                                             const int* __restrict__ pairs, i64 roff) {
    extern __shared__ double buf[];  // [npairs]
    const int np = pairs[0];
    if (np == 0) return;
    const i64 col = blockIdx.x;
    if (col >= ncols) return;
    double* a = A + col * lda;
    for (int q = threadIdx.x; q < np; q += blockDim.x) buf[q] = a[pairs[2 + 2 * q] + roff];
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += blockDim.x) a[pairs[1 + 2 * q] + roff] = buf[q];
}

__global__ void zero_i64_kernel(i64* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__global__ void zero_i32_kernel(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

// ===========================================================================
// Host-side driver, templated on the working precision (FP64 for lu, FP32 for
// the mixed-precision factorization)
// ===========================================================================
namespace {

int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

template <typename T>
struct GetrfCtx {
    oxi_handle_t h;
    hipStream_t s;
    T* Wm; i64 ldw;
    i64 m, n;
    i64* ipiv;
    i64* dinfo;
    unsigned long long tag0;   // (call sequence) << 32: tags of the leaf exchange records
    double* pub;
    int* pairs;
    int* flags;
    int prec;          // FP32 only: -1 FP32 trailing update, 0 BF16, 1 FP16, 2 FP8 e4m3
    bool tournament = false;   // FP32 only: tournament-pivoting leaves (panel_tp.hip)
    void* lowp_ws; size_t lowp_bytes;
    i64 pc0 = 0, pc1 = 0;   // the top-level panel being factored: columns [pc0, pc1)
};

template <typename T>
int gemm_nn(oxi_handle_t h, hipStream_t s, i64 m, i64 n, i64 k, T alpha, const T* A, i64 lda, const T* B, i64 ldb,
            T beta, T* C, i64 ldc);
template <>
int gemm_nn<double>(oxi_handle_t h, hipStream_t s, i64 m, i64 n, i64 k, double alpha, const double* A, i64 lda,
                    const double* B, i64 ldb, double beta, double* C, i64 ldc) {
    return oxi_launch_dgemm(h, s, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
template <>
int gemm_nn<float>(oxi_handle_t h, hipStream_t s, i64 m, i64 n, i64 k, float alpha, const float* A, i64 lda,
                   const float* B, i64 ldb, float beta, float* C, i64 ldc) {
    return oxi_launch_sgemm(h, s, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int trsm_ru(oxi_handle_t h, hipStream_t s, i64 rows, i64 n, const double* T, i64 ldt, double* X, i64 ldx) {
    return oxi_dtrsm_right(h, s, 1, 0, 1, rows, n, T, ldt, X, ldx);
}
int trsm_ru(oxi_handle_t h, hipStream_t s, i64 rows, i64 n, const float* T, i64 ldt, float* X, i64 ldx) {
    return oxi_strsm_right(h, s, 1, 0, 1, rows, n, T, ldt, X, ldx);
}

// Interchanges k1..k2 (0-based steps, inclusive; pivots are global 1-based rows)
// applied to A columns [c0, c1) = W rows [c0, c1).
template <typename T>
int swap_w(GetrfCtx<T>& g, i64 k1, i64 k2, i64 c0, i64 c1, int inverse) {
    if (k2 < k1 || c1 <= c0) return OXI_OK;
    const int cnt = int(k2 - k1 + 1);
    if (cnt > PERM_MAX_STEPS) return oxi_set_error(g.h, OXI_ERR_UNSUPPORTED, "laswp block of %d steps", cnt);
    hipLaunchKernelGGL(perm_pairs_kernel, dim3(1), dim3(256), 0, g.s, g.ipiv, i64(0), k1, cnt, inverse, g.pairs);
    OXI_TRY(launch_check(g.h, "perm_pairs"));
    const size_t shm = size_t(2 * cnt) * (PAIR_ROWS * sizeof(T) + 2 * sizeof(int));
    hipLaunchKernelGGL(apply_pairs_w_kernel<T>, dim3(unsigned(cdiv(c1 - c0, PAIR_ROWS))), dim3(256), shm, g.s, g.Wm,
                       g.ldw, c0, c1, g.pairs, i64(0));
    return launch_check(g.h, "apply_pairs_w");
}

template <typename T>
int leaf(GetrfCtx<T>& g, i64 c0, int w) {
    const i64 mrows = g.m - c0;
    // rows per workgroup: 256 (one per thread) up to 8192 rows (32
    // workgroups), then 512 (two per thread: the most FP64 rows that stay in
    // registers without spilling); FP32 rows are half the size and go to 1024
    // per workgroup beyond 16384 rows (more than 32 workgroups of 512)
    int kind = mrows > 8192 ? 1 : 0;
    if (sizeof(T) == 4 && mrows > 16384) kind = 2;
    const i64 R = kind == 0 ? 256 : kind == 1 ? 512 : 1024;
    const i64 P = cdiv(mrows, R);
    if (P > std::min<i64>(g.h->num_cus, OXI_LEAF_MAX_P))
        return oxi_set_error(g.h, OXI_ERR_UNSUPPORTED, "panel of %lld rows exceeds the leaf kernel capacity",
                             (long long)mrows);
    LeafArgs a{(void*)g.Wm, g.ldw, c0, w, mrows, g.ipiv, g.dinfo, g.tag0, g.pub, int(P), g.flags, g.pc0, g.pc1};
    if (sizeof(T) == 4 && g.tournament && kind <= 1 && P <= OXI_TP_MAX_P) return oxi_launch_lu_leaf_tp(g.h, g.s, a, kind);
    if (sizeof(T) == 8) OXI_TRY(oxi_launch_lu_leaf(g.h, g.s, a, kind));
    else OXI_TRY(oxi_launch_lu_leaf32(g.h, g.s, a, kind));
    return OXI_OK;
}

// getrf2 recursion on the panel A[c0:m, c0:c0+np] in W coordinates
// (src/lapack/factorizations.jl:114-152).  Each leaf applies its
// interchanges to all other columns of the top-level panel [g.pc0, g.pc1)
// (leaf_swaps_kernel), which covers the recursion's "left half's
// interchanges on the right half" and "right half's interchanges on the left
// half" at every level.
template <typename T>
int panel(GetrfCtx<T>& g, i64 c0, i64 np) {
    if (np <= OXI_LEAF_W) return leaf(g, c0, int(np));
    const i64 mp = g.m - c0;
    const i64 h = np / 2;
    OXI_TRY(panel(g, c0, h));
    // A12 <- L11^-1 A12 : W[c0+h : c0+np, c0 : c0+h] <- W[..] * (L11^T)^-1, L11^T = W[c0:c0+h, c0:c0+h] upper unit
    OXI_TRY(trsm_ru(g.h, g.s, np - h, h, g.Wm + c0 + c0 * g.ldw, g.ldw, g.Wm + (c0 + h) + c0 * g.ldw, g.ldw));
    if (mp > h) {
        // A22 -= A21 * A12  <=>  W[c0+h:c0+np, c0+h:m] -= W[c0+h:c0+np, c0:c0+h] * W[c0:c0+h, c0+h:m]
        OXI_TRY(gemm_nn<T>(g.h, g.s, np - h, mp - h, h, T(-1), g.Wm + (c0 + h) + c0 * g.ldw, g.ldw,
                           g.Wm + c0 + (c0 + h) * g.ldw, g.ldw, T(1), g.Wm + (c0 + h) + (c0 + h) * g.ldw, g.ldw));
        OXI_TRY(panel(g, c0 + h, np - h));
    }
    return OXI_OK;
}

// A top-level panel: columns [c0, c0 + np) of the factorization.
template <typename T>
int panel_top(GetrfCtx<T>& g, i64 c0, i64 np) {
    g.pc0 = c0;
    g.pc1 = c0 + np;
    return panel(g, c0, np);
}

// Trailing update of A columns [r0, r1) (W rows) with panel [j0, j1):
// W[r0:r1, j1:m] -= W[r0:r1, j0:j1] * W[j0:j1, j1:m]  (A22 -= L21 * U12).
int trailing_rows(GetrfCtx<double>& g, i64 j0, i64 j1, i64 r0, i64 r1) {
    return oxi_launch_dgemm(g.h, g.s, 0, 0, r1 - r0, g.m - j1, j1 - j0, -1.0, g.Wm + r0 + j0 * g.ldw, g.ldw,
                            g.Wm + j0 + j1 * g.ldw, g.ldw, 1.0, g.Wm + r0 + j1 * g.ldw, g.ldw);
}
int trailing_rows(GetrfCtx<float>& g, i64 j0, i64 j1, i64 r0, i64 r1) {
    if (g.prec < 0)
        return oxi_launch_sgemm(g.h, g.s, 0, 0, r1 - r0, g.m - j1, j1 - j0, -1.0f, g.Wm + r0 + j0 * g.ldw, g.ldw,
                                g.Wm + j0 + j1 * g.ldw, g.ldw, 1.0f, g.Wm + r0 + j1 * g.ldw, g.ldw);
    return oxi_lowp_update(g.h, g.s, g.prec, r1 - r0, g.m - j1, j1 - j0, -1.0f, g.Wm + r0 + j0 * g.ldw, g.ldw,
                           g.Wm + j0 + j1 * g.ldw, g.ldw, g.Wm + r0 + j1 * g.ldw, g.ldw, g.lowp_ws, g.lowp_bytes);
}

// The trailing updates that the optional profiling times: those that run on
// the FP64 GEMM's 128 x 128 kernel (the factorization's dominant kernel).
template <typename T>
bool timed_update(i64 rows, i64 cols) {
    return sizeof(T) == 8 ? oxi_dgemm_large_tiles(rows, cols) : true;
}

int choose_nb(i64 m, i64 n) {
    const i64 q = std::min(m, n);
    if (q >= 8192) return 256;
    if (q >= 1024) return 128;
    return 64;
}

// src: with a source matrix (lds), A only receives the factors (src is read
// once, by the change of layout); else A is factored in place.
template <typename T>
int getrf_t(oxi_handle_t h, i64 m, i64 n, T* A, i64 lda, i64* ipiv, i64 nb, i64* info, int prec,
            bool tournament = false, const T* src = nullptr, i64 lds = 0) {
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    const i64 q = std::min(m, n);
    const bool tuned = nb == 0;
    if (nb == 0) nb = choose_nb(m, n);
    if (nb > PERM_MAX_STEPS) nb = PERM_MAX_STEPS;
    // Tuned block size on large matrices: 512-wide panels while more than 3/8
    // of the columns remain (the trailing update bounds those steps and runs
    // faster with K = 512), 256 after (the panel bounds them).  Measured on
    // MI355X at n = 16384: 256 throughout 136 ms, 512 throughout 134 ms.
    // Wider early panels lose: 768 or 1024 (over 3/8 to 5/8 of the columns)
    // 117-123 ms against 115 ms for this schedule.
    const i64 nb_wide = (tuned && q >= 12288) ? std::min<i64>(2 * nb, PERM_MAX_STEPS) : nb;
    const i64 wide_until = q - (3 * q) / 8;
    auto nb_at = [&](i64 j) { return j < wide_until ? nb_wide : nb; };

    // workspace: [W if rectangular][pub][pairs main][pairs side][low-precision operands]
    const bool square = (m == n);
    const i64 ldw = square ? lda : n;
    const size_t wbytes = square ? 0 : size_t(n) * size_t(m) * sizeof(T);
    // leaf records: column jn < w of workgroup p < P at jn * P + p, the
    // diagonal rows at w * P + jn (panel_kernel.hpp), P <= OXI_LEAF_MAX_P
    const size_t pub_doubles = size_t(OXI_LEAF_W) * (size_t(OXI_LEAF_MAX_P) + 1) * OXI_PUB_STRIDE +
                               size_t(OXI_LEAF_W) * OXI_PUB_STRIDE;
    const size_t pairs_bytes = (size_t(4 * PERM_MAX_STEPS + 4) * 4 + 255) & ~size_t(255);
    const size_t off_pub = (wbytes + 255) & ~size_t(255);
    const size_t off_pairs = off_pub + ((pub_doubles * 8 + 255) & ~size_t(255));
    const size_t off_pairs2 = off_pairs + pairs_bytes;
    const size_t off_lowp = off_pairs2 + pairs_bytes;
    const size_t lowp_bytes = (sizeof(T) == 4 && prec >= 0) ? oxi_lowp_update_ws(prec, n, m, nb_wide) : 0;
    const size_t total = off_lowp + lowp_bytes;
    void* ws;
    OXI_TRY(oxi_workspace(h, total, &ws));
    char* base = (char*)ws;
    GetrfCtx<T> g;
    g.h = h; g.s = s;
    g.Wm = square ? A : (T*)base; g.ldw = ldw;
    g.m = m; g.n = n; g.ipiv = ipiv; g.dinfo = h->dev_info;
    g.pub = (double*)(base + off_pub);
    g.pairs = (int*)(base + off_pairs);
    g.tag0 = (++h->getrf_seq) << 32;
    g.flags = h->dev_flags;
    g.prec = prec;
    g.tournament = tournament;
    g.lowp_ws = base + off_lowp; g.lowp_bytes = lowp_bytes;

    static bool attr_set = false;
    if (!attr_set) {
        // the pair-staging kernel may use up to 2 * 512 pairs * (16 doubles + 2 ints) of LDS
        OXI_HIP(h, hipFuncSetAttribute((const void*)apply_pairs_w_kernel<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        OXI_HIP(h, hipFuncSetAttribute((const void*)apply_pairs_w_kernel<float>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(64), 0, s, g.dinfo, 1);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, g.flags, 1);
    OXI_TRY(launch_check(h, "zero"));

    static const bool lookahead = !getenv("OXIBLAS_NO_LOOKAHEAD");
    GetrfCtx<T> gs = g;               // panel context on the side stream
    gs.s = h->side;
    gs.pairs = (int*)(base + off_pairs2);
    bool panel_on_side = false;

    // A -> W.  The first panel reads only W's first nb0 rows, written by the
    // tile pairs of the first k0 tile rows: for a square matrix (with the
    // look-ahead) those pairs go first, the first panel then starts on the
    // side stream, and the rest of the transpose runs beside it.
    const int tiles = int(cdiv(n, TT));
    const i64 nblk = i64(tiles) * (tiles + 1) / 2;
    const i64 nb0 = std::min<i64>(nb_at(0), q);
    const bool early_panel = square && lookahead && q > 0;
    if (square && src) {
        // out of place: the first panel's rows of W first, then the rest
        const i64 c0 = early_panel ? nb0 : n;
        hipLaunchKernelGGL(transpose_kernel<T>, dim3(unsigned(cdiv(n, TT)), unsigned(cdiv(c0, TT))), dim3(256), 0, s,
                           src, lds, A, lda, n, c0);
        OXI_TRY(launch_check(h, "transpose"));
        if (early_panel) {
            OXI_HIP(h, hipEventRecord(h->ev_a, s));
            OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
            OXI_TRY(panel_top(gs, 0, nb0));
            OXI_HIP(h, hipEventRecord(h->ev_b, h->side));
            panel_on_side = true;
        }
        if (c0 < n)
            hipLaunchKernelGGL(transpose_kernel<T>, dim3(unsigned(cdiv(n, TT)), unsigned(cdiv(n - c0, TT))), dim3(256),
                               0, s, src + c0 * lds, lds, A + c0, lda, n, n - c0);
    } else if (square) {
        const i64 k0 = early_panel ? std::min<i64>(cdiv(nb0, TT), tiles) : tiles;
        const i64 ids1 = k0 * tiles - k0 * (k0 - 1) / 2;      // the pairs of tile rows < k0
        hipLaunchKernelGGL(transpose_square_kernel<T>, dim3(unsigned(ids1)), dim3(256), 0, s, A, lda, n, tiles,
                           i64(0));
        OXI_TRY(launch_check(h, "transpose"));
        if (early_panel) {
            OXI_HIP(h, hipEventRecord(h->ev_a, s));
            OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
            OXI_TRY(panel_top(gs, 0, nb0));
            OXI_HIP(h, hipEventRecord(h->ev_b, h->side));
            panel_on_side = true;
        }
        if (ids1 < nblk)
            hipLaunchKernelGGL(transpose_square_kernel<T>, dim3(unsigned(nblk - ids1)), dim3(256), 0, s, A, lda, n,
                               tiles, ids1);
    } else {
        hipLaunchKernelGGL(transpose_kernel<T>, dim3(unsigned(cdiv(m, TT)), unsigned(cdiv(n, TT))), dim3(256), 0, s,
                           src ? src : A, src ? lds : lda, g.Wm, ldw, m, n);
    }
    OXI_TRY(launch_check(h, "transpose"));

    // The factorization's main sequence runs on the bulk stream (kept off a
    // few CUs that the panel on the side stream then always finds free) and
    // returns to the caller's stream before the transpose back.
    const hipStream_t s_caller = s;
    if (h->bulk) {
        OXI_HIP(h, hipEventRecord(h->ev_c, s));
        OXI_HIP(h, hipStreamWaitEvent(h->bulk, h->ev_c, 0));
        s = h->bulk;
        g.s = s;
    }

    // Right-looking blocked LU with one panel of look-ahead: after panel j's
    // interchanges reach the columns to its right, the next panel's columns
    // are updated first and panel j+1 is factored on the side stream while
    // the main stream updates the rest of the trailing matrix.  The
    // interchanges of each panel on the columns to its left touch the L21
    // block that the update running beside the panel reads, so the main
    // stream applies them after that update, at the end of the step.
    //
    // The look-ahead block itself goes to the side stream too, when no
    // workspace is shared between the two streams' updates (FP64, and FP32
    // with FP32 products; the low-precision update stages its operands in
    // one buffer): the side stream takes panel j's interchanges, U12 block
    // and update on the next panel's columns [j1, j2) and then factors them,
    // while the main stream updates [j2, n), the columns of the panel after
    // next [j2, j3) first.  Event ev_a marks that first product (the side
    // stream's next look-ahead block reads its columns), ev_b the panel.
    // The main stream then runs the large products back to back; per step
    // it only waits for the panel.  The streams write disjoint columns; the
    // side stream's TRSMs have fewer than 1024 rows and leave the handle's
    // shared inverse buffer to the main stream.
    static const bool main_la = getenv("OXIBLAS_MAIN_LOOKAHEAD") != nullptr;
    const bool side_la = lookahead && !main_la && !(sizeof(T) == 4 && prec >= 0);
    if (side_la) {
        OXI_HIP(h, hipEventRecord(h->ev_a, s));   // the transpose: the side stream's first update reads it
        if (!panel_on_side) {
            OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
            OXI_TRY(panel_top(gs, 0, nb0));
            OXI_HIP(h, hipEventRecord(h->ev_b, h->side));
            panel_on_side = true;
        }
        for (i64 j0 = 0, jn = 0; j0 < q; j0 = jn) {
            const i64 jb = std::min<i64>(nb_at(j0), q - j0);
            const i64 j1 = j0 + jb;
            jn = j1;
            const i64 j2 = j1 < q ? std::min<i64>(j1 + nb_at(j1), q) : j1;   // next panel [j1, j2)
            const i64 j3 = j2 < q ? std::min<i64>(j2 + nb_at(j2), q) : j2;   // the one after [j2, j3)
            OXI_HIP(h, hipStreamWaitEvent(s, h->ev_b, 0));                    // panel [j0, j1)
            if (j2 > j1) {
                // side: the next panel's columns, then the next panel (it
                // waits for the main stream's update of [j1, j2) of the
                // previous step, or the transpose)
                OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
                OXI_TRY(swap_w(gs, j0, j1 - 1, j1, j2, 0));
                OXI_TRY(trsm_ru(h, gs.s, j2 - j1, jb, g.Wm + j0 + j0 * ldw, ldw, g.Wm + j1 + j0 * ldw, ldw));
                hipEvent_t pe = nullptr;
                if (timed_update<T>(j2 - j1, m - j1)) OXI_TRY(oxi_prof_begin(h, gs.s, &pe));
                OXI_TRY(trailing_rows(gs, j0, j1, j1, j2));
                OXI_TRY(oxi_prof_end(h, gs.s, pe, 2.0 * double(j2 - j1) * double(m - j1) * double(jb)));
                OXI_TRY(panel_top(gs, j1, j2 - j1));
                OXI_HIP(h, hipEventRecord(h->ev_b, h->side));
            }
            // main: columns [j2, n)
            if (j2 < n) {
                OXI_TRY(swap_w(g, j0, j1 - 1, j2, n, 0));
                OXI_TRY(trsm_ru(h, s, n - j2, jb, g.Wm + j0 + j0 * ldw, ldw, g.Wm + j2 + j0 * ldw, ldw));
                if (j1 < m) {
                    for (int part = 0; part < 2; ++part) {
                        const i64 r0 = part == 0 ? j2 : j3, r1 = part == 0 ? j3 : n;
                        if (r1 > r0) {
                            hipEvent_t pe = nullptr;
                            if (timed_update<T>(r1 - r0, m - j1)) OXI_TRY(oxi_prof_begin(h, s, &pe));
                            OXI_TRY(trailing_rows(g, j0, j1, r0, r1));
                            OXI_TRY(oxi_prof_end(h, s, pe, 2.0 * double(r1 - r0) * double(m - j1) * double(jb)));
                        }
                        if (part == 0) OXI_HIP(h, hipEventRecord(h->ev_a, s));
                    }
                }
            }
            if (j0 > 0) OXI_TRY(swap_w(g, j0, j1 - 1, 0, j0, 0));
        }
    } else {
        if (!panel_on_side) OXI_TRY(panel_top(g, 0, nb0));
        for (i64 j0 = 0, jn = 0; j0 < q; j0 = jn) {
            const i64 jb = std::min<i64>(nb_at(j0), q - j0);
            const i64 j1 = j0 + jb;
            jn = j1;
            if (panel_on_side) OXI_HIP(h, hipStreamWaitEvent(s, h->ev_b, 0));
            const bool next = j1 < q;                         // then j1 < m and j1 < n
            const i64 j2 = next && lookahead ? std::min<i64>(j1 + nb_at(j1), q) : j1;
            // this panel's interchanges on the look-ahead block first (the next
            // panel waits for them), on the rest of the trailing columns after
            // the next panel has been started
            OXI_TRY(swap_w(g, j0, j1 - 1, j1, j2 > j1 ? j2 : n, 0));
            if (j1 >= n) {
                if (j0 > 0) OXI_TRY(swap_w(g, j0, j1 - 1, 0, j0, 0));
                continue;
            }
            // look-ahead block: the next panel's columns [j1, j2)
            if (j2 > j1) {
                OXI_TRY(trsm_ru(h, s, j2 - j1, jb, g.Wm + j0 + j0 * ldw, ldw, g.Wm + j1 + j0 * ldw, ldw));
                hipEvent_t pe = nullptr;
                if (timed_update<T>(j2 - j1, m - j1)) OXI_TRY(oxi_prof_begin(h, s, &pe));
                OXI_TRY(trailing_rows(g, j0, j1, j1, j2));
                OXI_TRY(oxi_prof_end(h, s, pe, 2.0 * double(j2 - j1) * double(m - j1) * double(jb)));
                OXI_HIP(h, hipEventRecord(h->ev_a, s));
                OXI_HIP(h, hipStreamWaitEvent(h->side, h->ev_a, 0));
                OXI_TRY(panel_top(gs, j1, j2 - j1));
                OXI_HIP(h, hipEventRecord(h->ev_b, h->side));
                panel_on_side = true;
            }
            // the rest of the trailing matrix: columns [j2, n)
            if (j2 > j1) OXI_TRY(swap_w(g, j0, j1 - 1, j2, n, 0));
            if (j2 < n) {
                OXI_TRY(trsm_ru(h, s, n - j2, jb, g.Wm + j0 + j0 * ldw, ldw, g.Wm + j2 + j0 * ldw, ldw));
                if (j1 < m) {
                    hipEvent_t pe = nullptr;
                    if (timed_update<T>(n - j2, m - j1)) OXI_TRY(oxi_prof_begin(h, s, &pe));
                    OXI_TRY(trailing_rows(g, j0, j1, j2, n));
                    OXI_TRY(oxi_prof_end(h, s, pe, 2.0 * double(n - j2) * double(m - j1) * double(jb)));
                }
            }
            // this panel's interchanges on the columns to its left: the last
            // reader of those columns was the previous step's trailing update,
            // which ran before this point on this stream (per column the
            // interchanges still arrive in panel order)
            if (j0 > 0) OXI_TRY(swap_w(g, j0, j1 - 1, 0, j0, 0));
            if (next && !lookahead) {
                OXI_TRY(panel_top(g, j1, std::min<i64>(nb_at(j1), q - j1)));
                panel_on_side = false;
            }
        }
    }
    if (panel_on_side) OXI_HIP(h, hipStreamWaitEvent(s, h->ev_b, 0));

    if (s != s_caller) {
        OXI_HIP(h, hipEventRecord(h->ev_c, s));
        OXI_HIP(h, hipStreamWaitEvent(s_caller, h->ev_c, 0));
        s = s_caller;
    }

    // W -> A
    if (square)
        hipLaunchKernelGGL(transpose_square_kernel<T>, dim3(unsigned(nblk)), dim3(256), 0, s, A, lda, n, tiles, i64(0));
    else
        hipLaunchKernelGGL(transpose_kernel<T>, dim3(unsigned(cdiv(n, TT)), unsigned(cdiv(m, TT))), dim3(256), 0, s,
                           g.Wm, ldw, A, lda, n, m);
    OXI_TRY(launch_check(h, "transpose back"));
    i64 dinfo = 0;
    int flag = 0;
    OXI_HIP(h, hipMemcpyAsync(&dinfo, g.dinfo, sizeof(i64), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipMemcpyAsync(&flag, g.flags, sizeof(int), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    if (flag) return oxi_set_error(h, OXI_ERR_TIMEOUT, "getrf: a panel barrier timed out");
    *info = dinfo;
    return OXI_OK;
}

}  // namespace

int oxi_getrf32(oxi_handle_t h, i64 m, i64 n, float* A, i64 lda, i64* ipiv, i64 nb, i64* info, int prec,
                bool tournament) {
    return getrf_t<float>(h, m, n, A, lda, ipiv, nb, info, prec, tournament);
}

extern "C" int oxi_dgetrf(oxi_handle_t h, int64_t m, int64_t n, double* A, int64_t lda, int64_t* ipiv, int64_t nb,
                          int64_t* info) {
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (lda < std::max<i64>(1, m)) return -5;
    if (nb < 0) return -7;
    if (!info) return -8;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = 0;
    if (m == 0 || n == 0) return OXI_OK;
    if (!A) return -4;
    if (!ipiv) return -6;
    return getrf_t<double>(h, m, n, A, lda, ipiv, nb, info, -1);
}

extern "C" int oxi_dgetrf_copy(oxi_handle_t h, int64_t m, int64_t n, const double* A, int64_t lda, double* F,
                               int64_t ldf, int64_t* ipiv, int64_t nb, int64_t* info) {
    if (!h) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (lda < std::max<i64>(1, m)) return -5;
    if (ldf < std::max<i64>(1, m)) return -7;
    if (nb < 0) return -9;
    if (!info) return -10;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = 0;
    if (m == 0 || n == 0) return OXI_OK;
    if (!A) return -4;
    if (!F) return -6;
    if (!ipiv) return -8;
    // the element ranges of A and F must be disjoint
    const uintptr_t a0 = reinterpret_cast<uintptr_t>(A), a1 = reinterpret_cast<uintptr_t>(A + (lda * (n - 1) + m));
    const uintptr_t f0 = reinterpret_cast<uintptr_t>(F), f1 = reinterpret_cast<uintptr_t>(F + (ldf * (n - 1) + m));
    if (a0 < f1 && f0 < a1) return -6;
    return getrf_t<double>(h, m, n, F, ldf, ipiv, nb, info, -1, false, A, lda);
}

static int sgetrf_entry(oxi_handle_t h, int prec, int64_t m, int64_t n, float* A, int64_t lda, int64_t* ipiv,
                        int64_t nb, int64_t* info, bool tournament) {
    if (!h) return -1;
    if (prec < OXI_PREC_BF16 || prec > OXI_PREC_FP32) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (lda < std::max<i64>(1, m)) return -6;
    if (nb < 0) return -8;
    if (!info) return -9;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    *info = 0;
    if (m == 0 || n == 0) return OXI_OK;
    if (!A) return -5;
    if (!ipiv) return -7;
    return getrf_t<float>(h, m, n, A, lda, ipiv, nb, info, prec == OXI_PREC_FP32 ? -1 : prec, tournament);
}

extern "C" int oxi_sgetrf_lowp(oxi_handle_t h, int prec, int64_t m, int64_t n, float* A, int64_t lda, int64_t* ipiv,
                               int64_t nb, int64_t* info) {
    return sgetrf_entry(h, prec, m, n, A, lda, ipiv, nb, info, false);
}

extern "C" int oxi_sgetrf_tournament(oxi_handle_t h, int prec, int64_t m, int64_t n, float* A, int64_t lda,
                                     int64_t* ipiv, int64_t nb, int64_t* info) {
    return sgetrf_entry(h, prec, m, n, A, lda, ipiv, nb, info, true);
}

// ---------------------------------------------------------------------------
// laswp without allocation: the pair list of one block of interchanges and a
// status word live in `scratch`, a file-scope device array (each GPU holds its
// own copy until the process ends).  Host code never passes `scratch` itself
// to a kernel: its device address comes from hipGetSymbolAddress, once per
// device.  Calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr int LASWP_SCRATCH_INTS = 4 * PERM_MAX_STEPS + 4 + 64;   // pairs of one block + status words
static __device__ __align__(256) int scratch[LASWP_SCRATCH_INTS];
constexpr int LASWP_STATUS = 4 * PERM_MAX_STEPS + 4;              // index of the status word

namespace {
std::mutex g_scratch_mu;
int* g_scratch_ptr[64] = {};    // per device: the device address of `scratch`

int scratch_ptr(oxi_handle_t h, int** out) {
    if (h->device < 0 || h->device >= 64) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "device index %d", h->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return oxi_set_error(h, OXI_ERR_HIP, "laswp scratch address: %s", hipGetErrorString(e));
        g_scratch_ptr[h->device] = (int*)p;
    }
    *out = g_scratch_ptr[h->device];
    return OXI_OK;
}

__global__ void clear_status_kernel(int* status) {
    if (threadIdx.x == 0) *status = 0;
}
// status = 1 when a pivot of steps k1..k2 (1-based) lies outside 1..lda
__global__ void check_pivots_kernel(const i64* __restrict__ ipiv, i64 k1, i64 k2, i64 lda, int* status) {
    for (i64 j = k1 - 1 + i64(blockIdx.x) * blockDim.x + threadIdx.x; j < k2; j += i64(gridDim.x) * blockDim.x) {
        const i64 p = ipiv[j];
        if (p < 1 || p > lda) atomicOr(status, 1);
    }
}
}  // namespace

extern "C" int oxi_dlaswp_noalloc(oxi_handle_t h, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                                  const int64_t* ipiv, int inverse) {
    if (!h) return -1;
    if (ncols < 0) return -2;
    if (lda < 1) return -4;
    if (k1 < 1) return -5;
    if (k2 < k1 - 1) return -6;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (ncols == 0 || k2 < k1) return OXI_OK;
    if (!A) return -3;
    if (!ipiv) return -7;
    if (k2 > lda) return -6;
    if (lda > i64(0x7fffffff) || ncols > i64(0x7fffffff))
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "laswp_noalloc: more than 2^31 - 1 rows or columns");
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    int* sp = nullptr;
    OXI_TRY(scratch_ptr(h, &sp));
    int* status = sp + LASWP_STATUS;
    // every row a step names lies inside the leading dimension (checked on the
    // device; one word comes back)
    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(64), 0, s, status);
    hipLaunchKernelGGL(check_pivots_kernel, dim3(unsigned(std::min<i64>(cdiv(k2 - k1 + 1, 256), 1024))), dim3(256), 0,
                       s, ipiv, k1, k2, lda, status);
    OXI_TRY(launch_check(h, "laswp pivot check"));
    int bad = 0;
    OXI_HIP(h, hipMemcpyAsync(&bad, status, sizeof(int), hipMemcpyDeviceToHost, s));
    OXI_HIP(h, hipStreamSynchronize(s));
    if (bad) return oxi_set_error(h, -7, "laswp: a pivot of steps %lld..%lld lies outside 1..lda", (long long)k1,
                                  (long long)k2);
    const i64 total = k2 - k1 + 1;
    const i64 nblocks = cdiv(total, PERM_MAX_STEPS);
    for (i64 bi = 0; bi < nblocks; ++bi) {
        const i64 b = inverse ? nblocks - 1 - bi : bi;
        const i64 s1 = (k1 - 1) + b * PERM_MAX_STEPS;
        const int cnt = int(std::min<i64>(PERM_MAX_STEPS, (k2 - 1) - s1 + 1));
        hipLaunchKernelGGL(perm_pairs_kernel, dim3(1), dim3(256), 0, s, ipiv, i64(0), s1, cnt, inverse, sp);
        const size_t shm = size_t(2 * cnt) * sizeof(double);
        hipLaunchKernelGGL(apply_pairs_cm_kernel, dim3(unsigned(ncols)), dim3(256), shm, s, A, lda, ncols, sp, i64(0));
        OXI_TRY(launch_check(h, "laswp"));
    }
    return OXI_OK;
}

extern "C" int oxi_dlaswp(oxi_handle_t h, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                          const int64_t* ipiv, int inverse) {
    if (!h) return -1;
    if (ncols < 0) return -2;
    if (lda < 1) return -4;
    if (k1 < 1) return -5;
    if (k2 < k1 - 1) return -6;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (ncols == 0 || k2 < k1) return OXI_OK;
    if (!A) return -3;
    if (!ipiv) return -7;
    if (k2 > lda) return -6;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    {
        // every row a step names must lie inside the leading dimension
        std::vector<i64> piv(size_t(k2 - k1 + 1));
        OXI_HIP(h, hipMemcpyAsync(piv.data(), ipiv + (k1 - 1), piv.size() * 8, hipMemcpyDeviceToHost, s));
        OXI_HIP(h, hipStreamSynchronize(s));
        for (size_t t = 0; t < piv.size(); ++t)
            if (piv[t] < 1 || piv[t] > lda)
                return oxi_set_error(h, -7, "laswp: ipiv[%lld] = %lld outside 1..lda", (long long)(k1 + i64(t)),
                                     (long long)piv[t]);
    }
    void* ws;
    OXI_TRY(oxi_workspace(h, size_t(4 * PERM_MAX_STEPS + 4) * 4 + 256, &ws));
    int* pairs = (int*)ws;
    // process in blocks of at most PERM_MAX_STEPS steps, in order (reverse for inverse)
    const i64 total = k2 - k1 + 1;
    const i64 nblocks = cdiv(total, PERM_MAX_STEPS);
    for (i64 bi = 0; bi < nblocks; ++bi) {
        const i64 b = inverse ? nblocks - 1 - bi : bi;
        const i64 s1 = (k1 - 1) + b * PERM_MAX_STEPS;
        const int cnt = int(std::min<i64>(PERM_MAX_STEPS, (k2 - 1) - s1 + 1));
        hipLaunchKernelGGL(perm_pairs_kernel, dim3(1), dim3(256), 0, s, ipiv, i64(0), s1, cnt, inverse, pairs);
        const size_t shm = size_t(2 * cnt) * sizeof(double);
        hipLaunchKernelGGL(apply_pairs_cm_kernel, dim3(unsigned(ncols)), dim3(256), shm, s, A, lda, ncols, pairs, i64(0));
        OXI_TRY(launch_check(h, "laswp"));
    }
    return OXI_OK;
}

