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
// per column the workgroups publish their candidate (key, row, and the whole
// row), meet at a device-scope counter barrier (release/acquire at agent
// scope, bounded spin with a timeout flag so every wave exits), and each reads
// all candidates to pick the same winner.  With P == 1 the exchange stays in LDS.
#include "oxi_internal.hpp"

#include <vector>
#include <algorithm>

namespace {

// ---------------------------------------------------------------------------
// Transposes
// ---------------------------------------------------------------------------
constexpr int TT = 32;  // transpose tile

// out (n x m, ldo) <- in^T, in is m x n (ldi)
__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ in, i64 ldi, double* __restrict__ out,
                                                        i64 ldo, i64 m, i64 n) {
    __shared__ double tile[TT][TT + 1];
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
__global__ __launch_bounds__(256) void transpose_square_kernel(double* __restrict__ A, i64 lda, i64 n, int tiles) {
    __shared__ double t1[TT][TT + 1], t2[TT][TT + 1];
    // linear block id -> (bi, bj) with bi <= bj
    i64 id = blockIdx.x;
    int bi = 0;
    i64 rowlen = tiles;
    while (id >= rowlen) { id -= rowlen; ++bi; --rowlen; }
    const int bj = bi + int(id);
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const i64 r0 = i64(bi) * TT, c0 = i64(bj) * TT;
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
// perm_pairs_kernel: one workgroup simulates the interchanges k1..k2 (0-based
// steps of ipiv, 1-based rows) on the affected rows and emits (dst, src) pairs
// of the net permutation: new[dst] = old[src].  Rows k1..k2 occupy slots
// 0..cnt-1; pivot rows beyond k2 get extra slots through an LDS hash table.
// ---------------------------------------------------------------------------
constexpr int PERM_MAX_STEPS = 512;
constexpr int HASH = 2048;

__global__ __launch_bounds__(256) void perm_pairs_kernel(const i64* __restrict__ ipiv, i64 poff, i64 k1, int cnt,
                                                         int inverse, int* __restrict__ pairs /*[2*2*cnt+1]*/) {
    __shared__ int slot_row[2 * PERM_MAX_STEPS];
    __shared__ int slot_src[2 * PERM_MAX_STEPS];
    __shared__ int hkey[HASH], hval[HASH];
    __shared__ int nslots;
    const int t = threadIdx.x;
    for (int i = t; i < HASH; i += blockDim.x) hkey[i] = -1;
    for (int i = t; i < cnt; i += blockDim.x) { slot_row[i] = int(k1 + i); slot_src[i] = int(k1 + i); }
    if (t == 0) nslots = cnt;
    __syncthreads();
    if (t == 0) {
        int ns = cnt;
        auto find = [&](int row) -> int {
            if (row >= k1 && row < k1 + cnt) return int(row - k1);
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
            const int j = int(k1 + step);
            const int p = int(ipiv[k1 + step] - 1 - poff);
            if (p == j) continue;
            const int a = find(j), b = find(p);
            const int tmp = slot_src[a]; slot_src[a] = slot_src[b]; slot_src[b] = tmp;
        }
        nslots = ns;
    }
    __syncthreads();
    // compact the pairs that move
    if (t == 0) {
        int np = 0;
        for (int s = 0; s < nslots; ++s)
            if (slot_src[s] != slot_row[s]) { pairs[1 + 2 * np] = slot_row[s]; pairs[2 + 2 * np] = slot_src[s]; ++np; }
        pairs[0] = np;
    }
}

// Apply pairs to columns of W: W[c][dst] <- W[c][src] for every c in
// [c0, c1) (W column = A row, W row = A column).  A workgroup stages 16 W rows
// x all pairs through LDS so the reads complete before any write.
constexpr int PAIR_ROWS = 16;
__global__ __launch_bounds__(256) void apply_pairs_w_kernel(double* __restrict__ Wm, i64 ldw, i64 c0, i64 c1,
                                                            const int* __restrict__ pairs, i64 roff) {
    extern __shared__ double buf[];  // [npairs][PAIR_ROWS]
    const int np = pairs[0];
    if (np == 0) return;
    const i64 cbase = c0 + i64(blockIdx.x) * PAIR_ROWS;
    const int t = threadIdx.x;
    const int r = t % PAIR_ROWS, q0 = t / PAIR_ROWS, qs = blockDim.x / PAIR_ROWS;
    const i64 c = cbase + r;
    const bool ok = c < c1;
    for (int q = q0; q < np; q += qs) {
        const i64 src = pairs[2 + 2 * q] + roff;
        buf[q * PAIR_ROWS + r] = ok ? Wm[c + src * ldw] : 0.0;
    }
    __syncthreads();
    if (ok)
        for (int q = q0; q < np; q += qs) {
            const i64 dst = pairs[1 + 2 * q] + roff;
            Wm[c + dst * ldw] = buf[q * PAIR_ROWS + r];
        }
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

__global__ void zero_u32_kernel(unsigned* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__global__ void zero_i64_kernel(i64* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__global__ void zero_i32_kernel(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

// ===========================================================================
// Host-side building blocks
// ===========================================================================
namespace {

struct GetrfCtx {
    oxi_handle_t h;
    hipStream_t s;
    double* Wm; i64 ldw;
    i64 m, n;
    i64* ipiv;
    i64* dinfo;
    unsigned* bar; unsigned bar_base;
    double* pub;
    int* pairs;
    int* flags;
};

int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

// Interchanges k1..k2 (0-based steps, inclusive; pivots are global 1-based rows)
// applied to A columns [c0, c1) = W rows [c0, c1).
int swap_w(GetrfCtx& g, i64 k1, i64 k2, i64 c0, i64 c1, int inverse) {
    if (k2 < k1 || c1 <= c0) return OXI_OK;
    const int cnt = int(k2 - k1 + 1);
    if (cnt > PERM_MAX_STEPS) return oxi_set_error(g.h, OXI_ERR_UNSUPPORTED, "laswp block of %d steps", cnt);
    hipLaunchKernelGGL(perm_pairs_kernel, dim3(1), dim3(256), 0, g.s, g.ipiv, i64(0), k1, cnt, inverse, g.pairs);
    OXI_TRY(launch_check(g.h, "perm_pairs"));
    const size_t shm = size_t(2 * cnt) * PAIR_ROWS * sizeof(double);
    hipLaunchKernelGGL(apply_pairs_w_kernel, dim3(unsigned(cdiv(c1 - c0, PAIR_ROWS))), dim3(256), shm, g.s, g.Wm, g.ldw,
                       c0, c1, g.pairs, i64(0));
    return launch_check(g.h, "apply_pairs_w");
}

int leaf(GetrfCtx& g, i64 c0, int w) {
    const i64 mrows = g.m - c0;
    // rows per workgroup: 256 (one per thread) up to 2048 rows, then 512
    // (two per thread, the most that stays in registers without spilling)
    const int kind = mrows > 2048 ? 1 : 0;
    const i64 R = kind == 0 ? 256 : 512;
    const i64 P = cdiv(mrows, R);
    if (P > std::min(g.h->num_cus, 256))
        return oxi_set_error(g.h, OXI_ERR_UNSUPPORTED, "panel of %lld rows exceeds the leaf kernel capacity",
                             (long long)mrows);
    LeafArgs a{g.Wm, g.ldw, c0, w, mrows, g.ipiv, g.dinfo, g.bar, g.bar_base, g.pub, int(P), g.flags};
    OXI_TRY(oxi_launch_lu_leaf(g.h, g.s, a, kind));
    if (P > 1) g.bar_base += unsigned(P) * unsigned(w);
    return OXI_OK;
}

}  // namespace

// getrf2 recursion on the panel A[c0:m, c0:c0+np] in W coordinates
// (src/lapack/factorizations.jl:114-152).
static int panel(GetrfCtx& g, i64 c0, i64 np) {
    if (np <= OXI_LEAF_W) return leaf(g, c0, int(np));
    const i64 mp = g.m - c0;
    i64 h = np / 2;
    OXI_TRY(panel(g, c0, h));
    // left half's interchanges on the right half columns
    OXI_TRY(swap_w(g, c0, c0 + h - 1, c0 + h, c0 + np, 0));
    // A12 <- L11^-1 A12 : W[c0+h : c0+np, c0 : c0+h] <- W[..] * (L11^T)^-1, L11^T = W[c0:c0+h, c0:c0+h] upper unit
    OXI_TRY(oxi_dtrsm_right(g.h, g.s, 1, 0, 1, np - h, h, g.Wm + c0 + c0 * g.ldw, g.ldw, g.Wm + (c0 + h) + c0 * g.ldw,
                             g.ldw));
    if (mp > h) {
        // A22 -= A21 * A12  <=>  W[c0+h:c0+np, c0+h:m] -= W[c0+h:c0+np, c0:c0+h] * W[c0:c0+h, c0+h:m]
        OXI_TRY(oxi_launch_dgemm(g.h, g.s, 0, 0, np - h, mp - h, h, -1.0, g.Wm + (c0 + h) + c0 * g.ldw, g.ldw,
                                 g.Wm + c0 + (c0 + h) * g.ldw, g.ldw, 1.0, g.Wm + (c0 + h) + (c0 + h) * g.ldw, g.ldw));
        OXI_TRY(panel(g, c0 + h, np - h));
        // right half's interchanges on the left half columns
        OXI_TRY(swap_w(g, c0 + h, c0 + std::min(mp, np) - 1, c0, c0 + h, 0));
    }
    return OXI_OK;
}

static int choose_nb(i64 m, i64 n) {
    const i64 q = std::min(m, n);
    if (q >= 8192) return 256;
    if (q >= 1024) return 128;
    return 64;
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
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
    const i64 q = std::min(m, n);
    if (nb == 0) nb = choose_nb(m, n);
    if (nb > PERM_MAX_STEPS) nb = PERM_MAX_STEPS;

    // workspace: [W if rectangular][pub][pairs][bar]
    const bool square = (m == n);
    const i64 ldw = square ? lda : n;
    const size_t wbytes = square ? 0 : size_t(n) * size_t(m) * 8;
    const size_t pub_doubles = size_t(OXI_LEAF_W) * (size_t(256) + 1) * OXI_PUB_STRIDE + size_t(OXI_LEAF_W) * OXI_PUB_STRIDE;
    const size_t off_pub = (wbytes + 255) & ~size_t(255);
    const size_t off_pairs = off_pub + ((pub_doubles * 8 + 255) & ~size_t(255));
    const size_t off_bar = off_pairs + ((size_t(4 * PERM_MAX_STEPS + 4) * 4 + 255) & ~size_t(255));
    const size_t total = off_bar + 256;
    void* ws;
    OXI_TRY(oxi_workspace(h, total, &ws));
    char* base = (char*)ws;
    GetrfCtx g;
    g.h = h; g.s = s;
    g.Wm = square ? A : (double*)base; g.ldw = ldw;
    g.m = m; g.n = n; g.ipiv = ipiv; g.dinfo = h->dev_info;
    g.pub = (double*)(base + off_pub);
    g.pairs = (int*)(base + off_pairs);
    g.bar = (unsigned*)(base + off_bar); g.bar_base = 0;
    g.flags = h->dev_flags;

    static bool attr_set = false;
    if (!attr_set) {
        // the pair-staging kernel may use up to 2 * OXI max steps * 16 doubles of LDS
        OXI_HIP(h, hipFuncSetAttribute((const void*)apply_pairs_w_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
    }
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, g.bar, 64);
    hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(64), 0, s, g.dinfo, 1);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, g.flags, 1);
    OXI_TRY(launch_check(h, "zero"));

    // A -> W
    if (square) {
        const int tiles = int(cdiv(n, TT));
        const i64 nblk = i64(tiles) * (tiles + 1) / 2;
        hipLaunchKernelGGL(transpose_square_kernel, dim3(unsigned(nblk)), dim3(256), 0, s, A, lda, n, tiles);
    } else {
        hipLaunchKernelGGL(transpose_kernel, dim3(unsigned(cdiv(m, TT)), unsigned(cdiv(n, TT))), dim3(256), 0, s, A,
                           lda, g.Wm, ldw, m, n);
    }
    OXI_TRY(launch_check(h, "transpose"));

    for (i64 j0 = 0; j0 < q; j0 += nb) {
        const i64 jb = std::min<i64>(nb, q - j0);
        const i64 j1 = j0 + jb;
        OXI_TRY(panel(g, j0, jb));
        OXI_TRY(swap_w(g, j0, j1 - 1, 0, j0, 0));
        OXI_TRY(swap_w(g, j0, j1 - 1, j1, n, 0));
        if (j1 >= n) continue;
        // U12 <- L11^-1 A12 : W[j1:n, j0:j1] <- W[j1:n, j0:j1] * (W[j0:j1, j0:j1])^-1 (upper unit)
        OXI_TRY(oxi_dtrsm_right(h, s, 1, 0, 1, n - j1, jb, g.Wm + j0 + j0 * ldw, ldw, g.Wm + j1 + j0 * ldw, ldw));
        if (j1 >= m) continue;
        // A22 -= L21 U12 : W[j1:n, j1:m] -= W[j1:n, j0:j1] * W[j0:j1, j1:m]
        hipEvent_t pe = nullptr;
        OXI_TRY(oxi_prof_begin(h, s, &pe));
        OXI_TRY(oxi_launch_dgemm(h, s, 0, 0, n - j1, m - j1, jb, -1.0, g.Wm + j1 + j0 * ldw, ldw, g.Wm + j0 + j1 * ldw,
                                 ldw, 1.0, g.Wm + j1 + j1 * ldw, ldw));
        OXI_TRY(oxi_prof_end(h, s, pe, 2.0 * double(n - j1) * double(m - j1) * double(jb)));
    }

    // W -> A
    if (square) {
        const int tiles = int(cdiv(n, TT));
        const i64 nblk = i64(tiles) * (tiles + 1) / 2;
        hipLaunchKernelGGL(transpose_square_kernel, dim3(unsigned(nblk)), dim3(256), 0, s, A, lda, n, tiles);
    } else {
        hipLaunchKernelGGL(transpose_kernel, dim3(unsigned(cdiv(n, TT)), unsigned(cdiv(m, TT))), dim3(256), 0, s, g.Wm,
                           ldw, A, lda, n, m);
    }
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

// ===========================================================================
// laswp, trsm and getrs entry points (column-major operands)
// ===========================================================================
extern "C" int oxi_dlaswp(oxi_handle_t h, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                          const int64_t* ipiv, int inverse) {
    if (!h) return -1;
    if (ncols < 0) return -2;
    if (lda < 1) return -4;
    if (k1 < 1) return -5;
    if (k2 < k1 - 1) return -6;
    std::lock_guard<std::recursive_mutex> lk(h->guard);
    if (ncols == 0 || k2 < k1) return OXI_OK;
    (void)hipSetDevice(h->device);
    hipStream_t s = h->stream;
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

