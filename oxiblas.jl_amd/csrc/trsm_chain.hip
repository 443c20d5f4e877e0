// Triangular solves of the LU factors with several right-hand sides (up to
// OXI_CHAIN_MAX_RHS): the left TRSMs of ldiv!(F, B) and
// ldiv!(transpose(F), B) with B n x k (src/api/linalg_api.jl:168-184,
// trsm! of src/blas/level3.jl:309-544 for each triangle), as one launch per
// triangle on the matrix cores.
//
// The n rows are cut into blocks of CB = 64 rows in solve order (L and U^T:
// top down, U and L^T: bottom up); block b needs X of every earlier block.
// Within a block the rows stay in their physical order.
//
// Preparation (once per factor, oxi_trsm_chain_prepare): for both triangles
// the inverse V_b of every diagonal block and its couplings to the two
// blocks solved just before it, G_b = V_b T_{b,b-1} and H_b = V_b T_{b,b-2}.
// Then
//     X_b = V_b (B_b - sum_{c < b-2} T_{b,c} X_c) - H_b X_{b-2} - G_b X_{b-1},
// so only the product G_b X_{b-1} waits for block b-1, and H_b X_{b-2} for
// block b-2; the sum over the earlier blocks and its product with V_b are
// formed while those are still being solved.  (With only G_b, the last
// block of the sum and the V_b product fell on the chain: measured with
// clock stamps, most workgroups reached their wait for block b-1 after it
// had published.)  (Inverting the diagonal blocks changes the rounding
// of a substitution, not the result beyond it; the solve is deterministic.)
//
// The solve: a workgroup per block and group of CR = 32 right-hand sides
// (the groups are independent chains, one `ready` counter each).
// Workgroups take their (block, group) from an atomic ticket, so a
// workgroup only ever waits on workgroups that are already running.  Each
// of its four waves owns 16 rows of the block and both 16-column tiles of
// the group: every product is 2 x 16 v_mfma_f64_16x16x4 per wave.  The
// operands: the factor's blocks straight from memory (coalesced along the
// rows), X of earlier blocks staged in LDS.  X goes between workgroups with
// device-scope stores and loads; the publishing waves wait for their stores
// to be acknowledged before `ready` moves.  Every wait is bounded: a
// timeout sets a sticky error word, read by oxi_trsm_chain_check.
#include "oxi_internal.hpp"

namespace {

constexpr int CB = 64;       // rows per block
constexpr int CR = 32;       // right-hand sides per workgroup
constexpr int CNT = 256;     // threads per workgroup
constexpr int XBC = 1;       // published blocks of X applied per round
constexpr int CHAINS = OXI_CHAIN_MAX_RHS / CR;
constexpr unsigned CSPIN_LIMIT = 1u << 20;
static_assert(CNT == 4 * 64 && CB == 4 * 16 && CR == 2 * 16, "four waves of 16 rows, two 16-column tiles");
static_assert(CB * CR == 8 * CNT, "eight X elements per thread per staged block");

// Control words: the ticket and the chains' counters (per launch), and the
// error word, sticky across the launches of one solve.
struct ChainCtrl { unsigned ticket, err, ready[CHAINS]; };

typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma16(double a, double b, d4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void put_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First physical row of block b (rows of it) in solve order.
__device__ __forceinline__ i64 block_row0(bool bwd, i64 n, int b, int rows) {
    return bwd ? n - i64(b) * CB - rows : i64(b) * CB;
}

// V_b: the inverse of diagonal block b of op(T) (T or, with TRANS, its
// transpose), V[b][k * CB + i] = V_b(i, k) in the block's physical row
// order, zero outside its rows.  Thread j solves
// T_bb x = e_j by substitution (forward for the lower triangle, backward for
// the upper one) with the block and x in LDS.
template <bool UPPER, bool UNIT, bool TRANS>
__global__ __launch_bounds__(CB) void chain_inv_kernel(const double* __restrict__ M, i64 ld, i64 n,
                                                       double* __restrict__ V) {
    __shared__ double Ts[CB][CB + 1];   // Ts[r][c] = T_bb(r, c)
    __shared__ double Xs[CB][CB + 1];   // Xs[r][j]: column j of the inverse
    const int j = threadIdx.x, b = blockIdx.x;
    const int rows = int(min(i64(CB), n - i64(b) * CB));
    const i64 p0 = block_row0(UPPER, n, b, rows);
    for (int c = 0; c < rows; ++c)
        Ts[j][c] = j < rows ? (TRANS ? M[(p0 + c) + (p0 + j) * ld] : M[(p0 + j) + (p0 + c) * ld]) : 0.0;
    __syncthreads();
    for (int r = 0; r < CB; ++r) Xs[r][j] = 0.0;
    if (j < rows) {
        if (!UPPER) {
            Xs[j][j] = UNIT ? 1.0 : 1.0 / Ts[j][j];
            for (int r = j + 1; r < rows; ++r) {
                double s = 0.0;
                for (int m = j; m < r; ++m) s = __builtin_fma(Ts[r][m], Xs[m][j], s);
                Xs[r][j] = UNIT ? -s : -s / Ts[r][r];
            }
        } else {
            Xs[j][j] = UNIT ? 1.0 : 1.0 / Ts[j][j];
            for (int r = j - 1; r >= 0; --r) {
                double s = 0.0;
                for (int m = r + 1; m <= j; ++m) s = __builtin_fma(Ts[r][m], Xs[m][j], s);
                Xs[r][j] = UNIT ? -s : -s / Ts[r][r];
            }
        }
    }
    __syncthreads();
    double* out = V + size_t(b) * CB * CB;
    for (int c = 0; c < CB; ++c) out[c * CB + j] = Xs[j][c];
}

// G_b = V_b T_{b,b-d} for b >= d (d = 1: G_b, d = 2: H_b): G[b][q * CB + i]
// over physical rows i of block b and physical rows q of block b-d (a full
// block).
template <bool BWD, bool TRANS>
__global__ __launch_bounds__(CNT) void chain_g_kernel(const double* __restrict__ M, i64 ld, i64 n, int d,
                                                      const double* __restrict__ V, double* __restrict__ G) {
    __shared__ double Tb[CB][CB + 1];   // Tb[k][q] = op(T)(row k of block b, row q of block b-d)
    const int t = threadIdx.x, b = blockIdx.x + d;
    const int rows = int(min(i64(CB), n - i64(b) * CB));
    const i64 p0 = block_row0(BWD, n, b, rows), q0 = block_row0(BWD, n, b - d, CB);
    for (int e = t; e < CB * CB; e += CNT) {
        const int k = e % CB, q = e / CB;
        Tb[k][q] = k < rows ? (TRANS ? M[(q0 + q) + (p0 + k) * ld] : M[(p0 + k) + (q0 + q) * ld]) : 0.0;
    }
    __syncthreads();
    const double* vb = V + size_t(b) * CB * CB;
    const int i = t % CB, qg = t / CB;
    constexpr int QPT = CB / (CNT / CB);
    double acc[QPT];
#pragma unroll
    for (int u = 0; u < QPT; ++u) acc[u] = 0.0;
    for (int k = 0; k < rows; ++k) {
        const double v = vb[k * CB + i];
#pragma unroll
        for (int u = 0; u < QPT; ++u) acc[u] = __builtin_fma(v, Tb[k][qg * QPT + u], acc[u]);
    }
    double* g = G + size_t(b) * CB * CB;
#pragma unroll
    for (int u = 0; u < QPT; ++u) g[(qg * QPT + u) * CB + i] = acc[u];
}

__global__ void chain_ctrl_reset(ChainCtrl* c) {
    if (threadIdx.x == 0) {
        c->ticket = 0;
        for (int h = 0; h < CHAINS; ++h) c->ready[h] = 0;
    }
}
__global__ void chain_status_clear(ChainCtrl* c) {
    if (threadIdx.x == 0) {
        c->ticket = 0;
        c->err = 0;
        for (int h = 0; h < CHAINS; ++h) c->ready[h] = 0;
    }
}

// X of other workgroups is read with device-scope (sc1) buffer loads: the
// compiler keeps several in flight and waits only where a value is used.
// The resource starts at column j0 of B and covers 2^31 - 1 bytes; the
// offsets stay below that (checked on the host).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t x_rsrc(const double* B, i64 ldb, int j0) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)(B + i64(j0) * ldb), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ double load_sc1(__amdgpu_buffer_rsrc_t r, unsigned off_bytes) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off_bytes, 0, 16));
}

// Thread t's eight elements of an X block (rows q0 + t % CB, columns
// t / CB + 4r of the group, clamped into the group's ncol columns): loads
// only; the caller writes them to LDS.
__device__ __forceinline__ void load_x(__amdgpu_buffer_rsrc_t r, i64 ldb, i64 q0, int ncol, int t, double (&v)[8]) {
    const int q = t % CB, jq = t / CB;
#pragma unroll
    for (int u = 0; u < 8; ++u)
        v[u] = load_sc1(r, unsigned((q0 + q + i64(min(jq + 4 * u, ncol - 1)) * ldb) * 8));
}
__device__ __forceinline__ void store_x(double (*xs)[CR + 1], int ncol, int t, const double (&v)[8]) {
    const int q = t % CB, jq = t / CB;
#pragma unroll
    for (int u = 0; u < 8; ++u) xs[q][jq + 4 * u] = jq + 4 * u < ncol ? v[u] : 0.0;
}

// Wait until the chain's counter exceeds `need` (thread 0 polls, the block
// then reads it); returns the counter, or -1 after a timeout or another
// workgroup's failure (the same in every thread).
__device__ __forceinline__ int wait_ready(ChainCtrl* ctrl, int h, int need, int t, int* s_ready) {
    if (t == 0) {
        unsigned v, spins = 0;
        while ((v = __hip_atomic_load(&ctrl->ready[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) <= unsigned(need)) {
            if (__hip_atomic_load(&ctrl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > CSPIN_LIMIT) {
                __hip_atomic_store(&ctrl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *s_ready = v > unsigned(need) ? int(v) : -1;
    }
    __syncthreads();
    const int r = *s_ready;
    __syncthreads();
    return r;
}

template <bool BWD, bool TRANS>
__global__ __launch_bounds__(CNT, 2) void trsm_chain_kernel(const double* __restrict__ M, i64 ld, i64 n, double* B,
                                                         i64 ldb, int nrhs, int nchains, ChainCtrl* ctrl,
                                                         const double* __restrict__ V,
                                                         const double* __restrict__ G,
                                                         const double* __restrict__ H) {
    __shared__ double xs[2][CB][CR + 1];     // X of the blocks being applied, physical rows
    double (*rs)[CR + 1] = xs[1];            // B_b - sum, physical rows (after the blocks)
    __shared__ int s_tk, s_ready;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) s_tk = int(atomicAdd(&ctrl->ticket, 1u));
    __syncthreads();
    const int b = s_tk / nchains, h = s_tk % nchains;
    const int j0 = h * CR, ncol = min(CR, nrhs - j0);
    const int rows = int(min(i64(CB), n - i64(b) * CB));
    const i64 p0 = block_row0(BWD, n, b, rows);
    // MFMA operands: A[i][k] at lane 16k + i, B[k][j] at lane 16k + j;
    // accumulator r of lane 16a + j holds D[4r + a][j]
    const int fi = lane & 15, fk = lane >> 4;
    const int ra = wave * 16 + fi;               // this lane's A-operand row in the block
    const bool ra_ok = ra < rows;
    d4v acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};   // sum over chunks, per column tile
    d4v xb[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};    // X_b
    const __amdgpu_buffer_rsrc_t xr = x_rsrc(B, ldb, j0);
    // everything the block needs besides X of other blocks, loaded first so
    // that no memory round trip is left between the last published block
    // and this block's own X: its rows of B, and its operands of V_b, H_b
    // and G_b (read-only here)
    double bb[2][4], av[16], ah[16], ag[16];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = wave * 16 + 4 * r + fk, col = ct * 16 + fi;
            bb[ct][r] = row < rows && col < ncol ? B[(p0 + row) + i64(j0 + col) * ldb] : 0.0;
        }
    {
        const double* vb = V + size_t(b) * CB * CB;
        const double* hb = H + size_t(b) * CB * CB;
        const double* gb = G + size_t(b) * CB * CB;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            av[s] = vb[(4 * s + fk) * CB + ra];
            ah[s] = b > 1 ? -hb[(4 * s + fk) * CB + ra] : 0.0;
            ag[s] = b > 0 ? -gb[(4 * s + fk) * CB + ra] : 0.0;
        }
    }

    // 1. sum over the blocks c < b - 2 as they are published, up to XBC
    // published blocks per round, their X and factor blocks loaded together
    // (more than one per round measured no faster: the workgroups keep up)
    int avail = 0;
    const int cend = b - 2;
    for (int c = 0; c < cend;) {
        if (c >= avail) {
            avail = wait_ready(ctrl, h, c, t, &s_ready);
            if (avail < 0) return;
        }
        const int nbat = min(min(avail, cend) - c, XBC);
        double ak[XBC][16];
        {
            double v[XBC][8];
#pragma unroll
            for (int k = 0; k < XBC; ++k) {
                if (k < nbat) {
                    const i64 q0 = block_row0(BWD, n, c + k, CB);
                    load_x(xr, ldb, q0, ncol, t, v[k]);
#pragma unroll
                    for (int s = 0; s < 16; ++s)
                        ak[k][s] = !ra_ok ? 0.0
                                          : TRANS ? M[(q0 + 4 * s + fk) + (p0 + ra) * ld]
                                                  : M[(p0 + ra) + (q0 + 4 * s + fk) * ld];
                }
            }
#pragma unroll
            for (int k = 0; k < XBC; ++k)
                if (k < nbat) store_x(xs[k], ncol, t, v[k]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < XBC; ++k) {
            if (k < nbat) {
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    acc[0] = mfma16(ak[k][s], xs[k][4 * s + fk][fi], acc[0]);
                    acc[1] = mfma16(ak[k][s], xs[k][4 * s + fk][16 + fi], acc[1]);
                }
            }
        }
        __syncthreads();   // xs is refilled next
        c += nbat;
    }

    // 2. rs = B_b - sum, then X_b = V_b rs
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = wave * 16 + 4 * r + fk, col = ct * 16 + fi;
            rs[row][col] = row < rows && col < ncol ? bb[ct][r] - acc[ct][r] : 0.0;
        }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        xb[0] = mfma16(av[s], rs[4 * s + fk][fi], xb[0]);
        xb[1] = mfma16(av[s], rs[4 * s + fk][16 + fi], xb[1]);
    }

    // 3. X_b -= H_b X_{b-2}, then X_b -= G_b X_{b-1}: the products that
    // wait for the two blocks solved last (their operands loaded first;
    // X_{b-2} to xs[0], X_{b-1} to xs[1], which the V_b product has left)
    if (b > 1) {
        if (b - 2 >= avail) {
            avail = wait_ready(ctrl, h, b - 2, t, &s_ready);
            if (avail < 0) return;
        }
        {
            double v[8];
            load_x(xr, ldb, block_row0(BWD, n, b - 2, CB), ncol, t, v);
            store_x(xs[0], ncol, t, v);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            xb[0] = mfma16(ah[s], xs[0][4 * s + fk][fi], xb[0]);
            xb[1] = mfma16(ah[s], xs[0][4 * s + fk][16 + fi], xb[1]);
        }
    }
    if (b > 0) {
        if (b - 1 >= avail) {
            avail = wait_ready(ctrl, h, b - 1, t, &s_ready);
            if (avail < 0) return;
        } else {
            __syncthreads();   // every wave has finished the V_b product: xs[1] is free
        }
        {
            double v[8];
            load_x(xr, ldb, block_row0(BWD, n, b - 1, CB), ncol, t, v);
            store_x(xs[1], ncol, t, v);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            xb[0] = mfma16(ag[s], xs[1][4 * s + fk][fi], xb[0]);
            xb[1] = mfma16(ag[s], xs[1][4 * s + fk][16 + fi], xb[1]);
        }
    }

    // 4. X_b to B, acknowledged, then the chain's counter
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = wave * 16 + 4 * r + fk, col = ct * 16 + fi;
            if (row < rows && col < ncol) put_agent(&B[(p0 + row) + i64(j0 + col) * ldb], xb[ct][r]);
        }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores acknowledged
    __syncthreads();                 // every wave's
    if (t == 0) __hip_atomic_store(&ctrl->ready[h], unsigned(b + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

size_t slot_doubles(i64 n) { return size_t(cdiv(n, CB)) * CB * CB; }
double* slot_ptr(void* ws, i64 n, int k) { return (double*)((char*)ws + 256) + size_t(k) * slot_doubles(n); }

int launch_check(oxi_handle_t h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return oxi_set_error(h, OXI_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return OXI_OK;
}

}  // namespace

bool oxi_trsm_chain_fits(i64 n, i64 nrhs, i64 ldb) {
    // grid size, and the device-scope loads' 32-bit byte offsets
    return nrhs <= OXI_CHAIN_MAX_RHS && cdiv(n, CB) * cdiv(nrhs, CR) <= 0x7fffffff &&
           (n + i64(CR) * ldb) * 8 < (i64(1) << 31) - 1;
}

size_t oxi_trsm_chain_ws_bytes(i64 n) { return 256 + 6 * slot_doubles(std::max<i64>(n, 1)) * sizeof(double); }

int oxi_trsm_chain_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* LU, i64 lda, void* ws) {
    if (n <= 0) return OXI_OK;
    const unsigned nb = unsigned(cdiv(n, CB));
    // slots: V of the first and second triangle, then their G, then their H
    double *v0 = slot_ptr(ws, n, 0), *v1 = slot_ptr(ws, n, 1), *g0 = slot_ptr(ws, n, 2), *g1 = slot_ptr(ws, n, 3);
    double *h0 = slot_ptr(ws, n, 4), *h1 = slot_ptr(ws, n, 5);
    if (tk == 0) {          // L (unit, forward), then U (backward)
        hipLaunchKernelGGL((chain_inv_kernel<false, true, false>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v0);
        hipLaunchKernelGGL((chain_inv_kernel<true, false, false>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v1);
    } else if (tk == 1) {   // U^T (forward), then L^T (unit, backward)
        hipLaunchKernelGGL((chain_inv_kernel<false, false, true>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v0);
        hipLaunchKernelGGL((chain_inv_kernel<true, true, true>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v1);
    } else {                // Cholesky: L (forward), then L^T (backward)
        hipLaunchKernelGGL((chain_inv_kernel<false, false, false>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v0);
        hipLaunchKernelGGL((chain_inv_kernel<true, false, true>), dim3(nb), dim3(CB), 0, s, LU, lda, n, v1);
    }
    // the first triangle is op(T) = T except for U^T, the second T^T except for U
    for (int d = 1; d <= 2; ++d) {
        if (nb <= d) break;
        double* c0 = d == 1 ? g0 : h0;
        double* c1 = d == 1 ? g1 : h1;
        const dim3 grid(nb - unsigned(d));
        if (tk == 1)
            hipLaunchKernelGGL((chain_g_kernel<false, true>), grid, dim3(CNT), 0, s, LU, lda, n, d, v0, c0);
        else
            hipLaunchKernelGGL((chain_g_kernel<false, false>), grid, dim3(CNT), 0, s, LU, lda, n, d, v0, c0);
        if (tk == 0)
            hipLaunchKernelGGL((chain_g_kernel<true, false>), grid, dim3(CNT), 0, s, LU, lda, n, d, v1, c1);
        else
            hipLaunchKernelGGL((chain_g_kernel<true, true>), grid, dim3(CNT), 0, s, LU, lda, n, d, v1, c1);
    }
    return launch_check(h, "trsm chain preparation");
}

int oxi_trsm_chain_begin(oxi_handle_t h, hipStream_t s, void* ws) {
    hipLaunchKernelGGL(chain_status_clear, dim3(1), dim3(64), 0, s, static_cast<ChainCtrl*>(ws));
    return launch_check(h, "trsm chain status clear");
}

int oxi_trsm_chain(oxi_handle_t h, hipStream_t s, int tk, int second, i64 n, i64 nrhs, const double* LU, i64 lda,
                   double* B, i64 ldb, void* ws) {
    if (n <= 0 || nrhs <= 0) return OXI_OK;
    if (nrhs > OXI_CHAIN_MAX_RHS) return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "trsm chain: too many right-hand sides");
    const i64 nb = cdiv(n, CB);
    const int nchains = int(cdiv(nrhs, CR));
    if (!oxi_trsm_chain_fits(n, nrhs, ldb))
        return oxi_set_error(h, OXI_ERR_UNSUPPORTED, "trsm chain: n or the leading dimension too large");
    ChainCtrl* c = static_cast<ChainCtrl*>(ws);
    const double* V = slot_ptr(ws, n, second ? 1 : 0);
    const double* G = slot_ptr(ws, n, second ? 3 : 2);
    const double* H = slot_ptr(ws, n, second ? 5 : 4);
    hipLaunchKernelGGL(chain_ctrl_reset, dim3(1), dim3(64), 0, s, c);
    const dim3 grid(unsigned(nb * nchains)), block(CNT);
    // the first triangle runs forward, the second backward; op(T) as in the
    // preparation
    const bool trans = second ? tk != 0 : tk == 1;
    if (!trans && !second)
        hipLaunchKernelGGL((trsm_chain_kernel<false, false>), grid, block, 0, s, LU, lda, n, B, ldb, int(nrhs), nchains,
                           c, V, G, H);
    else if (!trans)
        hipLaunchKernelGGL((trsm_chain_kernel<true, false>), grid, block, 0, s, LU, lda, n, B, ldb, int(nrhs), nchains,
                           c, V, G, H);
    else if (!second)
        hipLaunchKernelGGL((trsm_chain_kernel<false, true>), grid, block, 0, s, LU, lda, n, B, ldb, int(nrhs), nchains,
                           c, V, G, H);
    else
        hipLaunchKernelGGL((trsm_chain_kernel<true, true>), grid, block, 0, s, LU, lda, n, B, ldb, int(nrhs), nchains,
                           c, V, G, H);
    return launch_check(h, "trsm chain");
}

int oxi_trsm_chain_check(oxi_handle_t h, hipStream_t s, const void* ws) {
    unsigned err = 0;
    OXI_HIP(h, hipMemcpyAsync(&err, &static_cast<const ChainCtrl*>(ws)->err, sizeof(unsigned), hipMemcpyDeviceToHost,
                              s));
    OXI_HIP(h, hipStreamSynchronize(s));
    if (err) return oxi_set_error(h, OXI_ERR_TIMEOUT, "trsm chain: block chain timed out");
    return OXI_OK;
}
