// Internal declarations shared by the HIP translation units of
// liboxiblas_mi355x.so.  Everything here is gfx950 (CDNA4) code.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/oxiblas_mi355x.h"

typedef int64_t i64;

// ---------------------------------------------------------------------------
// Handle: one device, one user stream, two internal streams for the panel /
// trailing-update overlap, and a grow-only device workspace.
// ---------------------------------------------------------------------------
constexpr int TRSM_INV_BLOCKS = 256;
struct oxi_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;       // caller's stream (null = default stream)
    hipStream_t side = nullptr;         // internal stream for look-ahead work
    // internal stream for getrf's trailing updates, kept off `panel_cus` CUs
    // (spread evenly over the XCDs) so that the panel running beside them
    // finds free CUs; null when no CU is kept free
    hipStream_t bulk = nullptr;
    // inverses of 64 x 64 diagonal blocks for tall right-side triangular
    // solves (trsm.hip), TRSM_INV_BLOCKS blocks of 64 x 64 doubles
    double* trsm_inv = nullptr;
    int panel_cus = 0;
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_c = nullptr;
    void* ws = nullptr;                 // device workspace (kernels' scratch)
    size_t ws_bytes = 0;
    void* ws2 = nullptr;                // second workspace for drivers that call the others
    size_t ws2_bytes = 0;
    void* ws3 = nullptr;                // workspace of the blocked left triangular solve
    size_t ws3_bytes = 0;
    int* dev_flags = nullptr;           // [0]: barrier timeout flag, [1..]: reserved
    i64* dev_info = nullptr;            // device-side info (first zero pivot)
    std::string last_error;
    std::recursive_mutex guard;         // one caller at a time per handle
    // profiling of the trailing-update GEMM inside getrf (optional): event
    // pairs recorded on the stream the GEMM runs on, resolved when read
    int prof_enabled = 0;
    struct ProfEvent { hipEvent_t a, b; double flops; };
    std::vector<ProfEvent> prof_events;
    double prof_gemm_ms = 0.0, prof_gemm_flops = 0.0;
    i64 prof_gemm_launches = 0;
    int num_cus = 256;
    unsigned long long getrf_seq = 0;   // getrf calls on this handle (tags the leaf exchange)
};

int oxi_set_error(oxi_handle_t h, int code, const char* fmt, ...);

#define OXI_HIP(h, expr)                                                                 \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return oxi_set_error((h), OXI_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                                 hipGetErrorString(_e), __FILE__, __LINE__);             \
    } while (0)

#define OXI_TRY(expr)              \
    do {                           \
        int _r = (expr);           \
        if (_r != OXI_OK) return _r; \
    } while (0)

// Grow the handle workspace to at least `bytes` (synchronises the streams
// before freeing the old block).
int oxi_workspace(oxi_handle_t h, size_t bytes, void** out);
// Second grow-only workspace, for drivers (mixed solve) whose buffers must
// survive the calls they make into the routines that use the first one.
int oxi_workspace2(oxi_handle_t h, size_t bytes, void** out);
// Third one, used only by the blocked left triangular solve (trsm.hip), which
// runs inside routines that hold the other two.
int oxi_workspace3(oxi_handle_t h, size_t bytes, void** out);

// ---------------------------------------------------------------------------
// Internal launchers (stream-explicit, no argument validation).
// ---------------------------------------------------------------------------
// C <- alpha*op(A)*op(B) + beta*C, FP64 MFMA; ta/tb: 0 = N, 1 = T.
int oxi_launch_dgemm(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                     const double* A, i64 lda, const double* B, i64 ldb, double beta, double* C, i64 ldc);
// C <- beta*C (beta == 0 writes zeros without reading C).
int oxi_launch_dscale(hipStream_t s, i64 m, i64 n, double beta, double* C, i64 ldc);
// A batch of equal-shaped products at fixed element strides (grid.y = product).
int oxi_launch_dgemm_strided(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                             const double* A, i64 lda, i64 sa, const double* B, i64 ldb, i64 sb, double beta,
                             double* C, i64 ldc, i64 sc, i64 batch);
// A batch of equal-shaped products given by device arrays of operand pointers.
int oxi_launch_dgemm_ptrs(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                          const double* const* A, i64 lda, const double* const* B, i64 ldb, double beta,
                          double* const* C, i64 ldc, i64 batch);
// C[triangle] <- alpha op(A) op(A)^T + beta C[triangle] (upper != 0: upper
// triangle; trans 0: A is n x k, 1: A is k x n); the other triangle untouched.
int oxi_launch_dsyrk(oxi_handle_t h, hipStream_t s, int upper, int trans, i64 n, i64 k, double alpha,
                     const double* A, i64 lda, double beta, double* C, i64 ldc);
// Cholesky of the lower triangle of A in place on stream s; *dinfo (device)
// receives 0 or the first failing column (1-based).  No synchronisation.
int oxi_potrf_device(oxi_handle_t h, hipStream_t s, i64 n, double* A, i64 lda, i64* dinfo);
// the same in FP32 (its trailing updates write both triangles)
int oxi_potrf_device(oxi_handle_t h, hipStream_t s, i64 n, float* A, i64 lda, i64* dinfo);
// whether oxi_launch_dgemm runs an m x n product on its 128 x 128 (16x16x4) kernel
bool oxi_dgemm_large_tiles(i64 m, i64 n);
// X[:, 0:64] <- X[:, 0:64] * V in place (V 64 x 64, column-major, ld 64)
int oxi_dgemm_right64_inplace(hipStream_t s, i64 rows, const double* V, double* X, i64 ldx);
// X[0:64, 0:ncols] <- op(V) X[0:64, 0:ncols] in place (V 64 x 64, ld 64; ta: V^T)
int oxi_dgemm_left64_inplace(hipStream_t s, int ta, i64 ncols, const double* V, double* X, i64 ldx);
// Left-side triangular solve, host-recursive with GEMM updates.
int oxi_dtrsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs,
                   const double* A, i64 lda, double* B, i64 ldb);
int oxi_dtrsm_right(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n,
                    const double* A, i64 lda, double* B, i64 ldb);
int oxi_strsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs,
                   const float* A, i64 lda, float* B, i64 ldb);
int oxi_strsm_right(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n,
                    const float* A, i64 lda, float* B, i64 ldb);
// FP32 GEMM (sgemm.hip)
int oxi_launch_sgemm(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, float alpha,
                     const float* A, i64 lda, const float* B, i64 ldb, float beta, float* C, i64 ldc);
// Low-precision trailing update (lowp_gemm.hip): prec 0 BF16, 1 FP16, 2 FP8 e4m3
int oxi_lowp_update(oxi_handle_t h, hipStream_t s, int prec, i64 M, i64 N, i64 K, float alpha, const float* XA,
                    i64 ldxa, const float* XB, i64 ldxb, float* C, i64 ldc, void* ws, size_t ws_bytes);
size_t oxi_lowp_update_ws(int prec, i64 M, i64 N, i64 K);
// 1-based permutation from LAPACK interchanges (host, synchronises s)
int oxi_perm_from_ipiv(oxi_handle_t h, hipStream_t s, const i64* ipiv, i64 n, std::vector<i64>& perm);
// getrs with a device permutation and an n x nrhs workspace Y (no synchronisation)
int oxi_getrs_perm_f64(oxi_handle_t h, hipStream_t s, int trans, i64 n, i64 nrhs, const double* LU, i64 lda,
                       const i64* dperm, double* B, i64 ldb, double* Y, void* ctrl);
int oxi_getrs_perm_f32(oxi_handle_t h, hipStream_t s, int trans, i64 n, i64 nrhs, const float* LU, i64 lda,
                       const i64* dperm, float* B, i64 ldb, float* Y, void* ctrl);
// Random butterfly transform and LU without pivoting (nopiv.hip).  R32 / R64:
// 4n butterfly entries (level 0 U, level 0 V, level 1 U, level 1 V).
int oxi_rbt_make(oxi_handle_t h, hipStream_t s, i64 n, float* R32, double* R64);
int oxi_rbt_matrix(oxi_handle_t h, hipStream_t s, i64 n, float* A, i64 lda, const float* R32);
// X <- U^T X (ut = 1) or X <- V X (ut = 0)
int oxi_rbt_vec(oxi_handle_t h, hipStream_t s, int ut, i64 n, i64 ncols, float* X, i64 ldx, const float* R32);
int oxi_rbt_vec(oxi_handle_t h, hipStream_t s, int ut, i64 n, i64 ncols, double* X, i64 ldx, const double* R64);
size_t oxi_getrf_nopiv_ws(int prec, i64 n);
int oxi_getrf_nopiv32(oxi_handle_t h, hipStream_t s, int prec, i64 n, float* A, i64 lda, i64* dinfo, void* ws,
                      size_t ws_bytes);
// FP32 getrf with a trailing update in prec (-1 FP32, 0 BF16, 1 FP16, 2 FP8)
// tournament: tournament-pivoting leaves where the panel allows them
int oxi_getrf32(oxi_handle_t h, i64 m, i64 n, float* A, i64 lda, i64* ipiv, i64 nb, i64* info, int prec,
                bool tournament = false);

// First zero on the diagonal of an n x n triangle (1-based, 0 if none);
// synchronises the stream.
// Single right-hand-side triangular solve against LU factors (trsv.hip).
// which: 0 = L unit forward, 1 = U backward, 2 = U^T forward, 3 = L^T unit
// backward.  ctrl is the trsv workspace below, prepared for these factors;
// the kernel's bounded waits record a timeout there, read by oxi_trsv_check
// (which synchronises).
int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const double* M, i64 ld, double* x, void* ctrl);
int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const float* M, i64 ld, float* x, void* ctrl);
int oxi_trsv(oxi_handle_t h, hipStream_t s, int which, i64 n, const float* M, i64 ld, double* x, void* ctrl);
int oxi_trsv_check(oxi_handle_t h, hipStream_t s, const void* ctrl);
// clears the status word of ctrl (once per solve, before its chain launches)
int oxi_trsv_begin(oxi_handle_t h, hipStream_t s, void* ctrl);
// The trsv workspace (ctrl above): control words and the inverses of the
// 128 x 128 diagonal blocks of the two triangles of a solve, in the vector
// type (tv bytes).  oxi_trsv_prepare fills them for the solve kind tk
// (0: L then U, 1: U^T then L^T) of the factors M; the vector pointer only
// selects the vector type.  Every later oxi_trsv on the same factors and kind
// uses them.
size_t oxi_trsv_ws_bytes(i64 n, size_t tv);
int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* M, i64 ld, double*, void* ws);
// the same (FP64) with the second triangle's part on the handle's side
// stream: its end is recorded in h->ev_b, which the second triangle's solve
// on s must wait for
int oxi_trsv_prepare_split(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* M, i64 ld, void* ws);
int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const float* M, i64 ld, float*, void* ws);
int oxi_trsv_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const float* M, i64 ld, double*, void* ws);
// LU solves with at most this many right-hand sides go column by column
// through oxi_trsv (when the caller supplies ctrl), more through trsm.
constexpr i64 OXI_TRSV_MAX_RHS = 4;
// FP64 LU solves with more right-hand sides, up to this many, go through
// the chained multi-RHS solve (trsm_chain.hip) when the caller supplies its
// workspace: control words and, for both triangles of the solve kind tk
// (0: L then U, 1: U^T then L^T, 2: a Cholesky factor's L (non-unit) then
// L^T), the inverses of the 64 x 64 diagonal
// blocks and each block's coupling to the block solved before it, filled
// once per factor and kind by oxi_trsm_chain_prepare.  oxi_trsm_chain_begin
// clears the status word once per solve; each oxi_trsm_chain solves one
// triangle (second 0: the first of the kind, 1: the second) in place in B;
// oxi_trsm_chain_check reads the status (synchronises).
constexpr i64 OXI_CHAIN_MAX_RHS = 64;
size_t oxi_trsm_chain_ws_bytes(i64 n);
// whether the chain takes this solve (else it returns OXI_ERR_UNSUPPORTED)
bool oxi_trsm_chain_fits(i64 n, i64 nrhs, i64 ldb);
int oxi_trsm_chain_prepare(oxi_handle_t h, hipStream_t s, int tk, i64 n, const double* LU, i64 lda, void* ws);
int oxi_trsm_chain_begin(oxi_handle_t h, hipStream_t s, void* ws);
int oxi_trsm_chain(oxi_handle_t h, hipStream_t s, int tk, int second, i64 n, i64 nrhs, const double* LU, i64 lda,
                   double* B, i64 ldb, void* ws);
int oxi_trsm_chain_check(oxi_handle_t h, hipStream_t s, const void* ws);

int oxi_first_zero_diag(oxi_handle_t h, hipStream_t s, const double* T, i64 ldt, i64 n, i64* where);

// ---------------------------------------------------------------------------
// Panel leaf of the LU (panel.hip)
// ---------------------------------------------------------------------------
constexpr int OXI_LEAF_W = 32;      // maximum leaf width
constexpr int OXI_PUB_STRIDE = 48;  // doubles per published (column, workgroup) record
constexpr int OXI_LEAF_MAX_P = 127; // workgroups of one leaf (LDS holds P + 1 records per column)

struct LeafArgs {
    void* Wm; i64 ldw;   // double* or float*
    i64 c0;          // first column (A) of the leaf = its first diagonal row
    int w;           // leaf width <= LEAF_W
    i64 mrows;       // rows c0 .. c0+mrows-1 of A
    i64* ipiv;       // global pivots (1-based rows), entries c0..c0+w-1 written
    i64* info;       // device info (first zero pivot, 1-based), 0 if none
    unsigned long long tag0;  // (getrf call sequence) << 32; column c's records carry tag0 + c
    double* pub;     // published records [w][P][PUB_STRIDE] + diag rows [w][PUB_STRIDE]
    int P;
    int* flags;      // flags[0] = 1 on barrier timeout
    i64 xlo, xhi;    // columns [xlo, xhi) of the enclosing panel: receive the leaf's interchanges
};

// kind 0: 256 threads x 1 row per workgroup; kind 1: 256 threads x 2 rows.
int oxi_launch_lu_leaf(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind);
// FP32 leaf with tournament pivoting (panel_tp.hip), kind 0 or 1, at most
// OXI_TP_MAX_P workgroups.
constexpr int OXI_TP_MAX_P = 32;
int oxi_launch_lu_leaf_tp(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind);
// FP32 leaf; kind 2: 256 threads x 4 rows.
int oxi_launch_lu_leaf32(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind);

// Bracket one trailing-update GEMM launch with profiling events (no-op when
// profiling is off); resolved by oxi_profile_read.
int oxi_prof_begin(oxi_handle_t h, hipStream_t s, hipEvent_t* a);
int oxi_prof_end(oxi_handle_t h, hipStream_t s, hipEvent_t a, double flops);

// ceil division
static inline i64 cdiv(i64 a, i64 b) { return (a + b - 1) / b; }
