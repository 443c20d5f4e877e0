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
struct oxi_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;       // caller's stream (null = default stream)
    hipStream_t side = nullptr;         // internal stream for look-ahead work
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    void* ws = nullptr;                 // device workspace
    size_t ws_bytes = 0;
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

// ---------------------------------------------------------------------------
// Internal launchers (stream-explicit, no argument validation).
// ---------------------------------------------------------------------------
// C <- alpha*op(A)*op(B) + beta*C, FP64 MFMA; ta/tb: 0 = N, 1 = T.
int oxi_launch_dgemm(oxi_handle_t h, hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, double alpha,
                     const double* A, i64 lda, const double* B, i64 ldb, double beta, double* C, i64 ldc);
// C <- beta*C (beta == 0 writes zeros without reading C).
int oxi_launch_dscale(hipStream_t s, i64 m, i64 n, double beta, double* C, i64 ldc);
// Left-side triangular solve, host-recursive with GEMM updates.
int oxi_dtrsm_left(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 nrhs,
                   const double* A, i64 lda, double* B, i64 ldb);
int oxi_dtrsm_right(oxi_handle_t h, hipStream_t s, int upper, int trans, int unit, i64 m, i64 n,
                    const double* A, i64 lda, double* B, i64 ldb);

// First zero on the diagonal of an n x n triangle (1-based, 0 if none);
// synchronises the stream.
int oxi_first_zero_diag(oxi_handle_t h, hipStream_t s, const double* T, i64 ldt, i64 n, i64* where);

// ---------------------------------------------------------------------------
// Panel leaf of the LU (panel.hip)
// ---------------------------------------------------------------------------
constexpr int OXI_LEAF_W = 32;      // maximum leaf width
constexpr int OXI_PUB_STRIDE = 48;  // doubles per published (column, workgroup) record

struct LeafArgs {
    double* Wm; i64 ldw;
    i64 c0;          // first column (A) of the leaf = its first diagonal row
    int w;           // leaf width <= LEAF_W
    i64 mrows;       // rows c0 .. c0+mrows-1 of A
    i64* ipiv;       // global pivots (1-based rows), entries c0..c0+w-1 written
    i64* info;       // device info (first zero pivot, 1-based), 0 if none
    unsigned* bar;   // barrier counter (monotonic over one getrf call)
    unsigned base;   // counter value at launch
    double* pub;     // published records [w][P][PUB_STRIDE] + diag rows [w][PUB_STRIDE]
    int P;
    int* flags;      // flags[0] = 1 on barrier timeout
};

// kind 0: 256 threads x 1 row per workgroup; kind 1: 256 threads x 2 rows.
int oxi_launch_lu_leaf(oxi_handle_t h, hipStream_t s, const LeafArgs& a, int kind);

// Bracket one trailing-update GEMM launch with profiling events (no-op when
// profiling is off); resolved by oxi_profile_read.
int oxi_prof_begin(oxi_handle_t h, hipStream_t s, hipEvent_t* a);
int oxi_prof_end(oxi_handle_t h, hipStream_t s, hipEvent_t a, double flops);

// ceil division
static inline i64 cdiv(i64 a, i64 b) { return (a + b - 1) / b; }
