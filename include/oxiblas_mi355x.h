/*
 * oxiblas_mi355x.h -- C ABI of liboxiblas_mi355x.so, the MI355X (gfx950)
 * implementation of OxiBLAS.jl's level-3 hot path: gemm!, the blocked LU with
 * partial pivoting (lu / lu! / lu_factor) and its solves (ldiv / F \ b,
 * transposed solves), the mixed-precision LU with iterative refinement and
 * the pieces the adjoints of x = A\B and gemm! need.
 *
 * Conventions (those of the reference's Julia API):
 *   - matrices are column-major FP64 (FP32 for the low-precision factors) in
 *     device memory, with a leading dimension >= max(1, rows);
 *   - pivots are int64, 1-based global row indices in LAPACK ipiv form, as
 *     `LUFactor.pivots` (src/lapack/factorizations.jl:5-10,118);
 *   - trans characters are 'N', 'T', 'C' ('C' equals 'T' for real data),
 *     as `_trans_kind` (src/blas/gemm.jl:156-161);
 *   - return value: OXI_OK (0); -i when argument i is invalid (nothing is
 *     written, as the reference validates before any mutation,
 *     src/blas/gemm.jl:673-677); a positive OXI_ERR_* code for a runtime
 *     failure.  A singular pivot is not an error of the call: it is reported
 *     through `info` (> 0, the 1-based column), as `LUFactor.info`
 *     (src/lapack/factorizations.jl:125-129); the binding turns it into
 *     SingularException(info) when check=true (src/api/linalg_api.jl:163-167).
 *   - every call is enqueued on the handle's stream; calls that return host
 *     values (info, iteration counts, errors) synchronise that stream.
 *   - a handle serialises its callers (the GemmWorkspace lock contract,
 *     src/blas/gemm.jl:37-44); separate handles run independently.
 */
#ifndef OXIBLAS_MI355X_H
#define OXIBLAS_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oxi_handle_s* oxi_handle_t;

enum {
    OXI_OK = 0,
    OXI_ERR_HIP = 1,          /* a HIP runtime call failed */
    OXI_ERR_ALLOC = 2,        /* device allocation failed */
    OXI_ERR_UNSUPPORTED = 3,  /* a size beyond what the kernels support */
    OXI_ERR_SINGULAR = 4,     /* zero on a non-unit triangular diagonal (trsm/trsv/getrs) */
    OXI_ERR_TIMEOUT = 5,      /* a device-side barrier gave up (never expected) */
    OXI_ERR_NOT_FINITE = 6,   /* non-finite input where finite input is required */
    OXI_ERR_OVERFLOW = 7,     /* low-precision conversion overflowed */
    OXI_ERR_NOT_POSDEF = 8    /* not positive definite (PosDefException): the message names the column */
};

/* ---- handle, stream and memory ----------------------------------------- */
/* Replaces the per-type GemmWorkspace pool (src/blas/gemm.jl:107-153): the
 * handle owns the device workspace and the streams of one device. */
int oxi_create(int device, oxi_handle_t* handle);
int oxi_destroy(oxi_handle_t handle);
/* Use `stream` (a hipStream_t; NULL = default stream) for every call. */
int oxi_set_stream(oxi_handle_t handle, void* stream);
int oxi_sync(oxi_handle_t handle);
/* Text of the last error recorded on this handle. */
const char* oxi_last_error(oxi_handle_t handle);
const char* oxi_version(void);
int oxi_malloc(oxi_handle_t handle, void** ptr, uint64_t bytes);
int oxi_free(oxi_handle_t handle, void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous. */
int oxi_memcpy(oxi_handle_t handle, void* dst, const void* src, uint64_t bytes, int kind);

/* ---- level 3 ----------------------------------------------------------- */
/* gemm!(C, A, B, alpha, beta; transA, transB)  (src/blas/gemm.jl:625-703):
 * C <- alpha*op(A)*op(B) + beta*C.  m==0||n==0 returns at once; alpha==0 or
 * k==0 scales C only and never reads A or B; beta==0 never reads C (NaN in C
 * is overwritten); an A or B overlapping C is snapshotted before C is written
 * (src/blas/gemm.jl:682-688).  FP64 MFMA (v_mfma_f64_16x16x4). */
int oxi_dgemm(oxi_handle_t handle, char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
              const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

/* gemm_strided_batched!(C, A, B, alpha, beta; transA, transB)
 * (src/blas/batched.jl:80-100): `batch` independent products
 * C_z <- alpha*op(A_z)*op(B_z) + beta*C_z with A_z = A + z*strideA (elements),
 * B_z and C_z likewise, in one launch (grid.y = z).  Outputs must be disjoint
 * (strideC >= the span of one C; -17 otherwise, batched.jl:104-117); an A or B
 * overlapping the outputs is copied first (:87-88).  The per-product rules
 * of oxi_dgemm apply (beta == 0 never reads C). */
int oxi_dgemm_strided_batched(oxi_handle_t handle, char transA, char transB, int64_t m, int64_t n, int64_t k,
                              double alpha, const double* A, int64_t lda, int64_t strideA, const double* B,
                              int64_t ldb, int64_t strideB, double beta, double* C, int64_t ldc, int64_t strideC,
                              int64_t batch);

/* gemm_batched!(Cs, As, Bs, alpha, beta; transA, transB)
 * (src/blas/batched.jl:49-78) for a batch of one shape: A, B, C are HOST
 * arrays of `batch` device pointers.  Outputs must be pairwise disjoint
 * (-13 otherwise, :66-68); an operand overlapping any output is copied first
 * (:58-65).  One launch (grid.y = product); synchronises the stream.  A batch
 * of several shapes is one call per shape. */
int oxi_dgemm_batched(oxi_handle_t handle, char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* const* A, int64_t lda, const double* const* B, int64_t ldb, double beta,
                      double* const* C, int64_t ldc, int64_t batch);

/* syrk!(C, A, alpha, beta; uplo, trans) and herk! for Float64
 * (src/blas/level3.jl:546-783): C[tri] <- alpha*op(A)*op(A)^T + beta*C[tri] on
 * the triangle `uplo` ('U'/'L') only, the other triangle untouched; trans 'N':
 * A is n x k, 'T'/'C': A is k x n.  alpha == 0 or k == 0 scales the triangle;
 * beta == 0 never reads C (_triangle_scale!, :546-553).  FP64 MFMA; tiles
 * outside the triangle do no work.  The other triangle keeps its values;
 * with beta != 0 its elements inside the 128 x 128 tiles on the diagonal
 * (n >= about 2700) are read and stored back unchanged, so it must not be
 * written concurrently by other work. */
int oxi_dsyrk(oxi_handle_t handle, char uplo, char trans, int64_t n, int64_t k, double alpha, const double* A,
              int64_t lda, double beta, double* C, int64_t ldc);

/* trsm!(B, A, alpha; side, uplo, trans, unit)  (src/blas/level3.jl:518-544):
 * B <- alpha*op(A)^-1*B (side 'L') or alpha*B*op(A)^-1 (side 'R').  A zero on a
 * non-unit diagonal returns OXI_ERR_SINGULAR before B is touched
 * (src/blas/level3.jl:476-483); *singular_index (may be NULL) gets its
 * 1-based position. */
int oxi_dtrsm(oxi_handle_t handle, char side, char uplo, char trans, char diag, int64_t m, int64_t n, double alpha,
              const double* A, int64_t lda, double* B, int64_t ldb, int64_t* singular_index);

/* _lu_swap! / laswp (src/lapack/factorizations.jl:80-88) and _rowswap!
 * (src/api/linalg_api.jl:100-110): interchanges k1..k2 (1-based, inclusive)
 * of ipiv on columns 0..ncols-1 of A, in reverse order when inverse != 0. */
int oxi_dlaswp(oxi_handle_t handle, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
               const int64_t* ipiv, int inverse);

/* oxi_dlaswp without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as oxi_dlaswp; the pair list of each block of at most
 * 512 interchanges and a status word live in a fixed 8.5 KB device array of
 * the library (one per GPU, held until the process ends) instead of the
 * handle workspace, so the call allocates nothing (rows and columns below
 * 2^31).  Calls that overlap on one GPU (different handles or streams) share
 * that array and race.  Synchronises the handle's stream once, for the
 * pivot range check. */
int oxi_dlaswp_noalloc(oxi_handle_t handle, int64_t ncols, double* A, int64_t lda, int64_t k1, int64_t k2,
                       const int64_t* ipiv, int inverse);

/* ---- LU with partial pivoting ------------------------------------------- */
/* lu!(A) / _lu_factor!(F; blocksize, check)  (src/lapack/factorizations.jl:30-77,
 * src/api/linalg_api.jl:218): A (m x n, device) is overwritten by the unit
 * lower and upper factors; ipiv (device, min(m,n) entries) receives the
 * 1-based global pivot rows; *info (host) receives 0 or the first zero pivot
 * column (1-based).  nb is the panel width (the reference's blocksize; 0 picks
 * the tuned default).  Pivots are the first row of largest magnitude, as
 * _lu_pivot (src/lapack/factorizations.jl:89-107). */
int oxi_dgetrf(oxi_handle_t handle, int64_t m, int64_t n, double* A, int64_t lda, int64_t* ipiv, int64_t nb,
               int64_t* info);

/* lu(A) = lu!(copy(A))  (src/api/linalg_api.jl:217): the same factorization
 * with A untouched and the factors written to F (m x n, ldf, device), which
 * must not overlap A (-6).  The copy is the factorization's own change of
 * layout, so it costs no separate pass over the matrix. */
int oxi_dgetrf_copy(oxi_handle_t handle, int64_t m, int64_t n, const double* A, int64_t lda, double* F,
                    int64_t ldf, int64_t* ipiv, int64_t nb, int64_t* info);

/* ldiv!(F, B) and ldiv!(transpose(F), B)  (src/api/linalg_api.jl:168-184):
 * solves op(A)*X = B in place with the factors of oxi_dgetrf.  trans 'N':
 * row interchanges, unit-lower then upper solve; 'T'/'C': upper-transposed,
 * unit-lower-transposed solve, then the inverse interchanges (the transposed
 * solve used by the adjoint B-bar = A^-T X-bar, src/lapack/expert_refinement.jl:2-23).
 * A zero on U's diagonal returns OXI_ERR_SINGULAR with B untouched. */
int oxi_dgetrs(oxi_handle_t handle, char trans, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
               const int64_t* ipiv, double* B, int64_t ldb);

/* ---- what is built on the LU factors -------------------------------------- */
/* LinearAlgebra.logabsdet(F::LU) / logabsdet_factor (src/api/linalg_api.jl:185-191,
 * src/lapack/dense_utilities.jl:177-187) on the factors of oxi_dgetrf:
 * *logabs = sum log|u_ii| (-Inf with a zero pivot), *sign = the product of the
 * pivots' signs, negated per interchange (0 with a zero pivot); det =
 * sign * exp(logabs).  One device reduction; synchronises. */
int oxi_dlogabsdet(oxi_handle_t handle, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv,
                   double* logabs, double* sign);

/* native_inverse(A) from its LU factors (src/lapack/dense_utilities.jl:14-36):
 * Ainv <- inv(U) inv(L) P, i.e. U^-1 by a left triangular solve, times L^-1 by
 * a right solve with the unit lower factor, then the column permutation.
 * Ainv must not overlap the factors (-6).  A zero on U's diagonal returns
 * OXI_ERR_SINGULAR (SingularException).  Uses n*n doubles of handle workspace. */
int oxi_dgetri(oxi_handle_t handle, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv, double* Ainv,
               int64_t ldainv);

/* backward_error(A, x, b)  (src/lapack/dense_utilities.jl:214-228): R = B - A X
 * (A m x n, X n x nrhs, B m x nrhs) and *error = max_ij |r_ij| / (|A||X| + |B|)_ij
 * (0/0 = 0, r/0 = Inf, a non-finite denominator Inf; NaN propagates).  R may be
 * NULL (the residual is then not returned). */
int oxi_dbackward_error(oxi_handle_t handle, int64_t m, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                        const double* X, int64_t ldx, const double* B, int64_t ldb, double* R, int64_t ldr,
                        double* error);

/* inverse_norm_estimate(F, T, n; maxiter)  (src/lapack/expert_refinement.jl:45-60):
 * the Hager-Higham estimate of ||A^-1||_1 from solves with the factors of
 * oxi_dgetrf and their transpose (an estimate, not a bound).  maxiter >= 1. */
int oxi_dinverse_norm_estimate(oxi_handle_t handle, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv,
                               int32_t maxiter, double* estimate);

typedef struct {
    int64_t iterations;             /* refinement steps taken */
    int32_t converged;              /* backward_error <= rtol */
    int32_t reserved;
    double backward_error;          /* componentwise, as _refinement_residual! */
    double residual_norm;           /* nrm2 of the last residual */
    double rcond_estimate;          /* min(1, 1 / (||A||_1 * estimate of ||A^-1||_1)); refine_solve only */
    double forward_error_estimate;  /* delta / (1 - delta), delta = est ||A^-1||_1 ||R||_1 / ||X||_1; refine_solve only */
} oxi_refine_info;

/* refine_solve(A, F, B; maxiter, rtol)  (src/lapack/expert_refinement.jl:89-104):
 * X = solve(F, B), then X += solve(F, B - A X) until the componentwise backward
 * error is <= rtol or maxiter steps; then the condition and forward-error
 * estimates.  F are the factors (LU, ipiv) of oxi_dgetrf of A.  The reference's
 * defaults are maxiter = 8, rtol = 8 eps. */
int oxi_drefine_solve(oxi_handle_t handle, int64_t n, int64_t nrhs, const double* A, int64_t lda, const double* LU,
                      int64_t ldlu, const int64_t* ipiv, const double* B, int64_t ldb, double* X, int64_t ldx,
                      int32_t maxiter, double rtol, oxi_refine_info* info);

/* expert_solve(A, B; maxiter, rtol, equilibrate)
 * (src/lapack/advanced_factorizations.jl:591-645): with equilibrate != 0, rows
 * are divided by their largest |entry| (a zero row returns OXI_ERR_SINGULAR,
 * SingularException(i)) and then columns by theirs (1 when zero); the scaled
 * matrix is factored into LU/ipiv (caller storage, the NamedTuple's `factor`),
 * X = solve(F, B ./ rows) ./ cols, and X is refined against the original A, B
 * with corrections solve(F, R ./ rows) ./ cols.  rows, cols: device arrays of
 * n doubles (the returned scales).  Defaults of the reference: maxiter = 8,
 * rtol = 8 eps, equilibrate = true. */
int oxi_dexpert_solve(oxi_handle_t handle, int64_t n, int64_t nrhs, const double* A, int64_t lda, const double* B,
                      int64_t ldb, double* X, int64_t ldx, int32_t maxiter, double rtol, int32_t equilibrate,
                      double* LU, int64_t ldlu, int64_t* ipiv, double* rows, double* cols, oxi_refine_info* info);

/* ---- Cholesky ----------------------------------------------------------- */
/* _cholesky_factor!(L) / cholesky_factor(A)  (src/lapack/factorizations.jl:206-286):
 * A (n x n, device) is overwritten by its lower Cholesky factor L (A = L L^T)
 * read from the lower triangle only; uplo must be 'L' (the reference factors
 * the lower triangle).  *info (host) receives 0, or the 1-based column of the
 * first non-positive or non-finite pivot (PosDefException(info) in the
 * binding; the factor is then partial).  On success the strict upper
 * triangle is zeroed (:256-260).  Recursive potrf2: 64-column leaves in LDS,
 * right-side TRSM and lower SYRK on the matrix cores. */
int oxi_dpotrf(oxi_handle_t handle, char uplo, int64_t n, double* A, int64_t lda, int64_t* info);

/* solve(F::CholeskyFactor, B)  (src/lapack/factorizations.jl:288-300):
 * B <- (L L^T)^-1 B in place with the factor of oxi_dpotrf.  A zero on L's
 * diagonal returns OXI_ERR_SINGULAR with B untouched. */
int oxi_dpotrs(oxi_handle_t handle, char uplo, int64_t n, int64_t nrhs, const double* L, int64_t ldl, double* B,
               int64_t ldb);

/* Solve plans: ldiv!(F, B) repeated on the same factors (the LU object keeps
 * its factors resident, src/api/linalg_api.jl:168-184).  A plan holds what
 * every solve with these factors prepares -- the row permutation on the
 * device, the first zero pivot, and the inverses of the 128 x 128 diagonal
 * blocks of both triangles (4 * n * 128 doubles for 'N' and 'T' solves, each
 * kind made by its first solve with at most 4 right-hand sides) -- so later
 * solves only run the solve.  LU and ipiv must stay unchanged while the plan
 * lives; a plan is used by one handle at a time; oxi_dgetrs_plan_destroy
 * frees it.  oxi_dgetrs_plan returns what oxi_dgetrs returns
 * (OXI_ERR_SINGULAR for a zero pivot). */
typedef struct oxi_getrs_plan_s* oxi_getrs_plan_t;
int oxi_dgetrs_plan_create(oxi_handle_t handle, int64_t n, const double* LU, int64_t lda, const int64_t* ipiv,
                           oxi_getrs_plan_t* plan);
int oxi_dgetrs_plan(oxi_handle_t handle, oxi_getrs_plan_t plan, char trans, int64_t nrhs, double* B, int64_t ldb);
int oxi_dgetrs_plan_destroy(oxi_handle_t handle, oxi_getrs_plan_t plan);

/* ---- adjoints ------------------------------------------------------------ */
/* Pullback of X = A \ B given the factors of oxi_dgetrf and the cotangent
 * Xbar: Bbar = A^-T Xbar (the transposed solve ldiv!(transpose(F), .),
 * src/api/linalg_api.jl:176-184, src/lapack/expert_refinement.jl:2-23) and,
 * when Abar is not NULL, Abar = -Bbar X^T (FP64 MFMA, rank nrhs).  The
 * reference has no AD rules; this is the ChainRules rrule of `\` built from
 * its transposed solve and gemm!.  The gemm! pullback (Abar = Cbar op(B)^T,
 * Bbar = op(A)^T Cbar) is two oxi_dgemm calls.  A zero on U's diagonal returns
 * OXI_ERR_SINGULAR. */
int oxi_dgesv_adjoint(oxi_handle_t handle, int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                      const int64_t* ipiv, const double* X, int64_t ldx, const double* Xbar, int64_t ldxbar,
                      double* Bbar, int64_t ldbbar, double* Abar, int64_t ldabar);

/* ---- mixed precision ---------------------------------------------------- */
/* Precision of the trailing products of the FP32 factorization. */
enum { OXI_PREC_BF16 = 0, OXI_PREC_FP16 = 1, OXI_PREC_FP8 = 2, OXI_PREC_FP32 = 3 };

/* gemm! for Float32 (src/blas/gemm.jl:625-703), FP32 MFMA. */
int oxi_sgemm(oxi_handle_t handle, char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
              const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc);

/* C (FP32) <- C + alpha * round(A) * round(B): the low-precision trailing
 * product on its own; A (m x k) and B (k x n) are FP32 and rounded to `prec`
 * (FP8 e4m3 with one power-of-two scale per 256-row block of A and per
 * 256-column block of B), products accumulated in FP32.  No reference
 * counterpart: the building block of the mixed form below. */
int oxi_gemm_lowp_update(oxi_handle_t handle, int prec, int64_t m, int64_t n, int64_t k, float alpha,
                         const float* A, int64_t lda, const float* B, int64_t ldb, float* C, int64_t ldc);

/* gemm! for Float16  (src/blas/extended_blas.jl:39-88): A, B and C hold IEEE
 * binary16 values (column-major); each element of op(A)*op(B) accumulates in
 * Float32 over all of K (FP16 MFMA, FP32 accumulation) and is rounded to
 * Float16 once: C = Float16(a*acc + b*Float32(C)), a and b being alpha and beta
 * rounded to Float16 first.  beta == 0 never reads C; alpha == 0 or k == 0
 * scales C only.  A or B overlapping C is read before C is written. */
int oxi_hgemm(oxi_handle_t handle, char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
              const void* A, int64_t lda, const void* B, int64_t ldb, float beta, void* C, int64_t ldc);

/* FP32 blocked LU with partial pivoting whose trailing updates run in `prec`
 * (the low-precision factor of mixed_solve, src/lapack/extended_lapack.jl:183-198,
 * with BF16/FP16/FP8 trailing products on top of the reference's Float32
 * factor).  Same conventions as oxi_dgetrf. */
int oxi_sgetrf_lowp(oxi_handle_t handle, int prec, int64_t m, int64_t n, float* A, int64_t lda, int64_t* ipiv,
                    int64_t nb, int64_t* info);

/* oxi_sgetrf_lowp with tournament pivoting (no reference counterpart; the
 * factor of oxi_dgesv_mixed with pivot 2): each 32-column leaf's pivots are
 * the partial-pivoting pivots of the candidate rows that partial pivoting
 * picks inside each block of 256 rows (512 beyond 8192 rows), so one exchange
 * per leaf replaces one per column; every row is then eliminated with those
 * pivot rows by the operations of partial pivoting.  Ties go to the row of
 * lowest position, NaN never wins.  Panels of more than 16384 rows pivot
 * partially.  Same arguments, info and errors as oxi_sgetrf_lowp. */
int oxi_sgetrf_tournament(oxi_handle_t handle, int prec, int64_t m, int64_t n, float* A, int64_t lda, int64_t* ipiv,
                          int64_t nb, int64_t* info);

/* FP32 solve with the factors of oxi_sgetrf_lowp (as oxi_dgetrs). */
int oxi_sgetrs(oxi_handle_t handle, char trans, int64_t n, int64_t nrhs, const float* LU, int64_t lda,
               const int64_t* ipiv, float* B, int64_t ldb);

typedef struct {
    int64_t iterations;          /* refinement steps (residual evaluations - 1) */
    int64_t inner_iterations;    /* GMRES iterations over all refinement steps */
    int32_t converged;           /* backward_error <= rtol */
    int32_t fallback;            /* 1: the low-precision factor failed, FP64 factor used */
    int32_t reason;              /* 0 converged, 1 iteration limit, 2 non-finite, 3 working-precision fallback */
    int32_t solver;              /* 0 IR, 1 GMRES-IR (the one used) */
    double backward_error;       /* componentwise max_i |r_i| / (|A||x| + |b|)_i, as _refinement_residual! */
    double normwise_backward_error; /* max|r| / (||A||_inf max|x| + max|b|) over all columns */
    double residual_norm;        /* nrm2 of the last residual B - A X (the reference's `residual`) */
    double factor_ms;            /* factorization time (device) */
    double refine_ms;            /* first solve + refinement time (device) */
} oxi_ir_info;

/* mixed_solve(A, B; maxiter, rtol)  (src/lapack/factorizations.jl:419-427,
 * src/lapack/extended_lapack.jl:114-226): radix equilibration, a low-precision
 * LU (oxi_sgetrf_lowp with `prec`), and refinement in FP64 on the GPU until the
 * componentwise backward error is <= rtol: solver 0 = classical iterative
 * refinement (as the reference), 1 = GMRES-IR (left-preconditioned GMRES on
 * the correction equation, one cycle of at most gmres_restart iterations per
 * refinement step and right-hand side, the preconditioner being the low
 * factors applied in FP64 arithmetic; for large condition numbers),
 * 2 = automatic (GMRES-IR once classical refinement stops halving the error);
 * gmres_restart 0 picks 30.  The residual,
 * the correction accumulation and X stay FP64.  A (n x n) is not modified.
 * A factorization that fails in low precision falls back to FP64
 * (extended_lapack.jl:195-211).  Returns OXI_ERR_NOT_FINITE for non-finite
 * input (:182).
 * pivot 0: the low-precision LU pivots partially, as the reference.  pivot 1
 * (no reference counterpart): A_eq is first transformed by two-level random
 * butterflies, U^T A_eq V, and factored without pivoting, which needs no
 * device-wide pivot search per column; each correction applies U^T before and
 * V after the triangular solves.  It applies to prec = OXI_PREC_FP32 (factors
 * rounded further lose too much to multipliers that pivoting no longer bounds);
 * n must be a multiple of 4 and the factor free of zero pivots.  Otherwise the
 * partially pivoted factor is used.  pivot 2 (no reference counterpart): the
 * low-precision LU chooses each 32-column leaf's pivots by a tournament over
 * its row blocks (communication-avoiding LU, Grigori, Demmel and Xiang 2011):
 * one exchange between the workgroups of a leaf instead of one per column;
 * panels of more than 16384 rows pivot partially. */
int oxi_dgesv_mixed(oxi_handle_t handle, int prec, int solver, int64_t n, int64_t nrhs, const double* A,
                    int64_t lda, const double* B, int64_t ldb, double* X, int64_t ldx, int32_t maxiter, double rtol,
                    int32_t gmres_restart, int32_t pivot, oxi_ir_info* info);

/* mixed_solve(A, B; factorization=:cholesky)  (src/lapack/extended_lapack.jl:175-226,
 * _mixed_factor :161): A (symmetric positive definite, lower triangle read by
 * the factorization, all of A by the residual) divided by its largest |entry|
 * (rows = that scale, cols = 1, :187-189), the FP32 Cholesky of the lower
 * triangle (oxi_dpotrf's recursion in FP32), falling back to FP64 when the
 * FP32 factor meets a non-positive pivot, then the reference's correction
 * step with L and L^T solves and classical refinement in FP64 until the
 * componentwise backward error is <= rtol.  prec must be OXI_PREC_FP32 (the
 * reference's low type).  OXI_ERR_NOT_POSDEF when the FP64 factor fails too
 * (PosDefException), OXI_ERR_NOT_FINITE for non-finite input. */
int oxi_dposv_mixed(oxi_handle_t handle, int prec, int64_t n, int64_t nrhs, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double* X, int64_t ldx, int32_t maxiter, double rtol,
                    oxi_ir_info* info);

/* equilibrate(A; radix)  (src/lapack/dense_utilities.jl:191-212): row scales
 * rows[i] = max_j |A[i,j]| (1 for a zero row), then column scales of A ./ rows,
 * and E = (A ./ rows) ./ cols (E may be A).  radix != 0 rounds each scale to
 * the power of two 2^round(log2(v)) (ties to even), clamped to the normal
 * range, so that the scaling is exact.  A non-finite entry returns
 * OXI_ERR_NOT_FINITE before E is written.  rows, cols: device arrays of m and
 * n doubles. */
int oxi_dequilibrate(oxi_handle_t handle, int64_t m, int64_t n, const double* A, int64_t lda, double* E,
                     int64_t lde, double* rows, double* cols, int radix);

/* ---- profiling ---------------------------------------------------------- */
/* When enabled, oxi_dgetrf brackets every trailing-update GEMM launch with
 * HIP events on the stream it runs on and accumulates its time and flops. */
int oxi_profile_enable(oxi_handle_t handle, int enable);
int oxi_profile_read(oxi_handle_t handle, double* gemm_ms, double* gemm_flops, int64_t* gemm_launches);

#ifdef __cplusplus
}
#endif

#endif /* OXIBLAS_MI355X_H */
