/*
 * kblas_batch_ext.h -- entry points with no counterpart in the reference.
 *
 * The reference has no non-uniform POTRF/POTRS/POSV and no mixed-precision
 * POSV (its non-uniform support stops at MAGMA-backed TRSM/SYRK/GEMM,
 * src/batch_triangular/Xtrsm_batch_drivers.cuh:277-367).  These follow the
 * reference's own `_vbatch` naming for non-uniform batched routines
 * (include/batch_qr.h:272, include/batch_transpose.h:33) and LAPACK dsposv's
 * ITER convention for the mixed-precision solver.
 *
 * Non-uniform ("vbatch") arguments: n, m, lda, ldb, ldx are DEVICE int arrays
 * of length batchCount; A, B, X are DEVICE arrays of device pointers; max_n /
 * max_m are host hints (0 = unknown).  The whole batch runs in one persistent
 * launch that pulls matrices from a size-sorted queue built on the device; the
 * queue lives in the handle's device workspace (query, allocate, call).
 */
#ifndef KBLAS_BATCH_EXT_H_
#define KBLAS_BATCH_EXT_H_

#include "kblas.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- non-uniform Cholesky family ---------------------------------------- */
void kblas_potrf_vbatch_wsquery(kblasHandle_t handle, int max_n, int batchCount);
void kblas_posv_vbatch_wsquery(kblasHandle_t handle, int max_n, int batchCount);

int kblasDpotrf_vbatch(kblasHandle_t handle, char uplo, int* n, int max_n,
                       double** A, int* lda, int batchCount, int* info_array);
int kblasSpotrf_vbatch(kblasHandle_t handle, char uplo, int* n, int max_n,
                       float** A, int* lda, int batchCount, int* info_array);
/* B[i] is m[i] x n[i] as in the uniform calls: side 'R' solves X*A = B with A
 * of order n[i] (m[i] right-hand sides, the reference's layout); side 'L'
 * solves A*X = B with A of order m[i] (n[i] right-hand sides, LAPACK's).     */
int kblasDpotrs_vbatch(kblasHandle_t handle, char side, char uplo, int* m, int* n,
                       int max_m, int max_n, double** A, int* lda, double** B, int* ldb,
                       int batchCount);
int kblasDposv_vbatch(kblasHandle_t handle, char side, char uplo, int* m, int* n,
                      int max_m, int max_n, double** A, int* lda, double** B, int* ldb,
                      int batchCount, int* info_array);
int kblasSposv_vbatch(kblasHandle_t handle, char side, char uplo, int* m, int* n,
                      int max_m, int max_n, float** A, int* lda, float** B, int* ldb,
                      int batchCount, int* info_array);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): kblasDposv_vbatch with its size-sorted queue kept in a fixed
 * device-resident scratch area of the library (one per GPU, held until the
 * process ends) instead of the handle's workspace, so it needs no workspace
 * query, allocates nothing and performs no host synchronisation.  At most
 * 1,048,576 matrices per call (KBLAS_InsufficientWorkspace above).  Calls
 * that overlap on one GPU (two streams or two host threads) share the area and
 * race; serialise them.  Same arguments and results as kblasDposv_vbatch.     */
int kblasDposv_vbatch_noalloc(kblasHandle_t handle, char side, char uplo, int* m, int* n,
                              int max_m, int max_n, double** A, int* lda, double** B, int* ldb,
                              int batchCount, int* info_array);

/* ---- mixed precision POSV (LAPACK dsposv semantics, batched) -------------
 * FP32 Cholesky and triangular sweeps, FP64 residual refinement until
 * ||r||inf <= ||x||inf * ||A||inf * eps64 * sqrt(n) (at most 30 steps), FP64
 * Cholesky fallback when the FP32 pivot is not positive or refinement stalls.
 * A is read-only unless the fallback ran (then A holds L, as in dsposv).
 * B is read-only; X receives the solution (X may equal B with ldx == ldb).
 * iter[i] >= 0: refinement steps used; -2: an entry of A, B or a residual
 * beyond FP32 range; -3: FP32 factor failed; -31: no convergence (all three
 * then solved in FP64); -1: order above 64, solved in FP64 directly (dsposv's
 * "full precision for implementation reasons").  info[i] as in dposv.
 * side / m / n exactly as kblasDposv_batch_strided.                           */
int kblasDsposv_batch_strided(kblasHandle_t handle, char side, char uplo,
                              const int m, const int n,
                              double* A, int lda, long strideA,
                              const double* B, int ldb, long strideB,
                              double* X, int ldx, long strideX,
                              int batchCount, int* info_array, int* iter_array);
int kblasDsposv_batch(kblasHandle_t handle, char side, char uplo,
                      const int m, const int n,
                      double** A_array, int lda, const double** B_array, int ldb,
                      double** X_array, int ldx,
                      int batchCount, int* info_array, int* iter_array);
/* Non-uniform: orders up to 64 refine in registers; orders above 64 refine
 * in the workgroup phase (FP32 blocked factor and solves, FP64 residuals on
 * the matrix cores) when the workspace was sized with
 * kblas_sposv_vbatch_wsquery(handle, max_nrhs, max_order, batchCount) and
 * max_m / max_n bound the sizes (orders up to 1024, 64 right-hand sides);
 * otherwise, or beyond those bounds, they take dsposv's full-precision path
 * (iter = -1).  kblas_posv_vbatch_wsquery's queue alone is the minimum.      */
void kblas_sposv_vbatch_wsquery(kblasHandle_t handle, int max_nrhs, int max_order, int batchCount);
int kblasDsposv_vbatch(kblasHandle_t handle, char side, char uplo, int* m, int* n,
                       int max_m, int max_n, double** A, int* lda, const double** B, int* ldb,
                       double** X, int* ldx, int batchCount, int* info_array, int* iter_array);

/* ---- kernel choice per size bucket (no reference counterpart: it replaces
 * the reference's hard-coded launch-shape tables, Xpotrf_batch_drivers.cuh:45-61,
 * Xtrsm_batch_drivers.cuh:72-83, and its compile-time tuning array,
 * src/kblas_struct.h:36-39) ---------------------------------------------
 * A process-wide table of (precision 'd'/'s', operation, order range) ->
 * kernel variant, written by kblas-gpu_amd/tune.py from measurements on the
 * device.  It is loaded on first use from $KBLAS_TUNING_FILE, else from
 * tuning.json next to the library; without an entry the built-in rule runs.
 * Operations: "potrf", "posv", "potrs", "sposv" (uniform batches) and
 * "vsplit" (non-uniform batches: variant = the order, 0..64, at and below
 * which a matrix is served by one wavefront instead of a workgroup).
 * Variants: "rg16" (n <= 16) and "rg32" (n <= 32): four matrices per
 * wavefront in registers; "rg4": POTRF, 32 < n <= 64, one matrix per 16-lane
 * row; "tile": n <= 64, one matrix per wavefront (matrix-core tiles);
 * "reg": sposv, n <= 32; "blk": a 256-thread workgroup per matrix, any n.
 * kblas_tuning_set appends an entry (later entries win) and rejects a variant
 * that cannot serve every order of its range; kblas_tuning_query returns the
 * variant index (0 = built-in rule; for "vsplit" the order, or -1).  Not
 * thread-safe: change the table before concurrent calls.                    */
int kblas_tuning_load(const char* path);
int kblas_tuning_set(char prec, const char* op, int n_lo, int n_hi, const char* variant);
void kblas_tuning_clear(void);
int kblas_tuning_query(char prec, const char* op, int n);
const char* kblas_tuning_variant_name(int variant);

#ifdef __cplusplus
}  /* extern "C" */

inline int kblas_potrf_batch(kblasHandle_t h, char uplo, int* n, int max_n, double** A, int* lda,
                             int bc, int* info)
{ return kblasDpotrf_vbatch(h, uplo, n, max_n, A, lda, bc, info); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, int* n, int max_n, float** A, int* lda,
                             int bc, int* info)
{ return kblasSpotrf_vbatch(h, uplo, n, max_n, A, lda, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                            double** A, int* lda, double** B, int* ldb, int bc, int* info)
{ return kblasDposv_vbatch(h, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                            float** A, int* lda, float** B, int* ldb, int bc, int* info)
{ return kblasSposv_vbatch(h, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, bc, info); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                             double** A, int* lda, double** B, int* ldb, int bc)
{ return kblasDpotrs_vbatch(h, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, bc); }
#endif

#endif /* KBLAS_BATCH_EXT_H_ */
