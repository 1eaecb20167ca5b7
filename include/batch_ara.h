/*
 * batch_ara.h -- the batched SPD helpers of KBLAS's randomized range finder
 * (ARA), MI355X build: the rank-revealing Cholesky of small Gram matrices, the
 * FP32 -> FP64 Gram product and the right triangular solve with its factor.
 *
 * Names, argument order and meaning follow the reference header
 * include/batch_ara.h:74-102,212-270 (implementations
 * src/batch_svd/batch_ara.cu:497-561, src/batch_svd/ara_util.cuh:163-275,
 * 345-408,560-808).  Every array argument is device memory: pointer arrays of
 * device matrices and device int arrays of sizes / leading dimensions, one
 * entry per operation.  Calls are asynchronous on the given stream (the
 * handle's stream for the handle-taking ones) and allocate nothing.  The ARA
 * driver itself (kblas_{s,d}ara_batch, the sampling and vector-count
 * helpers) is not part of this build.
 *
 * kblas_{s,d}ara_fused_potrf_batch: for each operation, the leading
 *   dim = op_samples[i] (clamped to bs) order of the symmetric A (only its
 *   lower triangle is read; the reference reads both, and its arithmetic
 *   uses only the lower one) is factored column by column without pivoting
 *   while the updated diagonal entry exceeds max(0, max diag(A)) * 2^-52; the
 *   number of columns factored goes to block_ranks[i]; R (dim x dim, ldr)
 *   receives the transpose of the lower factor (past the rank: of the
 *   unfactored Schur complement) with zeros below the diagonal, in FP32
 *   (sara) or FP64 (dara); diag_R[i*bs + k] is multiplied by R(k,k) for k
 *   below the rank and set to 0 from the rank on (k < bs).  dim <= 0: only
 *   block_ranks[i] = 0 is written.  bs must be 16 or 32
 *   (KBLAS_NotImplemented otherwise, as in the reference).
 * kblas_sara_mp_syrk_batch: B := A^T A (n x n, both triangles, FP64) of the
 *   m x n FP32 matrices A, max_n <= 32 (KBLAS_NotImplemented above, as in the
 *   reference); FP64 products of the FP32 entries are exact, sums in FP64.
 *   Orders beyond 32 in n[] are treated as 32.
 * kblas_dara_mp_syrk_batch: the same for FP64 A, any order (the batched
 *   GEMM B := 1 * A^T A + 0 * B, as the reference).
 * kblas_{s,d}ara_trsm_batch: B(:, 0:c) := B(:, 0:c) R^-1 for the
 *   rows_batch[i] x cols_batch[i] matrices B, c = min(cols, bs), R upper
 *   triangular (only its upper triangle is read), bs 16 or 32
 *   (KBLAS_NotImplemented otherwise).  Needs no workspace (the reference's
 *   32-column case takes pointer workspace and returns
 *   KBLAS_InsufficientWorkspace without it).  rows_batch[i] <= max_rows.
 */
#ifndef KBLAS_BATCH_ARA_H_
#define KBLAS_BATCH_ARA_H_

#include "kblas.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/batch_ara.h:74-82 */
int kblas_sara_trsm_batch(kblasHandle_t handle, float** B_batch, int* ldb_batch, float** A_batch, int* lda_batch,
                          int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs);
int kblas_dara_trsm_batch(kblasHandle_t handle, double** B_batch, int* ldb_batch, double** A_batch, int* lda_batch,
                          int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs);

/* include/batch_ara.h:84-92 */
int kblas_sara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, float** R_batch, int* ldr_batch,
                                 double* diag_R, int bs, int* block_ranks, int num_ops, hipStream_t stream);
int kblas_dara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, double** R_batch,
                                 int* ldr_batch, double* diag_R, int bs, int* block_ranks, int num_ops,
                                 hipStream_t stream);

/* include/batch_ara.h:94-102 */
int kblas_sara_mp_syrk_batch(kblasHandle_t handle, int* m, int* n, int max_m, int max_n, const float** A, int* lda,
                             double** B, int* ldb, int batchCount);
int kblas_dara_mp_syrk_batch(kblasHandle_t handle, int* m, int* n, int max_m, int max_n, const double** A,
                             int* lda, double** B, int* ldb, int batchCount);

#ifdef __cplusplus
}  /* extern "C" */

/* include/batch_ara.h:212-270 */
inline int kblas_ara_trsm_batch(kblasHandle_t handle, float** B_batch, int* ldb_batch, float** A_batch,
                                int* lda_batch, int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs)
{ return kblas_sara_trsm_batch(handle, B_batch, ldb_batch, A_batch, lda_batch, rows_batch, cols_batch, num_ops, max_rows, bs); }
inline int kblas_ara_trsm_batch(kblasHandle_t handle, double** B_batch, int* ldb_batch, double** A_batch,
                                int* lda_batch, int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs)
{ return kblas_dara_trsm_batch(handle, B_batch, ldb_batch, A_batch, lda_batch, rows_batch, cols_batch, num_ops, max_rows, bs); }
inline int kblas_ara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, float** R_batch,
                                       int* ldr_batch, double* diag_R, int bs, int* block_ranks, int num_ops,
                                       hipStream_t stream)
{ return kblas_sara_fused_potrf_batch(op_samples, A_batch, lda_batch, R_batch, ldr_batch, diag_R, bs, block_ranks, num_ops, stream); }
inline int kblas_ara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, double** R_batch,
                                       int* ldr_batch, double* diag_R, int bs, int* block_ranks, int num_ops,
                                       hipStream_t stream)
{ return kblas_dara_fused_potrf_batch(op_samples, A_batch, lda_batch, R_batch, ldr_batch, diag_R, bs, block_ranks, num_ops, stream); }
inline int kblas_ara_mp_syrk_batch(kblasHandle_t handle, int* m, int* n, int max_m, int max_n, const float** A,
                                   int* lda, double** B, int* ldb, int batchCount)
{ return kblas_sara_mp_syrk_batch(handle, m, n, max_m, max_n, A, lda, B, ldb, batchCount); }
inline int kblas_ara_mp_syrk_batch(kblasHandle_t handle, int* m, int* n, int max_m, int max_n, const double** A,
                                   int* lda, double** B, int* ldb, int batchCount)
{ return kblas_dara_mp_syrk_batch(handle, m, n, max_m, max_n, A, lda, B, ldb, batchCount); }
#endif

#endif /* KBLAS_BATCH_ARA_H_ */
