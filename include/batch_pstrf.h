/*
 * batch_pstrf.h -- batched pivoted Cholesky (P^T A P = R^T R), MI355X build.
 *
 * Names, argument order and meaning follow the reference header
 * include/batch_pstrf.h:20-120.  A: the upper triangle of each symmetric
 * positive semi-definite matrix is read; on exit the upper n x n triangle holds
 * R (past the detected rank: the unfactored Schur complement), the strict lower
 * triangle is zero.  piv: 0-based pivot order.  ranks: detected rank (the
 * factorization stops when the largest remaining diagonal entry is
 * <= n * eps * max(diag(A))).  n <= 64 (KBLAS_Error_WrongInput above, as in
 * the reference).  Device memory, asynchronous on the handle's stream.
 */
#ifndef KBLAS_BATCH_PSTRF_H_
#define KBLAS_BATCH_PSTRF_H_

#include "kblas.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/batch_pstrf.h:53,70 */
int kblasDpstrf_batch_strided(kblasHandle_t handle, int n, double* A_strided, int lda, int stride_a, int* piv_strided,
                              int stride_piv, int* ranks, int num_ops);
int kblasSpstrf_batch_strided(kblasHandle_t handle, int n, float* A_strided, int lda, int stride_a, int* piv_strided,
                              int stride_piv, int* ranks, int num_ops);
/* include/batch_pstrf.h:81,92 */
int kblasDpstrf_batch(kblasHandle_t handle, int n, double** A_array, int lda, int** piv_array, int* ranks,
                      int num_ops);
int kblasSpstrf_batch(kblasHandle_t handle, int n, float** A_array, int lda, int** piv_array, int* ranks,
                      int num_ops);

#ifdef __cplusplus
}  /* extern "C" */

/* include/batch_pstrf.h:111-120 */
inline int kblas_pstrf_batch(kblasHandle_t handle, int n, double** A_array, int lda, int** piv_array, int* ranks,
                             int num_ops)
{ return kblasDpstrf_batch(handle, n, A_array, lda, piv_array, ranks, num_ops); }
inline int kblas_pstrf_batch(kblasHandle_t handle, int n, float** A_array, int lda, int** piv_array, int* ranks,
                             int num_ops)
{ return kblasSpstrf_batch(handle, n, A_array, lda, piv_array, ranks, num_ops); }
inline int kblas_pstrf_batch(kblasHandle_t handle, int n, double* A_strided, int lda, int stride_a, int* piv_strided,
                             int stride_piv, int* ranks, int num_ops)
{ return kblasDpstrf_batch_strided(handle, n, A_strided, lda, stride_a, piv_strided, stride_piv, ranks, num_ops); }
inline int kblas_pstrf_batch(kblasHandle_t handle, int n, float* A_strided, int lda, int stride_a, int* piv_strided,
                             int stride_piv, int* ranks, int num_ops)
{ return kblasSpstrf_batch_strided(handle, n, A_strided, lda, stride_a, piv_strided, stride_piv, ranks, num_ops); }
#endif

#endif /* KBLAS_BATCH_PSTRF_H_ */
