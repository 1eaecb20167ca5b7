/*
 * kblas_batch.h -- batched SPD path of KBLAS (POTRF / POTRS / POSV and the
 * TRSM / SYRK / GEMM batches under them), MI355X (gfx950) build.
 *
 * Each declaration keeps the name, argument order and meaning of the reference
 * header include/kblas_batch.h (line numbers cited per group).  Matrices are
 * column-major device memory; "pointer-array" arguments are DEVICE arrays of
 * device pointers (as the reference's tests and kernels use them); strides are
 * in elements; calls are asynchronous on the handle's stream.
 *
 * Supported: uplo Lower and Upper (the reference returns NotImplemented for
 * Upper; here A = U^T U with U read from and written to the upper triangle);
 * TRSM diag NonUnit and Unit (Unit: NotImplemented in the reference);
 * POTRS/POSV side Right (X*A = B, the reference's only side) and Left
 * (A*X = B, LAPACK layout -- declared but NotImplemented in the reference).
 * Complex (C/Z) POTRF / POTRS / POSV / TRSM / SYRK / GEMM compute what the
 * reference's complex instantiations compute: complex symmetric arithmetic
 * (A = L L^T without conjugation, principal square roots, SYRK = A A^T),
 * trans 'C' a conjugate transpose for GEMM and TRSM (and rejected by SYRK);
 * a zero or non-finite pivot sets info.  The complex inverse family (LAUUM =
 * L^T L, POTRI = (L L^T)^-1) and TRMM ('C' conjugates the triangle) likewise;
 * their upper-triangle variants return KBLAS_NotImplemented.
 * Extension: info_array is written for every matrix (0, or the LAPACK index of
 * the first non-positive leading minor), which the reference documents but
 * never sets (src/batch_triangular/Xpotrf_batch_kernels.cuh:123-125).
 */
#ifndef KBLAS_BATCH_H_
#define KBLAS_BATCH_H_

#include <hip/hip_complex.h>

#include "kblas.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- workspace queries (reference: C++ linkage; here C linkage) ---------- */
/* include/kblas_batch.h:35-36 */
void kblas_gemm_batch_strided_wsquery(kblasHandle_t handle, int batchCount);
void kblas_gemm_batch_nonuniform_wsquery(kblasHandle_t handle);
/* include/kblas_batch.h:485-487 */
void kblas_syrk_batch_wsquery(kblasHandle_t handle, const int m, int batchCount);
void kblas_syrk_batch_nonuniform_wsquery(kblasHandle_t handle);
/* include/kblas_batch.h:773-788 */
void kblas_trsm_batch_wsquery(kblasHandle_t handle, char side, int m, int n, int batchCount);
void kblas_trsm_batch_strided_wsquery(kblasHandle_t handle, char side, int m, int n, int batchCount);
void kblas_trsm_batch_nonuniform_wsquery(kblasHandle_t handle);
/* include/kblas_batch.h:1380-1391 */
void kblas_potrf_batch_wsquery(kblasHandle_t handle, const int n, int batchCount);
void kblas_potrf_batch_strided_wsquery(kblasHandle_t handle, const int n, int batchCount);
/* include/kblas_batch.h:2077-2089 */
void kblas_potrs_batch_wsquery(kblasHandle_t handle, const int m, const int n, int batchCount);
void kblas_potrs_batch_strided_wsquery(kblasHandle_t handle, const int m, const int n, int batchCount);
/* include/kblas_batch.h:2772-2785 */
void kblas_posv_batch_wsquery(kblasHandle_t handle, char side, const int m, const int n, int batchCount);
void kblas_posv_batch_strided_wsquery(kblasHandle_t handle, char side, const int m, const int n, int batchCount);

/* ---- GEMM: include/kblas_batch.h:309 (pointer array), :408 (strided) ---- */
int kblasSgemm_batch(kblasHandle_t handle, char transA, char transB,
                      const int m, const int n, const int k, const float alpha,
                      const float** A_array, int lda, const float** B_array, int ldb,
                      const float beta, float** C_array, int ldc, int batchCount);
int kblasSgemm_batch_strided(kblasHandle_t handle, char transA, char transB,
                              const int m, const int n, const int k, const float alpha,
                              const float* A, int lda, long strideA,
                              const float* B, int ldb, long strideB,
                              const float beta, float* C, int ldc, long strideC, int batchCount);
int kblasDgemm_batch(kblasHandle_t handle, char transA, char transB,
                      const int m, const int n, const int k, const double alpha,
                      const double** A_array, int lda, const double** B_array, int ldb,
                      const double beta, double** C_array, int ldc, int batchCount);
int kblasDgemm_batch_strided(kblasHandle_t handle, char transA, char transB,
                              const int m, const int n, const int k, const double alpha,
                              const double* A, int lda, long strideA,
                              const double* B, int ldb, long strideB,
                              const double beta, double* C, int ldc, long strideC, int batchCount);
int kblasCgemm_batch(kblasHandle_t handle, char transA, char transB,
                      const int m, const int n, const int k, const hipFloatComplex alpha,
                      const hipFloatComplex** A_array, int lda, const hipFloatComplex** B_array, int ldb,
                      const hipFloatComplex beta, hipFloatComplex** C_array, int ldc, int batchCount);
int kblasCgemm_batch_strided(kblasHandle_t handle, char transA, char transB,
                              const int m, const int n, const int k, const hipFloatComplex alpha,
                              const hipFloatComplex* A, int lda, long strideA,
                              const hipFloatComplex* B, int ldb, long strideB,
                              const hipFloatComplex beta, hipFloatComplex* C, int ldc, long strideC, int batchCount);
int kblasZgemm_batch(kblasHandle_t handle, char transA, char transB,
                      const int m, const int n, const int k, const hipDoubleComplex alpha,
                      const hipDoubleComplex** A_array, int lda, const hipDoubleComplex** B_array, int ldb,
                      const hipDoubleComplex beta, hipDoubleComplex** C_array, int ldc, int batchCount);
int kblasZgemm_batch_strided(kblasHandle_t handle, char transA, char transB,
                              const int m, const int n, const int k, const hipDoubleComplex alpha,
                              const hipDoubleComplex* A, int lda, long strideA,
                              const hipDoubleComplex* B, int ldb, long strideB,
                              const hipDoubleComplex beta, hipDoubleComplex* C, int ldc, long strideC, int batchCount);

/* ---- SYRK: include/kblas_batch.h:631 (pointer array), :711 (strided) ---- */
int kblasSsyrk_batch(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                      const float alpha, const float** A_array, int lda,
                      const float beta, float** B_array, int ldb, int batchCount);
int kblasSsyrk_batch_strided(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                              const float alpha, const float* A, int lda, long strideA,
                              const float beta, float* B, int ldb, long strideB, int batchCount);
int kblasDsyrk_batch(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                      const double alpha, const double** A_array, int lda,
                      const double beta, double** B_array, int ldb, int batchCount);
int kblasDsyrk_batch_strided(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                              const double alpha, const double* A, int lda, long strideA,
                              const double beta, double* B, int ldb, long strideB, int batchCount);
int kblasCsyrk_batch(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                      const hipFloatComplex alpha, const hipFloatComplex** A_array, int lda,
                      const hipFloatComplex beta, hipFloatComplex** B_array, int ldb, int batchCount);
int kblasCsyrk_batch_strided(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                              const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long strideA,
                              const hipFloatComplex beta, hipFloatComplex* B, int ldb, long strideB, int batchCount);
int kblasZsyrk_batch(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                      const hipDoubleComplex alpha, const hipDoubleComplex** A_array, int lda,
                      const hipDoubleComplex beta, hipDoubleComplex** B_array, int ldb, int batchCount);
int kblasZsyrk_batch_strided(kblasHandle_t handle, char uplo, char trans, const int m, const int n,
                              const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long strideA,
                              const hipDoubleComplex beta, hipDoubleComplex* B, int ldb, long strideB, int batchCount);

/* ---- TRSM: include/kblas_batch.h:948 (pointer array), :1034 (strided) ---- */
int kblasStrsm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                      const int m, const int n, const float alpha,
                      const float** A_array, int lda, float** B_array, int ldb, int batchCount);
int kblasStrsm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                              const int m, const int n, const float alpha,
                              const float* A, int lda, long strideA,
                              float* B, int ldb, long strideB, int batchCount);
int kblasDtrsm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                      const int m, const int n, const double alpha,
                      const double** A_array, int lda, double** B_array, int ldb, int batchCount);
int kblasDtrsm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                              const int m, const int n, const double alpha,
                              const double* A, int lda, long strideA,
                              double* B, int ldb, long strideB, int batchCount);
int kblasCtrsm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                      const int m, const int n, const hipFloatComplex alpha,
                      const hipFloatComplex** A_array, int lda, hipFloatComplex** B_array, int ldb, int batchCount);
int kblasCtrsm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                              const int m, const int n, const hipFloatComplex alpha,
                              const hipFloatComplex* A, int lda, long strideA,
                              hipFloatComplex* B, int ldb, long strideB, int batchCount);
int kblasZtrsm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                      const int m, const int n, const hipDoubleComplex alpha,
                      const hipDoubleComplex** A_array, int lda, hipDoubleComplex** B_array, int ldb, int batchCount);
int kblasZtrsm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                              const int m, const int n, const hipDoubleComplex alpha,
                              const hipDoubleComplex* A, int lda, long strideA,
                              hipDoubleComplex* B, int ldb, long strideB, int batchCount);

/* ---- POTRF: include/kblas_batch.h:1486 (pointer array), :1549 (strided) ---- */
int kblasSpotrf_batch(kblasHandle_t handle, char uplo, const int n,
                       float** A_array, int lda, int batchCount, int* info_array);
int kblasSpotrf_batch_strided(kblasHandle_t handle, char uplo, const int n,
                               float* A, int lda, long strideA, int batchCount, int* info_array);
int kblasDpotrf_batch(kblasHandle_t handle, char uplo, const int n,
                       double** A_array, int lda, int batchCount, int* info_array);
int kblasDpotrf_batch_strided(kblasHandle_t handle, char uplo, const int n,
                               double* A, int lda, long strideA, int batchCount, int* info_array);
int kblasCpotrf_batch(kblasHandle_t handle, char uplo, const int n,
                       hipFloatComplex** A_array, int lda, int batchCount, int* info_array);
int kblasCpotrf_batch_strided(kblasHandle_t handle, char uplo, const int n,
                               hipFloatComplex* A, int lda, long strideA, int batchCount, int* info_array);
int kblasZpotrf_batch(kblasHandle_t handle, char uplo, const int n,
                       hipDoubleComplex** A_array, int lda, int batchCount, int* info_array);
int kblasZpotrf_batch_strided(kblasHandle_t handle, char uplo, const int n,
                               hipDoubleComplex* A, int lda, long strideA, int batchCount, int* info_array);

/* ---- POTRS: include/kblas_batch.h:2190 (pointer array), :2261 (strided) ---- */
int kblasSpotrs_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                       const float** A_array, int lda, float** B_array, int ldb, int batchCount);
int kblasSpotrs_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                               const float* A, int lda, long strideA,
                               float* B, int ldb, long strideB, int batchCount);
int kblasDpotrs_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                       const double** A_array, int lda, double** B_array, int ldb, int batchCount);
int kblasDpotrs_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                               const double* A, int lda, long strideA,
                               double* B, int ldb, long strideB, int batchCount);
int kblasCpotrs_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                       const hipFloatComplex** A_array, int lda, hipFloatComplex** B_array, int ldb, int batchCount);
int kblasCpotrs_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                               const hipFloatComplex* A, int lda, long strideA,
                               hipFloatComplex* B, int ldb, long strideB, int batchCount);
int kblasZpotrs_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                       const hipDoubleComplex** A_array, int lda, hipDoubleComplex** B_array, int ldb, int batchCount);
int kblasZpotrs_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                               const hipDoubleComplex* A, int lda, long strideA,
                               hipDoubleComplex* B, int ldb, long strideB, int batchCount);

/* ---- POSV: include/kblas_batch.h:2896 (pointer array), :2970 (strided) ---- */
int kblasSposv_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                      float** A_array, int lda, float** B_array, int ldb, int batchCount, int* info_array);
int kblasSposv_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                              float* A, int lda, long strideA, float* B, int ldb, long strideB,
                              int batchCount, int* info_array);
int kblasDposv_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                      double** A_array, int lda, double** B_array, int ldb, int batchCount, int* info_array);
int kblasDposv_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                              double* A, int lda, long strideA, double* B, int ldb, long strideB,
                              int batchCount, int* info_array);
int kblasCposv_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                      hipFloatComplex** A_array, int lda, hipFloatComplex** B_array, int ldb, int batchCount, int* info_array);
int kblasCposv_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                              hipFloatComplex* A, int lda, long strideA, hipFloatComplex* B, int ldb, long strideB,
                              int batchCount, int* info_array);
int kblasZposv_batch(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                      hipDoubleComplex** A_array, int lda, hipDoubleComplex** B_array, int ldb, int batchCount, int* info_array);
int kblasZposv_batch_strided(kblasHandle_t handle, char side, char uplo, const int m, const int n,
                              hipDoubleComplex* A, int lda, long strideA, hipDoubleComplex* B, int ldb, long strideB,
                              int batchCount, int* info_array);

/* ---- non-uniform GEMM / SYRK / TRSM ----------------------------------------
 * The reference exposes these only as C++ overloads backed by MAGMA vbatched
 * routines (include/kblas_batch.h:198-250, 518-546, 827-863;
 * src/batch_triangular/Xtrsm_batch_drivers.cuh:277-367).  Here they are native
 * and also have C names.  m, n, k, lda, ldb, ldc are DEVICE int arrays; the
 * max_* hints are host ints (0 = unknown; no host synchronisation either way). */
int kblasSgemm_vbatch(kblasHandle_t handle, char transA, char transB,
                       int* m, int* n, int* k, int max_m, int max_n, int max_k,
                       const float alpha, const float** A, int* lda, const float** B, int* ldb,
                       const float beta, float** C, int* ldc, int batchCount);
int kblasSsyrk_vbatch(kblasHandle_t handle, char uplo, char trans, int* m, int* n,
                       int max_m, int max_n, float alpha, float** A, int* lda,
                       float beta, float** B, int* ldb, int batchCount);
int kblasStrsm_vbatch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                       int* m, int* n, int max_m, int max_n, float alpha,
                       float** A, int* lda, float** B, int* ldb, int batchCount);
int kblasDgemm_vbatch(kblasHandle_t handle, char transA, char transB,
                       int* m, int* n, int* k, int max_m, int max_n, int max_k,
                       const double alpha, const double** A, int* lda, const double** B, int* ldb,
                       const double beta, double** C, int* ldc, int batchCount);
int kblasDsyrk_vbatch(kblasHandle_t handle, char uplo, char trans, int* m, int* n,
                       int max_m, int max_n, double alpha, double** A, int* lda,
                       double beta, double** B, int* ldb, int batchCount);
int kblasDtrsm_vbatch(kblasHandle_t handle, char side, char uplo, char trans, char diag,
                       int* m, int* n, int max_m, int max_n, double alpha,
                       double** A, int* lda, double** B, int* ldb, int batchCount);


/* ---- inverse family and TRMM (lower triangle) --------------------------------
 * Served: uplo = Lower (Upper returns KBLAS_NotImplemented, as in the reference);
 * TRTRI diag NonUnit and Unit; TRMM side Left and Right, trans N/T, diag
 * NonUnit and Unit (the reference serves Left / NonUnit only,
 * src/batch_triangular/Xtrmm_batch.cu:63-105).  Orders up to 1024 for the
 * inverse family (KBLAS_NotImplemented above).  info_array: 0, or for
 * TRTRI / POTRI the 1-based index of the first exactly-zero diagonal element
 * (LAPACK xTRTRI; A is then left unchanged), for POTI POTRF's info (the
 * inversion is skipped for that matrix).  No workspace is needed. */
/* include/kblas_batch.h:1100,1113 */
void kblas_trmm_batch_wsquery(kblasHandle_t handle, char side, int m, int n, int batchCount);
void kblas_trmm_batch_strided_wsquery(kblasHandle_t handle, char side, int m, int n, int batchCount);
/* include/kblas_batch.h:1611,1622 */
void kblas_lauum_batch_wsquery(kblasHandle_t handle, const int n, int batchCount);
void kblas_lauum_batch_strided_wsquery(kblasHandle_t handle, const int n, int batchCount);
/* include/kblas_batch.h:1840,1851 */
void kblas_trtri_batch_wsquery(kblasHandle_t handle, const int n, int batchCount);
void kblas_trtri_batch_strided_wsquery(kblasHandle_t handle, const int n, int batchCount);
/* include/kblas_batch.h:2321,2332 */
void kblas_potri_batch_wsquery(kblasHandle_t handle, const int n, int batchCount);
void kblas_potri_batch_strided_wsquery(kblasHandle_t handle, const int n, int batchCount);
/* include/kblas_batch.h:2547,2558 */
void kblas_poti_batch_wsquery(kblasHandle_t handle, const int n, int batchCount);
void kblas_poti_batch_strided_wsquery(kblasHandle_t handle, const int n, int batchCount);

/* include/kblas_batch.h:1186-1366 (TRMM: C := alpha op(A) B / alpha B op(A), in place in B) */
int kblasStrmm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                     const int n, const float alpha, const float** A_array, int lda, float** B_array, int ldb,
                     int batchCount);
int kblasStrmm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                             const int n, const float alpha, const float* A, int lda, long strideA, float* B,
                             int ldb, long strideB, int batchCount);
int kblasDtrmm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                     const int n, const double alpha, const double** A_array, int lda, double** B_array, int ldb,
                     int batchCount);
int kblasDtrmm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                             const int n, const double alpha, const double* A, int lda, long strideA, double* B,
                             int ldb, long strideB, int batchCount);
int kblasCtrmm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                     const int n, const hipFloatComplex alpha, const hipFloatComplex** A_array, int lda, hipFloatComplex** B_array, int ldb,
                     int batchCount);
int kblasCtrmm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                             const int n, const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long strideA, hipFloatComplex* B,
                             int ldb, long strideB, int batchCount);
int kblasZtrmm_batch(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                     const int n, const hipDoubleComplex alpha, const hipDoubleComplex** A_array, int lda, hipDoubleComplex** B_array, int ldb,
                     int batchCount);
int kblasZtrmm_batch_strided(kblasHandle_t handle, char side, char uplo, char trans, char diag, const int m,
                             const int n, const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long strideA, hipDoubleComplex* B,
                             int ldb, long strideB, int batchCount);
/* include/kblas_batch.h:1686-1826 (LAUUM: A := L^T L) */
int kblasSlauum_batch(kblasHandle_t handle, char uplo, const int n, float** A_array, int lda, int batchCount,
                      int* info_array);
int kblasSlauum_batch_strided(kblasHandle_t handle, char uplo, const int n, float* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasDlauum_batch(kblasHandle_t handle, char uplo, const int n, double** A_array, int lda, int batchCount,
                      int* info_array);
int kblasDlauum_batch_strided(kblasHandle_t handle, char uplo, const int n, double* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasClauum_batch(kblasHandle_t handle, char uplo, const int n, hipFloatComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasClauum_batch_strided(kblasHandle_t handle, char uplo, const int n, hipFloatComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasZlauum_batch(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasZlauum_batch_strided(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
/* include/kblas_batch.h:1915-2061 (TRTRI: A := L^-1) */
int kblasStrtri_batch(kblasHandle_t handle, char uplo, char diag, const int n, float** A_array, int lda, int batchCount,
                      int* info_array);
int kblasStrtri_batch_strided(kblasHandle_t handle, char uplo, char diag, const int n, float* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasDtrtri_batch(kblasHandle_t handle, char uplo, char diag, const int n, double** A_array, int lda, int batchCount,
                      int* info_array);
int kblasDtrtri_batch_strided(kblasHandle_t handle, char uplo, char diag, const int n, double* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasCtrtri_batch(kblasHandle_t handle, char uplo, char diag, const int n, hipFloatComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasCtrtri_batch_strided(kblasHandle_t handle, char uplo, char diag, const int n, hipFloatComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasZtrtri_batch(kblasHandle_t handle, char uplo, char diag, const int n, hipDoubleComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasZtrtri_batch_strided(kblasHandle_t handle, char uplo, char diag, const int n, hipDoubleComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
/* include/kblas_batch.h:2396-2533 (POTRI: A := (L L^T)^-1 from the factor L) */
int kblasSpotri_batch(kblasHandle_t handle, char uplo, const int n, float** A_array, int lda, int batchCount,
                      int* info_array);
int kblasSpotri_batch_strided(kblasHandle_t handle, char uplo, const int n, float* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasDpotri_batch(kblasHandle_t handle, char uplo, const int n, double** A_array, int lda, int batchCount,
                      int* info_array);
int kblasDpotri_batch_strided(kblasHandle_t handle, char uplo, const int n, double* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasCpotri_batch(kblasHandle_t handle, char uplo, const int n, hipFloatComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasCpotri_batch_strided(kblasHandle_t handle, char uplo, const int n, hipFloatComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasZpotri_batch(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasZpotri_batch_strided(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
/* include/kblas_batch.h:2622-2756 (POTI: A := A^-1 of an SPD A, POTRF then POTRI) */
int kblasSpoti_batch(kblasHandle_t handle, char uplo, const int n, float** A_array, int lda, int batchCount,
                      int* info_array);
int kblasSpoti_batch_strided(kblasHandle_t handle, char uplo, const int n, float* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasDpoti_batch(kblasHandle_t handle, char uplo, const int n, double** A_array, int lda, int batchCount,
                      int* info_array);
int kblasDpoti_batch_strided(kblasHandle_t handle, char uplo, const int n, double* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasCpoti_batch(kblasHandle_t handle, char uplo, const int n, hipFloatComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasCpoti_batch_strided(kblasHandle_t handle, char uplo, const int n, hipFloatComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);
int kblasZpoti_batch(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex** A_array, int lda, int batchCount,
                      int* info_array);
int kblasZpoti_batch_strided(kblasHandle_t handle, char uplo, const int n, hipDoubleComplex* A, int lda, long strideA,
                              int batchCount, int* info_array);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#ifdef __cplusplus
/* ---- C++ overloads (include/kblas_batch.h:87-251, 493-581, 794-897,
 *      1396-1452, 2094-2150, 2790-2854), inline over the C entry points ---- */
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const float alpha, const float** A, int lda, const float** B, int ldb,
                            const float beta, float** C, int ldc, int bc)
{ return kblasSgemm_batch(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const float alpha, const float* A, int lda, long sA, const float* B, int ldb, long sB,
                            const float beta, float* C, int ldc, long sC, int bc)
{ return kblasSgemm_batch_strided(h, ta, tb, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const float alpha, const float** A, int lda, const float beta, float** B, int ldb, int bc)
{ return kblasSsyrk_batch(h, uplo, trans, m, n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const float alpha, const float* A, int lda, long sA, const float beta, float* B, int ldb, long sB, int bc)
{ return kblasSsyrk_batch_strided(h, uplo, trans, m, n, alpha, A, lda, sA, beta, B, ldb, sB, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const float alpha, const float** A, int lda, float** B, int ldb, int bc)
{ return kblasStrsm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const float alpha, const float* A, int lda, long sA, float* B, int ldb, long sB, int bc)
{ return kblasStrsm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, float** A, int lda, int bc, int* info)
{ return kblasSpotrf_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, float* A, int lda, long sA, int bc, int* info)
{ return kblasSpotrf_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const float** A, int lda, float** B, int ldb, int bc)
{ return kblasSpotrs_batch(h, side, uplo, m, n, A, lda, B, ldb, bc); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const float* A, int lda, long sA, float* B, int ldb, long sB, int bc)
{ return kblasSpotrs_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            float** A, int lda, float** B, int ldb, int bc, int* info)
{ return kblasSposv_batch(h, side, uplo, m, n, A, lda, B, ldb, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            float* A, int lda, long sA, float* B, int ldb, long sB, int bc, int* info)
{ return kblasSposv_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc, info); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const double alpha, const double** A, int lda, const double** B, int ldb,
                            const double beta, double** C, int ldc, int bc)
{ return kblasDgemm_batch(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const double alpha, const double* A, int lda, long sA, const double* B, int ldb, long sB,
                            const double beta, double* C, int ldc, long sC, int bc)
{ return kblasDgemm_batch_strided(h, ta, tb, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const double alpha, const double** A, int lda, const double beta, double** B, int ldb, int bc)
{ return kblasDsyrk_batch(h, uplo, trans, m, n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const double alpha, const double* A, int lda, long sA, const double beta, double* B, int ldb, long sB, int bc)
{ return kblasDsyrk_batch_strided(h, uplo, trans, m, n, alpha, A, lda, sA, beta, B, ldb, sB, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const double alpha, const double** A, int lda, double** B, int ldb, int bc)
{ return kblasDtrsm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const double alpha, const double* A, int lda, long sA, double* B, int ldb, long sB, int bc)
{ return kblasDtrsm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, double** A, int lda, int bc, int* info)
{ return kblasDpotrf_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, double* A, int lda, long sA, int bc, int* info)
{ return kblasDpotrf_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const double** A, int lda, double** B, int ldb, int bc)
{ return kblasDpotrs_batch(h, side, uplo, m, n, A, lda, B, ldb, bc); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const double* A, int lda, long sA, double* B, int ldb, long sB, int bc)
{ return kblasDpotrs_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            double** A, int lda, double** B, int ldb, int bc, int* info)
{ return kblasDposv_batch(h, side, uplo, m, n, A, lda, B, ldb, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            double* A, int lda, long sA, double* B, int ldb, long sB, int bc, int* info)
{ return kblasDposv_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc, info); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const hipFloatComplex alpha, const hipFloatComplex** A, int lda, const hipFloatComplex** B, int ldb,
                            const hipFloatComplex beta, hipFloatComplex** C, int ldc, int bc)
{ return kblasCgemm_batch(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, const hipFloatComplex* B, int ldb, long sB,
                            const hipFloatComplex beta, hipFloatComplex* C, int ldc, long sC, int bc)
{ return kblasCgemm_batch_strided(h, ta, tb, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex** A, int lda, const hipFloatComplex beta, hipFloatComplex** B, int ldb, int bc)
{ return kblasCsyrk_batch(h, uplo, trans, m, n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, const hipFloatComplex beta, hipFloatComplex* B, int ldb, long sB, int bc)
{ return kblasCsyrk_batch_strided(h, uplo, trans, m, n, alpha, A, lda, sA, beta, B, ldb, sB, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return kblasCtrsm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return kblasCtrsm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return kblasCpotrf_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return kblasCpotrf_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return kblasCpotrs_batch(h, side, uplo, m, n, A, lda, B, ldb, bc); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return kblasCpotrs_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc, int* info)
{ return kblasCposv_batch(h, side, uplo, m, n, A, lda, B, ldb, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc, int* info)
{ return kblasCposv_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc, info); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, const hipDoubleComplex** B, int ldb,
                            const hipDoubleComplex beta, hipDoubleComplex** C, int ldc, int bc)
{ return kblasZgemm_batch(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                            const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, const hipDoubleComplex* B, int ldb, long sB,
                            const hipDoubleComplex beta, hipDoubleComplex* C, int ldc, long sC, int bc)
{ return kblasZgemm_batch_strided(h, ta, tb, m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, const hipDoubleComplex beta, hipDoubleComplex** B, int ldb, int bc)
{ return kblasZsyrk_batch(h, uplo, trans, m, n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, const hipDoubleComplex beta, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return kblasZsyrk_batch_strided(h, uplo, trans, m, n, alpha, A, lda, sA, beta, B, ldb, sB, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return kblasZtrsm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return kblasZtrsm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return kblasZpotrf_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potrf_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return kblasZpotrf_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return kblasZpotrs_batch(h, side, uplo, m, n, A, lda, B, ldb, bc); }
inline int kblas_potrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return kblasZpotrs_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc, int* info)
{ return kblasZposv_batch(h, side, uplo, m, n, A, lda, B, ldb, bc, info); }
inline int kblas_posv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                            hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc, int* info)
{ return kblasZposv_batch_strided(h, side, uplo, m, n, A, lda, sA, B, ldb, sB, bc, info); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, int* m, int* n, int* k,
                            int max_m, int max_n, int max_k, const float alpha, const float** A, int* lda,
                            const float** B, int* ldb, const float beta, float** C, int* ldc, int bc)
{ return kblasSgemm_vbatch(h, ta, tb, m, n, k, max_m, max_n, max_k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, int* m, int* n, int max_m, int max_n,
                            float alpha, float** A, int* lda, float beta, float** B, int* ldb, int bc)
{ return kblasSsyrk_vbatch(h, uplo, trans, m, n, max_m, max_n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, int* m, int* n,
                            int max_m, int max_n, float alpha, float** A, int* lda, float** B, int* ldb, int bc)
{ return kblasStrsm_vbatch(h, side, uplo, trans, diag, m, n, max_m, max_n, alpha, A, lda, B, ldb, bc); }
inline int kblas_gemm_batch(kblasHandle_t h, char ta, char tb, int* m, int* n, int* k,
                            int max_m, int max_n, int max_k, const double alpha, const double** A, int* lda,
                            const double** B, int* ldb, const double beta, double** C, int* ldc, int bc)
{ return kblasDgemm_vbatch(h, ta, tb, m, n, k, max_m, max_n, max_k, alpha, A, lda, B, ldb, beta, C, ldc, bc); }
inline int kblas_syrk_batch(kblasHandle_t h, char uplo, char trans, int* m, int* n, int max_m, int max_n,
                            double alpha, double** A, int* lda, double beta, double** B, int* ldb, int bc)
{ return kblasDsyrk_vbatch(h, uplo, trans, m, n, max_m, max_n, alpha, A, lda, beta, B, ldb, bc); }
inline int kblas_trsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, int* m, int* n,
                            int max_m, int max_n, double alpha, double** A, int* lda, double** B, int* ldb, int bc)
{ return kblasDtrsm_vbatch(h, side, uplo, trans, diag, m, n, max_m, max_n, alpha, A, lda, B, ldb, bc); }

/* ---- C++ overloads of the inverse family and TRMM (include/kblas_batch.h:1115-1184,
 * 1624-1684, 1853-1913, 2334-2394, 2560-2620) */
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const float alpha, const float** A, int lda, float** B, int ldb, int bc)
{ return kblasStrmm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const float alpha, const float* A, int lda, long sA, float* B, int ldb, long sB, int bc)
{ return kblasStrmm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const double alpha, const double** A, int lda, double** B, int ldb, int bc)
{ return kblasDtrmm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const double alpha, const double* A, int lda, long sA, double* B, int ldb, long sB, int bc)
{ return kblasDtrmm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return kblasCtrmm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return kblasCtrmm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return kblasZtrmm_batch(h, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, bc); }
inline int kblas_trmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                            const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return kblasZtrmm_batch_strided(h, side, uplo, trans, diag, m, n, alpha, A, lda, sA, B, ldb, sB, bc); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, float** A, int lda, int bc, int* info)
{ return kblasSlauum_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, float* A, int lda, long sA, int bc, int* info)
{ return kblasSlauum_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, double** A, int lda, int bc, int* info)
{ return kblasDlauum_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, double* A, int lda, long sA, int bc, int* info)
{ return kblasDlauum_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return kblasClauum_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return kblasClauum_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return kblasZlauum_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_lauum_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return kblasZlauum_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, float** A, int lda, int bc, int* info)
{ return kblasStrtri_batch(h, uplo, diag, n, A, lda, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, float* A, int lda, long sA, int bc, int* info)
{ return kblasStrtri_batch_strided(h, uplo, diag, n, A, lda, sA, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, double** A, int lda, int bc, int* info)
{ return kblasDtrtri_batch(h, uplo, diag, n, A, lda, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, double* A, int lda, long sA, int bc, int* info)
{ return kblasDtrtri_batch_strided(h, uplo, diag, n, A, lda, sA, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return kblasCtrtri_batch(h, uplo, diag, n, A, lda, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return kblasCtrtri_batch_strided(h, uplo, diag, n, A, lda, sA, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return kblasZtrtri_batch(h, uplo, diag, n, A, lda, bc, info); }
inline int kblas_trtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return kblasZtrtri_batch_strided(h, uplo, diag, n, A, lda, sA, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, float** A, int lda, int bc, int* info)
{ return kblasSpotri_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, float* A, int lda, long sA, int bc, int* info)
{ return kblasSpotri_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, double** A, int lda, int bc, int* info)
{ return kblasDpotri_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, double* A, int lda, long sA, int bc, int* info)
{ return kblasDpotri_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return kblasCpotri_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return kblasCpotri_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return kblasZpotri_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_potri_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return kblasZpotri_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, float** A, int lda, int bc, int* info)
{ return kblasSpoti_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, float* A, int lda, long sA, int bc, int* info)
{ return kblasSpoti_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, double** A, int lda, int bc, int* info)
{ return kblasDpoti_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, double* A, int lda, long sA, int bc, int* info)
{ return kblasDpoti_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return kblasCpoti_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return kblasCpoti_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return kblasZpoti_batch(h, uplo, n, A, lda, bc, info); }
inline int kblas_poti_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return kblasZpoti_batch_strided(h, uplo, n, A, lda, sA, bc, info); }
#endif /* __cplusplus */

#endif /* KBLAS_BATCH_H_ */
