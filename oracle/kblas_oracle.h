/*
 * kblas_oracle.h -- CPU restatement of KBLAS's batched SPD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry and the CPU-baseline leg of the benchmark may load it; the product
 * library (libkblas_amd.so) never links, loads or calls it.
 *
 * What it restates (reference = ecrc/kblas-gpu v3.0.0, paths relative to the
 * reference root):
 *   - POTRF recursion and split rule   src/batch_triangular/Xpotrf_batch_drivers.cuh:31-137,
 *                                       src/kblas_common.cu:241-255
 *   - POTRF register kernel numerics   src/batch_triangular/Xpotrf_batch_kernels.cuh:37-78
 *   - TRSM recursion (-1/alpha GEMM)   src/batch_triangular/Xtrsm_batch_drivers.cuh:55-272
 *   - TRSM register kernel numerics    src/batch_triangular/Xtrsm_batch_kernels.cuh:36-103
 *   - SYRK diagonal-tile numerics      src/batch_triangular/Xsyrk_batch_kernels.cuh:40-86
 *   - POTRS Right/Lower sequence       src/batch_triangular/Xpotrs_batch_drivers.cuh:79-174
 *   - POSV = POTRF then POTRS          src/batch_triangular/Xposv_batch_drivers.cuh:80-117
 *   - TRTRI / LAUUM / POTRI / TRMM     src/batch_triangular/X{trtri,lauum,trmm}_batch_kernels.cuh,
 *                                       Xpotri_batch_drivers.cuh:78-110
 *   - PSTRF (pivoted Cholesky)         src/batch_svd/batch_pstrf.cu:44-158
 *   - complex (C / Z) arithmetic        include/kblas_operators.h:150-255 (the
 *                                       reference's complex instantiations)
 *   - ARA rank-revealing POTRF, Gram   src/batch_svd/ara_util.cuh:163-275,345-408,
 *     product, right TRSM                560-640,651-764
 *   - FLOP model                       testing/flops.h:74-150
 *   - reference test data              testing/testing_helper.cu:361-403
 * Extensions with no reference counterpart (documented as such):
 *   - Left-side POTRS/POSV (LAPACK dpotrs layout), n = 1 handled directly.
 *   - info = first non-positive pivot (LAPACK), documented but never set by
 *     the reference (src/batch_triangular/Xpotrf_batch_kernels.cuh:123-125).
 *   - Mixed-precision POSV, a restatement of LAPACK dsposv (FP32 factor,
 *     FP64 residual refinement, FP64 fallback).
 *   - Counter-based seeded generators so CPU and GPU inputs are identical.
 *
 * All matrices are column-major; "Lower" is the only triangle referenced.
 */
#ifndef KBLAS_ORACLE_H
#define KBLAS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- counter-based generators ---------------- */
uint64_t orc_hash(uint64_t seed, uint64_t id, uint64_t k);
double   orc_uniform(uint64_t seed, uint64_t id, uint64_t k);   /* U(-1,1) */
double   orc_normal(uint64_t seed, uint64_t id, uint64_t k);    /* N(0,1)  */
double   orc_exp2(double x);                                    /* deterministic 2^x */

/* A = B*B^T + shift*I, B_ij ~ U(-1,1), full symmetric storage.               */
void orc_gen_spd_shift(uint64_t seed, uint64_t id, int n, double shift, double* A, int lda);
/* A = Q*diag(s)*Q^T with Q Haar orthogonal (Stewart's product of Householder
 * reflectors built from Gaussian vectors), s log-spaced from 1 down to 1/kappa,
 * log10(kappa) ~ U(lk_lo, lk_hi).  Returns kappa.                             */
double orc_gen_spd_cond(uint64_t seed, uint64_t id, int n, double lk_lo, double lk_hi,
                        double* A, int lda);
/* RHS block of rows x cols, entries N(0,1).                                  */
void orc_gen_rhs(uint64_t seed, uint64_t id, int rows, int cols, double* B, int ldb);
/* Reference test data: U(0,1) entries, A_ii += n, mirror lower->upper        */
void orc_make_hpd(int n, double* A, int lda);

/* ---------------- reference algorithm restatement ---------------- */
int  orc_reg_size(int n);            /* power of two test   */
int  orc_closest_reg_size(int n);    /* largest pow2 < n    */
/* Returns info (0 ok, j>0 first non-positive leading minor).                 */
int  orc_potrf(int n, double* A, int lda);
/* Inverse family and TRMM, lower triangle (info: first exactly-zero diagonal
 * element, 1-based, matrix unchanged).                                      */
int  orc_trtri(char diag, int n, double* A, int lda);
/* Pivoted Cholesky P^T A P = R^T R (upper), n <= 64; piv 0-based; returns the rank. */
int  orc_pstrf(int n, double* A, int lda, int* piv);
/* ARA helpers (src/batch_svd/ara_util.cuh): rank-revealing Cholesky of the
 * lower triangle of a dim x dim Gram matrix (dim clamped to bs; R = upper
 * factor, d[0..bs) scaled by diag(R) below the rank, zeroed from it; returns
 * the rank), the Gram product B = A^T A, and B(:, 0:min(cols,bs)) *= R^-1
 * with R upper.                                                              */
int  orc_ara_potrf(int dim, int bs, const double* G, int ldg, double* R, int ldr, double* d);
void orc_ara_syrk(int m, int n, const double* A, int lda, double* B, int ldb);
void orc_ara_trsm(int rows, int cols, int bs, double* B, int ldb, const double* R, int ldr);
void orc_lauum(int n, double* A, int lda);
int  orc_potri(int n, double* A, int lda);
void orc_trmm(char side, char trans, char diag, int m, int n, double alpha, const double* A, int lda,
              double* B, int ldb);
void orc_trsm(char side, char trans, int m, int n, double alpha,
              const double* A, int lda, double* B, int ldb);
void orc_syrk(char trans, int m, int k, double alpha, const double* A, int lda,
              double beta, double* B, int ldb);
void orc_gemm(char ta, char tb, int m, int n, int k, double alpha,
              const double* A, int lda, const double* B, int ldb,
              double beta, double* C, int ldc);
/* side 'R': X*A = B, B is m x n (m = #RHS); side 'L': A*X = B, B is m x n
 * with n = #RHS and A of order m.  A holds the Cholesky factor (lower).      */
void orc_potrs(char side, int m, int n, const double* A, int lda, double* B, int ldb);
int  orc_posv(char side, int m, int n, double* A, int lda, double* B, int ldb);

/* Mixed precision (LAPACK dsposv restated).  A is read-only; X receives the
 * solution.  Returns info; *iter = #refinement steps, or -(reason) when it
 * fell back to FP64 (-2: an entry of B, A or a residual beyond the FP32
 * range, -3: FP32 factor failed, -31: no convergence).                      */
int  orc_sposv(char side, int m, int n, const double* A, int lda,
               const double* B, int ldb, double* X, int ldx, int* iter);

/* Complex (C / Z) restatement, complex-symmetric arithmetic as the reference's
 * complex instantiations (interleaved real/imaginary doubles; scalars as
 * pairs; upper selects the upper triangle, read as the lower view).        */
void orc_zsqrt(double re, double im, double* out);
int  orc_zpotrf(int upper, int n, double* A, int lda);
void orc_zpotrs(int upper, int side_left, int n, int nrhs, const double* A, int lda, double* B, int ldb);
void orc_zgemm(char ta, char tb, int m, int n, int k, const double* alpha, const double* A, int lda,
               const double* B, int ldb, const double* beta, double* C, int ldc);
void orc_zsyrk(int upper, char trans, int n, int k, const double* alpha, const double* A, int lda,
               const double* beta, double* C, int ldc);
void orc_ztrsm(char side, int upper, char trans, int unit, int m, int n, const double* alpha,
               const double* A, int lda, double* B, int ldb);
int  orc_ztrtri(int unit, int n, double* A, int lda);
void orc_zlauum(int n, double* A, int lda);
int  orc_zpotri(int n, double* A, int lda);
void orc_ztrmm(char side, char trans, int unit, int m, int n, const double* alpha, const double* A, int lda,
               double* B, int ldb);

/* single-precision register/recursive restatement (S API checks).          */
int  orc_spotrf(int n, float* A, int lda);

/* ---------------- metrics ---------------- */
/* ||A - L L^T||_F / ||A||_F using the lower triangle of A and L.            */
double orc_potrf_berr(int n, const double* A, int lda, const double* L, int ldl);
/* max over right-hand sides of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
double orc_solve_berr(char side, int m, int n, const double* A, int lda,
                      const double* X, int ldx, const double* B, int ldb);

/* ---------------- FLOP model (testing/flops.h) ---------------- */
double orc_flops_potrf(int n);
double orc_flops_trsm(char side, int m, int n);
double orc_flops_potrs(char side, int m, int n);
double orc_flops_posv(char side, int m, int n);
double orc_flops_syrk(int k, int n);
double orc_flops_gemm(int m, int n, int k);

/* ---------------- batch loops (OpenMP over the batch) ---------------- */
int  orc_threads(void);
void orc_set_threads(int t);
/* strided batches; info may be NULL                                          */
void orc_potrf_batch(int n, double* A, int lda, long strideA, int batch, int* info);
void orc_posv_batch(char side, int m, int n, double* A, int lda, long strideA,
                    double* B, int ldb, long strideB, int batch, int* info);
void orc_sposv_batch(char side, int m, int n, const double* A, int lda, long strideA,
                     const double* B, int ldb, long strideB, double* X, int ldx, long strideX,
                     int batch, int* info, int* iter);

/* LAPACK (MKL/OpenBLAS loaded at run time) loops, the reference tests' own CPU
 * comparator (testing/batch_triangular/test_Xposv_batch.cpp:313-326).
 * Return 0 when no LAPACK library could be loaded.                          */
int  orc_lapack_available(void);
const char* orc_lapack_name(void);
int  orc_lapack_potrf_batch(int n, double* A, int lda, long strideA, int batch, int* info);
int  orc_lapack_posv_vbatch(int nrhs, const int* n, double* A, const long* offA, double* B,
                            const long* offB, int count, int* info);
void orc_posv_vbatch(int nrhs, const int* n, double* A, const long* offA, double* B, const long* offB,
                     int count, int* info);
int  orc_lapack_posv_batch(char side, int m, int n, double* A, int lda, long strideA,
                           double* B, int ldb, long strideB, int batch, int* info);

/* generator batch helpers (OpenMP)                                           */
void orc_gen_spd_shift_batch(uint64_t seed, int n, double shift, double* A, int lda,
                             long strideA, long first, int batch);
void orc_gen_spd_cond_batch(uint64_t seed, int n, double lk_lo, double lk_hi, double* A,
                            int lda, long strideA, long first, int batch, double* kappa);
void orc_gen_rhs_batch(uint64_t seed, int rows, int cols, double* B, int ldb, long strideB,
                       long first, int batch);

#ifdef __cplusplus
}
#endif
#endif
