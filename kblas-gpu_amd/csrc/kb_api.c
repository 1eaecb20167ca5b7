/* kb_api.c -- the kblas{S,D}{potrf,potrs,posv,trsm,syrk,gemm}_batch* C entry points,
 * and those of the inverse family (trtri, lauum, potri, poti), TRMM and PSTRF.
 *
 * Front-end behaviour mirrors the reference front-ends
 * (src/batch_triangular/X{potrf,potrs,posv,trsm,syrk,gemm}_batch.cu): workspace
 * sufficiency check (KBLAS_InsufficientWorkspace), Upper / Unit variants
 * rejected with KBLAS_NotImplemented, launch failures reported as
 * KBLAS_UnknownError.  Added here (the reference does no argument checking):
 * negative sizes, too-small leading dimensions / strides, NULL operands and
 * negative batch counts return KBLAS_Error_WrongInput before anything is
 * launched; batchCount == 0 or an empty matrix returns KBLAS_Success (info is
 * zeroed when the order is 0).  Nothing here allocates or synchronises.
 */
#include <stdlib.h>
#include <string.h>

#include "kb_internal.h"
#include "batch_pstrf.h"
#include "batch_ara.h"

#define IMAX(a, b) ((a) > (b) ? (a) : (b))

static int upl(char c) { return c == 'L' || c == 'l' ? 0 : (c == 'U' || c == 'u') ? 1 : -1; }
static int sideof(char c) { return c == 'L' || c == 'l' ? 0 : (c == 'R' || c == 'r') ? 1 : -1; }
static int transof(char c)
{
    if (c == 'N' || c == 'n') return 0;
    if (c == 'T' || c == 't' || c == 'C' || c == 'c') return 1;
    return -1;
}
static int diagof(char c) { return c == 'N' || c == 'n' ? 0 : (c == 'U' || c == 'u') ? 1 : -1; }
static int cplx(char prec) { return prec == 'c' || prec == 'z'; }
/* op code of the generic kernels: N 0, T 1, C 2 (complex; C is T for real data) */
static int opof(char prec, char c)
{
    const int t = transof(c);
    return (t == 1 && cplx(prec) && (c == 'C' || c == 'c')) ? 2 : t;
}
/* scalars as (real, imaginary) pairs */
#define RS(v) ((const double[2]){(double)(v), 0.0})
#define CS(v) ((const double[2]){(double)(v).x, (double)(v).y})

static kb_mat_t mat_strided(const void* p, int ld, long stride)
{
    kb_mat_t m;
    m.base = (void*)p; m.stride = stride; m.ptrs = NULL; m.lds = NULL; m.ld = ld; m.off = 0;
    return m;
}
static kb_mat_t mat_ptrs(const void* pp, int ld)
{
    kb_mat_t m;
    m.base = NULL; m.stride = 0; m.ptrs = (void* const*)pp; m.lds = NULL; m.ld = ld; m.off = 0;
    return m;
}
static kb_mat_t mat_vptrs(const void* pp, const int* lds)
{
    kb_mat_t m;
    m.base = NULL; m.stride = 0; m.ptrs = (void* const*)pp; m.lds = lds; m.ld = 0; m.off = 0;
    return m;
}

static int launched(int hip_err)
{
    if (hip_err == 0) return KBLAS_Success;
    if (hip_err == (int)hipErrorNotSupported) return KBLAS_NotImplemented;
    return KBLAS_UnknownError;
}

static void zero_info(kblasHandle_t h, int* info, int bc)
{
    if (info && bc > 0) kb_launch_fill_int(info, 0, bc, h->stream);
}

/* Kernel for a uniform batch of order n: the tuning table's variant for this
 * (precision, operation, order) when it has one, else the built-in rule
 * (register-resident kernels up to KB_REG_MAX_N, the blocked workgroup
 * kernel above). */
static int chol_dispatch(kblasHandle_t h, char prec, int mode, int n, int nrhs, int side_left, kb_mat_t A,
                         kb_mat_t B, int* info, int bc, int upper)
{
    /* the upper triangle and the complex types: the generic kernels */
    if (upper || cplx(prec)) return kb_launch_gchol(prec, mode, upper, n, nrhs, side_left, A, B, info, bc, h->stream);
    const int v = kb_tune_variant(prec, mode, n);
    if (mode != KB_MODE_POTRF && n > KB_REG_MAX_N && nrhs >= KB_TRSM_REC_NVEC) {
        /* many right-hand sides above the register range: the blocked
         * factorization, then the two solves by half-order recursion with
         * matrix-core GEMM updates (a failed factorization leaves that
         * matrix's B as it was) */
        if (mode == KB_MODE_POSV) {
            const int e = kb_launch_blk_chol(prec, KB_MODE_POTRF, n, 0, side_left, A, B, info, bc, h->stream);
            if (e) return e;
        }
        return kb_launch_trsm_rec_solve(prec, side_left, n, nrhs, A, B, mode == KB_MODE_POSV ? info : NULL, bc,
                                        h->stream);
    }
    if (v == KB_VAR_BLK || n > KB_REG_MAX_N)
        return kb_launch_blk_chol(prec, mode, n, nrhs, side_left, A, B, info, bc, h->stream);
    if (mode != KB_MODE_POTRF && n > 32 && nrhs >= KB_SOLVE_SPLIT_NRHS) {
        /* many right-hand sides: factor, then the matrix-core sweeps (a failed
         * factorization leaves that matrix's B as it was) */
        if (mode == KB_MODE_POSV) {
            const int pv = kb_tune_variant(prec, KB_MODE_POTRF, n);
            const int e = pv == KB_VAR_BLK
                              ? kb_launch_blk_chol(prec, KB_MODE_POTRF, n, 0, side_left, A, B, info, bc, h->stream)
                              : kb_launch_reg_chol_v(prec, KB_MODE_POTRF, n, 0, side_left, A, B, info, bc, pv,
                                                     h->stream);
            if (e) return e;
        }
        return kb_launch_trsm_solve(prec, side_left, n, nrhs, A, B, mode == KB_MODE_POSV ? info : NULL, bc,
                                    h->stream);
    }
    return kb_launch_reg_chol_v(prec, mode, n, nrhs, side_left, A, B, info, bc, v, h->stream);
}

/* ============================== POTRF ================================== */
static int potrf_core(kblasHandle_t h, char prec, char uplo, int n, kb_mat_t A,
                      long strideA, int strided, int bc, int* info)
{
    KB_CHECK_HANDLE(h);
    int u = upl(uplo);
    if (u < 0 || n < 0 || bc < 0 || A.ld < IMAX(1, n)) return KBLAS_Error_WrongInput;
    if (bc == 0) return KBLAS_Success;
    if (strided ? (!A.base || strideA < (long)A.ld * n) : !A.ptrs) return KBLAS_Error_WrongInput;
    if (n == 0) { zero_info(h, info, bc); return KBLAS_Success; }
    return launched(chol_dispatch(h, prec, KB_MODE_POTRF, n, 0, 0, A, A, info, bc, u));
}

int kblasDpotrf_batch(kblasHandle_t h, char uplo, const int n, double** A, int lda, int bc, int* info)
{ return potrf_core(h, 'd', uplo, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasSpotrf_batch(kblasHandle_t h, char uplo, const int n, float** A, int lda, int bc, int* info)
{ return potrf_core(h, 's', uplo, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasDpotrf_batch_strided(kblasHandle_t h, char uplo, const int n, double* A, int lda, long sA,
                              int bc, int* info)
{ return potrf_core(h, 'd', uplo, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasSpotrf_batch_strided(kblasHandle_t h, char uplo, const int n, float* A, int lda, long sA,
                              int bc, int* info)
{ return potrf_core(h, 's', uplo, n, mat_strided(A, lda, sA), sA, 1, bc, info); }

/* ========================= POTRS / POSV ================================= */
/* B is m x n.  side Right: X A = B, A of order n, m right-hand sides (the
 * reference's only side, src/batch_triangular/Xpotrs_batch_drivers.cuh:40-43).
 * side Left: A X = B, A of order m, n right-hand sides (LAPACK layout). */
static int solve_core(kblasHandle_t h, char prec, int mode, char side, char uplo, int m, int n,
                      kb_mat_t A, long strideA, kb_mat_t B, long strideB, int strided,
                      int bc, int* info)
{
    KB_CHECK_HANDLE(h);
    int sd = sideof(side), u = upl(uplo);
    if (sd < 0 || u < 0 || m < 0 || n < 0 || bc < 0) return KBLAS_Error_WrongInput;
    int order = sd == 0 ? m : n;
    int nrhs = sd == 0 ? n : m;
    if (A.ld < IMAX(1, order) || B.ld < IMAX(1, m)) return KBLAS_Error_WrongInput;
    if (bc == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || strideB < (long)B.ld * n || strideA < 0) return KBLAS_Error_WrongInput;
        if (mode == KB_MODE_POSV && strideA < (long)A.ld * order) return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    if (order == 0) { if (mode == KB_MODE_POSV) zero_info(h, info, bc); return KBLAS_Success; }
    if (nrhs == 0 && mode == KB_MODE_POTRS) return KBLAS_Success;
    return launched(chol_dispatch(h, prec, mode, order, nrhs, sd == 0, A, B, info, bc, u));
}

int kblasDpotrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                      const double** A, int lda, double** B, int ldb, int bc)
{ return solve_core(h, 'd', KB_MODE_POTRS, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, NULL); }
int kblasSpotrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                      const float** A, int lda, float** B, int ldb, int bc)
{ return solve_core(h, 's', KB_MODE_POTRS, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, NULL); }
int kblasDpotrs_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n,
                              const double* A, int lda, long sA, double* B, int ldb, long sB, int bc)
{ return solve_core(h, 'd', KB_MODE_POTRS, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, NULL); }
int kblasSpotrs_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n,
                              const float* A, int lda, long sA, float* B, int ldb, long sB, int bc)
{ return solve_core(h, 's', KB_MODE_POTRS, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, NULL); }

int kblasDposv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                     double** A, int lda, double** B, int ldb, int bc, int* info)
{ return solve_core(h, 'd', KB_MODE_POSV, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, info); }
int kblasSposv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                     float** A, int lda, float** B, int ldb, int bc, int* info)
{ return solve_core(h, 's', KB_MODE_POSV, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, info); }
int kblasDposv_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             double* A, int lda, long sA, double* B, int ldb, long sB, int bc, int* info)
{ return solve_core(h, 'd', KB_MODE_POSV, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, info); }
int kblasSposv_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n,
                             float* A, int lda, long sA, float* B, int ldb, long sB, int bc, int* info)
{ return solve_core(h, 's', KB_MODE_POSV, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, info); }

/* ======================= mixed-precision POSV ============================ */
static int sposv_core(kblasHandle_t h, char side, char uplo, int m, int n, kb_mat_t A, long sA,
                      kb_mat_t B, long sB, kb_mat_t X, long sX, int strided, int bc,
                      int* info, int* iter)
{
    KB_CHECK_HANDLE(h);
    int sd = sideof(side), u = upl(uplo);
    if (sd < 0 || u < 0 || m < 0 || n < 0 || bc < 0) return KBLAS_Error_WrongInput;
    int order = sd == 0 ? m : n;
    int nrhs = sd == 0 ? n : m;
    if (A.ld < IMAX(1, order) || B.ld < IMAX(1, m) || X.ld < IMAX(1, m)) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || !X.base || sA < (long)A.ld * order || sB < 0 ||
            sX < (long)X.ld * n) return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs || !X.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    if (order == 0 || nrhs == 0) {
        zero_info(h, info, bc);
        if (iter) kb_launch_fill_int(iter, 0, bc, h->stream);
        if (order == 0 || nrhs == 0) return KBLAS_Success;
    }
    if (order <= KB_REG_MAX_N)
        return launched(kb_launch_reg_sposv_v(order, nrhs, sd == 0, A, B, X, info, iter, bc,
                                              kb_tune_variant('d', KB_TUNE_SPOSV, order), h->stream));
    return launched(kb_launch_blk_sposv(order, nrhs, sd == 0, A, B, X, info, iter, bc, h->stream));
}

int kblasDsposv_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n,
                              double* A, int lda, long sA, const double* B, int ldb, long sB,
                              double* X, int ldx, long sX, int bc, int* info, int* iter)
{
    return sposv_core(h, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB,
                      mat_strided(X, ldx, sX), sX, 1, bc, info, iter);
}
int kblasDsposv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n,
                      double** A, int lda, const double** B, int ldb, double** X, int ldx,
                      int bc, int* info, int* iter)
{
    return sposv_core(h, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0,
                      mat_ptrs(X, ldx), 0, 0, bc, info, iter);
}

/* ================================ TRSM =================================== */
static int trsm_core(kblasHandle_t h, char prec, char side, char uplo, char trans, char diag,
                     int m, int n, const double* alpha, kb_mat_t A, long sA, kb_mat_t B, long sB,
                     int strided, int bc)
{
    KB_CHECK_HANDLE(h);
    int sd = sideof(side), u = upl(uplo), t = opof(prec, trans), dg = diagof(diag);
    if (sd < 0 || u < 0 || t < 0 || dg < 0 || m < 0 || n < 0 || bc < 0) return KBLAS_Error_WrongInput;
    int ka = sd == 0 ? m : n;
    if (A.ld < IMAX(1, ka) || B.ld < IMAX(1, m)) return KBLAS_Error_WrongInput;
    if (bc == 0 || m == 0 || n == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || sA < 0 || sB < (long)B.ld * n) return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    return launched(kb_launch_trsm_x(prec, sd == 0, u, t, dg, m, n, alpha, A, B, NULL, NULL, bc, h->stream));
}

int kblasDtrsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                     const int n, const double alpha, const double** A, int lda, double** B, int ldb, int bc)
{ return trsm_core(h, 'd', side, uplo, trans, diag, m, n, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasStrsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                     const int n, const float alpha, const float** A, int lda, float** B, int ldb, int bc)
{ return trsm_core(h, 's', side, uplo, trans, diag, m, n, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasDtrsm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                             const int n, const double alpha, const double* A, int lda, long sA,
                             double* B, int ldb, long sB, int bc)
{ return trsm_core(h, 'd', side, uplo, trans, diag, m, n, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasStrsm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                             const int n, const float alpha, const float* A, int lda, long sA,
                             float* B, int ldb, long sB, int bc)
{ return trsm_core(h, 's', side, uplo, trans, diag, m, n, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }

/* ================================ SYRK =================================== */
static int syrk_core(kblasHandle_t h, char prec, char uplo, char trans, int m, int n, const double* alpha,
                     kb_mat_t A, long sA, const double* beta, kb_mat_t B, long sB, int strided, int bc)
{
    KB_CHECK_HANDLE(h);
    int u = upl(uplo), t = opof(prec, trans);
    /* symmetric rank-k update: no conjugate-transposed form for complex data */
    if (u < 0 || t < 0 || t == 2 || m < 0 || n < 0 || bc < 0) return KBLAS_Error_WrongInput;
    if (A.ld < IMAX(1, t ? n : m) || B.ld < IMAX(1, m)) return KBLAS_Error_WrongInput;
    if (bc == 0 || m == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || sA < 0 || sB < (long)B.ld * m) return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    return launched(kb_launch_syrk_x(prec, u, t, m, n, alpha, A, beta, B, NULL, NULL, bc, h->stream));
}

int kblasDsyrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n, const double alpha,
                     const double** A, int lda, const double beta, double** B, int ldb, int bc)
{ return syrk_core(h, 'd', uplo, trans, m, n, RS(alpha), mat_ptrs(A, lda), 0, RS(beta), mat_ptrs(B, ldb), 0, 0, bc); }
int kblasSsyrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n, const float alpha,
                     const float** A, int lda, const float beta, float** B, int ldb, int bc)
{ return syrk_core(h, 's', uplo, trans, m, n, RS(alpha), mat_ptrs(A, lda), 0, RS(beta), mat_ptrs(B, ldb), 0, 0, bc); }
int kblasDsyrk_batch_strided(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                             const double alpha, const double* A, int lda, long sA,
                             const double beta, double* B, int ldb, long sB, int bc)
{ return syrk_core(h, 'd', uplo, trans, m, n, RS(alpha), mat_strided(A, lda, sA), sA, RS(beta), mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasSsyrk_batch_strided(kblasHandle_t h, char uplo, char trans, const int m, const int n,
                             const float alpha, const float* A, int lda, long sA,
                             const float beta, float* B, int ldb, long sB, int bc)
{ return syrk_core(h, 's', uplo, trans, m, n, RS(alpha), mat_strided(A, lda, sA), sA, RS(beta), mat_strided(B, ldb, sB), sB, 1, bc); }

/* ================================ GEMM =================================== */
static int gemm_core(kblasHandle_t h, char prec, char ta, char tb, int m, int n, int k, const double* alpha,
                     kb_mat_t A, long sA, kb_mat_t B, long sB, const double* beta, kb_mat_t C, long sC,
                     int strided, int bc)
{
    KB_CHECK_HANDLE(h);
    int a = opof(prec, ta), b = opof(prec, tb);
    if (a < 0 || b < 0 || m < 0 || n < 0 || k < 0) return KBLAS_Error_WrongInput;
    /* src/batch_triangular/Xgemm_batch_core.cuh:181-182 rejects batchCount < 1 */
    if (bc < 0) return KBLAS_Error_WrongInput;
    if (A.ld < IMAX(1, a ? k : m) || B.ld < IMAX(1, b ? n : k) || C.ld < IMAX(1, m))
        return KBLAS_Error_WrongInput;
    if (bc == 0 || m == 0 || n == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || !C.base || sA < 0 || sB < 0 || sC < (long)C.ld * n)
            return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs || !C.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    return launched(kb_launch_gemm_x(prec, a, b, m, n, k, alpha, A, B, beta, C, NULL, NULL, NULL, bc, h->stream));
}

int kblasDgemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                     const double alpha, const double** A, int lda, const double** B, int ldb,
                     const double beta, double** C, int ldc, int bc)
{ return gemm_core(h, 'd', ta, tb, m, n, k, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, RS(beta), mat_ptrs(C, ldc), 0, 0, bc); }
int kblasSgemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                     const float alpha, const float** A, int lda, const float** B, int ldb,
                     const float beta, float** C, int ldc, int bc)
{ return gemm_core(h, 's', ta, tb, m, n, k, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, RS(beta), mat_ptrs(C, ldc), 0, 0, bc); }
int kblasDgemm_batch_strided(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                             const double alpha, const double* A, int lda, long sA,
                             const double* B, int ldb, long sB, const double beta,
                             double* C, int ldc, long sC, int bc)
{ return gemm_core(h, 'd', ta, tb, m, n, k, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, RS(beta), mat_strided(C, ldc, sC), sC, 1, bc); }
int kblasSgemm_batch_strided(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k,
                             const float alpha, const float* A, int lda, long sA,
                             const float* B, int ldb, long sB, const float beta,
                             float* C, int ldc, long sC, int bc)
{ return gemm_core(h, 's', ta, tb, m, n, k, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, RS(beta), mat_strided(C, ldc, sC), sC, 1, bc); }

/* ============================ non-uniform ================================ */
/* Per-matrix sizes live on the device, so only host-visible arguments are
 * validated here; a matrix with a non-positive size is skipped by the
 * kernels (its info is 0). */
static int vcheck(kblasHandle_t h, int bc, const void* a, const void* b)
{
    if (!h) return KBLAS_Error_NotInitialized;
    if (bc < 0) return KBLAS_Error_WrongInput;
    if (bc > 0 && (!a || !b)) return KBLAS_Error_WrongInput;
    return KBLAS_Success;
}

int kblasDgemm_vbatch(kblasHandle_t h, char ta, char tb, int* m, int* n, int* k, int max_m, int max_n,
                      int max_k, const double alpha, const double** A, int* lda, const double** B,
                      int* ldb, const double beta, double** C, int* ldc, int bc)
{
    int st = vcheck(h, bc, A, C);
    if (st != KBLAS_Success) return st;
    int a = transof(ta), b = transof(tb);
    if (a < 0 || b < 0 || !m || !n || !k || !lda || !ldb || !ldc || !B) return KBLAS_Error_WrongInput;
    if (bc == 0) return KBLAS_Success;
    /* maxima bound the sizes (as the reference's MAGMA path assumes); 0 when unknown */
    const int bounded = max_m > 0 && max_n > 0 && max_k > 0;
    return launched(kb_launch_gemm('d', a, b, bounded ? max_m : 0, bounded ? max_n : 0, bounded ? max_k : 0, alpha,
                                   mat_vptrs(A, lda), mat_vptrs(B, ldb), beta, mat_vptrs(C, ldc), m, n, k, bc,
                                   h->stream));
}
int kblasSgemm_vbatch(kblasHandle_t h, char ta, char tb, int* m, int* n, int* k, int max_m, int max_n,
                      int max_k, const float alpha, const float** A, int* lda, const float** B,
                      int* ldb, const float beta, float** C, int* ldc, int bc)
{
    int st = vcheck(h, bc, A, C);
    if (st != KBLAS_Success) return st;
    int a = transof(ta), b = transof(tb);
    if (a < 0 || b < 0 || !m || !n || !k || !lda || !ldb || !ldc || !B) return KBLAS_Error_WrongInput;
    if (bc == 0) return KBLAS_Success;
    /* maxima bound the sizes (as the reference's MAGMA path assumes); 0 when unknown */
    const int bounded = max_m > 0 && max_n > 0 && max_k > 0;
    return launched(kb_launch_gemm('s', a, b, bounded ? max_m : 0, bounded ? max_n : 0, bounded ? max_k : 0, alpha,
                                   mat_vptrs(A, lda), mat_vptrs(B, ldb), beta, mat_vptrs(C, ldc), m, n, k, bc,
                                   h->stream));
}

static int syrk_v(kblasHandle_t h, char prec, char uplo, char trans, int* m, int* n, int max_m, int max_n,
                  double alpha, void* A, int* lda, double beta, void* B, int* ldb, int bc)
{
    int st = vcheck(h, bc, A, B);
    if (st != KBLAS_Success) return st;
    int u = upl(uplo), t = transof(trans);
    if (u < 0 || t < 0 || !m || !n || !lda || !ldb) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    const int bounded = max_m > 0 && max_n > 0;
    return launched(kb_launch_syrk(prec, t, bounded ? max_m : 0, bounded ? max_n : 0, alpha, mat_vptrs(A, lda), beta,
                                   mat_vptrs(B, ldb), m, n, bc, h->stream));
}
int kblasDsyrk_vbatch(kblasHandle_t h, char uplo, char trans, int* m, int* n, int max_m, int max_n,
                      double alpha, double** A, int* lda, double beta, double** B, int* ldb, int bc)
{ return syrk_v(h, 'd', uplo, trans, m, n, max_m, max_n, alpha, A, lda, beta, B, ldb, bc); }
int kblasSsyrk_vbatch(kblasHandle_t h, char uplo, char trans, int* m, int* n, int max_m, int max_n,
                      float alpha, float** A, int* lda, float beta, float** B, int* ldb, int bc)
{ return syrk_v(h, 's', uplo, trans, m, n, max_m, max_n, alpha, A, lda, beta, B, ldb, bc); }

static int trsm_v(kblasHandle_t h, char prec, char side, char uplo, char trans, char diag, int* m, int* n,
                  int max_m, int max_n, double alpha, void* A, int* lda, void* B, int* ldb, int bc)
{
    int st = vcheck(h, bc, A, B);
    if (st != KBLAS_Success) return st;
    int sd = sideof(side), u = upl(uplo), t = transof(trans), dg = diagof(diag);
    if (sd < 0 || u < 0 || t < 0 || dg < 0 || !m || !n || !lda || !ldb) return KBLAS_Error_WrongInput;
    if (u == 1 || dg == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    const int bounded = max_m > 0 && max_n > 0;
    return launched(kb_launch_trsm(prec, sd == 0, t, dg, bounded ? max_m : 0, bounded ? max_n : 0, alpha,
                                   mat_vptrs(A, lda), mat_vptrs(B, ldb), m, n, bc, h->stream));
}
int kblasDtrsm_vbatch(kblasHandle_t h, char side, char uplo, char trans, char diag, int* m, int* n,
                      int max_m, int max_n, double alpha, double** A, int* lda, double** B, int* ldb, int bc)
{ return trsm_v(h, 'd', side, uplo, trans, diag, m, n, max_m, max_n, alpha, A, lda, B, ldb, bc); }
int kblasStrsm_vbatch(kblasHandle_t h, char side, char uplo, char trans, char diag, int* m, int* n,
                      int max_m, int max_n, float alpha, float** A, int* lda, float** B, int* ldb, int bc)
{ return trsm_v(h, 's', side, uplo, trans, diag, m, n, max_m, max_n, alpha, A, lda, B, ldb, bc); }

/* ---- non-uniform Cholesky family: workspace = the size-sorted queue ---- */
void kb_ws_need_vbatch(kb_ws_state* ws, int batchCount)
{
    memset(ws, 0, sizeof(*ws));
    ws->d_data = kb_vchol_ws_bytes(batchCount > 0 ? batchCount : 0);
}

void kblas_potrf_vbatch_wsquery(kblasHandle_t h, int max_n, int batchCount)
{
    (void)max_n;
    if (!h) return;
    kb_ws_state need;
    kb_ws_need_vbatch(&need, batchCount);
    kb_ws_request(h, &need);
}
void kblas_posv_vbatch_wsquery(kblasHandle_t h, int max_n, int batchCount)
{ kblas_potrf_vbatch_wsquery(h, max_n, batchCount); }

/* Order at and below which a non-uniform matrix takes the wavefront phase:
 * the tuning table's "vsplit" entry for max_n (the mixed mode keeps
 * KB_REG_MAX_N: above it the mixed path is dsposv's full-precision one). */
static int vsplit(char prec, int mode, int max_n)
{
    if (prec == 'm' || mode == KB_MODE_SPOSV) return KB_REG_MAX_N;
    const int v = kb_tune_variant(prec, KB_TUNE_VSPLIT, max_n > 0 ? max_n : 0);
    return (v >= 0 && v <= KB_REG_MAX_N) ? v : KB_REG_MAX_N;
}

static int vchol(kblasHandle_t h, char prec, int mode, char side, char uplo, int* m, int* n, int max_m, int max_n,
                 void* A, int* lda, void* B, int* ldb, void* X, int* ldx, int bc, int* info, int* iter)
{
    if (!h) return KBLAS_Error_NotInitialized;
    int u = upl(uplo);
    int sd = mode == KB_MODE_POTRF ? 1 : sideof(side);
    if (u < 0 || sd < 0 || bc < 0 || !n || !A || !lda) return KBLAS_Error_WrongInput;
    if (mode != KB_MODE_POTRF && (!m || !B || !ldb)) return KBLAS_Error_WrongInput;
    if (mode == KB_MODE_SPOSV && (!X || !ldx)) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    kb_ws_state need;
    kb_ws_need_vbatch(&need, bc);
    if (!kb_ws_sufficient(&need, &h->allocated)) return KBLAS_InsufficientWorkspace;
    kb_mat_t Am = mat_vptrs(A, lda);
    kb_mat_t Bm = B ? mat_vptrs(B, ldb) : Am;
    kb_mat_t Xm = X ? mat_vptrs(X, ldx) : Bm;
    /* right-hand sides: the columns of B (side L) or its rows (side R) */
    const int max_nrhs = mode == KB_MODE_POTRF ? 0 : (sd == 0 ? max_n : max_m);
    return launched(kb_launch_vchol_split(prec, mode, m, n, sd == 0, Am, Bm, Xm, info, iter, bc, max_nrhs,
                                          h->d_data, vsplit(prec, mode, max_n), h->stream));
}

int kblasDpotrf_vbatch(kblasHandle_t h, char uplo, int* n, int max_n, double** A, int* lda, int bc, int* info)
{ return vchol(h, 'd', KB_MODE_POTRF, 'R', uplo, NULL, n, 0, max_n, A, lda, NULL, NULL, NULL, NULL, bc, info, NULL); }
int kblasSpotrf_vbatch(kblasHandle_t h, char uplo, int* n, int max_n, float** A, int* lda, int bc, int* info)
{ return vchol(h, 's', KB_MODE_POTRF, 'R', uplo, NULL, n, 0, max_n, A, lda, NULL, NULL, NULL, NULL, bc, info, NULL); }
int kblasDpotrs_vbatch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                       double** A, int* lda, double** B, int* ldb, int bc)
{ return vchol(h, 'd', KB_MODE_POTRS, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, NULL, NULL, bc, NULL, NULL); }
int kblasDposv_vbatch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                      double** A, int* lda, double** B, int* ldb, int bc, int* info)
{ return vchol(h, 'd', KB_MODE_POSV, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, NULL, NULL, bc, info, NULL); }
int kblasSposv_vbatch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                      float** A, int* lda, float** B, int* ldb, int bc, int* info)
{ return vchol(h, 's', KB_MODE_POSV, side, uplo, m, n, max_m, max_n, A, lda, B, ldb, NULL, NULL, bc, info, NULL); }
/* Mixed non-uniform POSV: the queue is required; when the workspace also
 * holds the FP32 slots (kblas_sposv_vbatch_wsquery) the orders above 64 run
 * the mixed method, else dsposv's full-precision path (ITER = -1). */
void kblas_sposv_vbatch_wsquery(kblasHandle_t h, int max_nrhs, int max_order, int batchCount)
{
    if (!h) return;
    kb_ws_state need;
    memset(&need, 0, sizeof need);
    need.d_data = kb_vsposv_ws_bytes(batchCount > 0 ? batchCount : 0, max_order, max_nrhs);
    kb_ws_request(h, &need);
}

int kblasDsposv_vbatch(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                       double** A, int* lda, const double** B, int* ldb, double** X, int* ldx, int bc,
                       int* info, int* iter)
{
    if (!h) return KBLAS_Error_NotInitialized;
    int u = upl(uplo);
    int sd = sideof(side);
    if (u < 0 || sd < 0 || bc < 0 || !n || !m || !A || !lda || !B || !ldb || !X || !ldx) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    kb_ws_state need;
    kb_ws_need_vbatch(&need, bc);
    if (!kb_ws_sufficient(&need, &h->allocated)) return KBLAS_InsufficientWorkspace;
    const int max_order = sd == 0 ? max_m : max_n, max_nrhs = sd == 0 ? max_n : max_m;
    return launched(kb_launch_vsposv(m, n, sd == 0, mat_vptrs(A, lda), mat_vptrs((void*)B, ldb), mat_vptrs(X, ldx),
                                     info, iter, bc, h->d_data, h->allocated.d_data, max_order, max_nrhs,
                                     h->stream));
}

/* kblasDposv_vbatch with the queue in the library's fixed per-device scratch
 * area instead of the handle's workspace: no allocation, no workspace query.
 * At most KB_NOALLOC_MAX_BATCH matrices (KBLAS_InsufficientWorkspace above). */
int kblasDposv_vbatch_noalloc(kblasHandle_t h, char side, char uplo, int* m, int* n, int max_m, int max_n,
                              double** A, int* lda, double** B, int* ldb, int bc, int* info)
{
    if (!h) return KBLAS_Error_NotInitialized;
    int u = upl(uplo);
    int sd = sideof(side);
    if (u < 0 || sd < 0 || bc < 0 || !n || !A || !lda || !m || !B || !ldb) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    if (bc > KB_NOALLOC_MAX_BATCH) {
        fprintf(stderr, "kblasDposv_vbatch_noalloc: batchCount %d exceeds the fixed scratch capacity of %d "
                        "matrices; use kblasDposv_vbatch with a workspace\n", bc, KB_NOALLOC_MAX_BATCH);
        return KBLAS_InsufficientWorkspace;
    }
    int err = 0;
    void* ws = kb_noalloc_scratch(&err);
    if (!ws) return kb_hip_status(err ? err : 1);
    kb_mat_t Am = mat_vptrs(A, lda);
    kb_mat_t Bm = mat_vptrs(B, ldb);
    return launched(kb_launch_vchol_split('d', KB_MODE_POSV, m, n, sd == 0, Am, Bm, Bm, info, NULL, bc,
                                          sd == 0 ? max_n : max_m, ws, vsplit('d', KB_MODE_POSV, max_n), h->stream));
}

/* ============ inverse family: TRTRI / LAUUM / POTRI / POTI ============== */
/* Front-ends of src/batch_triangular/X{trtri,lauum,potri,poti}_batch.cu: Upper
 * returns KBLAS_NotImplemented (Xpotri_batch_drivers.cuh:38-41,
 * Xlauum_batch_drivers.cuh:39-42, Xpoti_batch_drivers.cuh:39-42).  POTI is
 * POTRF then POTRI (Xpoti_batch_drivers.cuh:80-110); matrices whose
 * factorization failed keep POTRF's info and factor.  Orders above
 * KB_INV_MAX_N return KBLAS_NotImplemented. */
enum { INV_TRTRI, INV_LAUUM, INV_POTRI, INV_POTI };
static int inv_core(kblasHandle_t h, char prec, int what, char uplo, char diag, int n, kb_mat_t A,
                    long strideA, int strided, int bc, int* info)
{
    KB_CHECK_HANDLE(h);
    int u = upl(uplo), dg = what == INV_TRTRI ? diagof(diag) : 0;
    if (u < 0 || dg < 0 || n < 0 || bc < 0 || A.ld < IMAX(1, n)) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0) return KBLAS_Success;
    if (strided ? (!A.base || strideA < (long)A.ld * n) : !A.ptrs) return KBLAS_Error_WrongInput;
    if (n == 0) { zero_info(h, info, bc); return KBLAS_Success; }
    if (n > KB_INV_MAX_N) return KBLAS_NotImplemented;
    if (what == INV_POTI) {
        int st = chol_dispatch(h, prec, KB_MODE_POTRF, n, 0, 0, A, A, info, bc, 0);
        if (st) return launched(st);
        return launched(kb_launch_inv(prec, KB_INV_OP_POTRI, 0, n, A, info, info, bc, h->stream));
    }
    int op = what == INV_TRTRI ? KB_INV_OP_TRTRI : what == INV_LAUUM ? KB_INV_OP_LAUUM : KB_INV_OP_POTRI;
    return launched(kb_launch_inv(prec, op, dg, n, A, info, NULL, bc, h->stream));
}

#define KB_INV_ENTRY(P, PREC, T, NAME, WHAT)                                                                  \
    int kblas##P##NAME##_batch(kblasHandle_t h, char uplo, const int n, T** A, int lda, int bc, int* info)   \
    { return inv_core(h, PREC, WHAT, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }                      \
    int kblas##P##NAME##_batch_strided(kblasHandle_t h, char uplo, const int n, T* A, int lda, long sA,       \
                                       int bc, int* info)                                                     \
    { return inv_core(h, PREC, WHAT, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
KB_INV_ENTRY(D, 'd', double, lauum, INV_LAUUM)
KB_INV_ENTRY(S, 's', float, lauum, INV_LAUUM)
KB_INV_ENTRY(D, 'd', double, potri, INV_POTRI)
KB_INV_ENTRY(S, 's', float, potri, INV_POTRI)
KB_INV_ENTRY(D, 'd', double, poti, INV_POTI)
KB_INV_ENTRY(S, 's', float, poti, INV_POTI)
#undef KB_INV_ENTRY

int kblasDtrtri_batch(kblasHandle_t h, char uplo, char diag, const int n, double** A, int lda, int bc, int* info)
{ return inv_core(h, 'd', INV_TRTRI, uplo, diag, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasStrtri_batch(kblasHandle_t h, char uplo, char diag, const int n, float** A, int lda, int bc, int* info)
{ return inv_core(h, 's', INV_TRTRI, uplo, diag, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasDtrtri_batch_strided(kblasHandle_t h, char uplo, char diag, const int n, double* A, int lda, long sA,
                              int bc, int* info)
{ return inv_core(h, 'd', INV_TRTRI, uplo, diag, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasStrtri_batch_strided(kblasHandle_t h, char uplo, char diag, const int n, float* A, int lda, long sA,
                              int bc, int* info)
{ return inv_core(h, 's', INV_TRTRI, uplo, diag, n, mat_strided(A, lda, sA), sA, 1, bc, info); }

/* ================================ TRMM ==================================== */
/* src/batch_triangular/Xtrmm_batch.cu:63-105 serves Left / Lower / NonUnit
 * only; here Right and Unit are served too, Upper returns NotImplemented. */
static int trmm_core(kblasHandle_t h, char prec, char side, char uplo, char trans, char diag,
                     int m, int n, const double* alpha, kb_mat_t A, long sA, kb_mat_t B, long sB,
                     int strided, int bc)
{
    KB_CHECK_HANDLE(h);
    int sd = sideof(side), u = upl(uplo), t = opof(prec, trans), dg = diagof(diag);
    if (sd < 0 || u < 0 || t < 0 || dg < 0 || m < 0 || n < 0 || bc < 0) return KBLAS_Error_WrongInput;
    int ka = sd == 0 ? m : n;
    if (A.ld < IMAX(1, ka) || B.ld < IMAX(1, m)) return KBLAS_Error_WrongInput;
    if (u == 1) return KBLAS_NotImplemented;
    if (bc == 0 || m == 0 || n == 0) return KBLAS_Success;
    if (strided) {
        if (!A.base || !B.base || sA < 0 || sB < (long)B.ld * n) return KBLAS_Error_WrongInput;
    } else if (!A.ptrs || !B.ptrs) {
        return KBLAS_Error_WrongInput;
    }
    return launched(kb_launch_trmm(prec, sd == 0, t, dg, m, n, alpha, A, B, bc, h->stream));
}

int kblasDtrmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                     const double alpha, const double** A, int lda, double** B, int ldb, int bc)
{ return trmm_core(h, 'd', side, uplo, trans, diag, m, n, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasStrmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n,
                     const float alpha, const float** A, int lda, float** B, int ldb, int bc)
{ return trmm_core(h, 's', side, uplo, trans, diag, m, n, RS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasDtrmm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                             const int n, const double alpha, const double* A, int lda, long sA, double* B,
                             int ldb, long sB, int bc)
{ return trmm_core(h, 'd', side, uplo, trans, diag, m, n, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasStrmm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m,
                             const int n, const float alpha, const float* A, int lda, long sA, float* B,
                             int ldb, long sB, int bc)
{ return trmm_core(h, 's', side, uplo, trans, diag, m, n, RS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }

/* ================================ PSTRF ==================================== */
/* src/batch_svd/batch_pstrf.cu:160-223: n > 64 returns KBLAS_Error_WrongInput;
 * the strides are int, as declared there.  Added: negative sizes, small lda
 * or strides and NULL operands return KBLAS_Error_WrongInput; num_ops == 0 and
 * n == 0 return KBLAS_Success (ranks zeroed for n == 0). */
static int pstrf_core(kblasHandle_t h, char prec, int n, kb_mat_t A, long sA, int strided, int** pivp, int* pivs,
                      long spiv, int* ranks, int bc)
{
    KB_CHECK_HANDLE(h);
    if (n < 0 || n > 64 || bc < 0 || A.ld < IMAX(1, n)) return KBLAS_Error_WrongInput;
    if (bc == 0) return KBLAS_Success;
    if (strided ? (!A.base || sA < (long)A.ld * n || !pivs || spiv < n) : (!A.ptrs || !pivp))
        return KBLAS_Error_WrongInput;
    if (n == 0) { if (ranks) kb_launch_fill_int(ranks, 0, bc, h->stream); return KBLAS_Success; }
    return launched(kb_launch_pstrf(prec, n, A, (int* const*)pivp, pivs, spiv, ranks, bc, h->stream));
}

int kblasDpstrf_batch_strided(kblasHandle_t h, int n, double* A, int lda, int stride_a, int* piv, int stride_piv,
                              int* ranks, int num_ops)
{ return pstrf_core(h, 'd', n, mat_strided(A, lda, stride_a), stride_a, 1, NULL, piv, stride_piv, ranks, num_ops); }
int kblasSpstrf_batch_strided(kblasHandle_t h, int n, float* A, int lda, int stride_a, int* piv, int stride_piv,
                              int* ranks, int num_ops)
{ return pstrf_core(h, 's', n, mat_strided(A, lda, stride_a), stride_a, 1, NULL, piv, stride_piv, ranks, num_ops); }
int kblasDpstrf_batch(kblasHandle_t h, int n, double** A, int lda, int** piv, int* ranks, int num_ops)
{ return pstrf_core(h, 'd', n, mat_ptrs(A, lda), 0, 0, piv, NULL, 0, ranks, num_ops); }
int kblasSpstrf_batch(kblasHandle_t h, int n, float** A, int lda, int** piv, int* ranks, int num_ops)
{ return pstrf_core(h, 's', n, mat_ptrs(A, lda), 0, 0, piv, NULL, 0, ranks, num_ops); }

/* ============================ ARA SPD helpers ============================ */
/* src/batch_svd/batch_ara.cu:497-561, ara_util.cuh:417-466,560-640,765-808:
 * bs other than 16 / 32 and max_n > 32 (FP32 Gram product) return
 * KBLAS_NotImplemented as there.  Added: negative counts and NULL arrays
 * return KBLAS_Error_WrongInput; num_ops == 0 returns KBLAS_Success. */
static int ara_potrf(char rprec, int* dims, double** A, int* lda, void** R, int* ldr, double* diag, int bs,
                     int* ranks, int num_ops, hipStream_t stream)
{
    if (bs != 16 && bs != 32) return KBLAS_NotImplemented;
    if (num_ops < 0) return KBLAS_Error_WrongInput;
    if (num_ops == 0) return KBLAS_Success;
    if (!dims || !A || !lda || !R || !ldr || !diag || !ranks) return KBLAS_Error_WrongInput;
    return launched(kb_launch_ara_potrf(rprec, dims, A, lda, (void* const*)R, ldr, diag, bs, ranks, num_ops, stream));
}

int kblas_sara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, float** R_batch, int* ldr_batch,
                                 double* diag_R, int bs, int* block_ranks, int num_ops, hipStream_t stream)
{ return ara_potrf('s', op_samples, A_batch, lda_batch, (void**)R_batch, ldr_batch, diag_R, bs, block_ranks, num_ops, stream); }
int kblas_dara_fused_potrf_batch(int* op_samples, double** A_batch, int* lda_batch, double** R_batch,
                                 int* ldr_batch, double* diag_R, int bs, int* block_ranks, int num_ops,
                                 hipStream_t stream)
{ return ara_potrf('d', op_samples, A_batch, lda_batch, (void**)R_batch, ldr_batch, diag_R, bs, block_ranks, num_ops, stream); }

int kblas_sara_mp_syrk_batch(kblasHandle_t h, int* m, int* n, int max_m, int max_n, const float** A, int* lda,
                             double** B, int* ldb, int batchCount)
{
    KB_CHECK_HANDLE(h);
    (void)max_m;
    if (max_n > 32) return KBLAS_NotImplemented;
    if (batchCount < 0) return KBLAS_Error_WrongInput;
    if (batchCount == 0) return KBLAS_Success;
    if (!m || !n || !A || !lda || !B || !ldb) return KBLAS_Error_WrongInput;
    return launched(kb_launch_ara_syrk(m, n, A, lda, B, ldb, batchCount, h->stream));
}
int kblas_dara_mp_syrk_batch(kblasHandle_t h, int* m, int* n, int max_m, int max_n, const double** A, int* lda,
                             double** B, int* ldb, int batchCount)
{
    return kblasDgemm_vbatch(h, 'T', 'N', n, n, m, max_n, max_n, max_m, 1.0, A, lda, A, lda, 0.0, B, ldb,
                             batchCount);
}

static int ara_trsm(kblasHandle_t h, char prec, void** B, int* ldb, void** A, int* lda, int* rows, int* cols,
                    int num_ops, int max_rows, int bs)
{
    KB_CHECK_HANDLE(h);
    if (bs != 16 && bs != 32) return KBLAS_NotImplemented;
    if (num_ops < 0 || max_rows < 0) return KBLAS_Error_WrongInput;
    if (num_ops == 0 || max_rows == 0) return KBLAS_Success;
    if (!B || !ldb || !A || !lda || !rows || !cols) return KBLAS_Error_WrongInput;
    return launched(kb_launch_ara_trsm(prec, (void* const*)B, ldb, (void* const*)A, lda, rows, cols, bs, max_rows,
                                       num_ops, h->stream));
}
int kblas_sara_trsm_batch(kblasHandle_t h, float** B_batch, int* ldb_batch, float** A_batch, int* lda_batch,
                          int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs)
{ return ara_trsm(h, 's', (void**)B_batch, ldb_batch, (void**)A_batch, lda_batch, rows_batch, cols_batch, num_ops, max_rows, bs); }
int kblas_dara_trsm_batch(kblasHandle_t h, double** B_batch, int* ldb_batch, double** A_batch, int* lda_batch,
                          int* rows_batch, int* cols_batch, int num_ops, int max_rows, int bs)
{ return ara_trsm(h, 'd', (void**)B_batch, ldb_batch, (void**)A_batch, lda_batch, rows_batch, cols_batch, num_ops, max_rows, bs); }

/* ====================== workspace queries (all zero) ===================== */
/* The uniform paths run without any workspace: no pointer arrays are built
 * (the reference needs them for its cuBLAS GEMM steps,
 * src/workspace_queries.cu:188-238).  The queries are kept for source
 * compatibility of the query -> allocate -> call protocol. */
void kblas_gemm_batch_strided_wsquery(kblasHandle_t h, int bc) { (void)h; (void)bc; }
void kblas_gemm_batch_nonuniform_wsquery(kblasHandle_t h) { (void)h; }
void kblas_syrk_batch_wsquery(kblasHandle_t h, const int m, int bc) { (void)h; (void)m; (void)bc; }
void kblas_syrk_batch_nonuniform_wsquery(kblasHandle_t h) { (void)h; }
void kblas_trsm_batch_wsquery(kblasHandle_t h, char side, int m, int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }
void kblas_trsm_batch_strided_wsquery(kblasHandle_t h, char side, int m, int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }
void kblas_trsm_batch_nonuniform_wsquery(kblasHandle_t h) { (void)h; }
void kblas_potrf_batch_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_potrf_batch_strided_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_potrs_batch_wsquery(kblasHandle_t h, const int m, const int n, int bc) { (void)h; (void)m; (void)n; (void)bc; }
void kblas_potrs_batch_strided_wsquery(kblasHandle_t h, const int m, const int n, int bc) { (void)h; (void)m; (void)n; (void)bc; }
void kblas_posv_batch_wsquery(kblasHandle_t h, char side, const int m, const int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }
void kblas_posv_batch_strided_wsquery(kblasHandle_t h, char side, const int m, const int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }

void kblas_trtri_batch_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_trtri_batch_strided_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_lauum_batch_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_lauum_batch_strided_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_potri_batch_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_potri_batch_strided_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_poti_batch_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_poti_batch_strided_wsquery(kblasHandle_t h, const int n, int bc) { (void)h; (void)n; (void)bc; }
void kblas_trmm_batch_wsquery(kblasHandle_t h, char side, int m, int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }
void kblas_trmm_batch_strided_wsquery(kblasHandle_t h, char side, int m, int n, int bc) { (void)h; (void)side; (void)m; (void)n; (void)bc; }

/* ========================= complex precisions ============================ */
/* The C and Z entry points run the same kernel templates instantiated for
 * the complex types (the reference instantiates its templates for them too,
 * src/batch_triangular/Makefile:30-57): complex symmetric arithmetic as the
 * reference's operators define it, 'C' a conjugate transpose (GEMM, TRSM,
 * TRMM).  POTRF / POTRS / POSV run the generic Cholesky kernels. */
#define KB_NI return KBLAS_NotImplemented
int kblasCgemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k, const hipFloatComplex alpha, const hipFloatComplex** A, int lda, const hipFloatComplex** B, int ldb, const hipFloatComplex beta, hipFloatComplex** C, int ldc, int bc)
{ return gemm_core(h, 'c', ta, tb, m, n, k, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, CS(beta), mat_ptrs(C, ldc), 0, 0, bc); }
int kblasZgemm_batch(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k, const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, const hipDoubleComplex** B, int ldb, const hipDoubleComplex beta, hipDoubleComplex** C, int ldc, int bc)
{ return gemm_core(h, 'z', ta, tb, m, n, k, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, CS(beta), mat_ptrs(C, ldc), 0, 0, bc); }
int kblasCgemm_batch_strided(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k, const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, const hipFloatComplex* B, int ldb, long sB, const hipFloatComplex beta, hipFloatComplex* C, int ldc, long sC, int bc)
{ return gemm_core(h, 'c', ta, tb, m, n, k, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, CS(beta), mat_strided(C, ldc, sC), sC, 1, bc); }
int kblasZgemm_batch_strided(kblasHandle_t h, char ta, char tb, const int m, const int n, const int k, const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, const hipDoubleComplex* B, int ldb, long sB, const hipDoubleComplex beta, hipDoubleComplex* C, int ldc, long sC, int bc)
{ return gemm_core(h, 'z', ta, tb, m, n, k, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, CS(beta), mat_strided(C, ldc, sC), sC, 1, bc); }
int kblasCsyrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex** A, int lda, const hipFloatComplex beta, hipFloatComplex** B, int ldb, int bc)
{ return syrk_core(h, 'c', uplo, trans, m, n, CS(alpha), mat_ptrs(A, lda), 0, CS(beta), mat_ptrs(B, ldb), 0, 0, bc); }
int kblasZsyrk_batch(kblasHandle_t h, char uplo, char trans, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, const hipDoubleComplex beta, hipDoubleComplex** B, int ldb, int bc)
{ return syrk_core(h, 'z', uplo, trans, m, n, CS(alpha), mat_ptrs(A, lda), 0, CS(beta), mat_ptrs(B, ldb), 0, 0, bc); }
int kblasCsyrk_batch_strided(kblasHandle_t h, char uplo, char trans, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, const hipFloatComplex beta, hipFloatComplex* B, int ldb, long sB, int bc)
{ return syrk_core(h, 'c', uplo, trans, m, n, CS(alpha), mat_strided(A, lda, sA), sA, CS(beta), mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasZsyrk_batch_strided(kblasHandle_t h, char uplo, char trans, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, const hipDoubleComplex beta, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return syrk_core(h, 'z', uplo, trans, m, n, CS(alpha), mat_strided(A, lda, sA), sA, CS(beta), mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasCtrsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return trsm_core(h, 'c', side, uplo, trans, diag, m, n, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasZtrsm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return trsm_core(h, 'z', side, uplo, trans, diag, m, n, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasCtrsm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return trsm_core(h, 'c', side, uplo, trans, diag, m, n, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasZtrsm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return trsm_core(h, 'z', side, uplo, trans, diag, m, n, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasCpotrf_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return potrf_core(h, 'c', uplo, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasZpotrf_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return potrf_core(h, 'z', uplo, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasCpotrf_batch_strided(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return potrf_core(h, 'c', uplo, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasZpotrf_batch_strided(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return potrf_core(h, 'z', uplo, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasCpotrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n, const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return solve_core(h, 'c', KB_MODE_POTRS, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, NULL); }
int kblasZpotrs_batch(kblasHandle_t h, char side, char uplo, const int m, const int n, const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return solve_core(h, 'z', KB_MODE_POTRS, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, NULL); }
int kblasCpotrs_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n, const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return solve_core(h, 'c', KB_MODE_POTRS, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, NULL); }
int kblasZpotrs_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n, const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return solve_core(h, 'z', KB_MODE_POTRS, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, NULL); }
int kblasCposv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n, hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc, int* info)
{ return solve_core(h, 'c', KB_MODE_POSV, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, info); }
int kblasZposv_batch(kblasHandle_t h, char side, char uplo, const int m, const int n, hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc, int* info)
{ return solve_core(h, 'z', KB_MODE_POSV, side, uplo, m, n, mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc, info); }
int kblasCposv_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n, hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc, int* info)
{ return solve_core(h, 'c', KB_MODE_POSV, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, info); }
int kblasZposv_batch_strided(kblasHandle_t h, char side, char uplo, const int m, const int n, hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc, int* info)
{ return solve_core(h, 'z', KB_MODE_POSV, side, uplo, m, n, mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc, info); }
int kblasClauum_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'c', INV_LAUUM, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasClauum_batch_strided(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'c', INV_LAUUM, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasCpotri_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'c', INV_POTRI, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasCpotri_batch_strided(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'c', INV_POTRI, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasCpoti_batch(kblasHandle_t h, char uplo, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'c', INV_POTI, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasCpoti_batch_strided(kblasHandle_t h, char uplo, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'c', INV_POTI, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasCtrtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipFloatComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'c', INV_TRTRI, uplo, diag, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasCtrtri_batch_strided(kblasHandle_t h, char uplo, char diag, const int n, hipFloatComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'c', INV_TRTRI, uplo, diag, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasCtrmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex** A, int lda, hipFloatComplex** B, int ldb, int bc)
{ return trmm_core(h, 'c', side, uplo, trans, diag, m, n, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasCtrmm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipFloatComplex alpha, const hipFloatComplex* A, int lda, long sA, hipFloatComplex* B, int ldb, long sB, int bc)
{ return trmm_core(h, 'c', side, uplo, trans, diag, m, n, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
int kblasZlauum_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'z', INV_LAUUM, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasZlauum_batch_strided(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'z', INV_LAUUM, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasZpotri_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'z', INV_POTRI, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasZpotri_batch_strided(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'z', INV_POTRI, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasZpoti_batch(kblasHandle_t h, char uplo, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'z', INV_POTI, uplo, 'N', n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasZpoti_batch_strided(kblasHandle_t h, char uplo, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'z', INV_POTI, uplo, 'N', n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasZtrtri_batch(kblasHandle_t h, char uplo, char diag, const int n, hipDoubleComplex** A, int lda, int bc, int* info)
{ return inv_core(h, 'z', INV_TRTRI, uplo, diag, n, mat_ptrs(A, lda), 0, 0, bc, info); }
int kblasZtrtri_batch_strided(kblasHandle_t h, char uplo, char diag, const int n, hipDoubleComplex* A, int lda, long sA, int bc, int* info)
{ return inv_core(h, 'z', INV_TRTRI, uplo, diag, n, mat_strided(A, lda, sA), sA, 1, bc, info); }
int kblasZtrmm_batch(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex** A, int lda, hipDoubleComplex** B, int ldb, int bc)
{ return trmm_core(h, 'z', side, uplo, trans, diag, m, n, CS(alpha), mat_ptrs(A, lda), 0, mat_ptrs(B, ldb), 0, 0, bc); }
int kblasZtrmm_batch_strided(kblasHandle_t h, char side, char uplo, char trans, char diag, const int m, const int n, const hipDoubleComplex alpha, const hipDoubleComplex* A, int lda, long sA, hipDoubleComplex* B, int ldb, long sB, int bc)
{ return trmm_core(h, 'z', side, uplo, trans, diag, m, n, CS(alpha), mat_strided(A, lda, sA), sA, mat_strided(B, ldb, sB), sB, 1, bc); }
