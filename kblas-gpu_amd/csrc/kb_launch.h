/* kb_launch.h -- C interface between the host library (C) and the HIP kernels.
 *
 * Every launcher is asynchronous on the given stream, allocates nothing and
 * returns a hipError_t value (0 = hipSuccess).  Argument validation is done by
 * the host layer before these are called.
 */
#ifndef KB_LAUNCH_H
#define KB_LAUNCH_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A batch operand: strided (base + b*stride), pointer array (ptrs[b]) and a
 * uniform (ld) or per-matrix (lds[b]) leading dimension, plus an element
 * offset into every matrix (a sub-block; 0 for the caller's matrices).  All
 * device memory. */
typedef struct {
    void* base;
    long stride;
    void* const* ptrs;
    const int* lds;
    int ld;
    long off;
} kb_mat_t;

enum { KB_MODE_POTRF = 0, KB_MODE_POSV = 1, KB_MODE_POTRS = 2 };

/* Largest order served by the register-resident kernels. */
#define KB_REG_MAX_N 64

/* Register-resident uniform Cholesky family (n <= 64).
 * prec 'd' or 's'.  side_left: 1 = A X = B (B is n x nrhs), 0 = X A = B
 * (B is nrhs x n, the reference's Right side). */
int kb_launch_reg_chol(char prec, int mode, int n, int nrhs, int side_left,
                       kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s);

/* Kernel variants of the per-size-bucket choice (kb_tune.c). */
enum { KB_VAR_AUTO = 0, KB_VAR_RG16, KB_VAR_RG32, KB_VAR_RG4, KB_VAR_TILE, KB_VAR_REG, KB_VAR_BLK };

/* The same with an explicit kernel variant (KB_VAR_AUTO: the built-in rule;
 * rg16 n <= 16, rg32 n <= 32, rg4 POTRF 32 < n <= 64, tile n <= 64);
 * hipErrorInvalidValue for a variant that cannot serve n. */
int kb_launch_reg_chol_v(char prec, int mode, int n, int nrhs, int side_left,
                         kb_mat_t A, kb_mat_t B, int* info, int batch, int variant, hipStream_t s);

/* Mixed precision POSV (FP32 factor + FP64 refinement + FP64 fallback), n <= 64.
 * A read-only unless the fallback ran (then it holds L, as LAPACK dsposv).   */
int kb_launch_reg_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                        int* info, int* iter, int batch, hipStream_t s);
/* variant: 0 built-in, KB_VAR_REG (n <= 32), KB_VAR_TILE (n <= 64) */
int kb_launch_reg_sposv_v(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                          int* info, int* iter, int batch, int variant, hipStream_t s);

/* Blocked one-workgroup-per-matrix Cholesky family for any n (used for
 * n > KB_REG_MAX_N); same arguments as kb_launch_reg_chol. */
int kb_launch_blk_chol(char prec, int mode, int n, int nrhs, int side_left,
                       kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s);
int kb_launch_blk_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                        int* info, int* iter, int batch, hipStream_t s);

/* Batched GEMM C := alpha op(A) op(B) + beta C.  vm/vn/vk: optional device
 * arrays of per-matrix sizes (non-uniform); then m/n/k are upper bounds or 0. */
int kb_launch_gemm(char prec, int transA, int transB, int m, int n, int k, double alpha,
                   kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C,
                   const int* vm, const int* vn, const int* vk, int batch, hipStream_t s);
/* Batched lower SYRK B := alpha op(A) op(A)^T + beta B (B is m x m, k inner). */
int kb_launch_syrk(char prec, int trans, int m, int k, double alpha, kb_mat_t A,
                   double beta, kb_mat_t B, const int* vm, const int* vk, int batch, hipStream_t s);
/* Batched lower TRSM: op(A) X = alpha B (side_left) or X op(A) = alpha B. */
int kb_launch_trsm(char prec, int side_left, int trans, int unit, int m, int n, double alpha,
                   kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s);

/* Non-uniform Cholesky family in one persistent launch over a size-sorted
 * queue.  B[i] is vm[i] x vn[i]; the order of A[i] is vm[i] (side_left) or
 * vn[i] (right side); mode POTRF passes vm = NULL and vn[i] = order.
 * prec 'd', 's' or 'm' (mode SPOSV: FP32 factor + FP64 refinement, X receives
 * the solution, iter as in dsposv).  ws: device workspace of
 * kb_vchol_ws_bytes(batch) bytes.  max_nrhs: a bound on the right-hand-side
 * counts (0: unknown); up to 32 the workgroup phase runs passes of 32. */
enum { KB_MODE_SPOSV = 3 };
/* ints of the queue header (histogram, offsets, counters) before the order array */
#define KB_VCHOL_WS_HDR_INTS (2 * 4096 + 64)
size_t kb_vchol_ws_bytes(int batch);

/* Fixed per-device scratch area holding the queue for up to
 * KB_NOALLOC_MAX_BATCH matrices (kb_noalloc.hip).  kb_noalloc_init looks up
 * the current device's copy (hipSuccess or the lookup's error);
 * kb_noalloc_scratch returns its device address (NULL and *err on failure). */
#define KB_NOALLOC_MAX_BATCH (1 << 20)
int kb_noalloc_init(void);
void* kb_noalloc_scratch(int* err);
int kb_launch_vchol(char prec, int mode, const int* vm, const int* vn, int side_left,
                    kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info, int* iter, int batch,
                    int max_nrhs, void* ws, hipStream_t s);
/* The same with the order at and below which a matrix goes to the wavefront
 * phase (0..KB_REG_MAX_N; orders above it go to the workgroup phase).  The
 * mixed mode always splits at KB_REG_MAX_N. */
int kb_launch_vchol_split(char prec, int mode, const int* vm, const int* vn, int side_left,
                          kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info, int* iter, int batch,
                          int max_nrhs, void* ws, int split, hipStream_t s);

/* Non-uniform mixed POSV.  ws holds ws_bytes bytes: the queue, then (when
 * ws_bytes >= kb_vsposv_ws_bytes(batch, max_order, max_nrhs)) one FP32 slot
 * per workgroup of the phase for orders above KB_REG_MAX_N, which then runs
 * the mixed method; without the slots (or for orders / right-hand-side counts
 * above max_order / max_nrhs, 1024 / 64 at most) those orders take dsposv's
 * full-precision path (ITER = -1). */
size_t kb_vsposv_ws_bytes(int batch, int max_order, int max_nrhs);
int kb_launch_vsposv(const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                     int* info, int* iter, int batch, void* ws, size_t ws_bytes, int max_order, int max_nrhs,
                     hipStream_t s);

/* Batched inverse family on the lower triangle, one workgroup per matrix,
 * n <= KB_INV_MAX_N (hipErrorNotSupported above).  op: KB_INV_OP_TRTRI
 * (A := L^-1, unit: diagonal taken as 1 and left unread/unwritten),
 * KB_INV_OP_LAUUM (A := L^T L), KB_INV_OP_POTRI (A := (L L^T)^-1).
 * info[b]: first exactly-zero diagonal element of a non-unit triangle
 * (1-based, A left unchanged) or 0.  skip (optional): matrices with
 * skip[b] != 0 are left alone, info included (POTI after a failed POTRF). */
#define KB_INV_MAX_N 1024
enum { KB_INV_OP_TRTRI = 0, KB_INV_OP_LAUUM = 1, KB_INV_OP_POTRI = 2 };
int kb_launch_inv(char prec, int op, int unit, int n, kb_mat_t A, int* info, const int* skip, int batch,
                  hipStream_t s);
/* Batched lower TRMM (prec s, d, c, z; trans 2 conjugates L): B := alpha op(L) B (side_left, L of order m) or
 * B := alpha B op(L) (L of order n); B is m x n.  One line of B (a column
 * for side_left, else a row) is staged in 64 KB of LDS, so the order is at
 * most 16384 (FP32) / 8192 (FP64) (hipErrorNotSupported above). */
int kb_launch_trmm(char prec, int side_left, int trans, int unit, int m, int n, const double* alpha,
                   kb_mat_t A, kb_mat_t B, int batch, hipStream_t s);

/* Batched pivoted Cholesky P^T A P = R^T R (upper), n <= 64 (hipErrorNotSupported
 * above).  Pivots: pivp[b] (pointer array) or pivs + b * spiv (strided), 0-based. */
int kb_launch_pstrf(char prec, int n, kb_mat_t A, int* const* pivp, int* pivs, long spiv, int* ranks,
                    int batch, hipStream_t s);

/* Any precision ('s', 'd', 'c', 'z'): GEMM / SYRK / TRSM with scalars as
 * (real, imaginary) pairs, op 0 / 1 / 2 = none / transpose / conjugate
 * transpose, either triangle (upper), TRSM with a unit diagonal or not. */
int kb_launch_gemm_x(char prec, int transA, int transB, int m, int n, int k, const double* alpha,
                     kb_mat_t A, kb_mat_t B, const double* beta, kb_mat_t C, const int* vm,
                     const int* vn, const int* vk, int batch, hipStream_t s);
int kb_launch_syrk_x(char prec, int upper, int trans, int m, int k, const double* alpha, kb_mat_t A,
                     const double* beta, kb_mat_t B, const int* vm, const int* vk, int batch, hipStream_t s);
int kb_launch_trsm_x(char prec, int side_left, int upper, int trans, int unit, int m, int n,
                     const double* alpha, kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch,
                     hipStream_t s);
/* Real TRSM (prec 'd' / 's', trans 0 = N, else T): orders up to 16 by register
 * substitution with the triangle in LDS, up to 64 on the matrix cores, lower
 * orders above 64 by the blocked solve; hipErrorNotSupported for an upper
 * triangle above 64.  Non-uniform: per-matrix sizes in vm / vn, bounded by
 * m / n (NULL: uniform). */
int kb_launch_trsm_fast(char prec, int side_left, int upper, int trans, int unit, int m, int n, double alpha,
                        kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s);
/* (L L^T) X = B for 16 < n <= 64 ('d' / 's'): the two sweeps of the matrix-core
 * TRSM; matrices with skip[b] != 0 are left alone (skip may be NULL). */
int kb_launch_trsm_solve(char prec, int side_left, int n, int nrhs, kb_mat_t A, kb_mat_t B, const int* skip,
                         int batch, hipStream_t s);
/* Orders above 32 with at least this many right-hand sides: POSV / POTRS as a
 * factorization followed by kb_launch_trsm_solve. */
#define KB_SOLVE_SPLIT_NRHS 32
/* Orders above 64 with at least this many vectors: TRSM (lower, uniform) and
 * the solves of POSV / POTRS by the recursion into half-order solves and one
 * matrix-core GEMM per level (kb_trsm.hip). */
#define KB_TRSM_REC_NVEC 64
int kb_launch_trsm_rec_solve(char prec, int side_left, int n, int nrhs, kb_mat_t A, kb_mat_t B, const int* skip,
                             int batch, hipStream_t s);
/* Real GEMM / SYRK on the matrix cores (prec 'd' / 's', op 0 = N, else T;
 * syrk: C is m x m, B = A, tb = !ta, upper picks the triangle).  Non-uniform:
 * per-matrix sizes in vm / vn / vk, bounded by m / n / k (NULL: uniform).
 * hipErrorNotSupported when the batch is too large for one launch. */
int kb_launch_gemm_fast(char prec, int syrk, int upper, int ta, int tb, int m, int n, int k, double alpha,
                        kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C, const int* vm, const int* vn,
                        const int* vk, int batch, hipStream_t s);
/* The same with optional per-matrix skip flags (a matrix with skip[b] != 0
 * keeps its C). */
int kb_launch_gemm_skip(char prec, int syrk, int upper, int ta, int tb, int m, int n, int k, double alpha,
                        kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C, const int* vm, const int* vn,
                        const int* vk, const int* skip, int batch, hipStream_t s);
/* Generic POTRF / POTRS / POSV (mode KB_MODE_*) for any precision and either
 * triangle; info per matrix for POTRF / POSV. */
int kb_launch_gchol(char prec, int mode, int upper, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B,
                    int* info, int batch, hipStream_t s);

/* Randomized-range-finder SPD helpers (all operands per matrix through device
 * arrays).  Rank-revealing Cholesky of dims[b] x dims[b] Gram matrices
 * (dims clamped to bs <= 32), R written as FP32 (rprec 's') or FP64 ('d'). */
int kb_launch_ara_potrf(char rprec, const int* dims, double* const* G, const int* ldg, void* const* R,
                        const int* ldr, double* diag, int bs, int* ranks, int num_ops, hipStream_t s);
/* B := A^T A (n x n, n clamped to 32) of FP32 m x n matrices, FP64 result. */
int kb_launch_ara_syrk(const int* m, const int* n, const float* const* A, const int* lda,
                       double* const* B, const int* ldb, int batch, hipStream_t s);
/* B(:, 0:c) := B(:, 0:c) R^-1, R upper of order c = min(cols[b], bs), bs <= 32;
 * rows[b] <= max_rows rows. */
int kb_launch_ara_trsm(char prec, void* const* B, const int* ldb, void* const* R, const int* ldr,
                       const int* rows, const int* cols, int bs, int max_rows, int num_ops, hipStream_t s);

/* Fill count ints with v (a kernel, so it is safe inside captured graphs). */
int kb_launch_fill_int(int* p, int v, long count, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
