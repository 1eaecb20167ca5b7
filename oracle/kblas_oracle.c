/*
 * kblas_oracle.c -- CPU restatement of KBLAS's batched SPD path (checker only).
 *
 * TEST INFRASTRUCTURE.  Only tests/, the smoke entry and the benchmark's CPU
 * baseline leg use this file.  The product library never links it.
 *
 * Compile with -ffp-contract=off: every fused multiply-add below is written as
 * fma() on purpose, matching where the reference kernels use FMA(); all other
 * products and sums must stay separately rounded so the generators are
 * bit-reproducible against their HIP twins.
 */
#define _GNU_SOURCE
#include "kblas_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(A, lda, i, j) (A)[(long)(i) + (long)(j) * (long)(lda)]

/* ======================================================================= */
/* Counter-based generators                                                 */
/* ======================================================================= */

static inline uint64_t mix64(uint64_t x)
{
    /* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

uint64_t orc_hash(uint64_t seed, uint64_t id, uint64_t k)
{
    uint64_t h = mix64(seed ^ 0x6A09E667F3BCC909ULL);
    h = mix64(h + id * 0x9E3779B97F4A7C15ULL);
    h = mix64(h + k * 0xD1B54A32D192ED03ULL + 1ULL);
    return h;
}

/* 53-bit uniform in the open interval (0,1). */
static inline double u01(uint64_t h)
{
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

double orc_uniform(uint64_t seed, uint64_t id, uint64_t k)
{
    return 2.0 * u01(orc_hash(seed, id, k)) - 1.0;
}

/* Natural log for x > 0 from frexp + the atanh series; deterministic. */
static double orc_log(double x)
{
    int e;
    double m = frexp(x, &e);            /* x = m 2^e, m in [0.5,1) */
    if (m < 0.70710678118654752) { m = m * 2.0; e = e - 1; }
    double s = (m - 1.0) / (m + 1.0);   /* |s| <= 0.1716 */
    double s2 = s * s;
    double p = 1.0 / 25.0;
    for (int t = 23; t >= 1; t -= 2) p = fma(p, s2, 1.0 / (double)t);
    return (double)e * 0.69314718055994531 + 2.0 * s * p;
}

double orc_exp2(double x)
{
    double k = floor(x);
    double y = (x - k) * 0.69314718055994531;   /* in [0, ln 2) */
    double c[21];
    c[0] = 1.0;
    for (int i = 1; i <= 20; ++i) c[i] = c[i - 1] / (double)i;
    double p = c[20];
    for (int i = 19; i >= 0; --i) p = fma(p, y, c[i]);
    return ldexp(p, (int)k);
}

/* Marsaglia polar method on the counter stream (k, attempt). */
double orc_normal(uint64_t seed, uint64_t id, uint64_t k)
{
    for (uint64_t a = 0; a < 64; ++a) {
        double u = 2.0 * u01(orc_hash(seed, id, (k << 7) | (2 * a))) - 1.0;
        double v = 2.0 * u01(orc_hash(seed, id, (k << 7) | (2 * a + 1))) - 1.0;
        double s = u * u;
        s = fma(v, v, s);
        if (s >= 1.0 || s == 0.0) continue;
        return u * sqrt(-2.0 * orc_log(s) / s);
    }
    return 0.0;
}

void orc_gen_spd_shift(uint64_t seed, uint64_t id, int n, double shift, double* A, int lda)
{
    /* A_ij = sum_k B_ik B_jk (k ascending, fma), i >= j; A_ii += shift; mirror. */
    double* Bm = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i)
            Bm[i + (long)k * n] = orc_uniform(seed, id, (uint64_t)i + (uint64_t)k * (uint64_t)n);
    for (int j = 0; j < n; ++j) {
        for (int i = j; i < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s = fma(Bm[i + (long)k * n], Bm[j + (long)k * n], s);
            if (i == j) s = s + shift;
            IDX(A, lda, i, j) = s;
            IDX(A, lda, j, i) = s;
        }
    }
    free(Bm);
}

double orc_gen_spd_cond(uint64_t seed, uint64_t id, int n, double lk_lo, double lk_hi,
                        double* A, int lda)
{
    /* kappa and the spectrum */
    double uk = 0.5 * (orc_uniform(seed ^ 0x5A5A5A5AULL, id, 0) + 1.0);
    double lk = lk_lo + (lk_hi - lk_lo) * uk;            /* log10(kappa) */
    double l2k = lk * 3.3219280948873622;                 /* log2(kappa)  */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) IDX(A, lda, i, j) = 0.0;
    for (int i = 0; i < n; ++i) {
        double t;
        if (i == 0) t = 0.0;
        else if (i == n - 1) t = 1.0;
        else t = 0.5 * (orc_uniform(seed ^ 0x5A5A5A5AULL, id, (uint64_t)i + 1) + 1.0);
        IDX(A, lda, i, i) = orc_exp2(-l2k * t);
    }
    /* A <- H_p A H_p for p = n-2 .. 0, H_p acting on indices p..n-1 */
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    for (int p = n - 2; p >= 0; --p) {
        int m = n - p;
        for (int i = 0; i < m; ++i)
            v[i] = orc_normal(seed ^ 0xA5A5A5A5ULL, id, (uint64_t)p * (uint64_t)n + (uint64_t)i);
        double vv = 0.0;
        for (int i = 0; i < m; ++i) vv = fma(v[i], v[i], vv);
        if (vv == 0.0) continue;
        double tau = 2.0 / vv;
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s = fma(IDX(A, lda, p + i, p + j), v[j], s);
            w[i] = tau * s;
        }
        double vw = 0.0;
        for (int i = 0; i < m; ++i) vw = fma(v[i], w[i], vw);
        double alpha = (tau * vw) * 0.5;
        for (int i = 0; i < m; ++i) w[i] = fma(-alpha, v[i], w[i]);
        for (int j = 0; j < m; ++j) {
            for (int i = j; i < m; ++i) {
                double a = IDX(A, lda, p + i, p + j);
                a = fma(-v[i], w[j], a);
                a = fma(-w[i], v[j], a);
                IDX(A, lda, p + i, p + j) = a;
                IDX(A, lda, p + j, p + i) = a;
            }
        }
    }
    free(v);
    free(w);
    return orc_exp2(l2k);
}

void orc_gen_rhs(uint64_t seed, uint64_t id, int rows, int cols, double* B, int ldb)
{
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i)
            IDX(B, ldb, i, j) = orc_normal(seed, id, (uint64_t)i + (uint64_t)j * (uint64_t)rows);
}

/* testing/testing_helper.cu:394-403 (matrix_make_hpd) */
void orc_make_hpd(int n, double* A, int lda)
{
    for (int i = 0; i < n; ++i) {
        IDX(A, lda, i, i) += (double)n;
        for (int j = 0; j < i; ++j) IDX(A, lda, j, i) = IDX(A, lda, i, j);
    }
}

/* ======================================================================= */
/* Reference algorithm restatement                                           */
/* ======================================================================= */

/* src/kblas_common.cu:241-243 */
int orc_reg_size(int n) { return (n > 0) && !(n & (n - 1)); }

/* src/kblas_common.cu:244-255 */
int orc_closest_reg_size(int n)
{
    if (n <= 0) return 0;
    int r = 1;
    while (r < n) r <<= 1;
    return r >> 1;
}

static void split(int n, int* n1, int* n2)
{
    if (orc_reg_size(n)) { *n1 = *n2 = n / 2; }
    else { *n1 = orc_closest_reg_size(n); *n2 = n - *n1; }
}

/* Register POTRF, src/batch_triangular/Xpotrf_batch_kernels.cuh:52-69:
 * s = sqrt(a_jj); l_ij = a_ij / s (true division, i >= j);
 * a_ik = fma(l_ij, -l_kj, a_ik) for j < k <= i.
 * The non-positive-pivot test is the LAPACK rule; the reference comments it
 * out (Xpotrf_batch_kernels.cuh:123-125). */
static int potrf_reg(int n, double* A, int lda)
{
    for (int j = 0; j < n; ++j) {
        double d = IDX(A, lda, j, j);
        if (!(d > 0.0)) return j + 1;
        double s = sqrt(d);
        for (int i = j; i < n; ++i) IDX(A, lda, i, j) = IDX(A, lda, i, j) / s;
        for (int k = j + 1; k < n; ++k) {
            double lkj = IDX(A, lda, k, j);
            for (int i = k; i < n; ++i)
                IDX(A, lda, i, k) = fma(IDX(A, lda, i, j), -lkj, IDX(A, lda, i, k));
        }
    }
    return 0;
}

/* Recursion, src/batch_triangular/Xpotrf_batch_drivers.cuh:42-133:
 * n <= 16 -> register kernel; else POTRF(n1) ; TRSM(R,L,T,N, n2 x n1) ;
 * SYRK(L,N, n2, n1, -1, A21, +1, A22) ; POTRF(n2). */
int orc_potrf(int n, double* A, int lda)
{
    if (n <= 0) return 0;
    if (n <= 16) return potrf_reg(n, A, lda);
    int n1, n2;
    split(n, &n1, &n2);
    int info = orc_potrf(n1, A, lda);
    if (info) return info;
    orc_trsm('R', 'T', n2, n1, 1.0, A, lda, &IDX(A, lda, n1, 0), lda);
    orc_syrk('N', n2, n1, -1.0, &IDX(A, lda, n1, 0), lda, 1.0, &IDX(A, lda, n1, n1), lda);
    info = orc_potrf(n2, &IDX(A, lda, n1, n1), lda);
    return info ? info + n1 : 0;
}

/* Right-side register TRSM, src/batch_triangular/Xtrsm_batch_kernels.cuh:59-101.
 * Trans:  rows solved unscaled then multiplied by alpha on store.
 * NoTrans: rows scaled by alpha on load, solved backwards. */
static void trsm_reg_right(char trans, int m, int n, double alpha,
                           const double* A, int lda, double* B, int ldb)
{
    double b[64];
    for (int r = 0; r < m; ++r) {
        for (int i = 0; i < n; ++i)
            b[i] = (trans == 'T') ? IDX(B, ldb, r, i) : alpha * IDX(B, ldb, r, i);
        if (trans == 'T') {
            for (int j = 0; j < n; ++j) {
                b[j] = b[j] / IDX(A, lda, j, j);
                for (int i = j + 1; i < n; ++i) b[i] = fma(b[j], -IDX(A, lda, i, j), b[i]);
            }
            for (int i = 0; i < n; ++i) IDX(B, ldb, r, i) = alpha * b[i];
        } else {
            for (int j = n - 1; j >= 0; --j) {
                for (int i = j + 1; i < n; ++i) b[j] = fma(b[i], -IDX(A, lda, i, j), b[j]);
                b[j] = b[j] / IDX(A, lda, j, j);
            }
            for (int i = 0; i < n; ++i) IDX(B, ldb, r, i) = b[i];
        }
    }
}

/* Left-side register TRSM (op(L) X = alpha B), column by column. */
static void trsm_reg_left(char trans, int m, int n, double alpha,
                          const double* A, int lda, double* B, int ldb)
{
    double b[64];
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < m; ++i) b[i] = alpha * IDX(B, ldb, i, c);
        if (trans == 'N') {
            for (int j = 0; j < m; ++j) {
                b[j] = b[j] / IDX(A, lda, j, j);
                for (int i = j + 1; i < m; ++i) b[i] = fma(-IDX(A, lda, i, j), b[j], b[i]);
            }
        } else {
            for (int j = m - 1; j >= 0; --j) {
                for (int i = j + 1; i < m; ++i) b[j] = fma(-IDX(A, lda, i, j), b[i], b[j]);
                b[j] = b[j] / IDX(A, lda, j, j);
            }
        }
        for (int i = 0; i < m; ++i) IDX(B, ldb, i, c) = b[i];
    }
}

/* Recursion, src/batch_triangular/Xtrsm_batch_drivers.cuh:126-266. */
void orc_trsm(char side, char trans, int m, int n, double alpha,
              const double* A, int lda, double* B, int ldb)
{
    if (m <= 0 || n <= 0) return;
    if (side == 'R' && n <= 16) { trsm_reg_right(trans, m, n, alpha, A, lda, B, ldb); return; }
    if (side == 'L' && m <= 16) { trsm_reg_left(trans, m, n, alpha, A, lda, B, ldb); return; }
    if (side == 'R') {
        int n1, n2;
        split(n, &n1, &n2);
        if (trans == 'T') {
            orc_trsm(side, trans, m, n1, alpha, A, lda, B, ldb);
            orc_gemm('N', 'T', m, n2, n1, -1.0 / alpha, B, ldb, &IDX(A, lda, n1, 0), lda,
                     1.0, &IDX(B, ldb, 0, n1), ldb);
            orc_trsm(side, trans, m, n2, alpha, &IDX(A, lda, n1, n1), lda, &IDX(B, ldb, 0, n1), ldb);
        } else {
            orc_trsm(side, trans, m, n2, alpha, &IDX(A, lda, n1, n1), lda, &IDX(B, ldb, 0, n1), ldb);
            orc_gemm('N', 'N', m, n1, n2, -1.0, &IDX(B, ldb, 0, n1), ldb, &IDX(A, lda, n1, 0), lda,
                     alpha, B, ldb);
            orc_trsm(side, trans, m, n1, 1.0, A, lda, B, ldb);
        }
    } else {
        int m1, m2;
        split(m, &m1, &m2);
        if (trans == 'T') {
            orc_trsm(side, trans, m2, n, alpha, &IDX(A, lda, m1, m1), lda, &IDX(B, ldb, m1, 0), ldb);
            orc_gemm('T', 'N', m1, n, m2, -1.0, &IDX(A, lda, m1, 0), lda, &IDX(B, ldb, m1, 0), ldb,
                     alpha, B, ldb);
            orc_trsm(side, trans, m1, n, 1.0, A, lda, B, ldb);
        } else {
            orc_trsm(side, trans, m1, n, alpha, A, lda, B, ldb);
            orc_gemm('N', 'N', m2, n, m1, -1.0, &IDX(A, lda, m1, 0), lda, B, ldb,
                     alpha, &IDX(B, ldb, m1, 0), ldb);
            orc_trsm(side, trans, m2, n, 1.0, &IDX(A, lda, m1, m1), lda, &IDX(B, ldb, m1, 0), ldb);
        }
    }
}

/* C := alpha op(A) op(B) + beta C (BLAS semantics: beta == 0 does not read C).
 * The reference delegates this to cuBLAS (src/batch_triangular/Xgemm_batch_core.cuh:260-267). */
void orc_gemm(char ta, char tb, int m, int n, int k, double alpha,
              const double* A, int lda, const double* B, int ldb,
              double beta, double* C, int ldc)
{
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int p = 0; p < k; ++p) {
                double a = (ta == 'N') ? IDX(A, lda, i, p) : IDX(A, lda, p, i);
                double b = (tb == 'N') ? IDX(B, ldb, p, j) : IDX(B, ldb, j, p);
                s = fma(a, b, s);
            }
            IDX(C, ldc, i, j) = (beta == 0.0) ? alpha * s : alpha * s + beta * IDX(C, ldc, i, j);
        }
    }
}

/* Lower SYRK, src/batch_triangular/Xsyrk_batch_kernels.cuh:58-75:
 * b = beta*B; for every chunk of 8 columns of A: b = fma(alpha, s_chunk, b). */
void orc_syrk(char trans, int m, int k, double alpha, const double* A, int lda,
              double beta, double* B, int ldb)
{
    for (int j = 0; j < m; ++j) {
        for (int i = j; i < m; ++i) {
            double b = IDX(B, ldb, i, j) * beta;
            for (int p0 = 0; p0 < k; p0 += 8) {
                double s = 0.0;
                int pe = p0 + 8 < k ? p0 + 8 : k;
                for (int p = p0; p < pe; ++p) {
                    double ai = (trans == 'N') ? IDX(A, lda, i, p) : IDX(A, lda, p, i);
                    double aj = (trans == 'N') ? IDX(A, lda, j, p) : IDX(A, lda, p, j);
                    s = fma(ai, aj, s);
                }
                b = fma(alpha, s, b);
            }
            IDX(B, ldb, i, j) = b;
        }
    }
}

/* POTRS, src/batch_triangular/Xpotrs_batch_drivers.cuh:79-171 (Right/Lower):
 * TRSM(R,L,T,n1) ; GEMM(N,T,-1) ; TRSM(R,L,T,n2) ; TRSM(R,L,N,n2) ; GEMM(N,N,-1) ; TRSM(R,L,N,n1).
 * n == 1 (where the reference split gives n1 = n2 = 0) is solved directly.
 * Left side (LAPACK dpotrs layout) is an extension: TRSM(L,L,N) then TRSM(L,L,T). */
void orc_potrs(char side, int m, int n, const double* A, int lda, double* B, int ldb)
{
    if (m <= 0 || n <= 0) return;
    if (side == 'L') {
        orc_trsm('L', 'N', m, n, 1.0, A, lda, B, ldb);
        orc_trsm('L', 'T', m, n, 1.0, A, lda, B, ldb);
        return;
    }
    if (n == 1) {
        orc_trsm('R', 'T', m, 1, 1.0, A, lda, B, ldb);
        orc_trsm('R', 'N', m, 1, 1.0, A, lda, B, ldb);
        return;
    }
    int n1, n2;
    split(n, &n1, &n2);
    const double* A21 = &IDX(A, lda, n1, 0);
    const double* A22 = &IDX(A, lda, n1, n1);
    double* B2 = &IDX(B, ldb, 0, n1);
    orc_trsm('R', 'T', m, n1, 1.0, A, lda, B, ldb);
    orc_gemm('N', 'T', m, n2, n1, -1.0, B, ldb, A21, lda, 1.0, B2, ldb);
    orc_trsm('R', 'T', m, n2, 1.0, A22, lda, B2, ldb);
    orc_trsm('R', 'N', m, n2, 1.0, A22, lda, B2, ldb);
    orc_gemm('N', 'N', m, n1, n2, -1.0, B2, ldb, A21, lda, 1.0, B, ldb);
    orc_trsm('R', 'N', m, n1, 1.0, A, lda, B, ldb);
}

/* POSV, src/batch_triangular/Xposv_batch_drivers.cuh:83-114: POTRF then POTRS.
 * (The reference runs POTRS even on a failed factor; here the solve is skipped
 *  for that matrix, as LAPACK dposv does.)  side 'R': A is n x n; side 'L': m x m. */
int orc_posv(char side, int m, int n, double* A, int lda, double* B, int ldb)
{
    int order = (side == 'R') ? n : m;
    int info = orc_potrf(order, A, lda);
    if (info) return info;
    orc_potrs(side, m, n, A, lda, B, ldb);
    return 0;
}

/* ---------------- single precision (S API checks) ---------------- */
int orc_spotrf(int n, float* A, int lda)
{
    /* right-looking column Cholesky with the register-kernel arithmetic */
    for (int j = 0; j < n; ++j) {
        float d = IDX(A, lda, j, j);
        if (!(d > 0.0f)) return j + 1;
        float s = sqrtf(d);
        for (int i = j; i < n; ++i) IDX(A, lda, i, j) = IDX(A, lda, i, j) / s;
        for (int k = j + 1; k < n; ++k) {
            float lkj = IDX(A, lda, k, j);
            for (int i = k; i < n; ++i)
                IDX(A, lda, i, k) = fmaf(IDX(A, lda, i, j), -lkj, IDX(A, lda, i, k));
        }
    }
    return 0;
}

/* ---------------- mixed precision: LAPACK dsposv restated ---------------- */
static double* rhs_ptr(char side, double* B, int ldb, int r, long* inc)
{
    if (side == 'L') { *inc = 1; return B + (long)r * ldb; }
    *inc = ldb;
    return B + r;
}

static void spotrs_vec(int n, const float* L, float* x)
{
    for (int j = 0; j < n; ++j) {
        x[j] = x[j] / L[j + (long)j * n];
        for (int i = j + 1; i < n; ++i) x[i] = fmaf(-L[i + (long)j * n], x[j], x[i]);
    }
    for (int j = n - 1; j >= 0; --j) {
        for (int i = j + 1; i < n; ++i) x[j] = fmaf(-L[i + (long)j * n], x[i], x[j]);
        x[j] = x[j] / L[j + (long)j * n];
    }
}

/* LAPACK dsposv (reference LAPACK 3.x, SRC/dsposv.f; the CPU comparator is
 * MKL's build of it, and tests/golden/lapack_golden_mixed.npz holds its ITER
 * and X): CTE = ||A||_inf * eps * sqrt(n); B and the lower triangle of A are
 * converted to FP32 with DLAG2S / DLAT2S's overflow check (|v| > FLT_MAX gives
 * ITER = -2); SPOTRF (failure: ITER = -3); x = SPOTRS(b); then up to
 * ITERMAX = 30 times: r = b - A x (DSYMM, FP64), stop when every right-hand side
 * has ||r||_inf <= ||x||_inf * CTE (ITER = the number of corrections), else
 * convert r to FP32 (overflow: ITER = -2), solve, x += d.  No convergence gives
 * ITER = -31.  Every failure ends in the FP64 path: X := B, DPOTRF, DPOTRS.
 * One deliberate difference: a NaN in r counts as not converged (IDAMAX in
 * reference BLAS skips NaNs, so LAPACK can report convergence on a NaN
 * residual); NaN-free inputs take identical decisions. */
static int fits_fp32(double v)
{
    return !(v > 3.4028234663852886e38 || v < -3.4028234663852886e38);
}

int orc_sposv(char side, int m, int n, const double* A, int lda,
              const double* B, int ldb, double* X, int ldx, int* iter)
{
    int order = (side == 'R') ? n : m;
    int nrhs = (side == 'R') ? m : n;
    const int itermax = 30;
    if (order <= 0 || nrhs <= 0) { *iter = 0; return 0; }
    /* full symmetric copy of A from its lower triangle */
    double* Af = (double*)malloc(sizeof(double) * (size_t)order * order);
    for (int j = 0; j < order; ++j)
        for (int i = j; i < order; ++i) {
            Af[i + (long)j * order] = IDX(A, lda, i, j);
            Af[j + (long)i * order] = IDX(A, lda, i, j);
        }
    double anrm = 0.0;  /* infinity norm (DLANSY 'I') */
    for (int i = 0; i < order; ++i) {
        double s = 0.0;
        for (int j = 0; j < order; ++j) s += fabs(Af[i + (long)j * order]);
        if (s > anrm || s != s) anrm = s;
    }
    double cte = anrm * 1.1102230246251565e-16 * sqrt((double)order);
    float* L = (float*)malloc(sizeof(float) * (size_t)order * order);
    float* xs = (float*)malloc(sizeof(float) * (size_t)order);
    double* r = (double*)malloc(sizeof(double) * (size_t)order);
    int result_iter = 0, ok = 0;
    /* DLAG2S(B), then DLAT2S(lower A) */
    int ovf = 0;
    for (int c = 0; c < nrhs && !ovf; ++c) {
        long ib;
        const double* b = rhs_ptr(side, (double*)B, ldb, c, &ib);
        for (int i = 0; i < order; ++i)
            if (!fits_fp32(b[i * ib])) { ovf = 1; break; }
    }
    for (int j = 0; j < order && !ovf; ++j)
        for (int i = j; i < order; ++i)
            if (!fits_fp32(Af[i + (long)j * order])) { ovf = 1; break; }
    if (ovf) {
        result_iter = -2;
    } else {
        for (long t = 0; t < (long)order * order; ++t) L[t] = (float)Af[t];
        int sinfo = orc_spotrf(order, L, order);
        if (sinfo != 0) result_iter = -3;
    }
    if (result_iter == 0) {
        /* x0 = solve(b) */
        for (int c = 0; c < nrhs; ++c) {
            long ib, ix;
            const double* b = rhs_ptr(side, (double*)B, ldb, c, &ib);
            double* x = rhs_ptr(side, X, ldx, c, &ix);
            for (int i = 0; i < order; ++i) xs[i] = (float)b[i * ib];
            spotrs_vec(order, L, xs);
            for (int i = 0; i < order; ++i) x[i * ix] = (double)xs[i];
        }
        for (int it = 0; it <= itermax && !ok && result_iter == 0; ++it) {
            /* residual and convergence test over all right-hand sides */
            int conv = 1;
            for (int c = 0; c < nrhs; ++c) {
                long ib, ix;
                const double* b = rhs_ptr(side, (double*)B, ldb, c, &ib);
                double* x = rhs_ptr(side, X, ldx, c, &ix);
                double xn = 0.0, rn = 0.0;
                for (int i = 0; i < order; ++i) {
                    double s = b[i * ib];
                    for (int k = 0; k < order; ++k) s = fma(-Af[i + (long)k * order], x[k * ix], s);
                    r[i] = s;
                    if (fabs(s) > rn || s != s) rn = fabs(s);
                    if (fabs(x[i * ix]) > xn) xn = fabs(x[i * ix]);
                }
                if (!(rn <= xn * cte)) conv = 0;
            }
            if (conv) { result_iter = it; ok = 1; break; }
            if (it == itermax) { result_iter = -(itermax + 1); break; }
            for (int c = 0; c < nrhs && result_iter == 0; ++c) {
                long ib, ix;
                const double* b = rhs_ptr(side, (double*)B, ldb, c, &ib);
                double* x = rhs_ptr(side, X, ldx, c, &ix);
                for (int i = 0; i < order; ++i) {
                    double s = b[i * ib];
                    for (int k = 0; k < order; ++k) s = fma(-Af[i + (long)k * order], x[k * ix], s);
                    if (!fits_fp32(s)) { result_iter = -2; break; }   /* DLAG2S(R) */
                    xs[i] = (float)s;
                }
                if (result_iter) break;
                spotrs_vec(order, L, xs);
                for (int i = 0; i < order; ++i) x[i * ix] = x[i * ix] + (double)xs[i];
            }
        }
    }
    int info = 0;
    if (!ok) {
        /* FP64 fallback: dpotrf + dpotrs */
        double* Lf = (double*)malloc(sizeof(double) * (size_t)order * order);
        memcpy(Lf, Af, sizeof(double) * (size_t)order * order);
        info = orc_potrf(order, Lf, order);
        if (info == 0) {
            for (int c = 0; c < nrhs; ++c) {
                long ib, ix;
                const double* b = rhs_ptr(side, (double*)B, ldb, c, &ib);
                double* x = rhs_ptr(side, X, ldx, c, &ix);
                for (int i = 0; i < order; ++i) x[i * ix] = b[i * ib];
            }
            if (side == 'L') orc_potrs('L', order, nrhs, Lf, order, X, ldx);
            else orc_potrs('R', nrhs, order, Lf, order, X, ldx);
        }
        free(Lf);
    }
    *iter = result_iter;
    free(Af); free(L); free(xs); free(r);
    return info;
}

/* ======================================================================= */
/* Metrics (long double accumulation)                                        */
/* ======================================================================= */
double orc_potrf_berr(int n, const double* A, int lda, const double* L, int ldl)
{
    long double num = 0.0L, den = 0.0L;
    for (int j = 0; j < n; ++j) {
        for (int i = j; i < n; ++i) {
            long double s = 0.0L;
            for (int k = 0; k <= j; ++k) s += (long double)IDX(L, ldl, i, k) * IDX(L, ldl, j, k);
            long double d = (long double)IDX(A, lda, i, j) - s;
            long double a = IDX(A, lda, i, j);
            long double w = (i == j) ? 1.0L : 2.0L;
            num += w * d * d;
            den += w * a * a;
        }
    }
    if (den == 0.0L) return (double)sqrtl(num);
    return (double)sqrtl(num / den);
}

double orc_solve_berr(char side, int m, int n, const double* A, int lda,
                      const double* X, int ldx, const double* B, int ldb)
{
    int order = (side == 'R') ? n : m;
    int nrhs = (side == 'R') ? m : n;
    long double anrm = 0.0L;
    for (int i = 0; i < order; ++i) {
        long double s = 0.0L;
        for (int k = 0; k < order; ++k)
            s += fabsl((long double)(i >= k ? IDX(A, lda, i, k) : IDX(A, lda, k, i)));
        if (s > anrm) anrm = s;
    }
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        long double rn = 0.0L, xn = 0.0L, bn = 0.0L;
        for (int i = 0; i < order; ++i) {
            long double b = (side == 'L') ? IDX(B, ldb, i, c) : IDX(B, ldb, c, i);
            long double s = b;
            for (int k = 0; k < order; ++k) {
                long double a = (i >= k) ? IDX(A, lda, i, k) : IDX(A, lda, k, i);
                long double x = (side == 'L') ? IDX(X, ldx, k, c) : IDX(X, ldx, c, k);
                s -= a * x;
            }
            long double x = (side == 'L') ? IDX(X, ldx, i, c) : IDX(X, ldx, c, i);
            if (fabsl(s) > rn) rn = fabsl(s);
            if (fabsl(x) > xn) xn = fabsl(x);
            if (fabsl(b) > bn) bn = fabsl(b);
        }
        long double den = anrm * xn + bn;
        double e = (den > 0.0L) ? (double)(rn / den) : (double)rn;
        if (!(e <= worst)) worst = e;   /* NaN propagates */
    }
    return worst;
}

/* ======================================================================= */
/* FLOP model, testing/flops.h:74-150 (real arithmetic)                      */
/* ======================================================================= */
double orc_flops_potrf(int n)
{
    double x = n;
    double fmuls = x * (((1.0 / 6.0) * x + 0.5) * x + (1.0 / 3.0));
    double fadds = x * (((1.0 / 6.0) * x) * x - (1.0 / 6.0));
    return fmuls + fadds;
}
double orc_flops_trsm(char side, int m, int n)
{
    double a = (side == 'L') ? m : n, b = (side == 'L') ? n : m;
    return 0.5 * b * a * (a + 1.0) + 0.5 * b * a * (a - 1.0);
}
double orc_flops_potrs(char side, int m, int n) { return 2.0 * orc_flops_trsm(side, m, n); }
double orc_flops_posv(char side, int m, int n)
{
    int order = (side == 'R') ? n : m;
    return orc_flops_potrf(order) + orc_flops_potrs(side, m, n);
}
double orc_flops_syrk(int k, int n) { return (double)k * n * (n + 1.0); }
double orc_flops_gemm(int m, int n, int k) { return 2.0 * m * (double)n * k; }

/* ======================================================================= */
/* Batch loops                                                                */
/* ======================================================================= */
static int g_threads = 0;
int orc_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_num_procs();
#else
    return 1;
#endif
}
void orc_set_threads(int t) { g_threads = t; }

void orc_potrf_batch(int n, double* A, int lda, long strideA, int batch, int* info)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        int r = orc_potrf(n, A + (long)b * strideA, lda);
        if (info) info[b] = r;
    }
}

void orc_posv_batch(char side, int m, int n, double* A, int lda, long strideA,
                    double* B, int ldb, long strideB, int batch, int* info)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        int r = orc_posv(side, m, n, A + (long)b * strideA, lda, B + (long)b * strideB, ldb);
        if (info) info[b] = r;
    }
}

void orc_sposv_batch(char side, int m, int n, const double* A, int lda, long strideA,
                     const double* B, int ldb, long strideB, double* X, int ldx, long strideX,
                     int batch, int* info, int* iter)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        int it = 0;
        int r = orc_sposv(side, m, n, A + (long)b * strideA, lda, B + (long)b * strideB, ldb,
                          X + (long)b * strideX, ldx, &it);
        if (info) info[b] = r;
        if (iter) iter[b] = it;
    }
}

void orc_gen_spd_shift_batch(uint64_t seed, int n, double shift, double* A, int lda,
                             long strideA, long first, int batch)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b)
        orc_gen_spd_shift(seed, (uint64_t)(first + b), n, shift, A + (long)b * strideA, lda);
}

void orc_gen_spd_cond_batch(uint64_t seed, int n, double lk_lo, double lk_hi, double* A,
                            int lda, long strideA, long first, int batch, double* kappa)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        double k = orc_gen_spd_cond(seed, (uint64_t)(first + b), n, lk_lo, lk_hi,
                                    A + (long)b * strideA, lda);
        if (kappa) kappa[b] = k;
    }
}

void orc_gen_rhs_batch(uint64_t seed, int rows, int cols, double* B, int ldb, long strideB,
                       long first, int batch)
{
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b)
        orc_gen_rhs(seed, (uint64_t)(first + b), rows, cols, B + (long)b * strideB, ldb);
}

/* ======================================================================= */
/* LAPACK loaded at run time (the reference tests' CPU comparator)          */
/* ======================================================================= */
typedef void (*potrf_t)(const char*, const int*, double*, const int*, int*);
typedef void (*potrs_t)(const char*, const int*, const int*, const double*, const int*,
                        double*, const int*, int*);
typedef void (*trsm_t)(const char*, const char*, const char*, const char*, const int*,
                       const int*, const double*, const double*, const int*, double*, const int*);

static void* g_lapack = NULL;
static const char* g_lapack_name = "none";
static potrf_t p_potrf = NULL;
static potrs_t p_potrs = NULL;
static trsm_t p_trsm = NULL;
static int g_lapack_tried = 0;

static void lapack_load(void)
{
    if (g_lapack_tried) return;
    g_lapack_tried = 1;
    const char* names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so",
                           "libopenblas.so.0", "libopenblas.so", "liblapack.so.3", NULL};
    for (int i = 0; names[i]; ++i) {
        void* h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        potrf_t f = (potrf_t)dlsym(h, "dpotrf_");
        potrs_t s = (potrs_t)dlsym(h, "dpotrs_");
        trsm_t t = (trsm_t)dlsym(h, "dtrsm_");
        if (f && s && t) {
            /* one LAPACK thread per OpenMP thread, as the reference's -lmkl_sequential */
            int (*mkl_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
            if (mkl_layer) mkl_layer(1 /* MKL_THREADING_SEQUENTIAL */);
            void (*ob_threads)(int) = (void (*)(int))dlsym(h, "openblas_set_num_threads");
            if (ob_threads) ob_threads(1);
            g_lapack = h; p_potrf = f; p_potrs = s; p_trsm = t; g_lapack_name = names[i];
            return;
        }
        dlclose(h);
    }
}

int orc_lapack_available(void) { lapack_load(); return g_lapack != NULL; }
const char* orc_lapack_name(void) { lapack_load(); return g_lapack_name; }

int orc_lapack_potrf_batch(int n, double* A, int lda, long strideA, int batch, int* info)
{
    lapack_load();
    if (!g_lapack) return 0;
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        int r = 0;
        p_potrf("L", &n, A + (long)b * strideA, &lda, &r);
        if (info) info[b] = r;
    }
    return 1;
}

/* testing/batch_triangular/test_Xposv_batch.cpp:313-326: dpotrf then
 * dtrsm(R,L,T,N) and dtrsm(R,L,N,N) for side 'R'; dpotrs for side 'L'. */
int orc_lapack_posv_batch(char side, int m, int n, double* A, int lda, long strideA,
                          double* B, int ldb, long strideB, int batch, int* info)
{
    lapack_load();
    if (!g_lapack) return 0;
    const double one = 1.0;
#pragma omp parallel for schedule(static) num_threads(orc_threads())
    for (int b = 0; b < batch; ++b) {
        int r = 0;
        int order = (side == 'R') ? n : m;
        double* Ab = A + (long)b * strideA;
        double* Bb = B + (long)b * strideB;
        p_potrf("L", &order, Ab, &lda, &r);
        if (r == 0) {
            if (side == 'R') {
                p_trsm("R", "L", "T", "N", &m, &n, &one, Ab, &lda, Bb, &ldb);
                p_trsm("R", "L", "N", "N", &m, &n, &one, Ab, &lda, Bb, &ldb);
            } else {
                int rr = 0;
                p_potrs("L", &m, &n, Ab, &lda, Bb, &ldb, &rr);
            }
        }
        if (info) info[b] = r;
    }
    return 1;
}

/* The same loop over a non-uniform batch packed in one buffer: matrix b is
 * n[b] x n[b] at A + offA[b] (lda = n[b]), its right-hand sides (side 'R':
 * nrhs x n[b], ldb = nrhs) at B + offB[b]; dynamic schedule, largest first
 * when the caller orders them so. */
int orc_lapack_posv_vbatch(int nrhs, const int* n, double* A, const long* offA, double* B,
                           const long* offB, int count, int* info)
{
    lapack_load();
    if (!g_lapack) return 0;
    const double one = 1.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads())
    for (int b = 0; b < count; ++b) {
        int r = 0, nb = n[b], m = nrhs;
        double* Ab = A + offA[b];
        double* Bb = B + offB[b];
        p_potrf("L", &nb, Ab, &nb, &r);
        if (r == 0) {
            p_trsm("R", "L", "T", "N", &m, &nb, &one, Ab, &nb, Bb, &m);
            p_trsm("R", "L", "N", "N", &m, &nb, &one, Ab, &nb, Bb, &m);
        }
        if (info) info[b] = r;
    }
    return 1;
}

/* the oracle's own POSV over the same packed non-uniform batch */
void orc_posv_vbatch(int nrhs, const int* n, double* A, const long* offA, double* B, const long* offB,
                     int count, int* info)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads())
    for (int b = 0; b < count; ++b) {
        int r = orc_posv('R', nrhs, n[b], A + offA[b], n[b], B + offB[b], nrhs);
        if (info) info[b] = r;
    }
}

/* ======================= inverse family and TRMM ======================== */
/* TRTRI, lower: the row recurrence of the reference register kernel
 * (src/batch_triangular/Xtrtri_batch_kernels.cuh:37-79):
 *   X(r,j) = -(sum_{j<k<=r} X(r,k) A(k,j)) / A(j,j),  X(j,j) = 1 / A(j,j),
 * columns j descending, k ascending.  The reference uses it for n <= 16 and
 * splits larger orders with TRSMs (Xtrtri_batch_drivers.cuh:31-75), the same
 * quantity in another association order.  A non-unit triangle with an exactly
 * zero diagonal element returns its 1-based index and is left unchanged
 * (LAPACK dtrtri). */
int orc_trtri(char diag, int n, double* A, int lda)
{
    const int unit = diag == 'U' || diag == 'u';
    if (!unit)
        for (int j = 0; j < n; ++j)
            if (A[j + (long)j * lda] == 0.0) return j + 1;
    for (int j = n - 1; j >= 0; --j) {
        const double a = unit ? 1.0 : A[j + (long)j * lda];
        /* rows descending: row r's sum reads column j only at rows <= r */
        for (int r = n - 1; r > j; --r) {
            double s = 0.0;
            for (int k = j + 1; k < r; ++k) s = fma(A[r + (long)k * lda], A[k + (long)j * lda], s);
            s = unit ? s + A[r + (long)j * lda] : fma(A[r + (long)r * lda], A[r + (long)j * lda], s);
            A[r + (long)j * lda] = -s / a;
        }
        if (!unit) A[j + (long)j * lda] = 1.0 / a;
    }
    return 0;
}

/* LAUUM, lower: A := L^T L on the lower triangle, row i of the result
 * (LtL)(i,c) = sum_{k>=i} L(k,c) L(k,i), k ascending, rows ascending
 * (src/batch_triangular/Xlauum_batch_kernels.cuh:112-146). */
void orc_lauum(int n, double* A, int lda)
{
    for (int i = 0; i < n; ++i)
        for (int c = 0; c <= i; ++c) {   /* (i,c) is read by no later column of this row */
            double s = 0.0;
            for (int k = i; k < n; ++k) s = fma(A[k + (long)c * lda], A[k + (long)i * lda], s);
            A[i + (long)c * lda] = s;
        }
}

/* POTRI, lower: TRTRI then LAUUM (src/batch_triangular/Xpotri_batch_drivers.cuh:78-110). */
int orc_potri(int n, double* A, int lda)
{
    int info = orc_trtri('N', n, A, lda);
    if (info) return info;
    orc_lauum(n, A, lda);
    return 0;
}

/* TRMM, lower: B := alpha op(L) B (side 'L', L of order m) or alpha B op(L)
 * (side 'R', L of order n); each element an ascending-k dot product over the
 * triangle's nonzeros, scaled by alpha last; alpha = 0 gives B = 0 (BLAS).
 * src/batch_triangular/Xtrmm_batch_kernels.cuh:37-175 (left side). */
void orc_trmm(char side, char trans, char diag, int m, int n, double alpha, const double* A, int lda,
              double* B, int ldb)
{
    const int left = side == 'L' || side == 'l', tr = !(trans == 'N' || trans == 'n');
    const int unit = diag == 'U' || diag == 'u';
    double* C = (double*)malloc(sizeof(double) * (size_t)m * n + 1);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) C[i + (long)j * m] = B[i + (long)j * ldb];
#define L_(i, j) A[(i) + (long)(j) * lda]
#define C_(i, j) C[(i) + (long)(j) * m]
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            if (left && !tr) {
                for (int k = 0; k < i; ++k) s = fma(L_(i, k), C_(k, j), s);
                s = unit ? s + C_(i, j) : fma(L_(i, i), C_(i, j), s);
            } else if (left) {
                s = unit ? C_(i, j) : L_(i, i) * C_(i, j);
                for (int k = i + 1; k < m; ++k) s = fma(L_(k, i), C_(k, j), s);
            } else if (!tr) {
                s = unit ? C_(i, j) : L_(j, j) * C_(i, j);
                for (int k = j + 1; k < n; ++k) s = fma(C_(i, k), L_(k, j), s);
            } else {
                for (int k = 0; k < j; ++k) s = fma(C_(i, k), L_(j, k), s);
                s = unit ? s + C_(i, j) : fma(C_(i, j), L_(j, j), s);
            }
            B[i + (long)j * ldb] = alpha == 0.0 ? 0.0 : alpha * s;
        }
#undef L_
#undef C_
    free(C);
}

/* ================================ PSTRF ==================================== */
/* Pivot search of src/batch_svd/batch_pstrf.cu:44-66 (getMaxDiagEntry +
 * warp_max): 32 slots hold (value, index) with value -1 outside [k, n); for
 * n > 32 slot t takes entry t + 32 when strictly larger; then a butterfly over
 * masks 16, 8, 4, 2, 1 where a slot takes its partner's pair only when the
 * partner's value is strictly larger; slot 0's pair wins. */
static void pstrf_maxdiag(const double* S, int ld, int n, int k, double* val, int* idx)
{
    double v[32];
    int key[32];
    for (int t = 0; t < 32; ++t) {
        v[t] = (t >= k && t < n) ? S[t + (long)t * ld] : -1.0;
        key[t] = t;
        if (n > 32) {
            const int nx = t + 32;
            if (nx >= k && nx < n && S[nx + (long)nx * ld] > v[t]) { v[t] = S[nx + (long)nx * ld]; key[t] = nx; }
        }
    }
    for (int m = 16; m >= 1; m >>= 1) {
        double nv[32];
        int nk[32];
        for (int t = 0; t < 32; ++t) {
            const int o = t ^ m;
            nv[t] = v[t]; nk[t] = key[t];
            if (v[o] > v[t]) { nv[t] = v[o]; nk[t] = key[o]; }
        }
        for (int t = 0; t < 32; ++t) { v[t] = nv[t]; key[t] = nk[t]; }
    }
    *val = v[0];
    *idx = key[0];
}

/* Batched pivoted Cholesky P^T A P = R^T R (src/batch_svd/batch_pstrf.cu:73-158):
 * the upper triangle mirrored into the lower one, tol = max(diag) * n * eps
 * (eps = 2^-52), right-looking steps with symmetric row/column swaps, sqrt,
 * division of row k, trailing update a -= u * r as one fused multiply-add.
 * On exit: upper triangle = R (past the rank: the unfactored remainder),
 * strict lower = 0, piv 0-based, returns the rank.  n <= 64. */
int orc_pstrf(int n, double* A, int lda, int* piv)
{
    const int ld = n > 0 ? n : 1;
    double* S = (double*)calloc((size_t)ld * ld + 1, sizeof(double));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) S[i + (long)j * ld] = i > j ? A[j + (long)i * lda] : A[i + (long)j * lda];
    for (int i = 0; i < n; ++i) piv[i] = i;
    double dmax;
    int dummy;
    pstrf_maxdiag(S, ld, n, 0, &dmax, &dummy);
    const double tol = dmax * (double)n * 2.2204460492503131e-16;
    int rank = n;
    for (int k = 0; k < n; ++k) {
        double val;
        int p;
        pstrf_maxdiag(S, ld, n, k, &val, &p);
        /* p outside [k, n): the -1 placeholder beat a tolerance below -1 (not semi-definite) */
        if (val <= tol || p < k || p >= n) { rank = k; break; }
        if (p != k) {
            int ti = piv[k]; piv[k] = piv[p]; piv[p] = ti;
            for (int c = 0; c < n; ++c) { double t = S[k + (long)c * ld]; S[k + (long)c * ld] = S[p + (long)c * ld]; S[p + (long)c * ld] = t; }
            for (int r = 0; r < n; ++r) { double t = S[r + (long)k * ld]; S[r + (long)k * ld] = S[r + (long)p * ld]; S[r + (long)p * ld] = t; }
        }
        S[k + (long)k * ld] = sqrt(S[k + (long)k * ld]);
        for (int c = k + 1; c < n; ++c) S[k + (long)c * ld] = S[k + (long)c * ld] / S[k + (long)k * ld];
        for (int r = k + 1; r < n; ++r) {
            const double u = S[k + (long)r * ld];
            for (int c = k + 1; c < n; ++c) S[r + (long)c * ld] = fma(-u, S[k + (long)c * ld], S[r + (long)c * ld]);
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + (long)j * lda] = i > j ? 0.0 : S[i + (long)j * ld];
    free(S);
    return rank;
}

/* ARA rank-revealing Cholesky, src/batch_svd/ara_util.cuh:163-275
 * (ara_fused_potrf_kernel): dim = min(dim, bs); tol = max(0, max diag(G)) *
 * 2^-52 (:220, KBlasEpsilon<double>, src/kblas_gpu_util.ch:63); for k = 0..:
 * stop when the updated G(k,k) <= tol (:227), G(k,k) = sqrt, d[k] *= G(k,k)
 * (:234-235), G(i,k) /= G(k,k) for i > k (:240), G(i,c) -= G(c,k) G(i,k)
 * for i, c > k (:251, contracted to one FMA as the CUDA compiler does);
 * then d[k] = 0 from the rank on (k < bs, :261), R(i,c) = G(c,i) for i <= c,
 * 0 below (:272).  The lower triangle evolves on its own (the update of
 * G(i,c), c <= i, reads column k only), so only it is kept.  dim <= 0:
 * returns 0 and writes nothing else (:186-190).  Returns the rank. */
int orc_ara_potrf(int dim, int bs, const double* G, int ldg, double* R, int ldr, double* d)
{
    if (dim > bs) dim = bs;
    if (dim <= 0) return 0;
    double* S = (double*)calloc((size_t)dim * dim, sizeof(double));
    double mx = 0.0;
    for (int j = 0; j < dim; ++j) {
        for (int i = j; i < dim; ++i) S[i + (long)j * dim] = G[i + (long)j * ldg];
        if (G[j + (long)j * ldg] > mx) mx = G[j + (long)j * ldg];
    }
    const double tol = mx * 2.2204460492503131e-16;
    int k;
    for (k = 0; k < dim; ++k) {
        double* ck = S + (long)k * dim;
        if (ck[k] <= tol) break;
        ck[k] = sqrt(ck[k]);
        d[k] *= ck[k];
        for (int i = k + 1; i < dim; ++i) ck[i] = ck[i] / ck[k];
        for (int c = k + 1; c < dim; ++c)
            for (int i = c; i < dim; ++i) S[i + (long)c * dim] = fma(-ck[c], ck[i], S[i + (long)c * dim]);
    }
    for (int i = k; i < bs; ++i) d[i] = 0.0;
    for (int c = 0; c < dim; ++c)
        for (int i = 0; i < dim; ++i) R[i + (long)c * ldr] = i > c ? 0.0 : S[c + (long)i * dim];
    free(S);
    return k;
}

/* ARA Gram product B := A^T A (n x n, both triangles), A m x n:
 * src/batch_svd/ara_util.cuh:651-744 (ara_mp_syrk_batch_kernel, FP32 A,
 * :706,737): each B(i,j) sums A(r,i) A(r,j) over r = 0..m-1 in order in FP64
 * (the products of FP32 entries are exact); :755-764 (FP64 A) is the batched
 * GEMM 1 * A^T A + 0 * B, whose summation order is the library's. */
void orc_ara_syrk(int m, int n, const double* A, int lda, double* B, int ldb)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int r = 0; r < m; ++r) s = fma(A[r + (long)i * lda], A[r + (long)j * lda], s);
            B[i + (long)j * ldb] = s;
        }
}

/* ARA right triangular solve B(:, 0:c) := B(:, 0:c) R^-1, c = min(cols, bs),
 * R upper: src/batch_svd/ara_util.cuh:345-408 (ara_16x16_trsm_kernel, per
 * row: b_j -= R(k,j) b_k for k < j, then b_j /= R(j,j)) and :560-640 (the
 * 32-column case as two 16-column solves around B2 -= B1 R12, a batched
 * GEMM); restated with one sequential sum over k < j. */
void orc_ara_trsm(int rows, int cols, int bs, double* B, int ldb, const double* R, int ldr)
{
    if (cols > bs) cols = bs;
    for (int r = 0; r < rows; ++r)
        for (int j = 0; j < cols; ++j) {
            double s = B[r + (long)j * ldb];
            for (int k = 0; k < j; ++k) s = fma(-R[k + (long)j * ldr], B[r + (long)k * ldb], s);
            B[r + (long)j * ldb] = s / R[j + (long)j * ldr];
        }
}

/* ======================================================================= */
/* Complex precisions (C / Z)                                               */
/* ======================================================================= */
/* The reference instantiates its batched templates for cuFloatComplex /
 * cuDoubleComplex (src/batch_triangular/Makefile:30-57) with the operators of
 * include/kblas_operators.h: a * b and a + b without conjugation (:150-200),
 * a / b = (a conj(b)) / |b|^2 written out (:237-241), sqrt(x) from the modulus
 * and cos(arg) with the sign of Im(x) (:243-255).  So its complex POTRF is
 * the factorization A = L L^T of a complex symmetric matrix (its tests build
 * such matrices: testing/testing_helper.cu:404-414).  Restated here over
 * interleaved (real, imaginary) doubles; the fused multiply-add a b + c is
 * two FMA chains per part.  'C' (conjugate transpose) conjugates the operand
 * in GEMM and TRSM.  "Lower view": Lv(i,j) = A(i,j) (lower) or A(j,i) (upper). */
typedef struct { double x, y; } zc;

static zc zmk(double x, double y) { zc r; r.x = x; r.y = y; return r; }
static zc zmul(zc a, zc b) { return zmk(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y); }
static zc zdiv(zc a, zc b)
{
    const double d = b.x * b.x + b.y * b.y;
    return zmk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
static zc zfma(zc a, zc b, zc c) { return zmk(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))); }
static zc zneg(zc a) { return zmk(-a.x, -a.y); }
static zc zconj(zc a) { return zmk(a.x, -a.y); }

/* exported for the tests: the reference's principal square root */
void orc_zsqrt(double re, double im, double* out)
{
    const double r = hypot(re, im);
    if (r == 0.0) { out[0] = 0.0; out[1] = 0.0; return; }
    const double c = re / r;
    double y = sqrt(r * (1.0 - c) / 2.0);
    if (signbit(im)) y = -y;
    out[0] = sqrt(r * (c + 1.0) / 2.0);
    out[1] = y;
}

static zc* lv(zc* A, int lda, int upper, int i, int j)
{
    return upper ? A + j + (long)i * lda : A + i + (long)j * lda;
}

/* Column step of src/batch_triangular/Xpotrf_batch_kernels.cuh:52-69 on the
 * lower view; info = j + 1 at a zero or non-finite pivot (the reference has
 * no check).  Returns info. */
int orc_zpotrf(int upper, int n, double* Ad, int lda)
{
    zc* A = (zc*)Ad;
    for (int j = 0; j < n; ++j) {
        zc d = *lv(A, lda, upper, j, j);
        const double m = d.x * d.x + d.y * d.y;
        if (!(m > 0.0) || !(m < INFINITY)) return j + 1;
        double s2[2];
        orc_zsqrt(d.x, d.y, s2);
        const zc s = zmk(s2[0], s2[1]);
        *lv(A, lda, upper, j, j) = s;
        for (int i = j + 1; i < n; ++i) *lv(A, lda, upper, i, j) = zdiv(*lv(A, lda, upper, i, j), s);
        for (int i = j + 1; i < n; ++i) {
            const zc lij = *lv(A, lda, upper, i, j);
            for (int k = j + 1; k <= i; ++k)
                *lv(A, lda, upper, i, k) = zfma(zneg(lij), *lv(A, lda, upper, k, j), *lv(A, lda, upper, i, k));
        }
    }
    return 0;
}

/* (Lv Lv^T) X = B for the nrhs right-hand sides: side L -> columns of the
 * n x nrhs B, side R -> rows of the nrhs x n B (X A = B). */
void orc_zpotrs(int upper, int side_left, int n, int nrhs, const double* Ad, int lda, double* Bd, int ldb)
{
    zc* A = (zc*)Ad;
    zc* B = (zc*)Bd;
#define ZB(i, c) (side_left ? B + (i) + (long)(c) * ldb : B + (c) + (long)(i) * ldb)
    for (int c = 0; c < nrhs; ++c) {
        for (int i = 0; i < n; ++i) {
            zc v = *ZB(i, c);
            for (int k = 0; k < i; ++k) v = zfma(zneg(*lv(A, lda, upper, i, k)), *ZB(k, c), v);
            *ZB(i, c) = zdiv(v, *lv(A, lda, upper, i, i));
        }
        for (int i = n - 1; i >= 0; --i) {
            zc v = *ZB(i, c);
            for (int k = i + 1; k < n; ++k) v = zfma(zneg(*lv(A, lda, upper, k, i)), *ZB(k, c), v);
            *ZB(i, c) = zdiv(v, *lv(A, lda, upper, i, i));
        }
    }
#undef ZB
}

/* op(X)(i, j) for op 'N', 'T', 'C' */
static zc zop(const zc* X, int ld, char op, int i, int j)
{
    if (op == 'N' || op == 'n') return X[i + (long)j * ld];
    const zc v = X[j + (long)i * ld];
    return (op == 'C' || op == 'c') ? zconj(v) : v;
}

/* C := alpha op(A) op(B) + beta C (beta == 0: C not read) */
void orc_zgemm(char ta, char tb, int m, int n, int k, const double* alpha, const double* Ad, int lda,
               const double* Bd, int ldb, const double* beta, double* Cd, int ldc)
{
    const zc* A = (const zc*)Ad;
    const zc* B = (const zc*)Bd;
    zc* C = (zc*)Cd;
    const zc al = zmk(alpha[0], alpha[1]), be = zmk(beta[0], beta[1]);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            zc s = zmk(0.0, 0.0);
            for (int p = 0; p < k; ++p) s = zfma(zop(A, lda, ta, i, p), zop(B, ldb, tb, p, j), s);
            zc r = zmul(al, s);
            if (be.x != 0.0 || be.y != 0.0) r = zfma(be, C[i + (long)j * ldc], r);
            C[i + (long)j * ldc] = r;
        }
}

/* The lower (or upper) part of C := alpha op(A) op(A)^T + beta C (no conjugation). */
void orc_zsyrk(int upper, char trans, int n, int k, const double* alpha, const double* Ad, int lda,
               const double* beta, double* Cd, int ldc)
{
    const zc* A = (const zc*)Ad;
    zc* C = (zc*)Cd;
    const zc al = zmk(alpha[0], alpha[1]), be = zmk(beta[0], beta[1]);
    const char tt = (trans == 'N' || trans == 'n') ? 'T' : 'N';
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            if (upper ? i > j : i < j) continue;
            zc s = zmk(0.0, 0.0);
            for (int p = 0; p < k; ++p) s = zfma(zop(A, lda, trans, i, p), zop(A, lda, tt, p, j), s);
            zc r = zmul(al, s);
            if (be.x != 0.0 || be.y != 0.0) r = zfma(be, C[i + (long)j * ldc], r);
            C[i + (long)j * ldc] = r;
        }
}

/* op(T) X = alpha B (side L) or X op(T) = alpha B (side R), T triangular
 * (lower or upper storage), op 'N' / 'T' / 'C', unit diagonal or not; B is
 * m x n.  Substitution over the dense op(T), one right-hand side at a time. */
void orc_ztrsm(char side, int upper, char trans, int unit, int m, int n, const double* alpha,
               const double* Ad, int lda, double* Bd, int ldb)
{
    const zc* A = (const zc*)Ad;
    zc* B = (zc*)Bd;
    const int left = side == 'L' || side == 'l';
    const int ord = left ? m : n, nv = left ? n : m;
    const zc al = zmk(alpha[0], alpha[1]);
    /* dense op(T) (ord x ord), zeros outside the triangle */
    zc* T = (zc*)calloc((size_t)ord * ord + 1, sizeof(zc));
    for (int j = 0; j < ord; ++j)
        for (int i = 0; i < ord; ++i) {
            const int in = upper ? i <= j : i >= j;
            if (!in) continue;
            zc v = (unit && i == j) ? zmk(1.0, 0.0) : A[i + (long)j * lda];
            T[i + (long)j * ord] = v;
        }
    /* M := op(T) */
    zc* M = (zc*)calloc((size_t)ord * ord + 1, sizeof(zc));
    for (int j = 0; j < ord; ++j)
        for (int i = 0; i < ord; ++i) M[i + (long)j * ord] = zop(T, ord, trans, i, j);
    /* side L: M x = al b (x a column of B); side R: x M = al b (x a row) -> M^T x^T = al b^T */
    const int lowerM = upper ? (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c')
                             : !(trans == 'T' || trans == 't' || trans == 'C' || trans == 'c');
    /* effective system S y = r with S = M (side L) or M^T (side R) */
    const int lowerS = left ? lowerM : !lowerM;
    for (int v = 0; v < nv; ++v) {
#define ZX(i) (left ? B + (i) + (long)v * ldb : B + v + (long)(i) * ldb)
#define SS(i, j) (left ? M[(i) + (long)(j) * ord] : M[(j) + (long)(i) * ord])
        for (int i = 0; i < ord; ++i) *ZX(i) = zmul(al, *ZX(i));
        if (lowerS) {
            for (int i = 0; i < ord; ++i) {
                zc s = *ZX(i);
                for (int k = 0; k < i; ++k) s = zfma(zneg(SS(i, k)), *ZX(k), s);
                *ZX(i) = zdiv(s, SS(i, i));
            }
        } else {
            for (int i = ord - 1; i >= 0; --i) {
                zc s = *ZX(i);
                for (int k = i + 1; k < ord; ++k) s = zfma(zneg(SS(i, k)), *ZX(k), s);
                *ZX(i) = zdiv(s, SS(i, i));
            }
        }
#undef ZX
#undef SS
    }
    free(T);
    free(M);
}

/* Complex inverse family and TRMM on the lower triangle (the reference's
 * templates for C / Z, src/batch_triangular/X{trtri,lauum,potri,trmm}_batch*,
 * without conjugation; TRMM's 'C' conjugates the triangle): TRTRI A := L^-1
 * (info: first exactly-zero diagonal element of a non-unit triangle, A left
 * unchanged), LAUUM A := L^T L, POTRI A := (L L^T)^-1 = LAUUM(TRTRI(L)). */
int orc_ztrtri(int unit, int n, double* Ad, int lda)
{
    zc* A = (zc*)Ad;
    if (!unit)
        for (int j = 0; j < n; ++j) {
            const zc d = A[j + (long)j * lda];
            if (d.x == 0.0 && d.y == 0.0) return j + 1;
        }
    zc* X = (zc*)calloc((size_t)n * n + 1, sizeof(zc));
    for (int j = 0; j < n; ++j) {
        const zc djj = unit ? zmk(1.0, 0.0) : A[j + (long)j * lda];
        X[j + (long)j * n] = zdiv(zmk(1.0, 0.0), djj);
        for (int i = j + 1; i < n; ++i) {
            zc s = zmk(0.0, 0.0);
            for (int k = j; k < i; ++k) {
                const zc lik = A[i + (long)k * lda];
                s = zfma(lik, X[k + (long)j * n], s);
            }
            const zc dii = unit ? zmk(1.0, 0.0) : A[i + (long)i * lda];
            X[i + (long)j * n] = zdiv(zneg(s), dii);
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i)
            if (!(unit && i == j)) A[i + (long)j * lda] = X[i + (long)j * n];
    free(X);
    return 0;
}

void orc_zlauum(int n, double* Ad, int lda)
{
    zc* A = (zc*)Ad;
    zc* X = (zc*)calloc((size_t)n * n + 1, sizeof(zc));
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            zc s = zmk(0.0, 0.0);
            for (int k = i; k < n; ++k) s = zfma(A[k + (long)i * lda], A[k + (long)j * lda], s);
            X[i + (long)j * n] = s;
        }
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) A[i + (long)j * lda] = X[i + (long)j * n];
    free(X);
}

int orc_zpotri(int n, double* A, int lda)
{
    const int info = orc_ztrtri(0, n, A, lda);
    if (info) return info;
    orc_zlauum(n, A, lda);
    return 0;
}

/* B := alpha op(L) B (side L, L of order m) or alpha B op(L) (side R, order n) */
void orc_ztrmm(char side, char trans, int unit, int m, int n, const double* alpha, const double* Ad, int lda,
               double* Bd, int ldb)
{
    const zc* A = (const zc*)Ad;
    zc* B = (zc*)Bd;
    const int left = side == 'L' || side == 'l';
    const int ord = left ? m : n;
    const zc al = zmk(alpha[0], alpha[1]);
    zc* T = (zc*)calloc((size_t)ord * ord + 1, sizeof(zc));
    for (int j = 0; j < ord; ++j)
        for (int i = j; i < ord; ++i) T[i + (long)j * ord] = (unit && i == j) ? zmk(1.0, 0.0) : A[i + (long)j * lda];
    zc* R = (zc*)calloc((size_t)m * n + 1, sizeof(zc));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            zc s = zmk(0.0, 0.0);
            if (left)
                for (int k = 0; k < ord; ++k) s = zfma(zop(T, ord, trans, i, k), B[k + (long)j * ldb], s);
            else
                for (int k = 0; k < ord; ++k) s = zfma(B[i + (long)k * ldb], zop(T, ord, trans, k, j), s);
            R[i + (long)j * m] = zmul(al, s);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) B[i + (long)j * ldb] = R[i + (long)j * m];
    free(T);
    free(R);
}
