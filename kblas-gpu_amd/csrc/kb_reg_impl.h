// kb_reg_impl.h -- register-resident batched Cholesky family for n <= 64 (gfx950).
//
// Layout: one matrix per G-lane group (G = 16, 32 or 64 = padded order NB),
// lane r holds row r of the lower triangle in VGPRs.  Factorization is the
// right-looking column algorithm of the reference's register kernels
// (src/batch_triangular/Xpotrf_batch_kernels.cuh:37-78): per column j a
// broadcast of the pivot, one reciprocal square root, a scale, then one
// broadcast + FMA per trailing column.  Unlike the reference, which stops at
// n = 16 and recurses through ~28 launches at n = 64
// (src/batch_triangular/Xpotrf_batch_drivers.cuh:42-133,
//  src/batch_triangular/Xposv_batch_drivers.cuh:83-114), POTRF, POTRS and
// POSV run as ONE kernel per call: A is read once and (for POTRF/POSV) written
// once, the solve runs on the factor still in registers, and only the
// transposed factor for the backward sweep goes through LDS.
#pragma once
#include "kb_device.h"
#include "kb_launch.h"

template <typename T> KB_DEV T kb_fma(T a, T b, T c);
template <> KB_DEV double kb_fma<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> KB_DEV float kb_fma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// ---------------------------------------------------------------------------
// In-register Cholesky of the group's matrix.  a[k] = A(r,k) on entry for
// k <= r (entries above the diagonal are don't-care); on exit a[k] = L(r,k).
// Returns LAPACK info (0, or j+1 for the first non-positive pivot); after a
// failure the remaining columns are left as they were, like dpotrf.
// ---------------------------------------------------------------------------
template <typename T, int NB, int G, bool EX = false>
KB_DEV int reg_potrf(T (&a)[NB], int n, int r, T& my_diag, T& my_inv)
{
    // EX: n == NB, every guard below folds away.  A non-positive pivot only
    // records info (first one wins); the later columns then carry NaNs, the
    // columns before it are final -- the part of the factor LAPACK defines.
    int info = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (EX || j < n) {
            const T d = bcastk<G>(a[j], j);
            info = (info == 0 && !(d > T(0))) ? j + 1 : info;
            const T inv = kb_rsqrt<T>(d);
            const T s = d * inv;
            if (r == j) { my_diag = s; my_inv = inv; }
            a[j] = (r == j) ? s : a[j] * inv;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (k > j && (EX || k < n)) {
                    const T lk = bcastk<G>(a[j], k);
                    a[k] = kb_fma<T>(-a[j], lk, a[k]);
                }
            }
        }
    }
    return info;
}

// Forward sweep L y = b on RC vectors; a[] holds the strictly lower row
// (zeros at and above the diagonal), my_inv = 1/L(r,r).
template <typename T, int NB, int G, int RC, bool EX = false>
KB_DEV void reg_forward(const T (&a)[NB], T (&x)[RC], int n, T my_inv)
{
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        if (EX || k < n) {
            T ik = bcastk<G>(my_inv, k);
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                T yk = bcastk<G>(x[c], k) * ik;
                x[c] = kb_fma<T>(-a[k], yk, x[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < RC; ++c) x[c] *= my_inv;
}

// Backward sweep L^T x = y; Lt points at the group's LDS copy of L stored by
// rows with stride NB+1 and zeros on/above the diagonal, so lane r reads
// L(k,r) at Lt[k*(NB+1) + r] (consecutive lanes -> consecutive banks).
template <typename T, int NB, int G, int RC, bool EX = false>
KB_DEV void reg_backward(const T* Lt, T (&x)[RC], int n, int r, T my_inv)
{
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
        if (EX || k < n) {
            T l = Lt[k * (NB + 1) + r];
            T ik = bcastk<G>(my_inv, k);
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                T xk = bcastk<G>(x[c], k) * ik;
                x[c] = kb_fma<T>(-l, xk, x[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < RC; ++c) x[c] *= my_inv;
}

template <typename T, int NB, int G>
KB_DEV void lds_store_rows(T* Lt, const T (&a)[NB], int r)
{
#pragma unroll
    for (int k = 0; k < NB; ++k) Lt[r * (NB + 1) + k] = a[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Matrices per workgroup: bounded by 64 lanes / G and by ~34 KB of LDS.
template <typename T, int NB, int MODE>
constexpr int reg_mpb()
{
    if (MODE == KB_MODE_POTRF) return 256 / NB;
    int per = NB * (NB + 1) * (int)sizeof(T);
    int m = 34816 / per;
    if (m < 1) m = 1;
    if (m > 256 / NB) m = 256 / NB;
    // keep whole wavefronts: groups per block must fill 64 lanes
    int per_wave = 64 / NB;
    if (m < per_wave) m = per_wave;
    return m;
}


// Backward sweep reading L(k,r) straight from the factor in global memory
// (used where no LDS copy is kept: lanes read a row of L, which the same
// wavefront has just written and fenced).
template <typename T, int NB, int G, int RC>
KB_DEV void reg_backward_global(const T* Ap, int lda, T (&x)[RC], int n, int r, T my_inv)
{
#pragma unroll
    for (int k = NB - 1; k >= 0; --k) {
        if (k < n) {
            T l = (r < k) ? gld(Ap + k + (long)r * lda) : T(0);
            T ik = bcastk<G>(my_inv, k);
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                T xk = bcastk<G>(x[c], k) * ik;
                x[c] = kb_fma<T>(-l, xk, x[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < RC; ++c) x[c] *= my_inv;
}

// One matrix on one G-lane group (lane r = row r).  Lt: the group's LDS area
// of NB*(NB+1) elements, or nullptr to read the transposed factor from A.
template <typename T, int NB, int G, int MODE, int RC, bool EX = false>
KB_DEV void reg_chol_one(int n, int nrhs, int side_left, T* Ap, int lda, T* Bp, int ldb,
                         int* info_out, T* Lt, int r)
{
    T a[NB];
    T my_diag = T(1), my_inv = T(1);
    const bool rowin = EX || r < n;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        // branch-free: lanes above the diagonal (and padding lanes) load the
        // column's diagonal element, a line the lower part fetches anyway
        const bool in = k <= r && rowin && (EX || k < n);
        const int rr = in ? r : ((EX || k < n) ? k : 0);
        const int kk = (EX || k < n) ? k : 0;
        const T v = gld(Ap + rr + (long)kk * lda);
        a[k] = in ? v : T(0);
        if constexpr (MODE == KB_MODE_POTRS) {
            if (k == r) my_diag = in ? v : T(1);
        }
    }
    int inf = 0;
    if constexpr (MODE != KB_MODE_POTRS) {
        inf = reg_potrf<T, NB, G, EX>(a, n, r, my_diag, my_inv);
        // write L (lower triangle only; the strict upper part is never touched).
        // After a failed pivot the unfinished columns keep their partially
        // updated values, as dpotrf leaves them.
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k <= r && r < n) gst(Ap + r + (long)k * lda, a[k]);
        if (r == 0 && info_out) *info_out = inf;
    } else {
        my_inv = T(1) / my_diag;
    }
    if constexpr (MODE != KB_MODE_POTRF) {
        if (inf != 0) return;
        // strictly lower row for the sweeps
#pragma unroll
        for (int k = 0; k < NB; ++k) a[k] = (k < r) ? a[k] : T(0);
        if (Lt) lds_store_rows<T, NB, G>(Lt, a, r);
        else wave_global_sync();
        for (int c0 = 0; c0 < nrhs; c0 += RC) {
            T x[RC];
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                T v = T(0);
                if (r < n && c0 + c < nrhs) v = gld(rhs_elem(Bp, ldb, side_left, r, c0 + c));
                x[c] = v;
            }
            reg_forward<T, NB, G, RC, EX>(a, x, n, my_inv);
            if (Lt) reg_backward<T, NB, G, RC, EX>(Lt, x, n, r, my_inv);
            else reg_backward_global<T, NB, G, RC>(Ap, lda, x, n, r, my_inv);
#pragma unroll
            for (int c = 0; c < RC; ++c)
                if (r < n && c0 + c < nrhs) gst(rhs_elem(Bp, ldb, side_left, r, c0 + c), x[c]);
        }
    }
}

template <typename T, int NB, int MODE, int RC, bool EX>
__global__ void __launch_bounds__((reg_mpb<T, NB, MODE>() * NB))
k_reg_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    constexpr int G = NB;
    constexpr int MPB = reg_mpb<T, NB, MODE>();
    constexpr int LDSN = (MODE == KB_MODE_POTRF) ? 1 : MPB * NB * (NB + 1);
    __shared__ T lds[LDSN];

    const int r = threadIdx.x & (G - 1);
    const int g = threadIdx.x / G;
    const long mat = (long)blockIdx.x * MPB + g;
    if (mat >= batch) return;   // the whole group leaves together
    T* Bp = (MODE == KB_MODE_POTRF) ? nullptr : B.get(mat);
    const int ldb = (MODE == KB_MODE_POTRF) ? 1 : B.ldim(mat);
    reg_chol_one<T, NB, G, MODE, RC, EX>(n, nrhs, side_left, A.get(mat), A.ldim(mat), Bp, ldb,
                                     info ? info + mat : nullptr,
                                     (MODE == KB_MODE_POTRF) ? nullptr : lds + g * NB * (NB + 1), r);
}

// ---------------------------------------------------------------------------
// Mixed precision POSV (LAPACK dsposv restated for one matrix per group):
// FP32 Cholesky + FP32 sweeps, FP64 residual r = b - A x with the FP64 row of
// A kept in registers, stop when ||r||inf <= ||x||inf * ||A||inf * eps * sqrt(n)
// (at most 30 refinements), else FP64 Cholesky of the same registers.
// ---------------------------------------------------------------------------
constexpr int KB_IR_ITERMAX = 30;

template <int NB>
constexpr int sposv_mpb()
{
    int per = NB * (NB + 1) * 4;
    int m = 34816 / per;
    if (m > 256 / NB) m = 256 / NB;
    if (m < 64 / NB) m = 64 / NB;
    return m;
}

template <int NB, int G, int RC>
KB_DEV void reg_sposv_one(int n, int nrhs, int side_left, double* Ap, int lda, double* Bp, int ldb,
                          double* Xp, int ldx, int* info_out, int* iter_out, float* Lt, int r)
{
    // full row r of the symmetric matrix from its lower triangle
    double ad[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const bool in = r < n && k < n;
        const double v = gld_opaque(Ap + (in ? ((k <= r) ? r + (long)k * lda : k + (long)r * lda) : 0));
        ad[k] = in ? v : 0.0;
    }
    double rowsum = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) rowsum += __builtin_fabs(ad[k]);
    const double anrm = group_max<G>(rowsum);
    const double cte = anrm * 1.1102230246251565e-16 * __builtin_sqrt((double)n);

    // FP32 range of B, then of A (dsposv's DLAG2S / DLAT2S checks): ITER = -2
    int small = 1;
    for (int c0 = 0; c0 < nrhs; c0 += 4) {   // four loads in flight per step
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bool in = r < n && c0 + c < nrhs;
            v[c] = gld_opaque(rhs_elem(Bp, ldb, side_left, in ? r : 0, in ? c0 + c : 0));
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (r < n && c0 + c < nrhs) small &= !(__builtin_fabs(v[c]) > 3.4028234663852886e38);
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) small &= !(k <= r && __builtin_fabs(ad[k]) > 3.4028234663852886e38);
    const bool ovf = !group_all<G>(small);

    float a[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) a[k] = (k <= r) ? (float)ad[k] : 0.0f;
    float d32 = 1.0f, inv32 = 1.0f;
    int finfo = 0;
    if (!ovf) finfo = reg_potrf<float, NB, G>(a, n, r, d32, inv32);

    int iters = ovf ? -2 : (finfo != 0 ? -3 : 0);
    bool ok = !ovf && (finfo == 0);
    if (ok) {
#pragma unroll
        for (int k = 0; k < NB; ++k) a[k] = (k < r) ? a[k] : 0.0f;
        lds_store_rows<float, NB, G>(Lt, a, r);
        for (int c0 = 0; c0 < nrhs && ok; c0 += RC) {
            double b[RC], x[RC], res[RC];
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                const bool in = r < n && c0 + c < nrhs;
                const double lv = gld_opaque(rhs_elem(Bp, ldb, side_left, in ? r : 0, in ? c0 + c : 0));
                const double v = in ? lv : 0.0;
                b[c] = v;
                x[c] = 0.0;
                res[c] = v;
            }
            bool conv = false, rovf = false;
            int it = 0;
            for (; it <= KB_IR_ITERMAX; ++it) {
                int fits = 1;
#pragma unroll
                for (int c = 0; c < RC; ++c) fits &= !(__builtin_fabs(res[c]) > 3.4028234663852886e38);
                if (!group_all<G>(fits)) { rovf = true; break; }     // DLAG2S(R) overflow
                float z[RC];
#pragma unroll
                for (int c = 0; c < RC; ++c) z[c] = (float)res[c];
                reg_forward<float, NB, G, RC>(a, z, n, inv32);
                reg_backward<float, NB, G, RC>(Lt, z, n, r, inv32);
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    x[c] += (double)z[c];
                    res[c] = b[c];
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    if (k < n) {
#pragma unroll
                        for (int c = 0; c < RC; ++c) {
                            double xk = bcastk<G>(x[c], k);
                            res[c] = __builtin_fma(-ad[k], xk, res[c]);
                        }
                    }
                }
                int all = 1;
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    double rn = group_max<G>(__builtin_fabs(res[c]));
                    double xn = group_max<G>(__builtin_fabs(x[c]));
                    all &= (rn <= xn * cte) ? 1 : 0;
                }
                if (all) { conv = true; break; }
            }
            if (!conv) {
                ok = false;
                iters = rovf ? -2 : -(KB_IR_ITERMAX + 1);
                break;
            }
            if (it > iters) iters = it;
#pragma unroll
            for (int c = 0; c < RC; ++c)
                if (r < n && c0 + c < nrhs) gst(rhs_elem(Xp, ldx, side_left, r, c0 + c), x[c]);
        }
    }
    int inf = 0;
    if (!ok) {
        // FP64 fallback on the same registers; A receives L as in LAPACK dsposv
        double dd = 1.0, dinv = 1.0;
        inf = reg_potrf<double, NB, G>(ad, n, r, dd, dinv);
#pragma unroll
        for (int k = 0; k < NB; ++k)
            if (k <= r && r < n && k < n) gst(Ap + r + (long)k * lda, ad[k]);
        if (inf == 0) {
#pragma unroll
            for (int k = 0; k < NB; ++k) ad[k] = (k < r) ? ad[k] : 0.0;
            wave_global_sync();
            for (int c0 = 0; c0 < nrhs; c0 += RC) {
                double x[RC];
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    const bool in = r < n && c0 + c < nrhs;
                    const double lv = gld_opaque(rhs_elem(Bp, ldb, side_left, in ? r : 0, in ? c0 + c : 0));
                    x[c] = in ? lv : 0.0;
                }
                reg_forward<double, NB, G, RC>(ad, x, n, dinv);
                // backward sweep reads L(k,r) from the factor just written to A
#pragma unroll
                for (int k = NB - 1; k >= 0; --k) {
                    if (k < n) {
                        double l = (r < k) ? gld(Ap + k + (long)r * lda) : 0.0;
                        double ik = bcastk<G>(dinv, k);
#pragma unroll
                        for (int c = 0; c < RC; ++c) {
                            double xk = bcastk<G>(x[c], k) * ik;
                            x[c] = __builtin_fma(-l, xk, x[c]);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    x[c] *= dinv;
                    if (r < n && c0 + c < nrhs) gst(rhs_elem(Xp, ldx, side_left, r, c0 + c), x[c]);
                }
            }
        }
    }
    if (r == 0) {
        if (info_out) *info_out = inf;
        if (iter_out) *iter_out = iters;
    }
}

template <int NB, int RC>
__global__ void __launch_bounds__((sposv_mpb<NB>() * NB))
k_reg_sposv(int n, int nrhs, int side_left, BatchArg<double> A, BatchArg<double> B,
            BatchArg<double> X, int* info, int* iter, int batch)
{
    constexpr int G = NB;
    constexpr int MPB = sposv_mpb<NB>();
    __shared__ float lds[MPB * NB * (NB + 1)];
    const int r = threadIdx.x & (G - 1);
    const int g = threadIdx.x / G;
    const long mat = (long)blockIdx.x * MPB + g;
    if (mat >= batch) return;
    reg_sposv_one<NB, G, RC>(n, nrhs, side_left, A.get(mat), A.ldim(mat), B.get(mat), B.ldim(mat),
                             X.get(mat), X.ldim(mat), info ? info + mat : nullptr,
                             iter ? iter + mat : nullptr, lds + g * NB * (NB + 1), r);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <typename T, int NB, int MODE, int RC>
static int launch_chol(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                       int batch, hipStream_t s)
{
    constexpr int MPB = reg_mpb<T, NB, MODE>();
    dim3 grid((unsigned)((batch + MPB - 1) / MPB));
    dim3 block(MPB * NB);
    if (n == NB && NB <= 32)
        hipLaunchKernelGGL((k_reg_chol<T, NB, MODE, RC, true>), grid, block, 0, s, n, nrhs, side_left,
                           to_arg<T>(A), to_arg<T>(B), info, batch);
    else
        hipLaunchKernelGGL((k_reg_chol<T, NB, MODE, RC, false>), grid, block, 0, s, n, nrhs, side_left,
                           to_arg<T>(A), to_arg<T>(B), info, batch);
    return (int)hipGetLastError();
}

template <typename T, int NB>
static int dispatch_rc(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B,
                       int* info, int batch, hipStream_t s)
{
    if (mode == KB_MODE_POTRF) return launch_chol<T, NB, KB_MODE_POTRF, 1>(n, nrhs, side_left, A, B, info, batch, s);
    if (mode == KB_MODE_POSV) {
        if (nrhs <= 1) return launch_chol<T, NB, KB_MODE_POSV, 1>(n, nrhs, side_left, A, B, info, batch, s);
        return launch_chol<T, NB, KB_MODE_POSV, 4>(n, nrhs, side_left, A, B, info, batch, s);
    }
    if (nrhs <= 1) return launch_chol<T, NB, KB_MODE_POTRS, 1>(n, nrhs, side_left, A, B, info, batch, s);
    return launch_chol<T, NB, KB_MODE_POTRS, 4>(n, nrhs, side_left, A, B, info, batch, s);
}

template <int NB, int RC>
static int launch_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                        int* info, int* iter, int batch, hipStream_t s)
{
    constexpr int MPB = sposv_mpb<NB>();
    dim3 grid((unsigned)((batch + MPB - 1) / MPB));
    dim3 block(MPB * NB);
    hipLaunchKernelGGL((k_reg_sposv<NB, RC>), grid, block, 0, s, n, nrhs, side_left,
                       to_arg<double>(A), to_arg<double>(B), to_arg<double>(X), info, iter, batch);
    return (int)hipGetLastError();
}

template <int NB>
static int sposv_rc(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                    int* info, int* iter, int batch, hipStream_t s)
{
    if (nrhs <= 1) return launch_sposv<NB, 1>(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    return launch_sposv<NB, 4>(n, nrhs, side_left, A, B, X, info, iter, batch, s);
}

