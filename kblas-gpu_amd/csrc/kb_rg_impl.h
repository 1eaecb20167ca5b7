// kb_rg_impl.h -- batched Cholesky family (POTRF / POSV / POTRS) for n <= 32,
// one matrix per 16-lane DPP row of a wavefront (gfx950).
//
// Layout: lane l of a row holds matrix row l (slot 0) and, for the 32-wide
// kernel, row 16 + l (slot 1), every column of it up to the slot's last row,
// in VGPRs.  A wavefront carries four matrices.  The only cross-lane traffic
// is a read of another lane of the same row, and it is fused into the FMA that
// consumes it:
//     v_fmac_f64_dpp acc, src, mul row_newbcast:k   (acc += src[lane k] * mul)
// so one trailing-update term costs one VALU instruction.
//
// Factorization (right-looking, the reference's column step of
// src/batch_triangular/Xpotrf_batch_kernels.cuh:37-78 restated in the form
// A = Lu D^2 Lu^T, Lu unit lower):
//     step j:  d = A(j,j);  inv = 1/sqrt(d)
//              L(:,j)   = A(:,j) * inv           (written to memory at once)
//              -Lu(:,j) = A(:,j) * (-inv^2)      (kept for the update / solve)
//              A(i,k)  += -Lu(i,j) * A(k,j)      k > j, A(k,j) read from lane k
// Solves (POSV / POTRS, reference src/batch_triangular/Xposv_batch_drivers.cuh:83-114
// and Xpotrs_batch_drivers.cuh) with the same fused reads:
//     z = Lu^-1 b      (unit forward sweep, z(k) read from lane k)
//     w = D^-2 z
//     x = Lu^-T w      (unit backward sweep; lane i needs -Lu(k,i), a column of
//                       what it holds, so the rows are transposed once through
//                       LDS)
// Rows and columns at or beyond n are padded with the identity, so every
// kernel runs its full 16 / 32 steps without branches on n.
#pragma once
#include "kb_device.h"
#include "kb_launch.h"

// a(s, k) is kept only for columns up to the slot's last row
constexpr KB_DEV bool rg_live(int s, int k) { return k < 16 * (s + 1); }
// LDS copy of -Lu: row k holds the columns of its slot, 16 (k / 16 + 1), at an
// odd pitch so that the row-wise writes are free of bank conflicts.  Rows do
// not overlap, so the writes need no masking.
constexpr KB_DEV int rg_row_off(int k) { return k < 16 ? 17 * k : 17 * 16 + 33 * (k - 16); }
template <int NB> constexpr KB_DEV int rg_lds_per() { return NB == 16 ? 17 * 16 : 17 * 16 + 33 * 16; }

// ---------------------------------------------------------------------------
// load the lower triangle (upper part and padding never read)
// ---------------------------------------------------------------------------
template <typename T, int NB, bool EX>
KB_DEV void rg_load(T (&a)[NB / 16][NB], const T* Ap, int lda, int n, int l)
{
#pragma unroll
    for (int s = 0; s < NB / 16; ++s) {
        const int i = 16 * s + l;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (!rg_live(s, k)) continue;
            const bool colin = EX || k < n;
            const bool in = k <= i && (EX || i < n) && colin;
            // branch-free: other lanes fetch the column's diagonal (same line)
            const int kk = colin ? k : 0;
            const int rr = in ? i : kk;
            const T v = gld(Ap + rr + (long)kk * lda);
            a[s][k] = in ? v : ((i == k) ? T(1) : T(0));
        }
    }
}

// ---------------------------------------------------------------------------
// factorization.  Writes L (rows j..n-1 of column j) as soon as column j is
// final (wr: this row of lanes writes).  KEEP: leave -Lu (strictly lower, zero elsewhere) in a and D^-2 of the
// lane's own rows in dinv2 for the solves.  Returns the LAPACK info.
// ---------------------------------------------------------------------------
// lds_d / lds_inv (optional, written when wr): L as a dense NB x NB lower
// triangle at pitch NB + 1 (zero above the diagonal) and 1 / L(j,j).
template <typename T, int NB, bool EX, bool KEEP>
KB_DEV int rg_potrf(T (&a)[NB / 16][NB], T (&dinv2)[NB / 16], int n, int l, T* Ap, int lda, bool wr,
                    T* lds_d = nullptr, T* lds_inv = nullptr)
{
    constexpr int R = NB / 16;
    int info = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int sj = j >> 4, lj = j & 15;
        const T d = row_bcast(a[sj][j], lj);
        info = (info == 0 && !(d > T(0))) ? j + 1 : info;
        const T inv = kb_rsqrt<T>(d);
        const T ninv2 = -(inv * inv);
        T nl[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (s < sj) continue;
            const int i = 16 * s + l;
            const T lv = a[s][j] * inv;          // the diagonal gives d / sqrt(d)
            if (wr && i >= j && (EX || i < n)) gst(Ap + i + (long)j * lda, lv);
            nl[s] = a[s][j] * ninv2;
            if (lds_d && wr) lds_d[i * (NB + 1) + j] = (i >= j) ? lv : T(0);
        }
        if (lds_d && wr) {
#pragma unroll
            for (int s = 0; s < R; ++s)
                if (s < sj) lds_d[(16 * s + l) * (NB + 1) + j] = T(0);
            if (l == 0) lds_inv[j] = inv;
        }
        if (KEEP) dinv2[sj] = (l == lj) ? -ninv2 : dinv2[sj];
#pragma unroll
        for (int k = j + 1; k < NB; ++k) {
            const int sk = k >> 4;
#pragma unroll
            for (int s = 0; s < R; ++s)
                if (s >= sk) fmac_rowk<T>(a[s][k], a[sk][j], nl[s], k);
        }
        if (KEEP) {
#pragma unroll
            for (int s = 0; s < R; ++s) {
                if (s < sj) continue;
                a[s][j] = (s > sj || l > lj) ? nl[s] : T(0);
            }
        }
    }
    return info;
}

// POTRS: -Lu and D^-2 from a factor L already in a (loaded by rg_load);
// dg[s] = L(i,i) of the lane's row i (1 on padding rows).
template <typename T, int NB>
KB_DEV void rg_from_factor(T (&a)[NB / 16][NB], T (&dinv2)[NB / 16], const T (&dg)[NB / 16], int l)
{
    constexpr int R = NB / 16;
    T myinv[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        myinv[s] = T(1) / dg[s];
        dinv2[s] = myinv[s] * myinv[s];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int sk = k >> 4, lk = k & 15;
        const T ninvk = -row_bcast(myinv[sk], lk);
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (s < sk) continue;
            a[s][k] = (s > sk || l > lk) ? a[s][k] * ninvk : T(0);
        }
    }
}

// Store -Lu (zero on and above the diagonal) row by row into LDS.
template <typename T, int NB>
KB_DEV void rg_lds_rows(T* Lt, const T (&a)[NB / 16][NB], int l)
{
    constexpr int R = NB / 16;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        T* row = Lt + rg_row_off(16 * s + l);
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (rg_live(s, c)) row[c] = a[s][c];
    }
}

// ---------------------------------------------------------------------------
// solve for RC right-hand sides held as x[c][s] (row 16 s + l)
// ---------------------------------------------------------------------------
template <typename T, int NB, int RC>
KB_DEV void rg_solve(const T (&a)[NB / 16][NB], const T (&dinv2)[NB / 16], const T* Lt, T (&x)[RC][NB / 16],
                     int l)
{
    constexpr int R = NB / 16;
    // z = Lu^-1 b
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int sk = k >> 4;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            if (s < sk) continue;
#pragma unroll
            for (int c = 0; c < RC; ++c) fmac_rowk<T>(x[c][s], x[c][sk], a[s][k], k);
        }
    }
    // w = D^-2 z
#pragma unroll
    for (int s = 0; s < R; ++s)
#pragma unroll
        for (int c = 0; c < RC; ++c) x[c][s] *= dinv2[s];
    // x = Lu^-T w: lane row i adds -Lu(k, i) x(k) for k > i
#pragma unroll
    for (int k = NB - 1; k > 0; --k) {
        const int sk = k >> 4;
#pragma unroll
        for (int s = 0; s <= sk; ++s) {
            const T m = Lt[rg_row_off(k) + 16 * s + l];   // zero for rows i >= k
#pragma unroll
            for (int c = 0; c < RC; ++c) fmac_rowk<T>(x[c][s], x[c][sk], m, k);
        }
    }
}

// ---------------------------------------------------------------------------
// one matrix per 16-lane row; WPB wavefronts (4 WPB matrices) per workgroup
// ---------------------------------------------------------------------------
template <typename T, int NB, int MODE, int RC, bool EX, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_rg_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    constexpr int R = NB / 16;
    constexpr int LPER = rg_lds_per<NB>();
    constexpr int LDSN = (MODE == KB_MODE_POTRF) ? 1 : 4 * WPB * LPER;
    __shared__ T lds[LDSN];

    const int l = threadIdx.x & 15;
    const int g = threadIdx.x >> 4;               // matrix slot in the workgroup
    const long mat = (long)blockIdx.x * (4 * WPB) + g;
    // A row of lanes past the batch end works on the last matrix without
    // writing anything, so every row keeps executing the same instructions.
    const bool live = mat < batch;
    const long m = live ? mat : (long)batch - 1;

    T* Ap = A.get(m);
    const int lda = A.ldim(m);
    T a[R][NB];
    T dinv2[R];
#pragma unroll
    for (int s = 0; s < R; ++s) dinv2[s] = T(1);
    rg_load<T, NB, EX>(a, Ap, lda, n, l);

    int inf = 0;
    if constexpr (MODE != KB_MODE_POTRS) {
        inf = rg_potrf<T, NB, EX, MODE == KB_MODE_POSV>(a, dinv2, n, l, Ap, lda, live);
        if (live && l == 0 && info) info[mat] = inf;
    } else {
        T dg[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const int i = 16 * s + l;
            const bool in = EX || i < n;
            const T v = gld(Ap + (in ? i : 0) * (long)(lda + 1));
            dg[s] = in ? v : T(1);
        }
        rg_from_factor<T, NB>(a, dinv2, dg, l);
    }
    if constexpr (MODE != KB_MODE_POTRF) {
        T* Lt = lds + g * LPER;
        rg_lds_rows<T, NB>(Lt, a, l);
        wave_lds_sync();
        T* Bp = B.get(m);
        const int ldb = B.ldim(m);
        const bool ok = live && inf == 0;
        for (int c0 = 0; c0 < nrhs; c0 += RC) {
            T x[RC][R];
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const int i = 16 * s + l;
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    const bool in = (EX || i < n) && c0 + c < nrhs;
                    const T v = gld(rhs_elem(Bp, ldb, side_left, in ? i : 0, in ? c0 + c : 0));
                    x[c][s] = in ? v : T(0);
                }
            }
            rg_solve<T, NB, RC>(a, dinv2, Lt, x, l);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const int i = 16 * s + l;
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (ok && (EX || i < n) && c0 + c < nrhs) gst(rhs_elem(Bp, ldb, side_left, i, c0 + c), x[c][s]);
            }
        }
    }
}
