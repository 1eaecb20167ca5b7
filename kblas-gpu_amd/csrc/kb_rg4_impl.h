// kb_rg4_impl.h -- batched Cholesky family (POTRF / POSV / POTRS) for
// 32 < n <= 64, one matrix per 16-lane DPP row of a wavefront (gfx950).
//
// Lane l of a row holds matrix rows l, 16 + l, 32 + l and 48 + l (slots 0..3),
// as the n <= 32 kernels of kb_rg_impl.h do, and every cross-lane read is a
// `row_newbcast` DPP operand fused into the FMA that consumes it.  Four whole
// 64 x 64 triangles would need 320 VGPRs per lane, so the factorization walks
// the matrix in 16-column blocks (one slot wide):
//   block J:  load A(rows >= 16J, block J)                        (registers)
//             A(:, J) -= L(:, P) L(J, P)^T  for every earlier block P
//                          (left-looking; L(:, P) re-read from the output,
//                           which this wavefront has just written)
//             factor block J right-looking, column by column, and write it.
// The re-reads cost memory traffic (the per-XCD L2 cannot hold the batch in
// flight), so this kernel serves POTRF with 32 < n <= 48 and the MFMA-tiled
// kernel (kb_tile_impl.h), which keeps the factor in registers, the rest; the
// dispatch table is measured.  The solve routines below are kept with the
// kernel for its POSV/POTRS modes.
// That is the reference's recursive split POTRF -> TRSM -> SYRK -> POTRF
// (src/batch_triangular/Xpotrf_batch_drivers.cuh:91-131) carried out as one
// pass over 16-column blocks inside one kernel, with the SYRK/GEMM block
// updates done by fused DPP FMAs instead of separate launches.
//
// Solves (POSV / POTRS): forward L y = b and backward L^T x = y column by
// column; the column of L (forward) or the row of L read as a column
// (backward) is loaded from the factor in memory, the solution entry of
// row k comes from lane k by DPP.
#pragma once
#include "kb_rg_impl.h"

// Element (i, k) of the lower triangle of a column-major matrix, branch-free:
// other positions (upper part, or beyond n) read A(0,0) and return `other`.
template <typename T, bool EX>
KB_DEV T rg4_ld(const T* Ap, int lda, int n, int i, int k, bool want, T other)
{
    const bool in = want && (EX || (i < n && k < n));
    const T v = gld(Ap + (in ? i + (long)k * lda : 0));
    return in ? v : other;
}

// ---------------------------------------------------------------------------
// factorization: writes L, returns the LAPACK info; myinv[s] = 1 / L(i,i) of
// the lane's row i = 16 s + l (1 on padding rows).
// ---------------------------------------------------------------------------
template <typename T, bool EX>
KB_DEV int rg4_potrf(T* Ap, int lda, int n, int l, bool wr, T (&myinv)[4])
{
    int info = 0;
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        if (!EX && 16 * J >= n) break;             // identity padding: nothing left
        T acc[4][16];
        // ---- load block J (rows of slots J..3) ----
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s < J) continue;
            const int i = 16 * s + l;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int k = 16 * J + c;
                const T pad = (i == k) ? T(1) : T(0);
                acc[s][c] = rg4_ld<T, EX>(Ap, lda, n, i, k, k <= i, pad);
            }
        }
        // ---- left-looking update from the finished blocks ----
        // Runtime loop over the 16 J columns already factored, one column of
        // L in flight ahead of the FMAs that use it.
        if (J > 0) {
            T lp[4], lq[4];
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s >= J) lq[s] = rg4_ld<T, EX>(Ap, lda, n, 16 * s + l, 0, true, T(0));
#pragma nounroll
            for (int k = 0; k < 16 * J; ++k) {
#pragma unroll
                for (int s = 0; s < 4; ++s) lp[s] = lq[s];
                const int kn = (k + 1 < 16 * J) ? k + 1 : k;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (s >= J) lq[s] = rg4_ld<T, EX>(Ap, lda, n, 16 * s + l, kn, true, T(0));
#pragma unroll
                for (int c = 0; c < 16; ++c)
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        if (s >= J) fnmac_rowk<T>(acc[s][c], lp[J], lp[s], c);
            }
        }
        // ---- factor block J, right-looking ----
        // column offset j * lda advanced through an opaque update, so that the
        // 16 store addresses are not all computed (and held) up front
        long jo = (long)(16 * J) * lda;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int j = 16 * J + c;
            if (c > 0) {
                jo += lda;
                asm volatile("" : "+v"(jo));
            }
            const T d = row_bcast(acc[J][c], c);
            info = (info == 0 && !(d > T(0))) ? j + 1 : info;
            asm volatile("" : "+v"(info));
            const T inv = kb_rsqrt<T>(d);
            myinv[J] = (l == c) ? inv : myinv[J];
            T lv[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s < J) continue;
                const int i = 16 * s + l;
                lv[s] = acc[s][c] * inv;             // the diagonal gives d / sqrt(d)
                if (wr && i >= j && (EX || i < n)) gst(Ap + i + jo, lv[s]);
            }
#pragma unroll
            for (int cp = c + 1; cp < 16; ++cp)
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (s >= J) fnmac_rowk<T>(acc[s][cp], lv[J], lv[s], cp);
        }
        // the next block reads this one back from memory
        wave_global_sync();
    }
    return info;
}

// ---------------------------------------------------------------------------
// solve (L L^T) X = B for RC right-hand sides x[c][s] (row 16 s + l), with L
// in memory and myinv[s] = 1 / L(16 s + l, 16 s + l).
// ---------------------------------------------------------------------------
template <typename T, bool EX, int RC>
KB_DEV void rg4_solve(const T* Ap, int lda, int n, int l, const T (&myinv)[4], T (&x)[RC][4])
{
    // forward: y(k) = x(k) / L(k,k); x(i) -= L(i,k) y(k), i > k.  Column
    // k + 1 is loaded before the FMAs of step k.  The column offset advances
    // through an opaque register update and an empty asm orders the memory
    // accesses, so the compiler neither precomputes 64 addresses nor hoists
    // every load of the sweep to the top (both would exhaust the VGPRs).
    T lc[4], ln[4];
    long ko = 0;                                   // k * lda
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int i = 16 * s + l;
        const bool in = i > 0 && (EX || i < n);
        const T v = gld(Ap + (in ? i : 0));
        ln[s] = in ? v : T(0);
    }
#pragma unroll
    for (int k = 0; k < 64; ++k) {
        const int sk = k >> 4, lk = k & 15;
#pragma unroll
        for (int s = 0; s < 4; ++s) lc[s] = ln[s];
        if (k + 1 < 64) {
            ko += lda;
            asm volatile("" : "+v"(ko));
            const int s1 = (k + 1) >> 4;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 16 * s + l;
                if (s >= s1) {
                    const bool in = i > k + 1 && (EX || (i < n && k + 1 < n));
                    const T v = gld(Ap + (in ? i + ko : 0));
                    ln[s] = in ? v : T(0);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const T t = x[c][sk] * myinv[sk];
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s >= sk) fnmac_rowk<T>(x[c][s], t, lc[s], lk);
        }
        asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s) x[c][s] *= myinv[s];
    // backward: x(k) = y(k) / L(k,k); y(i) -= L(k,i) x(k), i < k.  Lane row i
    // reads L(k, i) at Ap + k + i lda: row pointers per slot, k counting down.
    long ro[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) ro[s] = (long)(16 * s + l) * lda;
    int kk = 63;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int i = 16 * s + l;
        const bool in = i < 63 && (EX || 63 < n);
        const T v = gld(Ap + (in ? kk + ro[s] : 0));
        ln[s] = in ? v : T(0);
    }
#pragma unroll
    for (int k = 63; k >= 0; --k) {
        const int sk = k >> 4, lk = k & 15;
#pragma unroll
        for (int s = 0; s < 4; ++s) lc[s] = ln[s];
        if (k > 0) {
            kk -= 1;
            asm volatile("" : "+v"(kk));
            const int s1 = (k - 1) >> 4;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 16 * s + l;
                if (s <= s1) {
                    const bool in = i < k - 1 && (EX || k - 1 < n);
                    const T v = gld(Ap + (in ? kk + ro[s] : 0));
                    ln[s] = in ? v : T(0);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const T t = x[c][sk] * myinv[sk];
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s <= sk) fnmac_rowk<T>(x[c][s], t, lc[s], lk);
        }
        asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int c = 0; c < RC; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s) x[c][s] *= myinv[s];
}

template <typename T, bool EX, int RC>
KB_DEV void rg4_load_rhs(T (&x)[RC][4], T* Bp, int ldb, int side_left, int n, int nrhs, int c0, int l)
{
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int i = 16 * s + l;
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const bool in = (EX || i < n) && c0 + c < nrhs;
            const T v = gld(rhs_elem(Bp, ldb, side_left, in ? i : 0, in ? c0 + c : 0));
            x[c][s] = in ? v : T(0);
        }
    }
}

template <typename T, int MODE, int RC, bool EX, int WPB>
__global__ void __launch_bounds__(64 * WPB)
k_rg4_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    const int l = threadIdx.x & 15;
    const int g = threadIdx.x >> 4;
    const long mat = (long)blockIdx.x * (4 * WPB) + g;
    // rows of lanes past the batch end repeat the last matrix without writing
    const bool live = mat < batch;
    const long m = live ? mat : (long)batch - 1;
    T* Ap = A.get(m);
    const int lda = A.ldim(m);

    T myinv[4] = {T(1), T(1), T(1), T(1)};
    int inf = 0;
    if constexpr (MODE != KB_MODE_POTRS) {
        inf = rg4_potrf<T, EX>(Ap, lda, n, l, live, myinv);
        if (live && l == 0 && info) info[mat] = inf;
    } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int i = 16 * s + l;
            myinv[s] = T(1) / rg4_ld<T, EX>(Ap, lda, n, i, i, true, T(1));
        }
    }
    if constexpr (MODE != KB_MODE_POTRF) {
        T* Bp = B.get(m);
        const int ldb = B.ldim(m);
        const bool ok = live && inf == 0;
        for (int c0 = 0; c0 < nrhs; c0 += RC) {
            T x[RC][4];
            rg4_load_rhs<T, EX, RC>(x, Bp, ldb, side_left, n, nrhs, c0, l);
            rg4_solve<T, EX, RC>(Ap, lda, n, l, myinv, x);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int i = 16 * s + l;
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    if (ok && (EX || i < n) && c0 + c < nrhs) gst(rhs_elem(Bp, ldb, side_left, i, c0 + c), x[c][s]);
            }
        }
    }
}

template <typename T, int MODE, int RC>
static int rg4_launch(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s)
{
    constexpr int WPB = 4;
    constexpr int MPB = 4 * WPB;
    dim3 grid((unsigned)((batch + MPB - 1) / MPB)), block(64 * WPB);
    hipLaunchKernelGGL((k_rg4_chol<T, MODE, RC, false, WPB>), grid, block, 0, s, n, nrhs, side_left,
                       to_arg<T>(A), to_arg<T>(B), info, batch);
    return (int)hipGetLastError();
}

// POTRF only: the solves run in the MFMA-tiled kernel (see kb_reg.hip)
template <typename T>
static int rg4_dispatch(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch,
                        hipStream_t s)
{
    if (mode != KB_MODE_POTRF) return (int)hipErrorInvalidValue;
    return rg4_launch<T, KB_MODE_POTRF, 1>(n, nrhs, side_left, A, B, info, batch, s);
}
