// kb_blk_impl.h -- blocked batched Cholesky family, one 256-thread workgroup
// per matrix, for orders above the register-resident range.
//
// Right-looking blocked algorithm with panel width NBK = 32, the same
// POTRF -> TRSM(R,L,T) -> SYRK(L,N,-1,+1) sequence the reference runs as a
// recursion of separate launches (src/batch_triangular/Xpotrf_batch_drivers.cuh:91-131),
// here inside one kernel so every stage of a matrix reuses the workgroup's
// LDS and the matrix's cache lines:
//   - the 32x32 diagonal block is factored by wave 0 in registers as three
//     16x16 tiles (4-column panels on DPP rows, rank-4 updates on the matrix
//     cores), written to memory and to LDS with its inverse diagonal;
//   - the panel below is solved one row per thread against the LDS block;
//   - panels go in pairs: the second panel's columns are updated with the
//     first, and the trailing lower triangle then receives one update of
//     depth 64 (both panels) on the matrix cores, in 64x64 tiles (one 64x64
//     slice staged in LDS, the other operand in registers), so the trailing
//     matrix is read and written once per 64 columns instead of once per 32.
// The solve (POSV/POTRS) runs blocked forward and backward sweeps over up to
// 64 right-hand sides per pass (so L is streamed once per sweep for up to 64
// right-hand sides): the diagonal blocks by DPP-row substitutions, the
// off-diagonal block products on the matrix cores.
#pragma once
#include "kb_device.h"
#include "kb_launch.h"
#include "kb_reg_impl.h"
#include "kb_rg_impl.h"
#include "kb_tile_impl.h"

constexpr int KB_BLK_NBK = 32;
constexpr int KB_BLK_KW = 2 * KB_BLK_NBK;   // depth of a trailing update (two panels)
constexpr int KB_BLK_WG = 256;
constexpr int KB_BLK_TILE = 64;
constexpr int KB_BLK_RC = 64;   // right-hand sides per pass of the solve

// The trailing update's staged slice is never live at the same time as the
// diagonal block, the panel transposes or the solve's staging, so they share
// its space (33 KB for FP64 instead of 44 KB).
template <typename T>
struct BlkSmem {
    union {
        T sB[KB_BLK_TILE * (KB_BLK_KW + 1)];   // a 64-row slice of the panels (trailing update)
        struct {
            T sY[KB_BLK_RC * (KB_BLK_NBK + 1)];   // the solve's staging of a block of the solution
            T sD[KB_BLK_NBK * (KB_BLK_NBK + 1)];  // the diagonal block
            T tpan[4][16][4];                     // panel transposes of the diagonal-block factorization
        };
    };
    T sInv[KB_BLK_NBK];
    int flag;
};

// B(i, c) of a right-hand-side block: side_left -> column c, else row c.
template <typename T>
KB_DEV T& rhs_at(T* B, int ldb, int side_left, int i, int c)
{
    return side_left ? B[i + (long)c * ldb] : B[c + (long)i * ldb];
}

// C (m2 x m2, lower; only its first ncol columns) -= L21 L21^T, L21 is
// m2 x kb (kb <= 64, ld lda), on the matrix cores.  C is walked in 64 x 64
// tiles.  Wave w computes rows 16w..16w+15 of a tile as four 16 x 16 products
// D' = L21_J(16 x kb) L21_I(16 x kb)^T, i.e. the transpose of the C block,
// whose accumulator layout puts consecutive C rows in consecutive lanes
// (coalesced column-major read-modify-write):
//   FP64: lane (g, i), register r -> C(16w + i, 16q + 4r + g)
//   FP32: lane (g, i), register r -> C(16w + i, 16q + 4g + r)
// The wave's own 16 rows of slice I (the B operand of every K step) stay in
// registers for the whole row of tiles; slice J is staged in LDS.
template <typename T>
KB_DEV void blk_syrk_update(const T* L21, T* C, int lda, int m2, int kb, BlkSmem<T>& S, int ncol)
{
    constexpr int KW = KB_BLK_KW, TL = KB_BLK_TILE;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));   // lane-dependent values formed per call, not kept live across
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int nt = (m2 + TL - 1) / TL;
    const int kb4 = (kb + 3) & ~3;
    for (int I = 0; I < nt; ++I) {
        const int gi = I * TL + 16 * wave + li;
        T bI[KW / 4];
#pragma unroll
        for (int u = 0; u < KW / 4; ++u) {
            const int p = 4 * u + lg;
            const bool in = gi < m2 && p < kb;
            const T v = gld(L21 + (in ? gi + (long)p * lda : 0));
            bI[u] = in ? v : T(0);
        }
        for (int J = 0; J <= I && J * TL < ncol; ++J) {
            for (int e = tid; e < TL * KW; e += KB_BLK_WG) {
                const int r = e % TL, p = e / TL;
                const int gj = J * TL + r;
                if (p < kb4) S.sB[r * (KW + 1) + p] = (gj < m2 && p < kb) ? gld(L21 + gj + (long)p * lda) : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (J == I && q > wave) continue;          // above the diagonal
                if (J * TL + 16 * q >= ncol) continue;     // beyond the columns asked for
                int col[4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    col[r] = J * TL + 16 * q + ((sizeof(T) == 8) ? 4 * r + lg : 4 * lg + r);
                // the C block first, so its latency overlaps the products
                T cv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const bool in = gi < m2 && col[r] <= gi && col[r] < ncol;
                    const T v = gld(C + (in ? gi + (long)col[r] * lda : 0));
                    cv[r] = in ? v : T(0);
                }
                typename TileV<T>::v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int u = 0; u < KW / 4; ++u) {
                    if (4 * u < kb) {
                        const T a = S.sB[(16 * q + li) * (KW + 1) + 4 * u + lg];
                        acc = mfma16x16x4(a, bI[u], acc);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (gi < m2 && col[r] <= gi && col[r] < ncol) gst(C + gi + (long)col[r] * lda, cv[r] - acc[r]);
            }
            __syncthreads();
        }
    }
}

// The diagonal block L(k0.., k0..) of the factor in two steps, so the solve
// can issue the loads of the next block before working on the current one:
// blk_diag_fetch reads this thread's elements (e = tid + 256 m) into
// registers, blk_diag_put writes them to LDS with the inverse diagonal.
constexpr int KB_BLK_DPT = KB_BLK_NBK * KB_BLK_NBK / KB_BLK_WG;   // elements per thread

template <typename T>
KB_DEV void blk_diag_fetch(T (&v)[KB_BLK_DPT], const T* L, int lda, int k0, int kb)
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));   // lane-dependent values formed per call, not kept live across
#pragma unroll
    for (int m = 0; m < KB_BLK_DPT; ++m) {
        const int e = tid + KB_BLK_WG * m;
        const int r = e % NBK, c = e / NBK;
        const bool in = r < kb && c <= r;
        const T x = gld(L + (in ? (k0 + r) + (long)(k0 + c) * lda : 0));
        v[m] = in ? x : T(0);
    }
}

template <typename T>
KB_DEV void blk_diag_put(const T (&v)[KB_BLK_DPT], int kb, BlkSmem<T>& S, int unit = 0)
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int m = 0; m < KB_BLK_DPT; ++m) {
        const int e = tid + KB_BLK_WG * m;
        const int r = e % NBK, c = e / NBK;
        S.sD[r * (NBK + 1) + c] = v[m];
        if (r == c) S.sInv[r] = (r < kb && !unit) ? T(1) / v[m] : T(1);
    }
    __syncthreads();
}

// One 32-column panel at k0 (kb columns, already updated by every panel to
// its left): its diagonal block factored by wave 0 (L_D also goes to LDS for
// the panel solve), then the rows below solved one row per thread,
// X L_D^T = A_panel.  Returns LAPACK info (0, or k0 + the failing column).
template <typename T>
KB_DEV int blk_panel(T* A, int lda, int n, int k0, int kb, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));   // lane-dependent values formed per panel, not kept live across
    const int wave = tid >> 6, lane = tid & 63;
    if (wave == 0) {
        // the 32 x 32 diagonal block as tiles (0,0), (1,0), (1,1): lane (i, g)
        // holds row i and columns 4g..4g+3 of each; rows and columns at or
        // beyond kb read as the identity
        T* D = A + k0 + (long)k0 * lda;
        const int i = lane & 15, g = lane >> 4;
        typename TileV<T>::v4 td[3];
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
            for (int J = 0; J <= I; ++J)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + i, col = 16 * J + 4 * g + r;
                    const bool in = row < kb && col <= row;
                    const T v = gld(D + (in ? row + (long)col * lda : 0));
                    td[tix(I, J)][r] = in ? v : ((row == col) ? T(1) : T(0));
                }
        T invd[2] = {T(1), T(1)};
        const int inf = tile_potrf<T, 2>(td, invd, kb, S.tpan, lane);
        // L_D to memory (lower part), to LDS for the panel solve below, and
        // 1/L(j, j) to sInv
#pragma unroll
        for (int I = 0; I < 2; ++I)
#pragma unroll
            for (int J = 0; J <= I; ++J)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + i, col = 16 * J + 4 * g + r;
                    if (row < kb && col <= row) gst(D + row + (long)col * lda, td[tix(I, J)][r]);
                    S.sD[row * (NBK + 1) + col] = (col <= row) ? td[tix(I, J)][r] : T(0);
                }
        if (g == 0) {
            S.sInv[i] = invd[0];
            S.sInv[16 + i] = invd[1];
        }
        if (lane == 0) S.flag = inf ? k0 + inf : 0;
    }
    __syncthreads();
    const int f = S.flag;
    if (f) return f;
    for (int i = k0 + kb + tid; i < n; i += KB_BLK_WG) {
        T x[NBK];
        // walk the row with a running pointer (made opaque each step), so the
        // 32 column addresses are not precomputed and kept live across rows
        T* p = A + i + (long)k0 * lda;
#pragma unroll
        for (int c = 0; c < NBK; ++c) {
            x[c] = (c < kb) ? gld(p) : T(0);
            p += lda;
            asm volatile("" : "+v"(p));
        }
#pragma unroll
        for (int j = 0; j < NBK; ++j) {
            if (j < kb) {
                x[j] = x[j] * S.sInv[j];
#pragma unroll
                for (int c = j + 1; c < NBK; ++c) x[c] = kb_fma<T>(-x[j], S.sD[c * (NBK + 1) + j], x[c]);
            }
            // keep each step's LDS reads with the step (hoisting them all would
            // spill the 32-element row)
            asm volatile("" ::: "memory");
        }
        p = A + i + (long)k0 * lda;
#pragma unroll
        for (int c = 0; c < NBK; ++c) {
            if (c < kb) gst(p, x[c]);
            p += lda;
            asm volatile("" : "+v"(p));
        }
    }
    __syncthreads();
    return 0;
}

// Row of a 16x16 matrix-core accumulator held in register q by lanes 16g..16g+15
// (the column is the lane's position in its row of 16): FP64 interleaves the
// rows, FP32 blocks them.
template <typename T>
KB_DEV int blk_drow(int q, int g)
{
    return (sizeof(T) == 8) ? 4 * q + g : 4 * g + q;
}

// Substitution with the staged diagonal block (sD, sInv) for the NS groups of
// 16 right-hand sides of a pass: DPP row (wave, lg) of the workgroup takes
// right-hand sides c = 16 s + 4 wave + lg, lane li of the row holds block rows
// li and 16 + li, and y(j) comes from lane j by DPP; the NS independent
// substitutions are interleaved and share the LDS reads of L.
// FWD: L_D y = b, else L_D^T x = y; the result also goes to sY (right-hand
// side major) for the off-diagonal update that follows.  b is read as sc * b.
template <typename T, int NS, bool FWD>
KB_DEV void blk_diag_subst(T* B, int ldb, int side_left, int k0, int kb, int c0, int rc, BlkSmem<T>& S, T* sY,
                           T sc = T(1))
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    // opaque to the optimizer, so the lane-dependent offsets below are formed
    // here and not hoisted out of the caller's loops (where they would stay
    // live through the factorization)
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    T y[NS][2], iv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) iv[h] = S.sInv[16 * h + li];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = 16 * s + 4 * wave + lg;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = 16 * h + li;
            const bool in = c < rc && r < kb;
            // unconditional load from a clamped address, then a select: every load of the
            // block is in flight at once (a load under a branch waits before the join)
            const T v = gld(rhs_elem(B, ldb, side_left, in ? k0 + r : 0, in ? c0 + c : 0));
            y[s][h] = in ? sc * v : T(0);
        }
    }
#pragma unroll
    for (int jj = 0; jj < NBK; ++jj) {
        const int j = FWD ? jj : NBK - 1 - jj;
        const T ij = S.sInv[j];
        T t[NS];     // lane j: y(j) / L(j, j), read through DPP (lane j is not updated at step j)
#pragma unroll
        for (int s = 0; s < NS; ++s) t[s] = y[s][j >> 4] * ij;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = 16 * h + li;
            if (FWD ? (16 * h + 15 > j) : (16 * h < j)) {
                const T l = FWD ? ((r > j) ? S.sD[r * (NBK + 1) + j] : T(0))
                                : ((r < j) ? S.sD[j * (NBK + 1) + r] : T(0));
#pragma unroll
                for (int s = 0; s < NS; ++s) fnmac_rowk<T>(y[s][h], t[s], l, j);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int c = 16 * s + 4 * wave + lg;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = 16 * h + li;
            const T v = y[s][h] * iv[h];
            if (c < rc && r < kb) {
                gst(rhs_elem(B, ldb, side_left, k0 + r, c0 + c), v);
                sY[c * (NBK + 1) + r] = v;
            }
        }
    }
}

// ns groups of 16 right-hand sides, ns <= CB (the widest pass instantiated)
template <typename T, bool FWD, int CB>
KB_DEV void blk_diag_subst_n(int ns, T* B, int ldb, int side_left, int k0, int kb, int c0, int rc, BlkSmem<T>& S,
                             T* sY, T sc = T(1))
{
    if (CB == 2 || ns <= 2) {
        if (ns == 1) blk_diag_subst<T, 1, FWD>(B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
        else blk_diag_subst<T, 2, FWD>(B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
        return;
    }
    if constexpr (CB > 2) {
        if (ns == 3) blk_diag_subst<T, 3, FWD>(B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
        else blk_diag_subst<T, 4, FWD>(B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
    }
}

// B(r, c) := sc B(r, c) - sum_k Lk(r, k) Y(k, c) for the rows r in [r_lo, r_hi), k < kb and
// the pass's right-hand sides, on the matrix cores.  Lk(r, k) = L(r, k0 + k)
// (TRANS false: the block column below the diagonal block, forward sweep) or
// L(k0 + k, r) (TRANS true: the block row left of it, backward sweep); Y is
// staged in sY (right-hand side major).  Each wave takes 16-row blocks: the
// L fragment (eight K steps of 4) and the B blocks of the NCB column blocks of
// 16 right-hand sides are loaded together, then NCB independent MFMA chains.
template <typename T, bool TRANS, int NCB>
KB_DEV void blk_update_rows(const T* L, int lda, T* B, int ldb, int side_left, int k0, int kb, int r_lo,
                            int r_hi, int c0, int rc, const T* sY, T sc = T(1))
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));   // lane-dependent offsets stay local (as in blk_diag_subst)
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    for (int r0 = r_lo + 16 * wave; r0 < r_hi; r0 += KB_BLK_WG / 4) {
        T a[NBK / 4];
#pragma unroll
        for (int q = 0; q < NBK / 4; ++q) {
            const int kk = 4 * q + lg;
            const int row = r0 + li;
            const bool ain = row < r_hi && kk < kb;
            const long off = TRANS ? (k0 + kk) + (long)row * lda : row + (long)(k0 + kk) * lda;
            const T av = gld(L + (ain ? off : 0));
            a[q] = ain ? av : T(0);
        }
        T bv[NCB][4];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const int cc = 16 * cb + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = r0 + blk_drow<T>(q, lg);
                const bool in = row < r_hi && cc < rc;
                const T v = gld(rhs_elem(B, ldb, side_left, in ? row : 0, in ? c0 + cc : c0));
                bv[cb][q] = in ? v : T(0);
            }
        }
        typename TileV<T>::v4 acc[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[cb] = typename TileV<T>::v4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < NBK / 4; ++q) {
            const int kk = 4 * q + lg;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const int cc = 16 * cb + li;
                const T b = (kk < kb && cc < rc) ? sY[cc * (NBK + 1) + kk] : T(0);
                acc[cb] = mfma16x16x4(a[q], b, acc[cb]);
            }
        }
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const int cc = 16 * cb + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = r0 + blk_drow<T>(q, lg);
                if (row < r_hi && cc < rc) gst(rhs_elem(B, ldb, side_left, row, c0 + cc), sc * bv[cb][q] - acc[cb][q]);
            }
        }
    }
}

template <typename T, bool TRANS, int CB>
KB_DEV void blk_update_rows_n(int ncb, const T* L, int lda, T* B, int ldb, int side_left, int k0, int kb, int r_lo,
                              int r_hi, int c0, int rc, const T* sY, T sc = T(1))
{
    if (CB == 2 || ncb <= 2) {
        if (ncb == 1) blk_update_rows<T, TRANS, 1>(L, lda, B, ldb, side_left, k0, kb, r_lo, r_hi, c0, rc, sY, sc);
        else blk_update_rows<T, TRANS, 2>(L, lda, B, ldb, side_left, k0, kb, r_lo, r_hi, c0, rc, sY, sc);
        return;
    }
    if constexpr (CB > 2) {
        if (ncb == 3) blk_update_rows<T, TRANS, 3>(L, lda, B, ldb, side_left, k0, kb, r_lo, r_hi, c0, rc, sY, sc);
        else blk_update_rows<T, TRANS, 4>(L, lda, B, ldb, side_left, k0, kb, r_lo, r_hi, c0, rc, sY, sc);
    }
}

// One block step of the forward sweep L Y = B: substitution with the
// diagonal block staged in sD / sInv, then the rows below updated with the
// block column of L (all right-hand sides, in passes of up to 16 CB: a
// narrower pass needs fewer registers, a wider one streams L fewer times).
// sc scales the right-hand sides as they are first read (the first step).
template <typename T, int CB = 4>
KB_DEV void blk_fwd_step(const T* L, int lda, int n, int k0, int kb, T* B, int ldb, int side_left, int nrhs,
                         BlkSmem<T>& S, T sc = T(1))
{
    constexpr int RC = 16 * CB;
    static_assert(RC <= KB_BLK_RC, "the staging holds KB_BLK_RC right-hand sides");
    T* sY = S.sY;   // RC x (NBK + 1) staging of the block's solution
    for (int c0 = 0; c0 < nrhs; c0 += RC) {
        const int rc = min(RC, nrhs - c0);
        const int ncb = (rc + 15) / 16;
        blk_diag_subst_n<T, true, CB>(ncb, B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
        __syncthreads();
        blk_update_rows_n<T, false, CB>(ncb, L, lda, B, ldb, side_left, k0, kb, k0 + kb, n, c0, rc, sY, sc);
        __syncthreads();
    }
}

// Forward sweep L Y = alpha B over 32-row blocks (right-looking), the next
// diagonal block fetched while the current one is used; unit: the diagonal
// of L is taken as 1 (not read).
template <typename T, int CB = 4>
KB_DEV void blk_solve_fwd(const T* L, int lda, int n, T* B, int ldb, int side_left, int nrhs, BlkSmem<T>& S,
                          T alpha = T(1), int unit = 0)
{
    constexpr int NBK = KB_BLK_NBK;
    T dnext[KB_BLK_DPT];
    blk_diag_fetch<T>(dnext, L, lda, 0, min(NBK, n));
    for (int k0 = 0; k0 < n; k0 += NBK) {
        const int kb = min(NBK, n - k0);
        blk_diag_put<T>(dnext, kb, S, unit);
        if (k0 + NBK < n) blk_diag_fetch<T>(dnext, L, lda, k0 + NBK, min(NBK, n - k0 - NBK));
        blk_fwd_step<T, CB>(L, lda, n, k0, kb, B, ldb, side_left, nrhs, S, k0 == 0 ? alpha : T(1));
    }
}

// Backward sweep L^T X = alpha Y over 32-row blocks (right-looking: after
// each block's substitution the rows above it are updated); unit as above.
template <typename T, int CB = 4>
KB_DEV void blk_solve_bwd(const T* L, int lda, int n, T* B, int ldb, int side_left, int nrhs, BlkSmem<T>& S,
                          T alpha = T(1), int unit = 0)
{
    constexpr int NBK = KB_BLK_NBK, RC = 16 * CB;
    T* sY = S.sY;
    T dnext[KB_BLK_DPT];
    const int nblk = (n + NBK - 1) / NBK;
    blk_diag_fetch<T>(dnext, L, lda, (nblk - 1) * NBK, n - (nblk - 1) * NBK);
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int k0 = bk * NBK;
        const int kb = min(NBK, n - k0);
        blk_diag_put<T>(dnext, kb, S, unit);
        if (bk > 0) blk_diag_fetch<T>(dnext, L, lda, k0 - NBK, NBK);
        const T sc = bk == nblk - 1 ? alpha : T(1);
        for (int c0 = 0; c0 < nrhs; c0 += RC) {
            const int rc = min(RC, nrhs - c0);
            const int ncb = (rc + 15) / 16;
            blk_diag_subst_n<T, false, CB>(ncb, B, ldb, side_left, k0, kb, c0, rc, S, sY, sc);
            __syncthreads();
            blk_update_rows_n<T, true, CB>(ncb, L, lda, B, ldb, side_left, k0, kb, 0, k0, c0, rc, sY, sc);
            __syncthreads();
        }
    }
}

// Solve (L L^T) X = B for nrhs vectors stored by rhs_at(); L lower in A.
template <typename T, int CB = 4>
KB_DEV void blk_solve(const T* L, int lda, int n, T* B, int ldb, int side_left, int nrhs, BlkSmem<T>& S)
{
    blk_solve_fwd<T, CB>(L, lda, n, B, ldb, side_left, nrhs, S);
    blk_solve_bwd<T, CB>(L, lda, n, B, ldb, side_left, nrhs, S);
}

// In-place lower Cholesky of an n x n matrix by the whole workgroup.
// Returns LAPACK info (identical in every thread).  With B given (POSV), the
// forward sweep L Y = B runs inside the factorization: each panel's block step
// follows the panel, while its diagonal block is still in LDS and the panel
// in cache; a factorization that fails at panel k leaves B forward-solved
// through the panels before k.
template <typename T, int CB = 4>
KB_DEV int blk_potrf(T* A, int lda, int n, BlkSmem<T>& S, T* B = nullptr, int ldb = 0, int side_left = 0,
                     int nrhs = 0)
{
    constexpr int NBK = KB_BLK_NBK;
    for (int k0 = 0; k0 < n; k0 += 2 * NBK) {
        const int kb0 = min(NBK, n - k0);
        int f = blk_panel<T>(A, lda, n, k0, kb0, S);
        if (f) return f;
        if (B) blk_fwd_step<T, CB>(A, lda, n, k0, kb0, B, ldb, side_left, nrhs, S);
        const int k1 = k0 + kb0;
        if (k1 >= n) break;
        const int kb1 = min(NBK, n - k1);
        // the second panel's columns, updated with the first
        blk_syrk_update<T>(A + k1 + (long)k0 * lda, A + k1 + (long)k1 * lda, lda, n - k1, kb0, S, kb1);
        __syncthreads();
        f = blk_panel<T>(A, lda, n, k1, kb1, S);
        if (f) return f;
        if (B) blk_fwd_step<T, CB>(A, lda, n, k1, kb1, B, ldb, side_left, nrhs, S);
        // the trailing matrix, with both panels at once (depth kb0 + kb1)
        const int t0 = k1 + kb1;
        if (t0 < n) blk_syrk_update<T>(A + t0 + (long)k0 * lda, A + t0 + (long)t0 * lda, lda, n - t0, kb0 + kb1, S, n - t0);
        __syncthreads();
    }
    return 0;
}

// CB = 2: passes of 32 right-hand sides, few enough registers for four
// waves per SIMD; CB = 4: passes of 64 (L streamed half as often), three.
template <typename T, int MODE, int CB>
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(CB == 2 ? 4 : 2)))
k_blk_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    __shared__ BlkSmem<T> S;
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        T* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        int inf = 0;
        if constexpr (MODE == KB_MODE_POTRF) {
            inf = blk_potrf<T, CB>(Ap, lda, n, S);
            if (threadIdx.x == 0 && info) info[mat] = inf;
        } else if constexpr (MODE == KB_MODE_POSV) {
            T* Bp = B.get(mat);
            const int ldb = B.ldim(mat);
            inf = blk_potrf<T, CB>(Ap, lda, n, S, Bp, ldb, side_left, nrhs);
            if (threadIdx.x == 0 && info) info[mat] = inf;
            if (inf == 0) blk_solve_bwd<T, CB>(Ap, lda, n, Bp, ldb, side_left, nrhs, S);
        } else {
            blk_solve<T, CB>(Ap, lda, n, B.get(mat), B.ldim(mat), side_left, nrhs, S);
        }
        __syncthreads();
    }
}

// Passes of 32 right-hand sides when the count allows (POTRF: none), else 64.
constexpr int KB_BLK_NARROW = 32;

template <typename T, int MODE>
static void launch_blk_mode(int n, int nrhs, int side_left, BatchArg<T> a, BatchArg<T> b, int* info, int batch,
                            int grid, hipStream_t s)
{
    if (MODE == KB_MODE_POTRF || nrhs <= KB_BLK_NARROW)
        hipLaunchKernelGGL((k_blk_chol<T, MODE, 2>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, a, b, info,
                           batch);
    else
        hipLaunchKernelGGL((k_blk_chol<T, MODE, 4>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, a, b, info,
                           batch);
}

// Batched lower TRSM for orders above the register range: one sweep of the
// blocked solve, FWD: L X = alpha B, else L^T X = alpha B, for the nrhs
// vectors of B (side_left: columns, else rows, i.e. X op(L) = alpha B as
// op(L)^T X^T = alpha B^T).
template <typename T, bool FWD, int CB>
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(CB == 2 ? 4 : 2)))
k_blk_trsm(int n0, int nrhs0, int side_left, int unit, T alpha, BatchArg<T> A, BatchArg<T> B, int batch,
           const int* vm, const int* vn)
{
    __shared__ BlkSmem<T> S;
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        int n = n0, nrhs = nrhs0;
        if (vm) {   // non-uniform: this matrix's order and right-hand-side count
            n = side_left ? vm[mat] : vn[mat];
            nrhs = side_left ? vn[mat] : vm[mat];
            if (n <= 0 || nrhs <= 0) continue;
        }
        if (alpha == T(0)) {
            // B := 0 without reading it (BLAS semantics for alpha = 0)
            T* Bp = B.get(mat);
            const int ldb = B.ldim(mat);
            for (long e = threadIdx.x; e < (long)n * nrhs; e += KB_BLK_WG)
                gst(rhs_elem(Bp, ldb, side_left, (int)(e % n), (int)(e / n)), T(0));
            continue;
        }
        if constexpr (FWD)
            blk_solve_fwd<T, CB>(A.get(mat), A.ldim(mat), n, B.get(mat), B.ldim(mat), side_left, nrhs, S, alpha, unit);
        else
            blk_solve_bwd<T, CB>(A.get(mat), A.ldim(mat), n, B.get(mat), B.ldim(mat), side_left, nrhs, S, alpha, unit);
        __syncthreads();
    }
}

template <typename T>
static int launch_blk_trsm(int fwd, int side_left, int unit, int n, int nrhs, T alpha, kb_mat_t A, kb_mat_t B,
                           const int* vm, const int* vn, int batch, hipStream_t s)
{
    const int grid = batch < 2048 ? batch : 2048;
    BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    const bool narrow = nrhs <= KB_BLK_NARROW;
    if (fwd && narrow)
        hipLaunchKernelGGL((k_blk_trsm<T, true, 2>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, unit,
                           alpha, a, b, batch, vm, vn);
    else if (fwd)
        hipLaunchKernelGGL((k_blk_trsm<T, true, 4>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, unit,
                           alpha, a, b, batch, vm, vn);
    else if (narrow)
        hipLaunchKernelGGL((k_blk_trsm<T, false, 2>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, unit,
                           alpha, a, b, batch, vm, vn);
    else
        hipLaunchKernelGGL((k_blk_trsm<T, false, 4>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, unit,
                           alpha, a, b, batch, vm, vn);
    return (int)hipGetLastError();
}

template <typename T>
static int launch_blk(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                      int batch, hipStream_t s)
{
    const int grid = batch < 2048 ? batch : 2048;
    BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    if (mode == KB_MODE_POTRF)
        launch_blk_mode<T, KB_MODE_POTRF>(n, nrhs, side_left, a, b, info, batch, grid, s);
    else if (mode == KB_MODE_POSV)
        launch_blk_mode<T, KB_MODE_POSV>(n, nrhs, side_left, a, b, info, batch, grid, s);
    else
        launch_blk_mode<T, KB_MODE_POTRS>(n, nrhs, side_left, a, b, info, batch, grid, s);
    return (int)hipGetLastError();
}
