// kb_blk_impl.h -- blocked batched Cholesky family, one 256-thread workgroup
// per matrix, for orders above the register-resident range.
//
// Right-looking blocked algorithm with panel width NBK = 32, the same
// POTRF -> TRSM(R,L,T) -> SYRK(L,N,-1,+1) sequence the reference runs as a
// recursion of separate launches (src/batch_triangular/Xpotrf_batch_drivers.cuh:91-131),
// here inside one kernel so every stage of a matrix reuses the workgroup's
// LDS and the matrix's cache lines:
//   - the 32x32 diagonal block is factored in VGPRs by wave 0 (row per lane,
//     v_readlane broadcasts), copied to LDS with its inverse diagonal;
//   - the panel below is solved one row per thread against the LDS block;
//   - the trailing lower triangle is updated in 64x64 tiles (4x4 per thread)
//     from two LDS-staged 64x32 slices of the panel.
// The solve (POSV/POTRS) runs blocked forward and backward sweeps: the
// diagonal block by one wave per right-hand side, the off-diagonal updates
// spread over (row, right-hand side) pairs.
#pragma once
#include "kb_device.h"
#include "kb_launch.h"
#include "kb_reg_impl.h"

constexpr int KB_BLK_NBK = 32;
constexpr int KB_BLK_WG = 256;
constexpr int KB_BLK_TILE = 64;
constexpr int KB_BLK_RC = 16;   // right-hand sides staged per pass of the solve

template <typename T>
struct BlkSmem {
    T sD[KB_BLK_NBK * (KB_BLK_NBK + 1)];
    T sInv[KB_BLK_NBK];
    T sA[KB_BLK_TILE * (KB_BLK_NBK + 1)];
    T sB[KB_BLK_TILE * (KB_BLK_NBK + 1)];
    int flag;
};

// B(i, c) of a right-hand-side block: side_left -> column c, else row c.
template <typename T>
KB_DEV T& rhs_at(T* B, int ldb, int side_left, int i, int c)
{
    return side_left ? B[i + (long)c * ldb] : B[c + (long)i * ldb];
}

// C (m2 x m2, lower) -= L21 L21^T, L21 is m2 x kb (ld lda).
template <typename T>
KB_DEV void blk_syrk_update(const T* L21, T* C, int lda, int m2, int kb, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK, TL = KB_BLK_TILE;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int nt = (m2 + TL - 1) / TL;
    for (int I = 0; I < nt; ++I) {
        for (int J = 0; J <= I; ++J) {
            // stage rows I*64.. and J*64.. of L21 (kb columns)
            for (int e = tid; e < TL * NBK; e += KB_BLK_WG) {
                const int r = e % TL, p = e / TL;
                const int gi = I * TL + r, gj = J * TL + r;
                S.sA[r * (NBK + 1) + p] = (gi < m2 && p < kb) ? L21[gi + (long)p * lda] : T(0);
                S.sB[r * (NBK + 1) + p] = (gj < m2 && p < kb) ? L21[gj + (long)p * lda] : T(0);
            }
            // the tile of C, fetched before the rank-kb product so that its
            // latency overlaps the FMAs
            T cv[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int gi = I * TL + ty + 16 * a, gj = J * TL + tx + 16 * b;
                    const bool in = gi < m2 && gj <= gi;
                    const T v = gld(C + (in ? gi + (long)gj * lda : 0));
                    cv[a][b] = in ? v : T(0);
                }
            __syncthreads();
            T acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
            for (int p = 0; p < kb; ++p) {
                T va[4], vb[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) va[a] = S.sA[(ty + 16 * a) * (NBK + 1) + p];
#pragma unroll
                for (int b = 0; b < 4; ++b) vb[b] = S.sB[(tx + 16 * b) * (NBK + 1) + p];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = kb_fma<T>(va[a], vb[b], acc[a][b]);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int gi = I * TL + ty + 16 * a, gj = J * TL + tx + 16 * b;
                    if (gi < m2 && gj <= gi) gst(C + gi + (long)gj * lda, cv[a][b] - acc[a][b]);
                }
            }
            __syncthreads();
        }
    }
}

// In-place lower Cholesky of an n x n matrix by the whole workgroup.
// Returns LAPACK info (identical in every thread).
template <typename T>
KB_DEV int blk_potrf(T* A, int lda, int n, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    for (int k0 = 0; k0 < n; k0 += NBK) {
        const int kb = min(NBK, n - k0);
        if (wave == 0) {
            T a[NBK];
            const int r = lane;
#pragma unroll
            for (int c = 0; c < NBK; ++c) {
                T v = T(0);
                if (r < kb && c <= r) v = A[(k0 + r) + (long)(k0 + c) * lda];
                a[c] = v;
            }
            T dg = T(1), inv = T(1);
            const int inf = reg_potrf<T, NBK, 64>(a, kb, r, dg, inv);
#pragma unroll
            for (int c = 0; c < NBK; ++c) {
                if (r < kb && c <= r) A[(k0 + r) + (long)(k0 + c) * lda] = a[c];
                if (r < NBK) S.sD[r * (NBK + 1) + c] = (c <= r && r < kb) ? a[c] : T(0);
            }
            if (r < NBK) S.sInv[r] = inv;
            if (lane == 0) S.flag = inf ? k0 + inf : 0;
        }
        __syncthreads();
        const int f = S.flag;
        if (f) return f;
        // panel below the diagonal block: X L_D^T = A_panel, one row per thread
        for (int i = k0 + kb + tid; i < n; i += KB_BLK_WG) {
            T x[NBK];
#pragma unroll
            for (int c = 0; c < NBK; ++c) x[c] = (c < kb) ? A[i + (long)(k0 + c) * lda] : T(0);
#pragma unroll
            for (int j = 0; j < NBK; ++j) {
                if (j < kb) {
                    x[j] = x[j] * S.sInv[j];
#pragma unroll
                    for (int c = j + 1; c < NBK; ++c) x[c] = kb_fma<T>(-x[j], S.sD[c * (NBK + 1) + j], x[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < NBK; ++c)
                if (c < kb) A[i + (long)(k0 + c) * lda] = x[c];
        }
        __syncthreads();
        const int t0 = k0 + kb;
        if (t0 < n) blk_syrk_update<T>(A + t0 + (long)k0 * lda, A + t0 + (long)t0 * lda, lda, n - t0, kb, S);
        __syncthreads();
    }
    return 0;
}

// Load the diagonal block L(k0.., k0..) of the factor into LDS with its inverse diagonal.
template <typename T>
KB_DEV void blk_stage_diag(const T* L, int lda, int k0, int kb, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    for (int e = threadIdx.x; e < NBK * NBK; e += KB_BLK_WG) {
        const int r = e % NBK, c = e / NBK;
        S.sD[r * (NBK + 1) + c] = (r < kb && c <= r) ? L[(k0 + r) + (long)(k0 + c) * lda] : T(0);
    }
    if (threadIdx.x < NBK) {
        const int r = threadIdx.x;
        S.sInv[r] = (r < kb) ? T(1) / L[(k0 + r) + (long)(k0 + r) * lda] : T(1);
    }
    __syncthreads();
}

// Solve (L L^T) X = B for nrhs vectors stored by rhs_at(); L lower in A.
template <typename T>
KB_DEV void blk_solve(const T* L, int lda, int n, T* B, int ldb, int side_left, int nrhs, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    T* sY = S.sA;   // KB_BLK_RC x (NBK + 1) staging of the current block's solution
    // ---------------- forward: L Y = B ----------------
    for (int k0 = 0; k0 < n; k0 += NBK) {
        const int kb = min(NBK, n - k0);
        blk_stage_diag<T>(L, lda, k0, kb, S);
        for (int c0 = 0; c0 < nrhs; c0 += KB_BLK_RC) {
            const int rc = min(KB_BLK_RC, nrhs - c0);
            for (int c = wave; c < rc; c += KB_BLK_WG / 64) {
                const int r = lane;
                T y = (r < kb) ? rhs_at(B, ldb, side_left, k0 + r, c0 + c) : T(0);
#pragma unroll
                for (int j = 0; j < NBK; ++j) {
                    if (j < kb) {
                        const T yj = bcastk<64>(y, j) * S.sInv[j];
                        if (r > j && r < NBK) y = kb_fma<T>(-S.sD[r * (NBK + 1) + j], yj, y);
                    }
                }
                if (r < kb) {
                    y = y * S.sInv[r];
                    rhs_at(B, ldb, side_left, k0 + r, c0 + c) = y;
                    sY[c * (NBK + 1) + r] = y;
                }
            }
            __syncthreads();
            // rows below: y_i -= L(i, block) * y_block
            const int rows = n - (k0 + kb);
            for (int e = tid; e < rows * rc; e += KB_BLK_WG) {
                const int i = k0 + kb + e % rows, c = e / rows;
                T s = T(0);
                for (int j = 0; j < kb; ++j) s = kb_fma<T>(L[i + (long)(k0 + j) * lda], sY[c * (NBK + 1) + j], s);
                T& bi = rhs_at(B, ldb, side_left, i, c0 + c);
                bi = bi - s;
            }
            __syncthreads();
        }
    }
    // ---------------- backward: L^T X = Y ----------------
    const int nblk = (n + NBK - 1) / NBK;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int k0 = bk * NBK;
        const int kb = min(NBK, n - k0);
        blk_stage_diag<T>(L, lda, k0, kb, S);
        for (int c0 = 0; c0 < nrhs; c0 += KB_BLK_RC) {
            const int rc = min(KB_BLK_RC, nrhs - c0);
            const int below = n - (k0 + kb);
            // y_p -= sum_{i below} L(i, k0+p) x_i
            for (int e = tid; e < kb * rc; e += KB_BLK_WG) {
                const int p = e % kb, c = e / kb;
                T s = T(0);
                const T* Lc = L + (long)(k0 + p) * lda;
                for (int i = k0 + kb; i < k0 + kb + below; ++i)
                    s = kb_fma<T>(Lc[i], rhs_at(B, ldb, side_left, i, c0 + c), s);
                T& bp = rhs_at(B, ldb, side_left, k0 + p, c0 + c);
                bp = bp - s;
            }
            __syncthreads();
            for (int c = wave; c < rc; c += KB_BLK_WG / 64) {
                const int r = lane;
                T y = (r < kb) ? rhs_at(B, ldb, side_left, k0 + r, c0 + c) : T(0);
#pragma unroll
                for (int j = NBK - 1; j >= 0; --j) {
                    if (j < kb) {
                        const T xj = bcastk<64>(y, j) * S.sInv[j];
                        if (r < j) y = kb_fma<T>(-S.sD[j * (NBK + 1) + r], xj, y);
                    }
                }
                if (r < kb) rhs_at(B, ldb, side_left, k0 + r, c0 + c) = y * S.sInv[r];
            }
            __syncthreads();
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(KB_BLK_WG)
k_blk_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    __shared__ BlkSmem<T> S;
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        T* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        int inf = 0;
        if constexpr (MODE != KB_MODE_POTRS) {
            inf = blk_potrf<T>(Ap, lda, n, S);
            if (threadIdx.x == 0 && info) info[mat] = inf;
        }
        if constexpr (MODE != KB_MODE_POTRF) {
            if (inf == 0) blk_solve<T>(Ap, lda, n, B.get(mat), B.ldim(mat), side_left, nrhs, S);
        }
        __syncthreads();
    }
}

template <typename T>
static int launch_blk(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                      int batch, hipStream_t s)
{
    const int grid = batch < 2048 ? batch : 2048;
    BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    if (mode == KB_MODE_POTRF)
        hipLaunchKernelGGL((k_blk_chol<T, KB_MODE_POTRF>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, a, b, info, batch);
    else if (mode == KB_MODE_POSV)
        hipLaunchKernelGGL((k_blk_chol<T, KB_MODE_POSV>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, a, b, info, batch);
    else
        hipLaunchKernelGGL((k_blk_chol<T, KB_MODE_POTRS>), dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, a, b, info, batch);
    return (int)hipGetLastError();
}
