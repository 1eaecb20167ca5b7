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
#include "kb_rg_impl.h"

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

// C (m2 x m2, lower) -= L21 L21^T, L21 is m2 x kb (ld lda), on the matrix
// cores.  C is walked in 64 x 64 tiles; the 64-row slices of L21 are staged in
// LDS (slice I once per row of tiles).  Wave w computes rows 16w..16w+15 of a
// tile as four 16 x 16 products D' = L21_J(16 x kb) L21_I(16 x kb)^T, i.e.
// the transpose of the C block, whose accumulator layout puts consecutive C
// rows in consecutive lanes (coalesced column-major read-modify-write):
//   FP64: lane (g, i), register r -> C(16w + i, 16q + 4r + g)
//   FP32: lane (g, i), register r -> C(16w + i, 16q + 4g + r)
template <typename T>
KB_DEV void blk_syrk_update(const T* L21, T* C, int lda, int m2, int kb, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK, TL = KB_BLK_TILE;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int nt = (m2 + TL - 1) / TL;
    for (int I = 0; I < nt; ++I) {
        for (int e = tid; e < TL * NBK; e += KB_BLK_WG) {
            const int r = e % TL, p = e / TL;
            const int gi = I * TL + r;
            S.sA[r * (NBK + 1) + p] = (gi < m2 && p < kb) ? gld(L21 + gi + (long)p * lda) : T(0);
        }
        for (int J = 0; J <= I; ++J) {
            if (J != I) {
                for (int e = tid; e < TL * NBK; e += KB_BLK_WG) {
                    const int r = e % TL, p = e / TL;
                    const int gj = J * TL + r;
                    S.sB[r * (NBK + 1) + p] = (gj < m2 && p < kb) ? gld(L21 + gj + (long)p * lda) : T(0);
                }
            }
            __syncthreads();
            const T* sJ = (J == I) ? S.sA : S.sB;
            const int gi = I * TL + 16 * wave + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (J == I && q > wave) continue;          // above the diagonal
                int col[4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    col[r] = J * TL + 16 * q + ((sizeof(T) == 8) ? 4 * r + lg : 4 * lg + r);
                // the C block first, so its latency overlaps the products
                T cv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const bool in = gi < m2 && col[r] <= gi;
                    const T v = gld(C + (in ? gi + (long)col[r] * lda : 0));
                    cv[r] = in ? v : T(0);
                }
                typename TileV<T>::v4 acc = {T(0), T(0), T(0), T(0)};
                for (int p0 = 0; p0 < kb; p0 += 4) {
                    const T a = sJ[(16 * q + li) * (NBK + 1) + p0 + lg];
                    const T b = S.sA[(16 * wave + li) * (NBK + 1) + p0 + lg];
                    acc = mfma16x16x4(a, b, acc);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (gi < m2 && col[r] <= gi) gst(C + gi + (long)col[r] * lda, cv[r] - acc[r]);
            }
            __syncthreads();
        }
    }
}

// Load the diagonal block L(k0.., k0..) of the factor into LDS with its inverse diagonal.
template <typename T>
KB_DEV void blk_stage_diag(const T* L, int lda, int k0, int kb, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    for (int e = threadIdx.x; e < NBK * NBK; e += KB_BLK_WG) {
        const int r = e % NBK, c = e / NBK;
        S.sD[r * (NBK + 1) + c] = (r < kb && c <= r) ? gld(L + (k0 + r) + (long)(k0 + c) * lda) : T(0);
    }
    if (threadIdx.x < NBK) {
        const int r = threadIdx.x;
        S.sInv[r] = (r < kb) ? T(1) / gld(L + (k0 + r) + (long)(k0 + r) * lda) : T(1);
    }
    __syncthreads();
}

// Solve (L L^T) X = B for nrhs vectors stored by rhs_at(); L lower in A.
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
            // the 32 x 32 diagonal block on one DPP row of wave 0 (lane l: rows
            // l and 16 + l), fused DPP FMAs; the other rows of the wave repeat
            // the work without writing
            T* D = A + k0 + (long)k0 * lda;
            T a[2][NBK];
            T dinv2[2] = {T(1), T(1)};
            rg_load<T, NBK, false>(a, D, lda, kb, lane & 15);
            // L_D also goes to LDS (sD, sInv) for the panel solve below
            const int inf = rg_potrf<T, NBK, false, false>(a, dinv2, kb, lane & 15, D, lda, lane < 16, S.sD, S.sInv);
            if (lane == 0) S.flag = inf ? k0 + inf : 0;
        }
        __syncthreads();
        const int f = S.flag;
        if (f) return f;
        // panel below the diagonal block: X L_D^T = A_panel, one row per thread
        for (int i = k0 + kb + tid; i < n; i += KB_BLK_WG) {
            T x[NBK];
#pragma unroll
            for (int c = 0; c < NBK; ++c) x[c] = (c < kb) ? gld(A + i + (long)(k0 + c) * lda) : T(0);
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
#pragma unroll
            for (int c = 0; c < NBK; ++c)
                if (c < kb) gst(A + i + (long)(k0 + c) * lda, x[c]);
        }
        __syncthreads();
        const int t0 = k0 + kb;
        if (t0 < n) blk_syrk_update<T>(A + t0 + (long)k0 * lda, A + t0 + (long)t0 * lda, lda, n - t0, kb, S);
        __syncthreads();
    }
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

template <typename T>
KB_DEV void blk_solve(const T* L, int lda, int n, T* B, int ldb, int side_left, int nrhs, BlkSmem<T>& S)
{
    constexpr int NBK = KB_BLK_NBK;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    T* sY = S.sA;   // KB_BLK_RC x (NBK + 1) staging of the current block's solution
    // ---------------- forward: L Y = B ----------------
    for (int k0 = 0; k0 < n; k0 += NBK) {
        const int kb = min(NBK, n - k0);
        blk_stage_diag<T>(L, lda, k0, kb, S);
        for (int c0 = 0; c0 < nrhs; c0 += KB_BLK_RC) {
            const int rc = min(KB_BLK_RC, nrhs - c0);
            // diagonal block: right-hand side c = 4 wave + DPP row, lane l of the
            // row holds rows l and 16 + l; y(j) comes from lane j by DPP
            {
                const int c = 4 * wave + lg;
                const bool act = c < rc;
                T y[2], iv[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = 16 * h + li;
                    const bool in = act && r < kb;
                    y[h] = in ? gld(rhs_elem(B, ldb, side_left, in ? k0 + r : 0, in ? c0 + c : 0)) : T(0);
                    iv[h] = S.sInv[r];
                }
#pragma unroll
                for (int j = 0; j < NBK; ++j) {
                    const T t = y[j >> 4] * S.sInv[j];      // lane j: y(j) / L(j,j)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int r = 16 * h + li;
                        if (16 * h + 15 > j) {
                            const T l = (r > j) ? S.sD[r * (NBK + 1) + j] : T(0);
                            fnmac_rowk<T>(y[h], t, l, j);
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = 16 * h + li;
                    const T v = y[h] * iv[h];
                    if (act && r < kb) {
                        gst(rhs_elem(B, ldb, side_left, k0 + r, c0 + c), v);
                        sY[c * (NBK + 1) + r] = v;
                    }
                }
            }
            __syncthreads();
            // rows below: B(i, c) -= sum_j L(i, k0 + j) Y(j, c) on the matrix cores:
            // 16-row blocks of the rows below (one per wave at a time), K = the
            // block's columns in steps of 4 (zero beyond kb), N = the rc
            // right-hand sides (zero beyond rc)
            const int rows = n - (k0 + kb);
            for (int rb = wave * 16; rb < rows; rb += KB_BLK_WG / 4) {
                const int r0 = k0 + kb + rb;
                T bv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + blk_drow<T>(q, lg);
                    const bool in = row < n && li < rc;
                    bv[q] = in ? gld(rhs_elem(B, ldb, side_left, row, c0 + li)) : T(0);
                }
                typename TileV<T>::v4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int q = 0; q < NBK / 4; ++q) {
                    const int kk = 4 * q + lg;
                    const int row = r0 + li;
                    const bool ain = row < n && kk < kb;
                    const T av = gld(L + (ain ? row + (long)(k0 + kk) * lda : 0));
                    const T a = ain ? av : T(0);
                    const T b = (kk < kb && li < rc) ? sY[li * (NBK + 1) + kk] : T(0);
                    acc = mfma16x16x4(a, b, acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + blk_drow<T>(q, lg);
                    if (row < n && li < rc) gst(rhs_elem(B, ldb, side_left, row, c0 + li), bv[q] - acc[q]);
                }
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
            // y_p -= sum_{i below} L(i, k0+p) x_i on the matrix cores: output rows
            // p in two 16-row halves of the block, N = the rc right-hand sides,
            // K = the rows below in steps of 4; wave w takes half (w & 1) and every
            // other K step (w >> 1), and the two partial sums meet in LDS
            {
                const int ph = wave & 1, kh = wave >> 1;
                const int pa = 16 * ph + li;
                typename TileV<T>::v4 acc = {T(0), T(0), T(0), T(0)};
                for (int r4 = 4 * kh; r4 < below; r4 += 8) {
                    const int r = k0 + kb + r4 + lg;
                    const bool ain = r < n && pa < kb;
                    const T av = gld(L + (ain ? r + (long)(k0 + pa) * lda : 0));
                    const T a = ain ? av : T(0);
                    const bool bin = r < n && li < rc;
                    const T b = bin ? gld(rhs_elem(B, ldb, side_left, bin ? r : 0, bin ? c0 + li : c0)) : T(0);
                    acc = mfma16x16x4(a, b, acc);
                }
                T* sR = S.sB;   // partial sums of waves 2 and 3: 2 x 4 x 64
                if (kh == 1) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) sR[(ph * 4 + q) * 64 + lane] = acc[q];
                }
                __syncthreads();
                if (kh == 0 && below > 0) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int po = 16 * ph + blk_drow<T>(q, lg);
                        if (po < kb && li < rc) {
                            T* bp = rhs_elem(B, ldb, side_left, k0 + po, c0 + li);
                            gst(bp, gld(bp) - (acc[q] + sR[(ph * 4 + q) * 64 + lane]));
                        }
                    }
                }
                __syncthreads();
            }
            {
                const int c = 4 * wave + lg;
                const bool act = c < rc;
                T y[2], iv[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = 16 * h + li;
                    const bool in = act && r < kb;
                    y[h] = in ? gld(rhs_elem(B, ldb, side_left, in ? k0 + r : 0, in ? c0 + c : 0)) : T(0);
                    iv[h] = S.sInv[r];
                }
#pragma unroll
                for (int j = NBK - 1; j >= 0; --j) {
                    const T t = y[j >> 4] * S.sInv[j];      // lane j: x(j)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int r = 16 * h + li;
                        if (16 * h < j) {
                            const T l = (r < j) ? S.sD[j * (NBK + 1) + r] : T(0);
                            fnmac_rowk<T>(y[h], t, l, j);
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int r = 16 * h + li;
                    if (act && r < kb) gst(rhs_elem(B, ldb, side_left, k0 + r, c0 + c), y[h] * iv[h]);
                }
            }
            __syncthreads();
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(2)))
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
