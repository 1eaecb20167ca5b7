// kb_tile_impl.h -- MFMA-tiled batched Cholesky family for 32 < n <= 64, one
// matrix per wavefront (gfx950).
//
// Storage: the lower triangle as ten 16x16 tiles T(I,J), I >= J.  Lane l holds
// row (l & 15) and columns 4*(l >> 4) .. 4*(l >> 4) + 3 of every tile, i.e. the
// accumulator layout of v_mfma_f32_16x16x4f32 for the tile's transpose:
//     D[i][j] lives at lane j + 16*(i/4), register i%4   =>   D = T^T
// (v_mfma_f64_16x16x4f64 interleaves the rows of D instead; the A operand is
// fed with permuted rows so the FP64 product lands in the same storage).
// Consequences:
//   * a 4-column panel (global columns 16J+4g .. 16J+4g+3) sits entirely in
//     DPP row g (lanes 16g..16g+15) of the tiles of tile-column J, so its
//     unblocked factorization needs only `row_newbcast` broadcasts;
//   * the rank-4 trailing update T(I',J') -= P(I') P(J')^T of that panel is
//     ONE 16x16x4 MFMA per tile: A operand = -P(J') and B operand = P(I'),
//     both in the "lane x + 16k holds P[x][k]" layout obtained from the
//     panel's DPP row through a 2 KB LDS transpose.
// The O(n^3) work therefore runs on the matrix cores (80 MFMAs at n = 64)
// while the VALU only does the 4-wide panels, instead of the reference's
// recursion of separate launches with cuBLAS GEMMs
// (src/batch_triangular/Xpotrf_batch_drivers.cuh:91-131).
//
// Solves (POSV/POTRS) use a second layout: right-hand side c of a group of
// four sits in DPP row c, lane 16c + i holding row 16K + i of block K, so the
// substitution broadcasts are again `row_newbcast`; each 16x16 factor tile is
// replicated per DPP row through a small LDS staging buffer.
#pragma once
#include "kb_device.h"
#include "kb_launch.h"
#include "kb_reg_impl.h"


constexpr KB_DEV int tix(int I, int J) { return I * (I + 1) / 2 + J; }

// Per-wave LDS: a 64 x 4 panel transpose buffer and a 16 x 17 tile staging buffer.
template <typename T>
struct TileLds {
    T pan[4][16][4];
    T tile[16][17];
    T inv[16];
    T y[4][64];      // right-hand-side blocks during the solve
    T invd[4][64];   // 1/L(16K+i,16K+i) per lane
};


// ---------------------------------------------------------------------------
// factorization
// ---------------------------------------------------------------------------
template <typename T>
KB_DEV void tile_load(typename TileV<T>::v4 (&t)[10], const T* A, int lda, int n, int lane)
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int I = 0; I < 4; ++I) {
#pragma unroll
        for (int J = 0; J <= I; ++J) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + i, col = 16 * J + 4 * g + r;
                const bool in = row < n && col <= row;
                const int rr = in ? row : 0, cc = in ? col : 0;
                const T ld = gld(A + rr + (long)cc * lda);
                // identity padding beyond n; the strict upper triangle is never read
                t[tix(I, J)][r] = in ? ld : ((row == col) ? T(1) : T(0));
            }
        }
    }
}

template <typename T>
KB_DEV void tile_store(const typename TileV<T>::v4 (&t)[10], T* A, int lda, int n, int lane)
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int I = 0; I < 4; ++I) {
#pragma unroll
        for (int J = 0; J <= I; ++J) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + i, col = 16 * J + 4 * g + r;
                if (row < n && col <= row) gst(A + row + (long)col * lda, t[tix(I, J)][r]);
            }
        }
    }
}

// In-register Cholesky of the (padded) 16NT x 16NT matrix (NT = 4: 64 x 64;
// NT = 2: a 32 x 32 block).  invd[K] receives 1/L(16K+i, 16K+i) in every lane
// with (lane & 15) == i.  Returns LAPACK info.
template <typename T, int NT = 4>
KB_DEV int tile_potrf(typename TileV<T>::v4 (&t)[NT * (NT + 1) / 2], T (&invd)[NT], int n, T (&pan)[4][16][4],
                      int lane)
{
    const int i = lane & 15, g = lane >> 4;
    int info = 0;
#pragma unroll
    for (int J = 0; J < NT; ++J) {
#pragma unroll
        for (int gp = 0; gp < 4; ++gp) {
            if (info) break;
            // ---- panel: global columns 16J + 4gp + c, held by DPP row gp ----
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int pl = 16 * gp + 4 * gp + c;          // lane holding the pivot
                const T d = bcastk<64>(t[tix(J, J)][c], pl);  // uniform
                if (!(d > T(0))) {
                    info = 16 * J + 4 * gp + c + 1;
                    break;
                }
                const T rs = kb_rsqrt<T>(d);
                const T s = d * rs;
                if (i == 4 * gp + c) invd[J] = rs;
                if (g == gp) {
#pragma unroll
                    for (int I = J; I < NT; ++I) {
                        T v = t[tix(I, J)][c] * rs;
                        if (I == J && i == 4 * gp + c) v = s;
                        t[tix(I, J)][c] = v;
                    }
#pragma unroll
                    for (int cp = c + 1; cp < 4; ++cp) {
#pragma unroll
                        for (int I = J; I < NT; ++I) {
                            // t(I,J)[cp] -= t(I,J)[c] * t(J,J)[c] of lane 4 gp + cp
                            T acc = t[tix(I, J)][cp];
                            fnmac_rowk<T>(acc, t[tix(J, J)][c], t[tix(I, J)][c], 4 * gp + cp);
                            t[tix(I, J)][cp] = acc;
                        }
                    }
                }
            }
            if (info) break;
            // ---- transpose the panel: lane x + 16k receives P(X)[x][k] ----
            if (g == gp) {
#pragma unroll
                for (int X = J; X < NT; ++X)
#pragma unroll
                    for (int k = 0; k < 4; ++k) pan[X][i][k] = t[tix(X, J)][k];
            }
            wave_lds_sync();
            // The accumulator row index i of the MFMA is the tile column.  For
            // the FP32 instruction D[i][j] lives at lane j + 16(i/4), register
            // i%4 (blocked, the tile storage as is); the FP64 instruction
            // interleaves rows, D[i][j] at lane j + 16(i%4), register i/4, so
            // its A operand row i must carry tile column sig(i) = 4(i%4) + i/4.
            const int ci = (sizeof(T) == 8) ? 4 * (i & 3) + (i >> 2) : i;
            T ts[NT], ta[NT];
#pragma unroll
            for (int X = 0; X < NT; ++X) {
                ts[X] = (X >= J) ? pan[X][i][g] : T(0);
                ta[X] = (X >= J) ? pan[X][ci][g] : T(0);
            }
            wave_lds_sync();
            // ---- rank-4 trailing update on the matrix cores ----
#pragma unroll
            for (int Jp = J; Jp < NT; ++Jp) {
                if (Jp == J && gp == 3) continue;             // tile column finished
                T a = -ta[Jp];
                if (Jp == J && ci < 4 * (gp + 1)) a = T(0);   // columns already factored
#pragma unroll
                for (int Ip = Jp; Ip < NT; ++Ip) t[tix(Ip, Jp)] = mfma16x16x4(a, ts[Ip], t[tix(Ip, Jp)]);
            }
        }
        if (info) break;
    }
    return info;
}

// ---------------------------------------------------------------------------
// solve: right-hand side c of the group of four in DPP row c
// ---------------------------------------------------------------------------
// Stage factor tile (I, J) from L in global memory (padding rows/columns at or
// beyond n read as the identity) into LDS and read row i (ROW) or column i of
// it into v[16].  STRICT (diagonal tiles) hands out the strictly lower part of
// the unit factor Lu = L D^-1 (column c scaled by 1/L(c,c); invk = this lane's
// 1/L(i,i)): row i gives Lu(i, j), column i gives Lu(j, i), so the sweeps need
// no per-step division.
template <typename T, bool ROW, bool STRICT>
KB_DEV void tile_stage(const T* L, int lda, int n, int I, int J, T (&v)[16], TileLds<T>& S, int lane,
                       T invk = T(1))
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gr = 16 * I + i, gc = 16 * J + 4 * g + r;
        const bool in = gr < n && gc < n && gc <= gr;
        // branch-free: always load a valid lower element, then select
        const int rr = in ? gr : 0, cc = in ? gc : 0;
        const T ld = gld(L + rr + (long)cc * lda);
        S.tile[i][4 * g + r] = in ? ld : ((gr == gc) ? T(1) : T(0));
    }
    if (STRICT && g == 0) S.inv[i] = invk;
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        T x = ROW ? S.tile[i][j] : S.tile[j][i];
        if (STRICT) x = ROW ? ((j < i) ? x * S.inv[j] : T(0)) : ((j > i) ? x * invk : T(0));
        v[j] = x;
    }
    wave_lds_sync();
}

// Factor tiles read from L in global memory (staged through LDS).
template <typename T>
struct GlobalTiles {
    const T* L;
    int lda, n;
    TileLds<T>* S;
    template <bool ROW, bool STRICT>
    KB_DEV void get(int I, int J, T (&v)[16], int lane, T invk) const
    {
        tile_stage<T, ROW, STRICT>(L, lda, n, I, J, v, *S, lane, invk);
    }
};

// Factor tiles taken from the wavefront's registers (the factorization's
// output, or a tile_load of L), staged through LDS one tile at a time.
template <typename T>
struct RegTiles {
    const typename TileV<T>::v4* t;
    TileLds<T>* S;
    template <bool ROW, bool STRICT>
    KB_DEV void get(int I, int J, T (&v)[16], int lane, T invk) const
    {
        const int i = lane & 15, g = lane >> 4;
        typename TileV<T>::v4 x;
        // uniform switch: the register tile is picked by a scalar branch (the
        // empty asm keeps the compiler from turning it into an indexed load
        // from a stack copy of t)
        switch (tix(I, J)) {
        case 0: x = t[0]; asm volatile("" : "+v"(x)); break;
        case 1: x = t[1]; asm volatile("" : "+v"(x)); break;
        case 2: x = t[2]; asm volatile("" : "+v"(x)); break;
        case 3: x = t[3]; asm volatile("" : "+v"(x)); break;
        case 4: x = t[4]; asm volatile("" : "+v"(x)); break;
        case 5: x = t[5]; asm volatile("" : "+v"(x)); break;
        case 6: x = t[6]; asm volatile("" : "+v"(x)); break;
        case 7: x = t[7]; asm volatile("" : "+v"(x)); break;
        case 8: x = t[8]; asm volatile("" : "+v"(x)); break;
        default: x = t[9]; asm volatile("" : "+v"(x)); break;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) S->tile[i][4 * g + r] = x[r];
        if (STRICT && g == 0) S->inv[i] = invk;
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            T y = ROW ? S->tile[i][j] : S->tile[j][i];
            if (STRICT) y = ROW ? ((j < i) ? y * S->inv[j] : T(0)) : ((j > i) ? y * invk : T(0));
            v[j] = y;
        }
        wave_lds_sync();
    }
};

// (L L^T) X = B for the group's four right-hand sides.  Right-hand side c of
// the group sits in DPP row c: y[K][16c + i] = row 16K + i.  The tile loops
// are runtime loops over LDS-resident blocks so that register use stays that
// of a single 16 x 16 step.  src.get<ROW,STRICT>(I, J, v, lane, 1/L(i,i))
// returns row / column i of factor tile (I,J) (STRICT: unit-scaled strictly
// lower part of a diagonal tile).
template <typename T, typename Src>
KB_DEV void tile_solve(const Src& src, T (&y)[4][64], const T (&invd)[4][64], int lane)
{
    T v[16];
    // forward: L y = b, L = Lu D with Lu unit lower: z = Lu^-1 b, y = D^-1 z
#pragma nounroll
    for (int K = 0; K < 4; ++K) {
        const T ik = invd[K][lane];
        T yk = y[K][lane];
        src.template get<true, true>(K, K, v, lane, ik);
#pragma unroll
        for (int j = 0; j < 16; ++j) fnmac_rowk<T>(yk, yk, v[j], j);
        yk *= ik;
        y[K][lane] = yk;
#pragma nounroll
        for (int I = K + 1; I < 4; ++I) {
            T yi = y[I][lane];
            src.template get<true, false>(I, K, v, lane, T(1));
#pragma unroll
            for (int j = 0; j < 16; ++j) fnmac_rowk<T>(yi, yk, v[j], j);
            y[I][lane] = yi;
        }
    }
    // backward: L^T x = y; on the diagonal tile L^T = D Lu^T, so scale by
    // 1/L(i,i) first, then the unit sweep with v[j] = Lu(j, i)
#pragma nounroll
    for (int K = 3; K >= 0; --K) {
        const T ik = invd[K][lane];
        T yk = y[K][lane];
#pragma nounroll
        for (int I = K + 1; I < 4; ++I) {
            const T xi = y[I][lane];
            src.template get<false, false>(I, K, v, lane, T(1));
#pragma unroll
            for (int j = 0; j < 16; ++j) fnmac_rowk<T>(yk, xi, v[j], j);
        }
        src.template get<false, true>(K, K, v, lane, ik);
        yk *= ik;
#pragma unroll
        for (int j = 15; j >= 0; --j) fnmac_rowk<T>(yk, yk, v[j], j);
        y[K][lane] = yk;
    }
}

template <typename T, int MODE>
KB_DEV void tile_chol_one(int n, int nrhs, int side_left, T* Ap, int lda, T* Bp, int ldb, int* info_out,
                          TileLds<T>& S, int lane)
{
    T invd[4] = {T(1), T(1), T(1), T(1)};
    int inf = 0;
    const int i = lane & 15;
    typename TileV<T>::v4 t[10];
    if constexpr (MODE != KB_MODE_POTRS) {
        tile_load<T>(t, Ap, lda, n, lane);
        inf = tile_potrf<T>(t, invd, n, S.pan, lane);
        tile_store<T>(t, Ap, lda, n, lane);
        if (lane == 0 && info_out) *info_out = inf;
    } else {
        // the factor, all ten tiles in one batch of loads
        tile_load<T>(t, Ap, lda, n, lane);
#pragma unroll
        for (int K = 0; K < 4; ++K) {
            const int row = 16 * K + i;
            invd[K] = (row < n) ? T(1) / gld(Ap + row + (long)row * lda) : T(1);
        }
    }
    if constexpr (MODE != KB_MODE_POTRF) {
        if (inf) return;
#pragma unroll
        for (int K = 0; K < 4; ++K) S.invd[K][lane] = invd[K];
        const int c = lane >> 4;
        for (int c0 = 0; c0 < nrhs; c0 += 4) {
            const bool act = c0 + c < nrhs;
#pragma unroll
            for (int K = 0; K < 4; ++K) {
                const int row = 16 * K + i;
                const bool in = act && row < n;
                const T ld = gld(rhs_elem(Bp, ldb, side_left, in ? row : 0, in ? c0 + c : 0));
                S.y[K][lane] = in ? ld : T(0);
            }
            wave_lds_sync();
            RegTiles<T> src{t, &S};
            tile_solve<T>(src, S.y, S.invd, lane);
            wave_lds_sync();
#pragma unroll
            for (int K = 0; K < 4; ++K) {
                const int row = 16 * K + i;
                if (act && row < n) gst(rhs_elem(Bp, ldb, side_left, row, c0 + c), S.y[K][lane]);
            }
        }
    }
}

// 4 wavefronts (4 matrices) per 256-thread workgroup
template <typename T, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_tile_chol(int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    __shared__ TileLds<T> S[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long mat = (long)blockIdx.x * 4 + wave;
    if (mat >= batch) return;
    T* Bp = (MODE == KB_MODE_POTRF) ? nullptr : B.get(mat);
    const int ldb = (MODE == KB_MODE_POTRF) ? 1 : B.ldim(mat);
    tile_chol_one<T, MODE>(n, nrhs, side_left, A.get(mat), A.ldim(mat), Bp, ldb, info ? info + mat : nullptr,
                           S[wave], lane);
}

template <typename T>
static int launch_tile(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch,
                       hipStream_t s)
{
    dim3 grid((unsigned)((batch + 3) / 4));
    BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    if (mode == KB_MODE_POTRF)
        hipLaunchKernelGGL((k_tile_chol<T, KB_MODE_POTRF>), grid, dim3(256), 0, s, n, nrhs, side_left, a, b, info, batch);
    else if (mode == KB_MODE_POSV)
        hipLaunchKernelGGL((k_tile_chol<T, KB_MODE_POSV>), grid, dim3(256), 0, s, n, nrhs, side_left, a, b, info, batch);
    else
        hipLaunchKernelGGL((k_tile_chol<T, KB_MODE_POTRS>), grid, dim3(256), 0, s, n, nrhs, side_left, a, b, info, batch);
    return (int)hipGetLastError();
}

