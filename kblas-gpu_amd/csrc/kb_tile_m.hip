// kb_tile_m.hip -- MFMA-tiled mixed-precision POSV (32 < n <= 64).
#include "kb_tile_impl.h"

// ---------------------------------------------------------------------------
// Mixed precision POSV (LAPACK dsposv restated), 32 < n <= 64, one matrix per
// wavefront: FP64 A kept in registers as all sixteen 16x16 tiles (the lower
// ten as loaded, the upper six built by transposition) for the residual, its
// FP32 copy factored with FP32 MFMA (2x the FP64 matrix-core rate), the unit
// FP32 factor kept in LDS for the refinement sweeps; on an FP32 overflow of A,
// B or a residual (ITER = -2), a non-positive FP32 pivot (-3) or no
// convergence after 30 steps (-31) the FP64 tiles are factored in place (A
// then holds L, as dsposv) and solved in FP64.
// ---------------------------------------------------------------------------
constexpr int KB_TILE_ITERMAX = 30;
constexpr double KB_FLT_MAX = 3.4028234663852886e38;

struct MixLds {
    union {
        struct {
            float pan[4][16][4];
            float l32[10][16][17];    // unit FP32 factor Lu = L D^-1, row-major tiles (diagonal: strictly lower)
            float d2[4][64];          // 1/L32(16K+i, 16K+i)^2 per lane
            double tile[16][17];      // FP64 tile transposes, then the solution blocks of the residual
        } m;
        TileLds<double> d;            // FP64 fallback
    };
};

KB_DEV double row16_max(double x)
{
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        double y = __shfl_xor(x, off, 64);
        x = (y > x || y != y) ? y : x;
    }
    return x;
}

// Full symmetric A in tile storage: tu[tix(I, J)] (I > J) receives tile (J, I)
// = T(I, J)^T and the diagonal tiles get their upper triangle mirrored.
KB_DEV void mix_symmetrize(kb_d4 (&ta)[10], kb_d4 (&tu)[10], double (*tl)[17], int lane)
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J) {
#pragma unroll
            for (int r = 0; r < 4; ++r) tl[i][4 * g + r] = ta[tix(I, J)][r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double y = tl[4 * g + r][i];
                if (I == J) {
                    if (4 * g + r > i) ta[tix(I, I)][r] = y;
                } else {
                    tu[tix(I, J)][r] = y;
                }
            }
            wave_lds_sync();
        }
}

// element (16I + i, 16J + 4g + r) of the full symmetric A
KB_DEV double mix_a(const kb_d4 (&ta)[10], const kb_d4 (&tu)[10], int I, int J, int r)
{
    return (I >= J) ? ta[tix(I, J)][r] : tu[tix(J, I)][r];
}

// ||A||_inf over the leading n x n block: row sums of |A| reduced over the
// four lanes that share a row, then the maximum over the rows.
KB_DEV double mix_anorm(const kb_d4 (&ta)[10], const kb_d4 (&tu)[10], int n, int lane)
{
    const int i = lane & 15;
    double anrm = 0.0;
#pragma unroll
    for (int I = 0; I < 4; ++I) {
        double s = 0.0;
#pragma unroll
        for (int J = 0; J < 4; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) s += __builtin_fabs(mix_a(ta, tu, I, J, r));
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (16 * I + i >= n) s = 0.0;     // identity padding
        anrm = (s > anrm || s != s) ? s : anrm;
    }
    return row16_max(anrm);
}

// One substitution step j of a unit triangular sweep as a single instruction
// block: the NO off-diagonal products of step j-1 (plain FMAs of t, which
// holds lane j-1's final value, into the accumulators a), the broadcast
// t := y(lane j) and the diagonal update y -= t * ld.  The FMAs (or an s_nop)
// supply the two wait states the DPP read of y needs after the previous
// block's write of y; the accumulators take positive sums.
template <int J>
KB_DEV void mix_step0(float& y, float& t, float ld)
{
    asm("s_nop 1\n\tv_mov_b32_dpp %1, %0 row_newbcast:%3 row_mask:0xf bank_mask:0xf\n\tv_fmac_f32_e64 %0, -%1, %2"
        : "+v"(y), "+v"(t)
        : "v"(ld), "n"(J));
}
template <int J>
KB_DEV void mix_step1(float& y, float& t, float ld, float& a1, float l1)
{
    asm("v_fmac_f32 %2, %4, %1\n\ts_nop 0\n\tv_mov_b32_dpp %1, %0 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_e64 %0, -%1, %3"
        : "+v"(y), "+v"(t), "+v"(a1)
        : "v"(ld), "v"(l1), "n"(J));
}
template <int J>
KB_DEV void mix_step2(float& y, float& t, float ld, float& a1, float l1, float& a2, float l2)
{
    asm("v_fmac_f32 %2, %5, %1\n\tv_fmac_f32 %3, %6, %1\n\tv_mov_b32_dpp %1, %0 row_newbcast:%7 row_mask:0xf "
        "bank_mask:0xf\n\tv_fmac_f32_e64 %0, -%1, %4"
        : "+v"(y), "+v"(t), "+v"(a1), "+v"(a2)
        : "v"(ld), "v"(l1), "v"(l2), "n"(J));
}
template <int J>
KB_DEV void mix_step3(float& y, float& t, float ld, float& a1, float l1, float& a2, float l2, float& a3, float l3)
{
    asm("v_fmac_f32 %2, %6, %1\n\tv_fmac_f32 %3, %7, %1\n\tv_fmac_f32 %4, %8, %1\n\tv_mov_b32_dpp %1, %0 "
        "row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\tv_fmac_f32_e64 %0, -%1, %5"
        : "+v"(y), "+v"(t), "+v"(a1), "+v"(a2), "+v"(a3)
        : "v"(ld), "v"(l1), "v"(l2), "v"(l3), "n"(J));
}
#define KB_MS0(i) case i: mix_step0<i>(y, t, ld); return;
#define KB_MS1(i) case i: mix_step1<i>(y, t, ld, a[0], l[0]); return;
#define KB_MS2(i) case i: mix_step2<i>(y, t, ld, a[0], l[0], a[1], l[1]); return;
#define KB_MS3(i) case i: mix_step3<i>(y, t, ld, a[0], l[0], a[1], l[1], a[2], l[2]); return;
// j becomes a constant once the sweeps are unrolled; the switch folds
template <int NO>
KB_DEV void mix_step(int j, float& y, float& t, float ld, float (&a)[3], const float (&l)[3])
{
    if constexpr (NO == 0) { switch (j & 15) { KB_REP16(KB_MS0, 0) } }
    else if constexpr (NO == 1) { switch (j & 15) { KB_REP16(KB_MS1, 0) } }
    else if constexpr (NO == 2) { switch (j & 15) { KB_REP16(KB_MS2, 0) } }
    else { switch (j & 15) { KB_REP16(KB_MS3, 0) } }
}
#undef KB_MS0
#undef KB_MS1
#undef KB_MS2
#undef KB_MS3

// Forward sweep over block K: unit substitution with the diagonal tile,
// interleaved with the products of the off-diagonal tiles (I, K), I > K,
// accumulated into acc[I]; then the D^-2 scaling.
template <int K>
KB_DEV void mix_fwd_block(float (&y)[4], float (&acc)[4], const MixLds& S, int lane)
{
    constexpr int NO = 3 - K;
    const int i = lane & 15;
    const float(*t0)[17] = S.m.l32[tix(K, K)];
    float yk = y[K] - acc[K];
    float t = 0.f;
    float a[3] = {0.f, 0.f, 0.f};
    float l[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j == 0) {
            mix_step<0>(0, yk, t, t0[i][0], a, l);
        } else {
#pragma unroll
            for (int o = 0; o < NO; ++o) l[o] = S.m.l32[tix(K + 1 + o, K)][i][j - 1];
            mix_step<NO>(j, yk, t, t0[i][j], a, l);
        }
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[K + 1 + o] += a[o] + S.m.l32[tix(K + 1 + o, K)][i][15] * t;
    y[K] = yk * S.m.d2[K][lane];
}

// Backward sweep over block K: unit substitution with the transposed
// diagonal tile, interleaved with the products of the transposed tiles
// (K, I), I < K, accumulated into acc[I].
template <int K>
KB_DEV void mix_bwd_block(float (&y)[4], float (&acc)[4], const MixLds& S, int lane)
{
    constexpr int NO = K;
    const int i = lane & 15;
    const float(*t0)[17] = S.m.l32[tix(K, K)];
    float yk = y[K] - acc[K];
    float t = 0.f;
    float a[3] = {0.f, 0.f, 0.f};
    float l[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 15; j >= 0; --j) {
        if (j == 15) {
            mix_step<0>(15, yk, t, t0[15][i], a, l);
        } else {
#pragma unroll
            for (int o = 0; o < NO; ++o) l[o] = S.m.l32[tix(K, o)][j + 1][i];
            mix_step<NO>(j, yk, t, t0[j][i], a, l);
        }
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] += a[o] + S.m.l32[tix(K, o)][0][i] * t;
    y[K] = yk;
}

// z := (Lu D^2 Lu^T)^-1 z in FP32 for the group's four right-hand sides
// (lane 16c + i holds row 16K + i of right-hand side c in y[K]): a unit forward
// sweep, the D^-2 scaling, a unit backward sweep.
KB_DEV void mix_solve32(float (&y)[4], const MixLds& S, int lane)
{
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    mix_fwd_block<0>(y, acc, S, lane);
    mix_fwd_block<1>(y, acc, S, lane);
    mix_fwd_block<2>(y, acc, S, lane);
    mix_fwd_block<3>(y, acc, S, lane);
#pragma unroll
    for (int K = 0; K < 4; ++K) acc[K] = 0.f;
    mix_bwd_block<3>(y, acc, S, lane);
    mix_bwd_block<2>(y, acc, S, lane);
    mix_bwd_block<1>(y, acc, S, lane);
    mix_bwd_block<0>(y, acc, S, lane);
}

// r := b - A x in FP64.  The solution blocks go through LDS (xs[J][row][c]);
// lane (i, g) forms the partial products of its columns 4g..4g+3 for all four
// right-hand sides (sixteen independent accumulators), and a two-step
// exchange across the lane groups (xor 32, xor 16) sums them so that lane
// 16c + i ends with row 16I + i of right-hand side c.
KB_DEV void mix_residual(const kb_d4 (&ta)[10], const kb_d4 (&tu)[10], const double (&xr)[4],
                         const double (&br)[4], double (&rr)[4], double (*xs)[16][4], int lane)
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int K = 0; K < 4; ++K) xs[K][i][g] = xr[K];
    wave_lds_sync();
    double q[4][4];
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int c = 0; c < 4; ++c) q[I][c] = 0.0;
#pragma unroll
    for (int J = 0; J < 4; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double x4[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) x4[c] = xs[J][4 * g + r][c];
#pragma unroll
            for (int I = 0; I < 4; ++I) {
                const double a = mix_a(ta, tu, I, J, r);
#pragma unroll
                for (int c = 0; c < 4; ++c) q[I][c] = __builtin_fma(a, x4[c], q[I][c]);
            }
        }
    wave_lds_sync();
    const bool b1 = (g & 2) != 0, b0 = (g & 1) != 0;
#pragma unroll
    for (int I = 0; I < 4; ++I) {
        double h[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double send = b1 ? q[I][c] : q[I][2 + c];
            const double keep = b1 ? q[I][2 + c] : q[I][c];
            h[c] = keep + __shfl_xor(send, 32, 64);
        }
        const double send = b0 ? h[0] : h[1];
        const double keep = b0 ? h[1] : h[0];
        rr[I] = br[I] - (keep + __shfl_xor(send, 16, 64));
    }
}

__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
k_tile_sposv(int n, int nrhs, int side_left, BatchArg<double> A, BatchArg<double> B, BatchArg<double> X,
             int* info, int* iter, int batch)
{
    __shared__ MixLds SS[2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long mat = (long)blockIdx.x * 2 + wave;
    if (mat >= batch) return;
    MixLds& S = SS[wave];
    double* Ap = A.get(mat);
    const int lda = A.ldim(mat);
    double* Bp = B.get(mat);
    const int ldb = B.ldim(mat);
    double* Xp = X.get(mat);
    const int ldx = X.ldim(mat);
    const int i = lane & 15, g = lane >> 4;

    kb_d4 ta[10];
    tile_load<double>(ta, Ap, lda, n, lane);

    // FP32 range of B, then of A (dsposv's DLAG2S / DLAT2S checks): ITER = -2
    bool big = false;
    for (int c0 = 0; c0 < nrhs; c0 += 4) {
#pragma unroll
        for (int K = 0; K < 4; ++K) {
            const int row = 16 * K + i;
            const bool in = c0 + g < nrhs && row < n;
            const double v = gld(rhs_elem(Bp, ldb, side_left, in ? row : 0, in ? c0 + g : 0));
            big = big || (in && __builtin_fabs(v) > KB_FLT_MAX);
        }
    }
#pragma unroll
    for (int t = 0; t < 10; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) big = big || __builtin_fabs(ta[t][r]) > KB_FLT_MAX;
    const bool ovf = __builtin_amdgcn_ballot_w64(big) != 0;

    // FP32 factorization; the unit factor Lu = L D^-1 goes to LDS
    int iters = 0;
    bool ok = false;
    if (ovf) {
        iters = -2;
    } else {
        kb_f4 tf[10];
#pragma unroll
        for (int t = 0; t < 10; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) tf[t][r] = (float)ta[t][r];
        float inv32[4] = {1.f, 1.f, 1.f, 1.f};
        const int finfo = tile_potrf<float>(tf, inv32, n, S.m.pan, lane);
        ok = (finfo == 0);
        if (ok) {
#pragma unroll
            for (int K = 0; K < 4; ++K) S.m.d2[K][lane] = inv32[K] * inv32[K];
#pragma unroll
            for (int I = 0; I < 4; ++I)
#pragma unroll
                for (int J = 0; J <= I; ++J)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = 4 * g + r;
                        // column j scaled by 1/L(16J+j, 16J+j), which lane j holds
                        float v = tf[tix(I, J)][r] * __shfl(inv32[J], j, 64);
                        if (I == J && j >= i) v = 0.0f;
                        S.m.l32[tix(I, J)][i][j] = v;
                    }
        } else {
            iters = -3;
        }
    }
    wave_lds_sync();

    kb_d4 tu[10];
    mix_symmetrize(ta, tu, S.m.tile, lane);
    const double anrm = mix_anorm(ta, tu, n, lane);
    const double cte = anrm * 1.1102230246251565e-16 * __builtin_sqrt((double)n);
    double(*xs)[16][4] = (double(*)[16][4]) & S.m.tile[0][0];

    const int c = lane >> 4;
    for (int c0 = 0; c0 < nrhs && ok; c0 += 4) {
        const bool act = c0 + c < nrhs;
        // x = 0, r = b; block K of right-hand side c of the group: row 16K + i in lane 16c + i
        double xr[4], rr[4], br[4];
#pragma unroll
        for (int K = 0; K < 4; ++K) {
            const int row = 16 * K + i;
            const bool in = act && row < n;
            const double ld = gld(rhs_elem(Bp, ldb, side_left, in ? row : 0, in ? c0 + c : 0));
            xr[K] = 0.0;
            br[K] = in ? ld : 0.0;
            rr[K] = br[K];
        }
        bool conv = false;
        int it = 0;
        for (; it <= KB_TILE_ITERMAX; ++it) {
            // x += (L32 L32^T)^-1 r ; r = b - A x
            bool rbig = false;
#pragma unroll
            for (int K = 0; K < 4; ++K) rbig = rbig || __builtin_fabs(rr[K]) > KB_FLT_MAX;
            if (__builtin_amdgcn_ballot_w64(rbig) != 0) break;      // DLAG2S(R) overflow
            float z[4];
#pragma unroll
            for (int K = 0; K < 4; ++K) z[K] = (float)rr[K];
            mix_solve32(z, S, lane);
#pragma unroll
            for (int K = 0; K < 4; ++K) xr[K] += (double)z[K];
            mix_residual(ta, tu, xr, br, rr, xs, lane);
            // convergence: ||r||inf <= ||x||inf * cte for every right-hand side of the group
            double rn = 0.0, xn = 0.0;
#pragma unroll
            for (int K = 0; K < 4; ++K) {
                const double ar = __builtin_fabs(rr[K]), ax = __builtin_fabs(xr[K]);
                rn = (ar > rn || ar != ar) ? ar : rn;
                xn = ax > xn ? ax : xn;
            }
            rn = row16_max(rn);
            xn = row16_max(xn);
            const bool mine = !act || (rn <= xn * cte);
            if (__builtin_amdgcn_ballot_w64(!mine) == 0) { conv = true; break; }
        }
        if (!conv) {
            ok = false;
            iters = (it > KB_TILE_ITERMAX) ? -(KB_TILE_ITERMAX + 1) : -2;
            break;
        }
        if (it > iters) iters = it;
#pragma unroll
        for (int K = 0; K < 4; ++K) {
            const int row = 16 * K + i;
            if (act && row < n) gst(rhs_elem(Xp, ldx, side_left, row, c0 + c), xr[K]);
        }
    }

    int inf = 0;
    if (!ok) {
        wave_lds_sync();
        // the lower tiles of ta are A's (the mirrored upper parts of the
        // diagonal tiles never feed the lower triangle and are not stored)
        double invd[4] = {1.0, 1.0, 1.0, 1.0};
        inf = tile_potrf<double>(ta, invd, n, S.d.pan, lane);
        tile_store<double>(ta, Ap, lda, n, lane);
        if (inf == 0) {
            wave_global_sync();
#pragma unroll
            for (int K = 0; K < 4; ++K) S.d.invd[K][lane] = invd[K];
            for (int c0 = 0; c0 < nrhs; c0 += 4) {
                const bool act = c0 + c < nrhs;
#pragma unroll
                for (int K = 0; K < 4; ++K) {
                    const int row = 16 * K + i;
                    const bool in = act && row < n;
                    const double ld = gld(rhs_elem(Bp, ldb, side_left, in ? row : 0, in ? c0 + c : 0));
                    S.d.y[K][lane] = in ? ld : 0.0;
                }
                wave_lds_sync();
                GlobalTiles<double> src{Ap, lda, n, &S.d};
                tile_solve<double>(src, S.d.y, S.d.invd, lane);
                wave_lds_sync();
#pragma unroll
                for (int K = 0; K < 4; ++K) {
                    const int row = 16 * K + i;
                    if (act && row < n) gst(rhs_elem(Xp, ldx, side_left, row, c0 + c), S.d.y[K][lane]);
                }
            }
        }
    }
    if (lane == 0) {
        if (info) info[mat] = inf;
        if (iter) iter[mat] = iters;
    }
}

static int launch_tile_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info,
                             int* iter, int batch, hipStream_t s)
{
    dim3 grid((unsigned)((batch + 1) / 2));
    hipLaunchKernelGGL(k_tile_sposv, grid, dim3(128), 0, s, n, nrhs, side_left, to_arg<double>(A),
                       to_arg<double>(B), to_arg<double>(X), info, iter, batch);
    return (int)hipGetLastError();
}

extern "C" int kb_tile_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info,
                             int* iter, int batch, hipStream_t s)
{
    return launch_tile_sposv(n, nrhs, side_left, A, B, X, info, iter, batch, s);
}
