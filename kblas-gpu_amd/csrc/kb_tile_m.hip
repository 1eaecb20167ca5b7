// kb_tile_m.hip -- MFMA-tiled mixed-precision POSV (32 < n <= 64).
#include "kb_tile_impl.h"

// ---------------------------------------------------------------------------
// Mixed precision POSV (LAPACK dsposv restated), 32 < n <= 64, one matrix per
// wavefront: FP64 A kept as ten register tiles for the residual, its FP32 copy
// factored with FP32 MFMA (2x the FP64 matrix-core rate), the FP32 factor
// kept in LDS for the refinement sweeps; on a non-positive FP32 pivot or no
// convergence after 30 steps the FP64 tiles are factored in place (A then
// holds L, as dsposv) and solved in FP64.
// ---------------------------------------------------------------------------
constexpr int KB_TILE_ITERMAX = 30;

struct MixLds {
    union {
        struct {
            float pan[4][16][4];
            float l32[10][16][17];    // FP32 factor tiles, row-major
            float inv32[4][64];       // 1/L32(16K+i, 16K+i) per lane
            float z[4][64];           // FP32 correction blocks
            double tile[16][17];      // FP64 A-tile staging for the residual
            double rs[64];            // row sums of |A|
        } m;
        TileLds<double> d;            // FP64 fallback
    };
};

// FP32 factor tiles read straight from LDS; the diagonal tiles are stored
// already unit-scaled (strictly lower part, entry (i,j) times 1/L(j,j), zero
// elsewhere), as the sweeps consume them.
struct LdsTiles32 {
    MixLds* S;
    template <bool ROW, bool STRICT>
    KB_DEV void get(int I, int J, float (&v)[16], int lane, float invk) const
    {
        (void)invk;
        const int i = lane & 15;
        const float(*t)[17] = S->m.l32[tix(I, J)];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = ROW ? t[i][j] : t[j][i];
    }
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

__global__ void __launch_bounds__(128)
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

    // ||A||_inf from the lower triangle (row sums of |A| through LDS atomics)
    S.m.rs[lane] = 0.0;
    wave_lds_sync();
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = 0; J <= I; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + i, col = 16 * J + 4 * g + r;
                if (row < n && col <= row) {
                    const double a = __builtin_fabs(ta[tix(I, J)][r]);
                    atomicAdd(&S.m.rs[row], a);
                    if (col < row) atomicAdd(&S.m.rs[col], a);
                }
            }
    wave_lds_sync();
    double anrm = S.m.rs[lane];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double y = __shfl_xor(anrm, off, 64);
        anrm = y > anrm ? y : anrm;
    }
    const double cte = anrm * 1.1102230246251565e-16 * __builtin_sqrt((double)n);

    // FP32 factorization
    int iters = 0;
    bool ok;
    {
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
            for (int K = 0; K < 4; ++K) S.m.inv32[K][lane] = inv32[K];
            wave_lds_sync();
#pragma unroll
            for (int I = 0; I < 4; ++I)
#pragma unroll
                for (int J = 0; J <= I; ++J)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int j = 4 * g + r;
                        float v = tf[tix(I, J)][r];
                        // diagonal tiles: unit lower part, 1/L(16J+j, 16J+j) sits in lane j
                        if (I == J) v = (j < i) ? v * S.m.inv32[J][j] : 0.0f;
                        S.m.l32[tix(I, J)][i][j] = v;
                    }
        } else {
            iters = -3;
        }
    }
    wave_lds_sync();

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
            // z = (L32 L32^T)^-1 r
#pragma unroll
            for (int K = 0; K < 4; ++K) S.m.z[K][lane] = (float)rr[K];
            wave_lds_sync();
            LdsTiles32 src{&S};
            tile_solve<float>(src, S.m.z, S.m.inv32, lane);
            wave_lds_sync();
            // x += z ; r = b - A x
#pragma unroll
            for (int K = 0; K < 4; ++K) {
                xr[K] += (double)S.m.z[K][lane];
                rr[K] = br[K];
            }
#pragma unroll
            for (int I = 0; I < 4; ++I) {
#pragma unroll
                for (int J = 0; J <= I; ++J) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) S.m.tile[i][4 * g + r] = ta[tix(I, J)][r];
                    wave_lds_sync();
                    double v[16];
                    {   // r_I -= A_IJ x_J   (row i of the tile)
#pragma unroll
                        for (int j = 0; j < 16; ++j) v[j] = (I == J && j > i) ? S.m.tile[j][i] : S.m.tile[i][j];
#pragma unroll
                        for (int j = 0; j < 16; ++j) fnmac_rowk<double>(rr[I], xr[J], v[j], j);
                    }
                    if (I != J) {   // r_J -= A_IJ^T x_I   (column i of the tile)
#pragma unroll
                        for (int j = 0; j < 16; ++j) v[j] = S.m.tile[j][i];
#pragma unroll
                        for (int j = 0; j < 16; ++j) fnmac_rowk<double>(rr[J], xr[I], v[j], j);
                    }
                    wave_lds_sync();
                }
            }
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
        if (!conv) { ok = false; break; }
        if (it > iters) iters = it;
#pragma unroll
        for (int K = 0; K < 4; ++K) {
            const int row = 16 * K + i;
            if (act && row < n) gst(rhs_elem(Xp, ldx, side_left, row, c0 + c), xr[K]);
        }
    }

    int inf = 0;
    if (!ok) {
        if (iters != -3) iters = -(KB_TILE_ITERMAX + 1);
        wave_lds_sync();
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
