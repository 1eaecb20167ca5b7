// kb_sposv_big.h -- mixed-precision POSV (LAPACK dsposv restated) for the
// orders above the register range of a non-uniform batch: one 256-thread
// workgroup per matrix, persistent over the size-sorted queue.
//
// Per matrix (dsposv, SRC/dsposv.f): B and the lower triangle of A are
// converted to FP32 with the range checks of DLAG2S / DLAT2S (ITER = -2),
// ||A||inf is formed on the way (CTE = ||A||inf eps sqrt(n)); the FP32 copy is
// factored by the blocked workgroup routine (ITER = -3 on a non-positive
// pivot); x0 = solve32(b); then, at most 30 times, r = b - A x in FP64 on the
// matrix cores, stop when every right-hand side has ||r||inf <= ||x||inf CTE
// (ITER = the number of corrections), else r to FP32 (overflow: ITER = -2),
// solve32, x += d.  Every failure ends in the FP64 path (X := B, A := L,
// solve; A then holds L).  The FP32 factor and the FP32 right-hand sides live
// in a workspace slot owned by the workgroup (max_n^2 + max_n * max_nrhs
// floats); orders or right-hand-side counts beyond the slot (or the limits
// below) take the FP64 path directly with ITER = -1, dsposv's "full precision
// for implementation reasons".
#pragma once
#include "kb_blk_impl.h"

constexpr int KB_SB_MAX_N = 1024;      // orders served by the mixed path
constexpr int KB_SB_MAX_RHS = 64;      // right-hand sides served by the mixed path
constexpr double KB_SB_FLT_MAX = 3.4028234663852886e38;
constexpr int KB_SB_ITERMAX = 30;

struct SbSmem {
    union {
        BlkSmem<float> f;                  // FP32 factorization and solves
        BlkSmem<double> d;                 // residual staging, FP64 path
    };
    unsigned long long rn[KB_SB_MAX_RHS];  // max |r| per right-hand side (bits of a non-negative double)
    unsigned long long xn[KB_SB_MAX_RHS];  // max |x| per right-hand side
    double red[KB_BLK_WG / 64];
};

KB_DEV unsigned long long sb_bits(double x) { return __builtin_bit_cast(unsigned long long, __builtin_fabs(x)); }
KB_DEV double sb_val(unsigned long long b) { return __builtin_bit_cast(double, b); }

// Workgroup-wide maximum (NaN wins).
KB_DEV double sb_block_max(double v, SbSmem& S)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v || o != o) ? o : v;
    }
    if (lane == 0) S.red[wave] = v;
    __syncthreads();
    double m = S.red[0];
#pragma unroll
    for (int w = 1; w < KB_BLK_WG / 64; ++w) {
        const double o = S.red[w];
        m = (o > m || o != o) ? o : m;
    }
    __syncthreads();
    return m;
}

// A32 := FP32 copy of the lower triangle of A (ld n); returns ||A||inf of
// the symmetric A.  Thread t takes rows t, t + 256, ...: row i's lower part
// (A(i, 0..i), read across the threads a column at a time, coalesced) and
// its upper part as column i below the diagonal (one contiguous run per
// thread), so every thread does about n reads per row.  big is set when an
// entry exceeds FP32 range.
constexpr int SB_U = 4;   // elements per thread whose loads go out together in the element-wise loops

KB_DEV double sb_convert_norm(const double* A, int lda, int n, float* A32, bool& big, SbSmem& S)
{
    double rmax = 0.0;
    for (int i = threadIdx.x; i < n; i += KB_BLK_WG) {
        double s = 0.0;
        const double* p = A + i;
        for (int j = 0; j <= i; ++j) {
            const double v = gld(p + (long)j * lda);
            gst(A32 + i + (long)j * n, (float)v);
            big = big || __builtin_fabs(v) > KB_SB_FLT_MAX;
            s += __builtin_fabs(v);
        }
        const double* q = A + i + (long)i * lda;
        for (int k = 1; k < n - i; ++k) s += __builtin_fabs(gld(q + k));
        rmax = (s > rmax || s != s) ? s : rmax;
    }
    return sb_block_max(rmax, S);
}

// Z := FP32 copy of the right-hand sides B (n x nrhs, ld n); big as above.
KB_DEV void sb_convert_rhs(const double* B, int ldb, int side_left, int n, int nrhs, float* Z, bool& big)
{
    const int tot = n * nrhs;
    for (int e0 = threadIdx.x; e0 < tot; e0 += SB_U * KB_BLK_WG) {
        double v[SB_U];
#pragma unroll
        for (int u = 0; u < SB_U; ++u) {
            const int e = e0 + u * KB_BLK_WG;
            const bool in = e < tot;
            v[u] = gld(rhs_elem(const_cast<double*>(B), ldb, side_left, in ? e % n : 0, in ? e / n : 0));
        }
#pragma unroll
        for (int u = 0; u < SB_U; ++u) {
            const int e = e0 + u * KB_BLK_WG;
            if (e < tot) {
                big = big || __builtin_fabs(v[u]) > KB_SB_FLT_MAX;
                gst(Z + e % n + (long)(e / n) * n, (float)v[u]);
            }
        }
    }
}

// X := (first ? 0 : X) + Z, and S.xn[c] := max_i |x(i, c)|.
KB_DEV void sb_update_x(double* X, int ldx, int side_left, int n, int nrhs, const float* Z, bool first, SbSmem& S)
{
    // SB_U elements per thread and step, their loads issued together (from
    // clamped addresses) before any store
    const int tot = n * nrhs;
    for (int e0 = threadIdx.x; e0 < tot; e0 += SB_U * KB_BLK_WG) {
        double xo[SB_U];
        float z[SB_U];
#pragma unroll
        for (int u = 0; u < SB_U; ++u) {
            const int e = e0 + u * KB_BLK_WG;
            const bool in = e < tot;
            const int i = in ? e % n : 0, c = in ? e / n : 0;
            xo[u] = first ? 0.0 : gld(rhs_elem(X, ldx, side_left, i, c));
            z[u] = gld(Z + i + (long)c * n);
        }
#pragma unroll
        for (int u = 0; u < SB_U; ++u) {
            const int e = e0 + u * KB_BLK_WG;
            if (e < tot) {
                const int i = e % n, c = e / n;
                const double x = xo[u] + (double)z[u];
                gst(rhs_elem(X, ldx, side_left, i, c), x);
                atomicMax(&S.xn[c], sb_bits(x));
            }
        }
    }
}

// Z := FP32 copy of R = B - A X for the symmetric A (lower triangle stored),
// S.rn[c] := max_i |r(i, c)|; ovf set when an |r| exceeds FP32 range.
// Wave w computes the 16-row blocks w, w + 4, ... of R as 16x16x4 FP64
// matrix-core products over k: the A operand (rows of the block, 4 columns)
// read from memory by symmetry, the B operand from 32-row chunks of X staged
// in LDS (right-hand-side major, odd pitch); the accumulator rows are
// interleaved (register q, lane group g -> row 4q + g).
KB_DEV void sb_residual(const double* A, int lda, int n, const double* B, int ldb, const double* X, int ldx,
                        int side_left, int nrhs, float* Z, SbSmem& S, bool& ovf)
{
    constexpr int NBK = KB_BLK_NBK;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lg = lane >> 4;
    const int ncb = (nrhs + 15) / 16;
    const int nb = (n + 15) / 16;
    double* sY = S.d.sB;   // 64 x 33: within the slice area, nothing else is live
    for (int rb0 = 0; rb0 < nb; rb0 += KB_BLK_WG / 64) {
        const int r0 = 16 * (rb0 + wave);
        kb_d4 acc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = kb_d4{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += NBK) {
            const int kb = min(NBK, n - k0);
            static_assert(NBK * KB_SB_MAX_RHS % KB_BLK_WG == 0, "whole steps");
#pragma unroll
            for (int u = 0; u < NBK * KB_SB_MAX_RHS / KB_BLK_WG; ++u) {
                const int e = tid + u * KB_BLK_WG;
                const int k = e % NBK, c = e / NBK;
                const bool in = k < kb && c < nrhs;
                const double v = gld(rhs_elem(const_cast<double*>(X), ldx, side_left, in ? k0 + k : 0, in ? c : 0));
                sY[c * (NBK + 1) + k] = in ? v : 0.0;
            }
            __syncthreads();
            if (r0 < n) {
                const int row = r0 + li;
#pragma unroll
                for (int q = 0; q < NBK / 4; ++q) {
                    const int k = k0 + 4 * q + lg;
                    const bool in = row < n && k < n;
                    const long off = (row >= k) ? row + (long)k * lda : k + (long)row * lda;
                    const double av = gld(A + (in ? off : 0));
                    const double a = in ? av : 0.0;
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        if (cb < ncb) {
                            const double b = sY[(16 * cb + li) * (NBK + 1) + 4 * q + lg];
                            acc[cb] = mfma16x16x4(a, b, acc[cb]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (r0 < n) {
            double bv[4][4];   // B's block, loaded together from clamped addresses
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int col = 16 * cb + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + 4 * q + lg;
                    const bool in = cb < ncb && row < n && col < nrhs;
                    bv[cb][q] = gld(rhs_elem(const_cast<double*>(B), ldb, side_left, in ? row : 0, in ? col : 0));
                }
            }
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int col = 16 * cb + li;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = r0 + 4 * q + lg;
                    if (cb < ncb && row < n && col < nrhs) {
                        const double r = bv[cb][q] - acc[cb][q];
                        ovf = ovf || __builtin_fabs(r) > KB_SB_FLT_MAX;
                        atomicMax(&S.rn[col], sb_bits(r));
                        gst(Z + row + (long)col * n, (float)r);
                    }
                }
            }
        }
    }
}

// One matrix of the mixed path (all threads of the workgroup).  Returns
// (info, iter) through the pointers (written by thread 0).
KB_DEV void sb_sposv_one(long mat, int n, int nrhs, int side_left, double* Ap, int lda, const double* Bp, int ldb,
                         double* Xp, int ldx, float* A32, float* Z, int* info, int* iter, SbSmem& S)
{
    // FP32 range of B, then of A; ||A||inf
    bool big = false;
    sb_convert_rhs(Bp, ldb, side_left, n, nrhs, Z, big);
    const double anrm = sb_convert_norm(Ap, lda, n, A32, big, S);
    const double cte = anrm * 1.1102230246251565e-16 * __builtin_sqrt((double)n);
    int iters = 0;
    bool ok = !__syncthreads_or(big);
    if (!ok) {
        iters = -2;
    } else {
        const int f32 = blk_potrf<float>(A32, n, n, S.f);
        if (f32 != 0) {
            ok = false;
            iters = -3;
        }
    }
    for (int it = 0; ok; ++it) {
        // x += (L32 L32^T)^-1 r  (r = b at it = 0)
        blk_solve<float>(A32, n, n, Z, n, 1, nrhs, S.f);
        if (threadIdx.x < KB_SB_MAX_RHS) {
            S.rn[threadIdx.x] = 0ull;
            S.xn[threadIdx.x] = 0ull;
        }
        __syncthreads();
        sb_update_x(Xp, ldx, side_left, n, nrhs, Z, it == 0, S);
        __syncthreads();
        // r = b - A x, its FP32 copy in Z
        bool ovf = false;
        sb_residual(Ap, lda, n, Bp, ldb, Xp, ldx, side_left, nrhs, Z, S, ovf);
        const bool rovf = __syncthreads_or(ovf);
        bool conv = true;
        for (int c = 0; c < nrhs; ++c) conv = conv && (sb_val(S.rn[c]) <= sb_val(S.xn[c]) * cte);
        if (conv) {
            iters = it;
            break;
        }
        if (rovf || it == KB_SB_ITERMAX) {
            ok = false;
            iters = rovf ? -2 : -(KB_SB_ITERMAX + 1);
        }
        __syncthreads();
    }
    int inf = 0;
    if (!ok) {
        // FP64: X := B, A := L, solve
        __syncthreads();
        copy_rhs<double>(Bp, ldb, Xp, ldx, side_left, n, nrhs, threadIdx.x, KB_BLK_WG);
        __syncthreads();
        inf = blk_potrf<double>(Ap, lda, n, S.d, Xp, ldx, side_left, nrhs);
        if (inf == 0) blk_solve_bwd<double>(Ap, lda, n, Xp, ldx, side_left, nrhs, S.d);
    }
    if (threadIdx.x == 0) {
        if (info) info[mat] = inf;
        if (iter) iter[mat] = iters;
    }
}
