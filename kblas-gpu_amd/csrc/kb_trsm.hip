// kb_trsm.hip -- batched real TRSM (uniform batches), the shapes the
// reference's register kernels and TRSM recursion serve
// (src/batch_triangular/Xtrsm_batch_drivers.cuh:54-272,
// Xtrsm_batch_kernels.cuh:36-898): op(T) X = alpha B (side left) or
// X op(T) = alpha B (side right), T lower or upper, unit diagonal or not.
//
// Every case is one triangular solve per right-hand-side vector (a column of
// B for side left, a row for side right) with the lower view Lv of T
// (Lv = T, or T^T for an upper T): forward (Lv x = alpha b) or backward
// (Lv^T x = alpha b).
//
// Orders up to 16 (k_trsm_small): a wavefront works on P = 64 / G groups of
// G lanes; each group takes one matrix and G of its vectors, one vector per
// lane held in registers (K values, K = the order rounded up to 8 or 16),
// and stages the matrix's triangle in LDS, packed by columns with the
// reciprocal of each diagonal element in its place.  The substitution reads
// the triangle as LDS broadcasts within a group (the P groups' triangles
// start on different banks), so the arithmetic is K^2 / 2 FMAs per lane with
// no cross-lane traffic.  Small orders put many matrices in one wavefront
// (order 8: eight matrices of eight vectors), so the launch is not dominated
// by idle lanes.
//
// Orders 17..64 (k_trsm_mfma, below): 16 x 16 blocks on the matrix cores.
//
// Orders above 64 (lower T): one sweep of the blocked Cholesky solve
// (kb_blk_impl.h): diagonal 32 x 32 blocks by DPP-row substitution, the
// off-diagonal block products on the matrix cores.
#include "kb_device.h"
#include "kb_launch.h"

extern "C" int kb_blk_trsm_d(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                             kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s);
extern "C" int kb_blk_trsm_s(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                             kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s);

namespace {

// packed column j of a K-order lower triangle starts at ts_col(K, j)
constexpr int ts_col(int K, int j) { return j * K - j * (j - 1) / 2; }
// per-group LDS stride: odd, so the groups of a wavefront start on different banks
constexpr int ts_pk(int K) { return K * (K + 1) / 2 + 1; }

// most groups per wavefront: about 16 KB of LDS per wavefront
template <typename T, int K>
constexpr int ts_pmax()
{
    constexpr int p = 16384 / (ts_pk(K) * (int)sizeof(T));
    return p >= 64 ? 64 : p >= 32 ? 32 : p >= 16 ? 16 : p >= 8 ? 8 : p >= 4 ? 4 : p >= 2 ? 2 : 1;
}

constexpr int TS_U = 16;   // staging loads in flight per lane

template <typename T>
KB_DEV T ts_fma(T a, T b, T c)
{
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}

// One 64-lane workgroup per P work items (item = matrix * chunks + chunk of G
// vectors), grid-stride over the items.  k <= K is the order, nvec the number
// of vectors, lg = log2(G).
template <typename T, int K, bool FWD>
__global__ void __launch_bounds__(64)
k_trsm_small(int k, int nvec, int lg, int chunks, long items, int side_left, int upper, int unit, T alpha,
             BatchArg<T> A, BatchArg<T> B, const int* vm, const int* vn)
{
    extern __shared__ __align__(16) unsigned char ts_smem[];
    constexpr int PK = ts_pk(K);
    const int lane = threadIdx.x;
    const int G = 1 << lg, P = 64 >> lg;
    const int g = lane >> lg, li = lane & (G - 1);
    T* Ls = reinterpret_cast<T*>(ts_smem) + g * PK;
    const long nwg = (items + P - 1) / P;
    for (long wg = blockIdx.x; wg < nwg; wg += gridDim.x) {
        const long it = wg * P + g;
        const bool live = it < items;
        const long mat = live ? it / chunks : 0;
        const int vec = (int)(live ? it - mat * chunks : 0) * G + li;
        // this group's order and vector count (non-uniform: the matrix's own, at most K)
        int kg = k, nvg = nvec;
        if (vm && live) {
            const int mm = vm[mat], nn = vn[mat];
            kg = side_left ? mm : nn;
            nvg = side_left ? nn : mm;
            if (kg <= 0 || nvg <= 0) kg = nvg = 0;
        }
        const bool act = live && vec < nvg;
        // this lane's vector, loaded first so the loads overlap the staging below
        // both matrices' addresses first (in pointer-array mode they are loads, and a
        // wait for them would also wait for every load issued before)
        T* Bp = B.get(mat);
        const long ldb = B.ldim(mat);
        const T* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        T* xp = Bp + (act ? (side_left ? (long)vec * ldb : (long)vec) : 0);
        const long es = side_left ? 1 : ldb;
        // Unconditional loads from clamped addresses, so they are all in flight at once
        // (a select on each value lets the compiler sink every load into its own
        // branch).  Rows k..K-1 and idle lanes hold values that no stored row
        // depends on: the padded triangle is zero below row k and the sweeps never
        // read those rows into rows < k.
        // (alpha is applied after the staging, so these loads need not be waited for first)
        T x[K];
        if (alpha != T(0)) {
#pragma unroll
            for (int i = 0; i < K; ++i) {
                // the address is opaque to the optimizer: otherwise it sees that a clamped
                // load repeats the first one and moves each load under a branch again
                const T* q = xp + (act && i < kg ? i * es : 0);
                asm volatile("" : "+v"(q));
                x[i] = gld(q);
            }
        }
        if (live) {
            // the lower view, packed by columns, rows k..K-1 zero: the group's
            // lanes walk the K x K square (element e = li + G t, row e % K,
            // column e / K), TS_U loads in flight per lane
            const int cnt = (K * K) >> lg;
            for (int t0 = 0; t0 < cnt; t0 += TS_U) {
                T v[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const int e = li + ((t0 + u) << lg);
                    const int i = e % K, j = e / K;
                    const bool rd = t0 + u < cnt && i >= j && i < kg && !(unit && i == j);
                    const T x = gld(Ap + (rd ? (upper ? j + (long)i * lda : i + (long)j * lda) : 0));
                    v[u] = rd ? x : T(0);
                }
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const int e = li + ((t0 + u) << lg);
                    const int i = e % K, j = e / K;
                    if (t0 + u < cnt && i >= j && j < kg) Ls[ts_col(K, j) + i - j] = v[u];
                }
            }
        }
        __syncthreads();
        if (live) {
            // reciprocal diagonal in place (1 for a unit diagonal)
            for (int d = li; d < kg; d += G) Ls[ts_col(K, d)] = unit ? T(1) : T(1) / Ls[ts_col(K, d)];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < K; ++i) x[i] = alpha != T(0) ? alpha * x[i] : T(0);
        if constexpr (FWD) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j < kg) {
                    const T* c = Ls + ts_col(K, j);
                    x[j] *= c[0];
#pragma unroll
                    for (int i = j + 1; i < K; ++i) x[i] = ts_fma(-c[i - j], x[j], x[i]);
                }
            }
        } else {
#pragma unroll
            for (int j = K - 1; j >= 0; --j) {
                if (j < kg) {
                    x[j] *= Ls[ts_col(K, j)];
#pragma unroll
                    for (int i = 0; i < j; ++i) x[i] = ts_fma(-Ls[ts_col(K, i) + j - i], x[j], x[i]);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int i = 0; i < K; ++i)
                if (i < kg) gst(xp + i * es, x[i]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Orders 17..64 on the matrix cores.  The working triangle W (forward: W = Lv,
// lower; backward: W = Lv^T, upper) is staged once per workgroup as NBK x NBK
// blocks of 16 x 16 (the blocks of W's triangle only, each column-major), the
// order padded to 16 NBK with the identity; then each diagonal block is
// replaced by its inverse (one column per thread, by substitution).  Every
// wavefront takes 32 vectors (two column blocks of 16) and solves block row
// by block row in W's order:
//     R_i = alpha B_i - sum_j W_ij X_j   (j before i),   X_i = W_ii^-1 R_i,
// each product a chain of four 16x16x4 products with the W fragment from LDS
// as the first operand and the right-hand block as the second.  The
// accumulator layout of a 16 x 16 product (register e of lane (c, g) holds row
// kmap(e, g), column c) is exactly the second-operand layout of the next
// product's K step e, so R_i and X_j feed the products from registers.
// ---------------------------------------------------------------------------
// row of a 16 x 16 accumulator held in register e of lane group g (FP64 / FP32 layouts)
template <typename T>
KB_DEV int tm_kmap(int e, int g) { return sizeof(T) == 8 ? 4 * e + g : 4 * g + e; }

// LDS slot of W's block (i, j) (forward: j <= i; backward: i <= j)
KB_DEV int tm_slot(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

constexpr int TM_NCB = 2;   // column blocks of 16 vectors per wavefront

template <typename T, int NBK, bool FWD>
__global__ void __launch_bounds__(256)
k_trsm_mfma(int k0, int nvec0, int chunks, long items, int side_left, int upper, int unit, T alpha,
            BatchArgO<T> A, BatchArgO<T> B, const int* skip, const int* vm, const int* vn)
{
    typedef typename TileV<T>::v4 v4;
    constexpr int NSLOT = NBK * (NBK + 1) / 2;
    __shared__ T sW[NSLOT * 256];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    for (long it = blockIdx.x; it < items; it += gridDim.x) {
        const long mat = it / chunks;
        if (skip && skip[mat]) continue;   // (the whole workgroup: one matrix per item)
        // this matrix's order and vector count (non-uniform: its own, at most 16 NBK)
        int k = k0, nvec = nvec0;
        if (vm) {
            const int mm = vm[mat], nn = vn[mat];
            k = side_left ? mm : nn;
            nvec = side_left ? nn : mm;
            if (k <= 0 || nvec <= 0) continue;
        }
        const int nb = (k + 15) >> 4;   // row blocks holding the order
        const int chunk = (int)(it - mat * chunks);
        const T* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        // this wavefront's vectors: alpha B, in the accumulator layout, loaded
        // first so the loads overlap the staging and the inversion below
        const int v0 = (chunk * (nthr >> 6) + wave) * (16 * TM_NCB);
        T* Bp = B.get(mat);
        const long ldb = B.ldim(mat);
        auto bptr = [&](int r, int vec) {
            return Bp + (side_left ? r + (long)vec * ldb : vec + (long)r * ldb);
        };
        v4 X[NBK][TM_NCB];
#pragma unroll
        for (int i = 0; i < NBK; ++i) {
#pragma unroll
            for (int cb = 0; cb < TM_NCB; ++cb) {
                const int vec = v0 + 16 * cb + li;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * i + tm_kmap<T>(e, lg);
                    // unconditional load from a clamped address (B's first element otherwise), so
                    // all the loads are in flight together
                    const bool in = i < nb && r < k && vec < nvec;
                    const T v = gld(in ? bptr(r, vec) : Bp);
                    X[i][cb][e] = (in && alpha != T(0)) ? alpha * v : T(0);
                }
            }
        }
        // stage W's blocks; the loads of TS_U elements per thread in flight together
        for (int e0 = 0; e0 < NSLOT * 256; e0 += nthr * TS_U) {
            T v[TS_U];
#pragma unroll
            for (int u = 0; u < TS_U; ++u) {
                const int e = e0 + u * nthr + tid;
                const int slot = e >> 8, rr = e & 15, cc = (e >> 4) & 15;
                int bi = 0;
                while ((bi + 1) * (bi + 2) / 2 <= slot) ++bi;
                const int bj = slot - bi * (bi + 1) / 2;       // bj <= bi
                // forward: block (bi, bj) of the lower W; backward: block (bj, bi) of the upper W
                const int r = 16 * (FWD ? bi : bj) + rr, c = 16 * (FWD ? bj : bi) + cc;
                // W(r, c) = Lv(r, c) (forward) or Lv(c, r) (backward); Lv(p, q) (p >= q) = A(p, q) or A(q, p)
                const int p = FWD ? r : c, q = FWD ? c : r;
                const bool tri = FWD ? r >= c : r <= c;
                const bool rd = e < NSLOT * 256 && tri && p < k && !(unit && p == q);
                const T x = gld(Ap + (rd ? (upper ? q + (long)p * lda : p + (long)q * lda) : 0));
                v[u] = (r == c && (p >= k || unit)) ? T(1) : rd ? x : T(0);
            }
#pragma unroll
            for (int u = 0; u < TS_U; ++u) {
                const int e = e0 + u * nthr + tid;
                if (e < NSLOT * 256) sW[e] = v[u];
            }
        }
        __syncthreads();
        // invert the diagonal blocks: thread t < 16 NBK takes column t % 16 of block t / 16
        {
            T x[16];
            const int b = tid >> 4, c = tid & 15;
            const bool on = tid < 16 * NBK && b < nb;
            const T* D = sW + (on ? tm_slot(b, b) : 0) * 256;
#pragma unroll
            for (int r = 0; r < 16; ++r) x[r] = T(0);
            if (on) {
                if constexpr (FWD) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        T sacc = (r == c) ? T(1) : T(0);
#pragma unroll
                        for (int q = 0; q < r; ++q) sacc = ts_fma(-D[r + 16 * q], x[q], sacc);
                        x[r] = (r >= c) ? sacc / D[r + 16 * r] : T(0);
                    }
                } else {
#pragma unroll
                    for (int r = 15; r >= 0; --r) {
                        T sacc = (r == c) ? T(1) : T(0);
#pragma unroll
                        for (int q = r + 1; q < 16; ++q) sacc = ts_fma(-D[r + 16 * q], x[q], sacc);
                        x[r] = (r <= c) ? sacc / D[r + 16 * r] : T(0);
                    }
                }
            }
            __syncthreads();
            if (on) {
                T* Dw = sW + tm_slot(b, b) * 256;
#pragma unroll
                for (int r = 0; r < 16; ++r) Dw[r + 16 * c] = x[r];
            }
            __syncthreads();
        }
#pragma unroll
        for (int ii = 0; ii < NBK; ++ii) {
            const int i = FWD ? ii : NBK - 1 - ii;
            if (i < nb) {
                v4 acc[TM_NCB];
#pragma unroll
                for (int cb = 0; cb < TM_NCB; ++cb) acc[cb] = X[i][cb];
#pragma unroll
                for (int jj = 0; jj < NBK; ++jj) {
                    const int j = FWD ? jj : NBK - 1 - jj;
                    if ((FWD ? j < i : j > i) && j < nb) {
                        const T* Wb = sW + tm_slot(i, j) * 256;
#pragma unroll
                        for (int s2 = 0; s2 < 4; ++s2) {
                            const T w = -Wb[li + 16 * tm_kmap<T>(s2, lg)];
#pragma unroll
                            for (int cb = 0; cb < TM_NCB; ++cb) acc[cb] = mfma16x16x4(w, X[j][cb][s2], acc[cb]);
                        }
                    }
                }
                const T* Di = sW + tm_slot(i, i) * 256;
#pragma unroll
                for (int cb = 0; cb < TM_NCB; ++cb) X[i][cb] = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
                    const T w = Di[li + 16 * tm_kmap<T>(s2, lg)];
#pragma unroll
                    for (int cb = 0; cb < TM_NCB; ++cb) X[i][cb] = mfma16x16x4(w, acc[cb][s2], X[i][cb]);
                }
#pragma unroll
                for (int cb = 0; cb < TM_NCB; ++cb) {
                    const int vec = v0 + 16 * cb + li;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = 16 * i + tm_kmap<T>(e, lg);
                        if (r < k && vec < nvec) gst(bptr(r, vec), X[i][cb][e]);
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <typename T, int NBK>
int tm_launch(bool fwd, int k, int nvec, int side_left, int upper, int unit, T alpha, kb_mat_t A, kb_mat_t B,
              int batch, hipStream_t s, const int* skip = nullptr, const int* vm = nullptr, const int* vn = nullptr)
{
    const int per_wave = 16 * TM_NCB;
    const int waves = nvec >= 4 * per_wave ? 4 : (nvec + per_wave - 1) / per_wave;
    const int chunks = (nvec + waves * per_wave - 1) / (waves * per_wave);
    const long items = (long)batch * chunks;
    const unsigned grid = (unsigned)(items < (1L << 22) ? items : (1L << 22));
    BatchArgO<T> a = to_arg_off<T>(A), b = to_arg_off<T>(B);
    if (fwd)
        hipLaunchKernelGGL((k_trsm_mfma<T, NBK, true>), dim3(grid), dim3(64 * waves), 0, s, k, nvec, chunks, items,
                           side_left, upper, unit, alpha, a, b, skip, vm, vn);
    else
        hipLaunchKernelGGL((k_trsm_mfma<T, NBK, false>), dim3(grid), dim3(64 * waves), 0, s, k, nvec, chunks,
                           items, side_left, upper, unit, alpha, a, b, skip, vm, vn);
    return (int)hipGetLastError();
}

template <typename T, int K>
int ts_launch(bool fwd, int k, int nvec, int side_left, int upper, int unit, T alpha, kb_mat_t A, kb_mat_t B,
              int batch, const int* vm, const int* vn, hipStream_t s)
{
    constexpr int PM = ts_pmax<T, K>();
    int lg = 0;
    while ((1 << lg) < nvec && lg < 6) ++lg;     // G: the vector count rounded up to a power of two, <= 64
    while ((64 >> lg) > PM) ++lg;                // at most PM groups per wavefront
    const int G = 1 << lg, P = 64 >> lg;
    const int chunks = (nvec + G - 1) / G;
    const long items = (long)batch * chunks;
    const long nwg = (items + P - 1) / P;
    const unsigned grid = (unsigned)(nwg < (1L << 22) ? nwg : (1L << 22));
    const size_t lds = (size_t)P * ts_pk(K) * sizeof(T);
    BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    if (fwd)
        hipLaunchKernelGGL((k_trsm_small<T, K, true>), dim3(grid), dim3(64), lds, s, k, nvec, lg, chunks, items,
                           side_left, upper, unit, alpha, a, b, vm, vn);
    else
        hipLaunchKernelGGL((k_trsm_small<T, K, false>), dim3(grid), dim3(64), lds, s, k, nvec, lg, chunks, items,
                           side_left, upper, unit, alpha, a, b, vm, vn);
    return (int)hipGetLastError();
}

template <typename T>
int ts_dispatch(bool fwd, int k, int nvec, int side_left, int upper, int unit, T alpha, kb_mat_t A, kb_mat_t B,
                int batch, const int* vm, const int* vn, hipStream_t s)
{
    if (k <= 8) return ts_launch<T, 8>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, vm, vn, s);
    if (k <= 16) return ts_launch<T, 16>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, vm, vn, s);
    if (k <= 32) return tm_launch<T, 2>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s, nullptr, vm, vn);
    return tm_launch<T, 4>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s, nullptr, vm, vn);
}

// Lower triangle of order k > 64 with many vectors: split the order at k1,
// solve one part (alpha applied), update the other part of B with one batched
// GEMM on the matrix cores, solve the other part; the parts recurse until
// each leaf, of order 33..64 (the split balances the leaves), runs the
// matrix-core solve.  In terms of the vectors y of B (side_left: columns,
// else rows):
//   forward  L y = alpha b:    y1 = L11^-1 alpha b1;   b2 := alpha b2 - L21 y1;    y2 = L22^-1 b2
//   backward L^T y = alpha b:  y2 = L22^-T alpha b2;   b1 := alpha b1 - L21^T y2;  y1 = L11^-T b1
// with the update as B2 := alpha B2 - L21 Y1 (left) or B2 := alpha B2 - Y1 L21^T
// (right), and the backward forms B1 := alpha B1 - L21^T Y2 / - Y2 L21.  The
// reference runs the same recursion as separate TRSM and GEMM launches
// (src/batch_triangular/Xtrsm_batch_drivers.cuh:145-171).  skip: optional
// per-matrix flags, a matrix with skip[b] != 0 keeps its B.
template <typename T>
int tr_rec(bool fwd, int side_left, int unit, int k, int nvec, T alpha, kb_mat_t A, kb_mat_t B, int batch,
           const int* skip, hipStream_t s)
{
    if (k <= KB_REG_MAX_N) {
        if (k <= 32) return tm_launch<T, 2>(fwd, k, nvec, side_left, 0, unit, alpha, A, B, batch, s, skip);
        return tm_launch<T, 4>(fwd, k, nvec, side_left, 0, unit, alpha, A, B, batch, s, skip);
    }
    const int parts = (k + KB_REG_MAX_N - 1) / KB_REG_MAX_N;
    const int leaf = (k + parts - 1) / parts;
    const int k1 = (parts / 2) * leaf, k2 = k - k1;
    const long lda = A.ld, ldb = B.ld;
    kb_mat_t L11 = A, L21 = A, L22 = A, B1 = B, B2 = B;
    L21.off += k1;
    L22.off += k1 + (long)k1 * lda;
    B2.off += side_left ? (long)k1 : (long)k1 * ldb;
    const char prec = sizeof(T) == 8 ? 'd' : 's';
    int e;
    if (fwd) {
        if ((e = tr_rec<T>(true, side_left, unit, k1, nvec, alpha, L11, B1, batch, skip, s))) return e;
        e = side_left ? kb_launch_gemm_skip(prec, 0, 0, 0, 0, k2, nvec, k1, -1.0, L21, B1, (double)alpha, B2, nullptr,
                                            nullptr, nullptr, skip, batch, s)
                      : kb_launch_gemm_skip(prec, 0, 0, 0, 1, nvec, k2, k1, -1.0, B1, L21, (double)alpha, B2, nullptr,
                                            nullptr, nullptr, skip, batch, s);
        if (e) return e;
        return tr_rec<T>(true, side_left, unit, k2, nvec, T(1), L22, B2, batch, skip, s);
    }
    if ((e = tr_rec<T>(false, side_left, unit, k2, nvec, alpha, L22, B2, batch, skip, s))) return e;
    e = side_left ? kb_launch_gemm_skip(prec, 0, 0, 1, 0, k1, nvec, k2, -1.0, L21, B2, (double)alpha, B1, nullptr,
                                        nullptr, nullptr, skip, batch, s)
                  : kb_launch_gemm_skip(prec, 0, 0, 0, 0, nvec, k1, k2, -1.0, B2, L21, (double)alpha, B1, nullptr,
                                        nullptr, nullptr, skip, batch, s);
    if (e) return e;
    return tr_rec<T>(false, side_left, unit, k1, nvec, T(1), L11, B1, batch, skip, s);
}

}  // namespace

// Real uniform TRSM ('d' or 's'; trans 0 = N, 1 or 2 = T).  B is m x n.
// hipErrorNotSupported for what these kernels do not serve (an upper
// triangle of order above 64): the caller then runs the generic kernel.
extern "C" int kb_launch_trsm_fast(char prec, int side_left, int upper, int trans, int unit, int m, int n,
                                   double alpha, kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch,
                                   hipStream_t s)
{
    if (batch < 1 || m < 1 || n < 1) return 0;
    const int k = side_left ? m : n, nvec = side_left ? n : m;
    // U = Lv^T: an upper triangle flips the transposition
    const bool et = upper ? (trans == 0) : (trans != 0);
    const bool fwd = side_left ? !et : et;
    if (k > KB_REG_MAX_N) {
        if (upper) return (int)hipErrorNotSupported;
        if (!vm && !vn && !A.lds && !B.lds && nvec >= KB_TRSM_REC_NVEC && alpha != 0.0) {
            if (prec == 'd') return tr_rec<double>(fwd, side_left, unit, k, nvec, alpha, A, B, batch, nullptr, s);
            if (prec == 's')
                return tr_rec<float>(fwd, side_left, unit, k, nvec, (float)alpha, A, B, batch, nullptr, s);
        }
        if (prec == 'd') return kb_blk_trsm_d(fwd, side_left, unit, k, nvec, alpha, A, B, vm, vn, batch, s);
        if (prec == 's') return kb_blk_trsm_s(fwd, side_left, unit, k, nvec, alpha, A, B, vm, vn, batch, s);
        return (int)hipErrorInvalidValue;
    }
    if (prec == 'd') return ts_dispatch<double>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, vm, vn, s);
    if (prec == 's')
        return ts_dispatch<float>(fwd, k, nvec, side_left, upper, unit, (float)alpha, A, B, batch, vm, vn, s);
    return (int)hipErrorInvalidValue;
}

namespace {
template <typename T>
int tm_solve(int side_left, int n, int nrhs, kb_mat_t A, kb_mat_t B, const int* skip, int batch, hipStream_t s)
{
    int e;
    if (n <= 32) {
        e = tm_launch<T, 2>(true, n, nrhs, side_left, 0, 0, T(1), A, B, batch, s, skip);
        return e ? e : tm_launch<T, 2>(false, n, nrhs, side_left, 0, 0, T(1), A, B, batch, s, skip);
    }
    e = tm_launch<T, 4>(true, n, nrhs, side_left, 0, 0, T(1), A, B, batch, s, skip);
    return e ? e : tm_launch<T, 4>(false, n, nrhs, side_left, 0, 0, T(1), A, B, batch, s, skip);
}
}  // namespace

// (L L^T) X = B with the lower factor L of order 16 < n <= 64 ('d' or 's'):
// the forward and the backward sweep of the matrix-core TRSM; matrices with
// skip[b] != 0 (a failed factorization) are left as they are.
// (L L^T) X = B with the lower factor L of order n > 64 and at least
// KB_TRSM_REC_NVEC right-hand sides ('d' or 's'): the forward and the backward
// solve by the recursion above; matrices with skip[b] != 0 keep their B.
extern "C" int kb_launch_trsm_rec_solve(char prec, int side_left, int n, int nrhs, kb_mat_t A, kb_mat_t B,
                                        const int* skip, int batch, hipStream_t s)
{
    if (batch < 1 || nrhs < 1) return 0;
    if (n <= KB_REG_MAX_N || A.lds || B.lds) return (int)hipErrorInvalidValue;
    int e;
    if (prec == 'd') {
        if ((e = tr_rec<double>(true, side_left, 0, n, nrhs, 1.0, A, B, batch, skip, s))) return e;
        return tr_rec<double>(false, side_left, 0, n, nrhs, 1.0, A, B, batch, skip, s);
    }
    if (prec == 's') {
        if ((e = tr_rec<float>(true, side_left, 0, n, nrhs, 1.0f, A, B, batch, skip, s))) return e;
        return tr_rec<float>(false, side_left, 0, n, nrhs, 1.0f, A, B, batch, skip, s);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_trsm_solve(char prec, int side_left, int n, int nrhs, kb_mat_t A, kb_mat_t B,
                                    const int* skip, int batch, hipStream_t s)
{
    if (batch < 1 || nrhs < 1) return 0;
    if (n <= 16 || n > KB_REG_MAX_N) return (int)hipErrorInvalidValue;
    if (prec == 'd') return tm_solve<double>(side_left, n, nrhs, A, B, skip, batch, s);
    if (prec == 's') return tm_solve<float>(side_left, n, nrhs, A, B, skip, batch, s);
    return (int)hipErrorInvalidValue;
}
