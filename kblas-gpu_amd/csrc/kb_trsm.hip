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
// Orders up to 64 (k_trsm_small): a wavefront works on P = 64 / G groups of
// G lanes; each group takes one matrix and G of its vectors, one vector per
// lane held in registers (K values, K = the order rounded up to a power of
// two), and stages the matrix's triangle in LDS, packed by columns with the
// reciprocal of each diagonal element in its place.  The substitution reads
// the triangle as LDS broadcasts within a group (the P groups' triangles
// start on different banks), so the arithmetic is K^2 / 2 FMAs per lane with
// no cross-lane traffic.  Small orders put many matrices in one wavefront
// (order 8: eight matrices of eight vectors), so the launch is not dominated
// by idle lanes.
//
// Orders above 64 (lower T): one sweep of the blocked Cholesky solve
// (kb_blk_impl.h): diagonal 32 x 32 blocks by DPP-row substitution, the
// off-diagonal block products on the matrix cores.
#include "kb_device.h"
#include "kb_launch.h"

extern "C" int kb_blk_trsm_d(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                             kb_mat_t B, int batch, hipStream_t s);
extern "C" int kb_blk_trsm_s(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                             kb_mat_t B, int batch, hipStream_t s);

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
             BatchArg<T> A, BatchArg<T> B)
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
        const bool act = live && vec < nvec;
        if (live) {
            // the lower view, packed by columns, rows k..K-1 zero: the group's
            // lanes walk the K x K square (element e = li + G t, row e % K,
            // column e / K), TS_U loads in flight per lane
            const T* Ap = A.get(mat);
            const int lda = A.ldim(mat);
            const int cnt = (K * K) >> lg;
            for (int t0 = 0; t0 < cnt; t0 += TS_U) {
                T v[TS_U];
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const int e = li + ((t0 + u) << lg);
                    const int i = e % K, j = e / K;
                    const bool rd = t0 + u < cnt && i >= j && i < k && !(unit && i == j);
                    v[u] = rd ? gld(Ap + (upper ? j + (long)i * lda : i + (long)j * lda)) : T(0);
                }
#pragma unroll
                for (int u = 0; u < TS_U; ++u) {
                    const int e = li + ((t0 + u) << lg);
                    const int i = e % K, j = e / K;
                    if (t0 + u < cnt && i >= j && j < k) Ls[ts_col(K, j) + i - j] = v[u];
                }
            }
        }
        __syncthreads();
        if (live) {
            // reciprocal diagonal in place (1 for a unit diagonal)
            for (int d = li; d < k; d += G) Ls[ts_col(K, d)] = unit ? T(1) : T(1) / Ls[ts_col(K, d)];
        }
        __syncthreads();
        T* Bp = B.get(mat);
        const long ldb = B.ldim(mat);
        T* xp = Bp + (act ? (side_left ? (long)vec * ldb : (long)vec) : 0);
        const long es = side_left ? 1 : ldb;
        T x[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            T v = T(0);
            if (act && i < k) v = alpha * gld(xp + i * es);
            x[i] = v;
        }
        if constexpr (FWD) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if (j < k) {
                    const T* c = Ls + ts_col(K, j);
                    x[j] *= c[0];
#pragma unroll
                    for (int i = j + 1; i < K; ++i) x[i] = ts_fma(-c[i - j], x[j], x[i]);
                }
            }
        } else {
#pragma unroll
            for (int j = K - 1; j >= 0; --j) {
                if (j < k) {
                    x[j] *= Ls[ts_col(K, j)];
#pragma unroll
                    for (int i = 0; i < j; ++i) x[i] = ts_fma(-Ls[ts_col(K, i) + j - i], x[j], x[i]);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int i = 0; i < K; ++i)
                if (i < k) gst(xp + i * es, x[i]);
        }
        __syncthreads();
    }
}

template <typename T, int K>
int ts_launch(bool fwd, int k, int nvec, int side_left, int upper, int unit, T alpha, kb_mat_t A, kb_mat_t B,
              int batch, hipStream_t s)
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
                           side_left, upper, unit, alpha, a, b);
    else
        hipLaunchKernelGGL((k_trsm_small<T, K, false>), dim3(grid), dim3(64), lds, s, k, nvec, lg, chunks, items,
                           side_left, upper, unit, alpha, a, b);
    return (int)hipGetLastError();
}

template <typename T>
int ts_dispatch(bool fwd, int k, int nvec, int side_left, int upper, int unit, T alpha, kb_mat_t A, kb_mat_t B,
                int batch, hipStream_t s)
{
    if (k <= 8) return ts_launch<T, 8>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s);
    if (k <= 16) return ts_launch<T, 16>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s);
    if (k <= 32) return ts_launch<T, 32>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s);
    return ts_launch<T, 64>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s);
}

}  // namespace

// Real uniform TRSM ('d' or 's'; trans 0 = N, 1 or 2 = T).  B is m x n.
// hipErrorNotSupported for what these kernels do not serve (an upper
// triangle of order above 64): the caller then runs the generic kernel.
extern "C" int kb_launch_trsm_fast(char prec, int side_left, int upper, int trans, int unit, int m, int n,
                                   double alpha, kb_mat_t A, kb_mat_t B, int batch, hipStream_t s)
{
    if (batch < 1 || m < 1 || n < 1) return 0;
    const int k = side_left ? m : n, nvec = side_left ? n : m;
    // U = Lv^T: an upper triangle flips the transposition
    const bool et = upper ? (trans == 0) : (trans != 0);
    const bool fwd = side_left ? !et : et;
    if (k > KB_REG_MAX_N) {
        if (upper) return (int)hipErrorNotSupported;
        if (prec == 'd') return kb_blk_trsm_d(fwd, side_left, unit, k, nvec, alpha, A, B, batch, s);
        if (prec == 's') return kb_blk_trsm_s(fwd, side_left, unit, k, nvec, alpha, A, B, batch, s);
        return (int)hipErrorInvalidValue;
    }
    if (prec == 'd') return ts_dispatch<double>(fwd, k, nvec, side_left, upper, unit, alpha, A, B, batch, s);
    if (prec == 's') return ts_dispatch<float>(fwd, k, nvec, side_left, upper, unit, (float)alpha, A, B, batch, s);
    return (int)hipErrorInvalidValue;
}
