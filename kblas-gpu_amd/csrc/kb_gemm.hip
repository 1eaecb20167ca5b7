// kb_gemm.hip -- batched real GEMM and SYRK on the matrix cores (uniform
// batches), the products the reference delegates to cuBLAS
// (src/batch_triangular/Xgemm_batch_core.cuh:170-313,492-634) and its SYRK
// recursion (Xsyrk_batch_drivers.cuh:32-119,125-433):
//   GEMM  C := alpha op(A) op(B) + beta C
//   SYRK  the lower (or upper) triangle of C := alpha op(A) op(A)^T + beta C
//
// GEMM with m, n >= 64: 64 x 64 workgroup tiles through LDS (k_gemm_lds below).
// Otherwise, and for SYRK: one wavefront per TM x TM output tile (TM = 16 or 32), the tile built from
// 16 x 16 x 4 matrix-core products (FP64 or FP32).  The product is formed
// transposed, D = op(B)^T op(A)^T, so that the accumulator's lanes run down
// the columns of C and the read-modify-write of C is coalesced.  Operand
// fragments are loaded straight from memory, 4 or 8 K steps deep, with the next
// chunk's loads issued before the current chunk's products.
//
// Launch order: workgroup w goes to XCD w mod 8, so the tiles of one matrix
// are numbered 8 apart (matrix b on XCD b mod 8) and dispatched together:
// every tile of a matrix reads its A and B panels through the same L2 while
// they are still there.
#include "kb_device.h"
#include "kb_launch.h"

namespace {

// in ? *p : 0, with the load issued unconditionally (p is clamped to a valid
// address by the caller) and its address opaque to the optimizer, which would
// otherwise see that the clamped loads repeat one another and put each load
// under a branch of its own (one wait per element)
template <typename T>
KB_DEV T gm_masked(bool in, const T* p)
{
    asm volatile("" : "+v"(p));
    const T v = gld(p);
    return in ? v : T(0);
}

// the load alone (the caller applies the mask where the value is first used)
template <typename T>
KB_DEV T gm_raw(const T* p)
{
    asm volatile("" : "+v"(p));
    return gld(p);
}

// K steps of 4 per chunk: two chunks of fragments in flight fit in about 128
// registers (four waves per SIMD) at 8 steps, except FP64 32 x 32 tiles (4)
template <typename T, int TM>
constexpr int gm_ku() { return (TM == 32 && sizeof(T) == 8) ? 4 : 8; }

template <typename T, int NB, int KU, bool TA, bool TB>
KB_DEV void gm_load(const T* Ap, int lda, const T* Bp, int ldb, int m, int n, int k, int I0, int J0, int p0, int li,
                    int lg, T (&a)[KU][NB], T (&b)[KU][NB])
{
#pragma unroll
    for (int s = 0; s < KU; ++s) {
        const int kk = p0 + 4 * s + lg;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int r = I0 + 16 * q + li;   // row of op(A)
            const bool ia = r < m && kk < k;
            a[s][q] = gm_masked(ia, Ap + (ia ? (TA ? kk + (long)r * lda : r + (long)kk * lda) : 0));
            const int c = J0 + 16 * q + li;   // column of op(B)
            const bool ib = c < n && kk < k;
            b[s][q] = gm_masked(ib, Bp + (ib ? (TB ? c + (long)kk * ldb : kk + (long)c * ldb) : 0));
        }
    }
}

template <typename T, int TM, bool TA, bool TB>
__global__ void __launch_bounds__(64)
k_gemm_mfma(int m, int n, int k, T alpha, BatchArgO<T> A, BatchArgO<T> B, T beta, BatchArgO<T> C, int batch,
            int tiles, int tm, int syrk, int upper, const int* vm, const int* vn, const int* vk, const int* skip)
{
    constexpr int NB = TM / 16, KU = gm_ku<T, TM>();
    typedef typename TileV<T>::v4 v4;
    const int lane = threadIdx.x, li = lane & 15, lg = lane >> 4;
    // workgroup -> (matrix, tile): matrices in groups of 8, tile-major inside a group
    const long w = blockIdx.x;
    const long grp = w / (8L * tiles);
    const int r8 = (int)(w - grp * 8L * tiles);
    const int t = r8 >> 3;
    const long mat = grp * 8 + (r8 & 7);
    if (mat >= batch || (skip && skip[mat])) return;   // skip: C left as it is
    int I, J;
    if (syrk) {
        // lower tiles (I, J <= I) in row order; upper: the transposed tile
        I = 0;
        while ((I + 1) * (I + 2) / 2 <= t) ++I;
        J = t - I * (I + 1) / 2;
        if (upper) {
            const int u = I;
            I = J;
            J = u;
        }
    } else {
        I = t % tm;
        J = t / tm;
    }
    const int I0 = I * TM, J0 = J * TM;
    if (vm) {
        // non-uniform: this matrix's sizes (the grid covers the largest); tiles outside it have no work
        m = vm[mat];
        n = syrk ? m : vn[mat];
        k = vk[mat];
        if (m <= 0 || n <= 0 || I0 >= m || J0 >= n) return;
    }
    const T* Ap = A.get(mat);
    const T* Bp = B.get(mat);
    T* Cp = C.get(mat);
    const int lda = A.ldim(mat), ldb = B.ldim(mat), ldc = C.ldim(mat);
    v4 acc[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) acc[x][y] = v4{T(0), T(0), T(0), T(0)};
    if (alpha != T(0) && k > 0) {
        T a[KU][NB], b[KU][NB], an[KU][NB], bn[KU][NB];
        gm_load<T, NB, KU, TA, TB>(Ap, lda, Bp, ldb, m, n, k, I0, J0, 0, li, lg, a, b);
        for (int p0 = 0; p0 < k; p0 += 4 * KU) {
            const bool more = p0 + 4 * KU < k;
            if (more) gm_load<T, NB, KU, TA, TB>(Ap, lda, Bp, ldb, m, n, k, I0, J0, p0 + 4 * KU, li, lg, an, bn);
#pragma unroll
            for (int s = 0; s < KU; ++s)
#pragma unroll
                for (int x = 0; x < NB; ++x)        // x: block of C rows (op(A) rows)
#pragma unroll
                    for (int y = 0; y < NB; ++y)    // y: block of C columns
                        acc[x][y] = mfma16x16x4(b[s][y], a[s][x], acc[x][y]);
            __builtin_amdgcn_sched_barrier(0);
            if (more) {
#pragma unroll
                for (int s = 0; s < KU; ++s)
#pragma unroll
                    for (int q = 0; q < NB; ++q) {
                        a[s][q] = an[s][q];
                        b[s][q] = bn[s][q];
                    }
            }
        }
    }
    // acc[x][y] lane (li, lg) register e: C(I0 + 16x + li, J0 + 16y + col(e, lg))
#pragma unroll
    for (int x = 0; x < NB; ++x) {
        const int r = I0 + 16 * x + li;
#pragma unroll
        for (int y = 0; y < NB; ++y) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = J0 + 16 * y + ((sizeof(T) == 8) ? 4 * e + lg : 4 * lg + e);
                const bool in = r < m && c < n && (!syrk || (upper ? c >= r : c <= r));
                if (in) {
                    T* cp = Cp + r + (long)c * ldc;
                    T v = alpha * acc[x][y][e];
                    if (beta != T(0)) v = beta * gld(cp) + v;
                    gst(cp, v);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Larger products (m, n >= 64): one 256-thread workgroup per 64 x 64 tile of
// C, each wavefront a 32 x 32 quarter.  The operands go through LDS in K
// chunks of 16: every thread loads 4 elements of each operand per chunk with
// the matrix's contiguous index running across the threads (coalesced for
// every transposition), stores them in the fragment-friendly layout
// s[k][row] (rows contiguous), and the next chunk's loads are in flight while
// the current chunk's products run (two LDS buffers, one barrier per chunk).
// ---------------------------------------------------------------------------
constexpr int GL_TM = 64, GL_KC = 16, GL_LD = GL_TM + 2;
// smallest m and n for the workgroup tile (smaller products, and SYRK, keep the wavefront tiles)
constexpr int GM_LDS_MIN = 64;

template <typename T, bool TA, bool TB>
KB_DEV void gl_load(const T* Ap, int lda, const T* Bp, int ldb, int m, int n, int k, int I0, int J0, int p0, int tid,
                    T (&ra)[4], T (&rb)[4])
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u;
        // op(A)(r, kk): A[r + kk lda] (rows contiguous) or A[kk + r lda] (k contiguous)
        const int ra_r = TA ? e / GL_KC : e % GL_TM, ra_k = TA ? e % GL_KC : e / GL_TM;
        const int r = I0 + ra_r, kk = p0 + ra_k;
        const bool ia = r < m && kk < k;
        ra[u] = gm_raw(Ap + (ia ? (TA ? kk + (long)r * lda : r + (long)kk * lda) : 0));
        // op(B)(kk, c): B[kk + c ldb] (k contiguous) or B[c + kk ldb] (columns contiguous)
        const int rb_c = TB ? e % GL_TM : e / GL_KC, rb_k = TB ? e / GL_TM : e % GL_KC;
        const int c = J0 + rb_c, kb = p0 + rb_k;
        const bool ib = c < n && kb < k;
        rb[u] = gm_raw(Bp + (ib ? (TB ? c + (long)kb * ldb : kb + (long)c * ldb) : 0));
    }
}

// the loaded values (raw, from clamped addresses) with the masks of gl_load applied
template <typename T, bool TA, bool TB>
KB_DEV void gl_store(T (*sA)[GL_LD], T (*sB)[GL_LD], int m, int n, int k, int I0, int J0, int p0, int tid,
                     const T (&ra)[4], const T (&rb)[4])
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u;
        const int ra_r = TA ? e / GL_KC : e % GL_TM, ra_k = TA ? e % GL_KC : e / GL_TM;
        sA[ra_k][ra_r] = (I0 + ra_r < m && p0 + ra_k < k) ? ra[u] : T(0);
        const int rb_c = TB ? e % GL_TM : e / GL_KC, rb_k = TB ? e / GL_TM : e % GL_KC;
        sB[rb_k][rb_c] = (J0 + rb_c < n && p0 + rb_k < k) ? rb[u] : T(0);
    }
}

template <typename T, bool TA, bool TB>
__global__ void __launch_bounds__(256)
k_gemm_lds(int m, int n, int k, T alpha, BatchArgO<T> A, BatchArgO<T> B, T beta, BatchArgO<T> C, int batch,
           int tiles, int tm, int syrk, int upper, const int* vm, const int* vn, const int* vk, const int* skip)
{
    typedef typename TileV<T>::v4 v4;
    __shared__ T sA[2][GL_KC][GL_LD];
    __shared__ T sB[2][GL_KC][GL_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const long w = blockIdx.x;
    const long grp = w / (8L * tiles);
    const int r8 = (int)(w - grp * 8L * tiles);
    const int t = r8 >> 3;
    const long mat = grp * 8 + (r8 & 7);
    if (mat >= batch || (skip && skip[mat])) return;   // skip: C left as it is
    int I, J;
    if (syrk) {
        I = 0;
        while ((I + 1) * (I + 2) / 2 <= t) ++I;
        J = t - I * (I + 1) / 2;
        if (upper) {
            const int u = I;
            I = J;
            J = u;
        }
    } else {
        I = t % tm;
        J = t / tm;
    }
    const int I0 = I * GL_TM, J0 = J * GL_TM;
    if (vm) {   // non-uniform, as in k_gemm_mfma (the whole workgroup leaves together)
        m = vm[mat];
        n = syrk ? m : vn[mat];
        k = vk[mat];
        if (m <= 0 || n <= 0 || I0 >= m || J0 >= n) return;
    }
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);   // this wavefront's quarter
    const T* Ap = A.get(mat);
    const T* Bp = B.get(mat);
    T* Cp = C.get(mat);
    const int lda = A.ldim(mat), ldb = B.ldim(mat), ldc = C.ldim(mat);
    v4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{T(0), T(0), T(0), T(0)};
    if (alpha != T(0) && k > 0) {
        T ra[4], rb[4];
        gl_load<T, TA, TB>(Ap, lda, Bp, ldb, m, n, k, I0, J0, 0, tid, ra, rb);
        gl_store<T, TA, TB>(sA[0], sB[0], m, n, k, I0, J0, 0, tid, ra, rb);
        __syncthreads();
        const int nq = (k + GL_KC - 1) / GL_KC;
        for (int q = 0; q < nq; ++q) {
            const int cur = q & 1;
            const bool more = q + 1 < nq;
            if (more) gl_load<T, TA, TB>(Ap, lda, Bp, ldb, m, n, k, I0, J0, (q + 1) * GL_KC, tid, ra, rb);
#pragma unroll
            for (int s2 = 0; s2 < GL_KC / 4; ++s2) {
                const int kk = 4 * s2 + lg;
                T fa[2], fb[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) fa[x] = sA[cur][kk][wr + 16 * x + li];
#pragma unroll
                for (int y = 0; y < 2; ++y) fb[y] = sB[cur][kk][wc + 16 * y + li];
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[x][y] = mfma16x16x4(fb[y], fa[x], acc[x][y]);
            }
            // the next chunk's values are used only after this chunk's products
            __builtin_amdgcn_sched_barrier(0);
            if (more) gl_store<T, TA, TB>(sA[cur ^ 1], sB[cur ^ 1], m, n, k, I0, J0, (q + 1) * GL_KC, tid, ra, rb);
            __syncthreads();
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int r = I0 + wr + 16 * x + li;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = J0 + wc + 16 * y + ((sizeof(T) == 8) ? 4 * e + lg : 4 * lg + e);
                const bool in = r < m && c < n && (!syrk || (upper ? c >= r : c <= r));
                if (in) {
                    T* cp = Cp + r + (long)c * ldc;
                    T v = alpha * acc[x][y][e];
                    if (beta != T(0)) v = beta * gld(cp) + v;
                    gst(cp, v);
                }
            }
        }
    }
}

template <typename T, bool TA, bool TB>
int gl_launch(int m, int n, int k, T alpha, kb_mat_t A, kb_mat_t B, T beta, kb_mat_t C, int batch, int syrk,
              int upper, const int* vm, const int* vn, const int* vk, const int* skip, hipStream_t s)
{
    const int tm = (m + GL_TM - 1) / GL_TM, tn = (n + GL_TM - 1) / GL_TM;
    const int tiles = syrk ? tm * (tm + 1) / 2 : tm * tn;
    const long wgs = (batch + 7) / 8 * 8L * tiles;
    if (wgs >= (1L << 24)) return (int)hipErrorNotSupported;   // grid of at most 2^32 work-items
    hipLaunchKernelGGL((k_gemm_lds<T, TA, TB>), dim3((unsigned)wgs), dim3(256), 0, s, m, n, k, alpha, to_arg_off<T>(A),
                       to_arg_off<T>(B), beta, to_arg_off<T>(C), batch, tiles, tm, syrk, upper, vm, vn, vk, skip);
    return (int)hipGetLastError();
}

template <typename T, int TM, bool TA, bool TB>
int gm_launch(int m, int n, int k, T alpha, kb_mat_t A, kb_mat_t B, T beta, kb_mat_t C, int batch, int syrk,
              int upper, const int* vm, const int* vn, const int* vk, const int* skip, hipStream_t s)
{
    const int tm = (m + TM - 1) / TM, tn = (n + TM - 1) / TM;
    const int tiles = syrk ? tm * (tm + 1) / 2 : tm * tn;
    const long groups = (batch + 7) / 8;
    const long wgs = groups * 8L * tiles;
    if (wgs >= (1L << 26)) return (int)hipErrorNotSupported;   // grid of at most 2^32 work-items
    hipLaunchKernelGGL((k_gemm_mfma<T, TM, TA, TB>), dim3((unsigned)wgs), dim3(64), 0, s, m, n, k, alpha,
                       to_arg_off<T>(A), to_arg_off<T>(B), beta, to_arg_off<T>(C), batch, tiles, tm, syrk, upper, vm, vn, vk,
                       skip);
    return (int)hipGetLastError();
}

template <typename T, int TM>
int gm_trans(int ta, int tb, int m, int n, int k, T alpha, kb_mat_t A, kb_mat_t B, T beta, kb_mat_t C, int batch,
             int syrk, int upper, const int* vm, const int* vn, const int* vk, const int* skip, hipStream_t s)
{
    if constexpr (TM == GL_TM) {
        if (!ta && !tb) return gl_launch<T, false, false>(m, n, k, alpha, A, B, beta, C,
            batch, syrk, upper, vm, vn, vk, skip, s);
        if (!ta && tb) return gl_launch<T, false, true>(m, n, k, alpha, A, B, beta, C,
            batch, syrk, upper, vm, vn, vk, skip, s);
        if (ta && !tb) return gl_launch<T, true, false>(m, n, k, alpha, A, B, beta, C,
            batch, syrk, upper, vm, vn, vk, skip, s);
        return gl_launch<T, true, true>(m, n, k, alpha, A, B, beta, C, batch, syrk, upper, vm, vn, vk, skip, s);
    }
    if (!ta && !tb) return gm_launch<T, TM, false, false>(m, n, k, alpha, A, B, beta, C,
        batch, syrk, upper, vm, vn, vk, skip, s);
    if (!ta && tb) return gm_launch<T, TM, false, true>(m, n, k, alpha, A, B, beta, C,
        batch, syrk, upper, vm, vn, vk, skip, s);
    if (ta && !tb) return gm_launch<T, TM, true, false>(m, n, k, alpha, A, B, beta, C,
        batch, syrk, upper, vm, vn, vk, skip, s);
    return gm_launch<T, TM, true, true>(m, n, k, alpha, A, B, beta, C, batch, syrk, upper, vm, vn, vk, skip, s);
}

template <typename T>
int gm_dispatch(int ta, int tb, int m, int n, int k, T alpha, kb_mat_t A, kb_mat_t B, T beta, kb_mat_t C,
                int batch, int syrk, int upper, const int* vm, const int* vn, const int* vk, const int* skip,
                hipStream_t s)
{
    if (m <= 16 || n <= 16) return gm_trans<T, 16>(ta, tb, m, n, k, alpha, A, B, beta, C,
        batch, syrk, upper, vm, vn, vk, skip, s);
    // SYRK keeps the wavefront tiles: they skip the quarters above the diagonal,
    // which a 64 x 64 workgroup tile computes and discards (measured slower)
    if (syrk || m < GM_LDS_MIN || n < GM_LDS_MIN)
        return gm_trans<T, 32>(ta, tb, m, n, k, alpha, A, B, beta, C, batch, syrk, upper, vm, vn, vk, skip, s);
    return gm_trans<T, GL_TM>(ta, tb, m, n, k, alpha, A, B, beta, C, batch, syrk, upper, vm, vn, vk, skip, s);
}

}  // namespace

// Real GEMM / SYRK on the matrix cores ('d' or 's'; op 0 = N, else T).
// SYRK (syrk != 0): C is m x m, B must be A and tb = !ta; upper selects the
// triangle.  Non-uniform batches pass the per-matrix sizes in vm / vn / vk
// (device arrays) and their maxima in m / n / k, which must bound them.
// hipErrorNotSupported when the grid would be too large (the caller then runs
// the general kernel).
// (skip: optional per-matrix flags, a matrix with skip[b] != 0 keeps its C.)
extern "C" int kb_launch_gemm_skip(char prec, int syrk, int upper, int ta, int tb, int m, int n, int k,
                                   double alpha, kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C, const int* vm,
                                   const int* vn, const int* vk, const int* skip, int batch, hipStream_t s)
{
    if (batch < 1 || m < 1 || n < 1) return 0;
    ta = ta != 0;
    tb = tb != 0;
    if (prec == 'd')
        return gm_dispatch<double>(ta, tb, m, n, k, alpha, A, B, beta, C, batch, syrk, upper, vm, vn, vk, skip, s);
    if (prec == 's')
        return gm_dispatch<float>(ta, tb, m, n, k, (float)alpha, A, B, (float)beta, C, batch, syrk, upper, vm, vn,
                                  vk, skip, s);
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_gemm_fast(char prec, int syrk, int upper, int ta, int tb, int m, int n, int k,
                                   double alpha, kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C, const int* vm,
                                   const int* vn, const int* vk, int batch, hipStream_t s)
{
    return kb_launch_gemm_skip(prec, syrk, upper, ta, tb, m, n, k, alpha, A, B, beta, C, vm, vn, vk, nullptr, batch,
                               s);
}
