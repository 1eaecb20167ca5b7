// kb_pstrf.hip -- batched pivoted Cholesky of symmetric positive semi-definite
// matrices, P^T A P = R^T R with R upper triangular, n <= 64
// (src/batch_svd/batch_pstrf.cu:73-190, include/batch_pstrf.h:20-92).
//
// One wave per matrix (two for n <= 32, below).  The upper triangle is read, mirrored into the lower
// one in LDS (so each step can update whole rows), and factored right-looking
// with diagonal pivoting: at step k the largest remaining diagonal entry of
// the Schur complement is swapped to position k (rows and columns, the
// factored rows included), R(k,k) = sqrt, row k of R = row k / R(k,k), and the
// trailing block loses the outer product of row k.  The factorization stops
// when the largest remaining diagonal entry is <= n * eps * max(diag(A))
// (the detected rank).  On exit the upper n x n triangle holds R (the trailing
// block past the rank holds its unfactored Schur complement), the strict lower
// triangle is zero, piv holds the 0-based pivot order and ranks the rank.
//
// The pivot search keeps the reference's reduction so ties resolve the same
// way: lanes 0..31 hold candidate (value, index) pairs (entries past n or
// before k count as -1; for n > 32 lane t also takes entry t + 32 when
// strictly larger), then a butterfly over masks 16, 8, 4, 2, 1 in which a lane
// takes its partner's pair only when its value is strictly larger; lane 0's
// pair is the pivot.
#include "kb_device.h"
#include "kb_launch.h"

template <typename T> struct PsEps;
template <> struct PsEps<double> { static constexpr double v = 2.2204460492503131e-16; };
template <> struct PsEps<float> { static constexpr float v = 1.1920928955078125e-07f; };

template <typename T>
KB_DEV void ps_maxdiag(const T* S, int LD, int n, int k, int t, T& val, int& idx, int width)
{
    T v = (t >= k && t < n) ? S[t + t * LD] : T(-1);
    int key = t;
    if (n > 32) {
        const int nx = t + 32;
        if (nx >= k && nx < n) {
            const T w = S[nx + nx * LD];
            if (w > v) { v = w; key = nx; }
        }
    }
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
        const T ov = __shfl_xor(v, m, 32);
        const int ok = __shfl_xor(key, m, 32);
        if (ov > v) { v = ov; key = ok; }
    }
    val = __shfl(v, 0, width);   // slot 0 of the matrix's lane segment
    idx = __shfl(key, 0, width);
}

// Row t of the n x n block G (ld lda) into S (ld LD), in chunks of 8 columns
// whose loads go out together: unconditional, from clamped addresses kept
// opaque to the optimizer (which would otherwise put each load under its own
// branch and wait for it).  on = false: nothing is stored (G must still be a
// valid matrix).
template <typename T>
KB_DEV void ps_stage(T* S, int LD, const T* G, int lda, int n, int t, bool on)
{
    for (int j0 = 0; j0 < n; j0 += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            const bool in = on && t < n && j < n;
            const T* q = G + (in ? t + (long)j * lda : 0);
            asm volatile("" : "+v"(q));
            v[u] = gld(q);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + u;
            if (on && t < n && j < n) S[t + j * LD] = v[u];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(64)
k_pstrf(int n, BatchArg<T> A, int* const* pivp, int* pivs, long spiv, int* ranks, int batch)
{
    extern __shared__ __align__(16) unsigned char kb_ps_lds[];
    T* S = reinterpret_cast<T*>(kb_ps_lds);   // n x (n + 1): sized to the order for occupancy
    const int LD = n + 1;
    const long b = blockIdx.x;
    if (b >= batch) return;
    const int t = threadIdx.x;
    T* G = A.get(b);
    const int lda = A.ldim(b);
    int* piv = pivp ? pivp[b] : pivs + b * spiv;
    ps_stage<T>(S, LD, G, lda, n, t, true);
    __syncthreads();
    for (int j = 0; j < n; ++j)
        if (t < n && t > j) S[t + j * LD] = S[j + t * LD];   // lower := upper^T
    __syncthreads();
    T dmax;
    int dummy;
    ps_maxdiag<T>(S, LD, n, 0, t, dmax, dummy, 64);
    const T tol = dmax * T(n) * PsEps<T>::v;
    int mypiv = t, rank = n;
    for (int k = 0; k < n; ++k) {
        T val;
        int p;
        ps_maxdiag<T>(S, LD, n, k, t, val, p, 64);
        // (p outside [k, n) only when the -1 placeholder beats a tolerance below -1: not semi-definite)
        if (val <= tol || p < k || p >= n) { rank = k; break; }
        if (p != k) {
            const int pk = __shfl(mypiv, k, 64), pp = __shfl(mypiv, p, 64);
            if (t == k) mypiv = pp;
            if (t == p) mypiv = pk;
            if (t < n) {   // rows k and p, lane t = column
                const T a = S[k + t * LD], c = S[p + t * LD];
                S[k + t * LD] = c;
                S[p + t * LD] = a;
            }
            __syncthreads();
            if (t < n) {   // columns k and p, lane t = row
                const T a = S[t + k * LD], c = S[t + p * LD];
                S[t + k * LD] = c;
                S[t + p * LD] = a;
            }
            __syncthreads();
        }
        if (t == k) S[k + k * LD] = sqrt(S[k + k * LD]);
        __syncthreads();
        if (t > k && t < n) S[k + t * LD] = S[k + t * LD] / S[k + k * LD];
        __syncthreads();
        if (t > k && t < n) {
            const T u = S[k + t * LD];
#pragma unroll 8
            for (int c = k + 1; c < n; ++c) S[t + c * LD] = fma(-u, S[k + c * LD], S[t + c * LD]);
        }
        __syncthreads();
    }
    if (t == 0 && ranks) ranks[b] = rank;
    if (t < n) piv[t] = mypiv;
    for (int j = 0; j < n; ++j)
        if (t < n) G[t + (long)j * lda] = t > j ? T(0) : S[t + j * LD];
}

// n <= 32: two matrices per wave, one per 32-lane half (the pivot search's
// butterfly already works within 32-lane segments).  Each half stops at its
// own rank: a finished half keeps taking part in the barriers and does no work.
template <typename T>
__global__ void __launch_bounds__(64)
k_pstrf2(int n, BatchArg<T> A, int* const* pivp, int* pivs, long spiv, int* ranks, int batch)
{
    extern __shared__ __align__(16) unsigned char kb_ps_lds[];
    const int LD = n + 1;
    const int g = threadIdx.x >> 5, t = threadIdx.x & 31;
    const long b = (long)blockIdx.x * 2 + g;
    const bool live = b < batch;
    T* S = reinterpret_cast<T*>(kb_ps_lds) + g * n * LD;
    const long bv = live ? b : batch - 1;   // a valid matrix for the clamped loads
    T* G = A.get(bv);
    const int lda = A.ldim(bv);
    ps_stage<T>(S, LD, G, lda, n, t, live);
    __syncthreads();
    for (int j = 0; j < n; ++j)
        if (live && t < n && t > j) S[t + j * LD] = S[j + t * LD];
    __syncthreads();
    T dmax;
    int dummy;
    ps_maxdiag<T>(S, LD, live ? n : 0, 0, t, dmax, dummy, 32);   // n <= 32: one 32-lane segment per matrix
    const T tol = dmax * T(n) * PsEps<T>::v;
    int mypiv = t, rank = n;
    bool done = !live;
    for (int k = 0; k < n; ++k) {
        T val;
        int p;
        ps_maxdiag<T>(S, LD, live ? n : 0, k, t, val, p, 32);
        if (!done && (val <= tol || p < k || p >= n)) { rank = k; done = true; }
        const bool act = !done;
        const bool swap = act && p != k;
        const int pk = __shfl(mypiv, k, 32), pp = __shfl(mypiv, p & 31, 32);
        if (swap && t == k) mypiv = pp;
        if (swap && t == p) mypiv = pk;
        if (swap && t < n) {
            const T a = S[k + t * LD], c = S[p + t * LD];
            S[k + t * LD] = c;
            S[p + t * LD] = a;
        }
        __syncthreads();
        if (swap && t < n) {
            const T a = S[t + k * LD], c = S[t + p * LD];
            S[t + k * LD] = c;
            S[t + p * LD] = a;
        }
        __syncthreads();
        if (act && t == k) S[k + k * LD] = sqrt(S[k + k * LD]);
        __syncthreads();
        if (act && t > k && t < n) S[k + t * LD] = S[k + t * LD] / S[k + k * LD];
        __syncthreads();
        if (act && t > k && t < n) {
            const T u = S[k + t * LD];
#pragma unroll 8
            for (int c = k + 1; c < n; ++c) S[t + c * LD] = fma(-u, S[k + c * LD], S[t + c * LD]);
        }
        __syncthreads();
    }
    if (!live) return;
    int* piv = pivp ? pivp[b] : pivs + b * spiv;
    if (t == 0 && ranks) ranks[b] = rank;
    if (t < n) piv[t] = mypiv;
    for (int j = 0; j < n; ++j)
        if (t < n) G[t + (long)j * lda] = t > j ? T(0) : S[t + j * LD];
}

extern "C" int kb_launch_pstrf(char prec, int n, kb_mat_t A, int* const* pivp, int* pivs, long spiv, int* ranks,
                               int batch, hipStream_t s)
{
    if (n > 64) return (int)hipErrorNotSupported;
    const size_t lds = (size_t)n * (n + 1) * (prec == 'd' ? sizeof(double) : sizeof(float));
    if (n <= 32) {
        if (prec == 'd')
            hipLaunchKernelGGL(k_pstrf2<double>, dim3((batch + 1) / 2), dim3(64), 2 * lds, s, n, to_arg<double>(A),
                               pivp, pivs, spiv, ranks, batch);
        else
            hipLaunchKernelGGL(k_pstrf2<float>, dim3((batch + 1) / 2), dim3(64), 2 * lds, s, n, to_arg<float>(A),
                               pivp, pivs, spiv, ranks, batch);
        return (int)hipGetLastError();
    }
    if (prec == 'd')
        hipLaunchKernelGGL(k_pstrf<double>, dim3(batch), dim3(64), lds, s, n, to_arg<double>(A), pivp, pivs, spiv,
                           ranks, batch);
    else
        hipLaunchKernelGGL(k_pstrf<float>, dim3(batch), dim3(64), lds, s, n, to_arg<float>(A), pivp, pivs, spiv,
                           ranks, batch);
    return (int)hipGetLastError();
}
