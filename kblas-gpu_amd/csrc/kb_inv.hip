// kb_inv.hip -- batched triangular inverse family and TRMM (lower triangle).
//
//   TRTRI  A := L^-1                 src/batch_triangular/Xtrtri_batch_kernels.cuh:37-79
//   LAUUM  A := L^T L (lower part)   src/batch_triangular/Xlauum_batch_kernels.cuh:112-146
//   POTRI  A := (L L^T)^-1 = LAUUM(TRTRI(L))
//                                    src/batch_triangular/Xpotri_batch_drivers.cuh:78-110
//   TRMM   B := alpha op(L) B  or  alpha B op(L)
//                                    src/batch_triangular/Xtrmm_batch_kernels.cuh:37-175
//
// Orders up to 64: k_inv_reg, 64 / NL matrices per wave, registers (below).
// Larger orders: k_inv, one workgroup per matrix, one thread per row (TRTRI)
// or column (LAUUM) of the matrix, so n <= KB_INV_MAX_N.  The matrix is staged in LDS when its
// n x n block fits in KB_INV_LDS bytes and is worked on in place in HBM
// (through the CU's cache) otherwise.  The element recurrences and their
// summation orders are the reference register kernels' (TRTRI row by row:
// X(r,j) = -(sum_{j<k<=r} X(r,k) A(k,j)) / A(j,j) with j descending;
// LAUUM row i of the result: sum_{k>=i} A(k,c) A(k,i), k ascending), which
// lets each step read only values that no thread writes in that step, so
// one barrier per step orders the reads before the writes.
//
// TRMM: the triangle (when it fits) and a panel of B (whole columns for the
// left side, whole rows for the right side) are copied to LDS, then every
// thread forms output elements from the copies and writes them to B, so the
// in-place update needs no ordering between elements.
#include "kb_cplx.h"
#include "kb_device.h"
#include "kb_launch.h"

// The kernels are templates over the element type; the complex (C, Z)
// entry points instantiate them with the types of kb_cplx.h (arithmetic
// without conjugation, as the reference's complex instantiations: LAUUM is
// L^T L, POTRI (L L^T)^-1; TRMM's trans 2 conjugates the triangle).

#define KB_INV_LDS (60 * 1024)    // k_inv: dynamic part (a few static bytes beside it)
#define KB_TRMM_LDS (64 * 1024)

// Singularity check of a non-unit triangle (LAPACK xTRTRI: the first exactly
// zero diagonal element, 1-based; nothing is computed when one exists).
template <typename T>
KB_DEV int inv_zero_pivot(const T* M, int ld, int n, int r, int* sflag)
{
    if (r == 0) *sflag = n + 1;
    __syncthreads();
    if (r < n && M[r + (long)r * ld] == T(0)) atomicMin(sflag, r + 1);
    __syncthreads();
    const int z = *sflag;
    return z <= n ? z : 0;
}

// In-place inverse of the lower triangle, thread r owns row r.
template <typename T>
KB_DEV void inv_trtri(T* M, int ld, int n, int unit, int r)
{
    for (int j = n - 1; j >= 0; --j) {
        const T a = unit ? T(1) : M[j + (long)j * ld];
        T s = T(0);
        if (r > j && r < n) {
            const T* xr = M + r;              // row r of X (columns > j done)
            const T* aj = M + (long)j * ld;   // column j of A (still original)
#pragma unroll 8
            for (int k = j + 1; k < r; ++k) s = fma(xr[(long)k * ld], aj[k], s);
            s = unit ? s + aj[r] : fma(xr[(long)r * ld], aj[r], s);
        }
        __syncthreads();
        if (r == j && !unit) M[j + (long)j * ld] = T(1) / a;
        else if (r > j && r < n) M[r + (long)j * ld] = -s / a;
    }
    __syncthreads();
}

// In-place L^T L of the lower triangle, thread c owns column c.
template <typename T>
KB_DEV void inv_lauum(T* M, int ld, int n, int c)
{
    for (int i = 0; i < n; ++i) {
        T s = T(0);
        if (c <= i) {
            const T* ac = M + (long)c * ld;
            const T* ai = M + (long)i * ld;
#pragma unroll 8
            for (int k = i; k < n; ++k) s = fma(ac[k], ai[k], s);
        }
        __syncthreads();
        if (c <= i) M[i + (long)c * ld] = s;
    }
    __syncthreads();
}

template <typename T>
KB_DEV int inv_run(int op, int unit, T* M, int ld, int n, int r, int* sflag)
{
    int info = 0;
    if (op != KB_INV_OP_LAUUM) {
        if (!unit) info = inv_zero_pivot<T>(M, ld, n, r, sflag);
        if (info) return info;
        inv_trtri<T>(M, ld, n, unit, r);
    }
    if (op != KB_INV_OP_TRTRI) inv_lauum<T>(M, ld, n, r);
    return info;
}

template <typename T>
__global__ void __launch_bounds__(1024)
k_inv(int op, int unit, int n, BatchArg<T> A, int* info, const int* skip, int staged, int batch)
{
    extern __shared__ __align__(16) unsigned char kb_inv_lds[];
    __shared__ int sflag;
    const long b = blockIdx.x;
    if (b >= batch) return;
    if (skip && skip[b] != 0) return;          // POTI: factorization failed, info already set
    T* G = A.get(b);
    const int lda = A.ldim(b);
    const int r = threadIdx.x, nt = blockDim.x;
    int inf;
    if (staged) {
        T* S = reinterpret_cast<T*>(kb_inv_lds);
        const int ls = n + 1;   // odd leading dimension: column-wise reads spread over the banks
        for (unsigned e = r; e < (unsigned)(n * n); e += nt) {   // staged sizes fit 32 bits
            const int i = (int)(e % (unsigned)n), j = (int)(e / (unsigned)n);
            if (i >= j) S[i + (long)j * ls] = G[i + (long)j * lda];
        }
        __syncthreads();
        inf = inv_run<T>(op, unit, S, ls, n, r, &sflag);
        if (inf == 0)
            for (unsigned e = r; e < (unsigned)(n * n); e += nt) {   // staged sizes fit 32 bits
                const int i = (int)(e % (unsigned)n), j = (int)(e / (unsigned)n);
                if (i >= j && !(unit && i == j)) G[i + (long)j * lda] = S[i + (long)j * ls];
            }
    } else {
        inf = inv_run<T>(op, unit, G, lda, n, r, &sflag);
    }
    if (info && r == 0) info[b] = inf;
}

// ------------------------------------------------- register form, n <= 64 ----
// One wave holds 64 / NL matrices (NL = 16, 32 or 64 lanes each).  The original
// triangle is staged once in LDS (zero outside n x n) and only read after
// that, by broadcast: TRTRI keeps row r of X = L^-1 in lane r's registers and
// needs column j of A at step j; LAUUM keeps column c of X in lane c and needs
// column i of X (written once to LDS) at step i.  No step waits for another
// lane, so there is no barrier between steps.  TRTRI's sums run from the far
// end (k descending), so the term that needs the previous step's result is
// added last and successive steps' products overlap; the pivot's reciprocal
// (refined hardware estimate) scales the sum instead of a division.
template <typename T, int NL>
__global__ void __launch_bounds__(64)
k_inv_reg(int op, int unit, int n, BatchArg<T> A, int* info, const int* skip, int batch)
{
    constexpr int MPW = 64 / NL, LD = NL + 1;
    __shared__ T tile[MPW][NL * LD];
    const int lane = threadIdx.x, g = lane / NL, r = lane % NL;
    const long b = (long)blockIdx.x * MPW + g;
    const bool live = b < batch && !(skip && skip[b] != 0);
    T* S = tile[g];
    const long bv = b < batch ? b : batch - 1;   // a valid matrix for the clamped loads below
    T* G = A.get(bv);
    const int lda = A.ldim(bv);
    // the triangle in chunks of 16 columns: each chunk's loads issued together
    // (unconditional, from clamped addresses kept opaque to the optimizer, which
    // would otherwise put every load under its own branch and wait for it)
#pragma unroll
    for (int j0 = 0; j0 < NL; j0 += 16) {
        T v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int j = j0 + u;
            const bool in = live && r < n && j < n && r >= j;
            const T* q = G + (in ? r + (long)j * lda : 0);
            asm volatile("" : "+v"(q));
            if constexpr (__is_same(T, double) || __is_same(T, float)) v[u] = gld(q);
            else v[u] = *q;   // the complex types: a plain load
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int j = j0 + u;
            const bool in = live && r < n && j < n && r >= j;
            S[r + j * LD] = in ? v[u] : T(0);
        }
    }
    __syncthreads();
    // first exactly-zero diagonal element of this lane group's matrix (non-unit TRTRI / POTRI)
    const bool zero = op != KB_INV_OP_LAUUM && !unit && r < n && S[r + r * LD] == T(0);
    const unsigned long long zm = (__ballot(zero) >> (g * NL)) & (NL == 64 ? ~0ull : ((1ull << NL) - 1));
    const int inf = zm ? __builtin_ctzll(zm) + 1 : 0;
    T x[NL];
    if (op != KB_INV_OP_LAUUM) {
#pragma unroll
        for (int k = 0; k < NL; ++k) x[k] = (k < r) ? S[r + k * LD] : (k == r ? (unit ? T(1) : S[r + k * LD]) : T(0));
#pragma unroll
        for (int j = NL - 1; j >= 0; --j) {
            // Steps past the order change no live row; the uniform branch also
            // keeps each step's LDS reads in its own block (scheduled there).
            if (j < n) {
                const T a = unit ? T(1) : S[j + j * LD];
                T s = T(0);
#pragma unroll
                for (int k = NL - 1; k > j; --k) s = fma(x[k], S[k + j * LD], s);   // x[k] = 0 above row r
                const T ia = kb_rcp<T>(a);   // one reciprocal per step, shared by the group's lanes
                const T nv = (r == j) ? (unit ? T(1) : ia) : -s * ia;
                x[j] = (r >= j) ? nv : x[j];
            }
        }
        if (op == KB_INV_OP_TRTRI) {
            if (live && inf == 0 && r < n)
#pragma unroll
                for (int k = 0; k < NL; ++k)
                    if (k < r || (k == r && !unit)) G[r + (long)k * lda] = x[k];
            if (live && info && r == 0) info[b] = inf;
            return;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NL; ++k) S[r + k * LD] = (r < n) ? x[k] : T(0);   // X, rows beyond n zero
        __syncthreads();
    }
    // LAUUM on the triangle in S: lane c holds column c, row i of the result is
    // sum_{k>=i} X(k,c) X(k,i), k ascending.
#pragma unroll
    for (int k = 0; k < NL; ++k) x[k] = S[k + r * LD];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        if (i < n) {   // rows past the order are not stored
            T s = T(0);
#pragma unroll
            for (int k = i; k < NL; ++k) s = fma(x[k], S[k + i * LD], s);
            x[i] = s;   // row i is not read by later rows
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NL; ++i) S[i + r * LD] = x[i];   // transpose through LDS for row-wise stores
    __syncthreads();
    if (live && inf == 0 && r < n)
#pragma unroll
        for (int k = 0; k < NL; ++k)
            if (k <= r) G[r + (long)k * lda] = S[r + k * LD];
    if (live && info && r == 0) info[b] = inf;
}

// ---------------------------------------------------------------- TRMM ----
// Element (r, c) of alpha op(L) B (left, L of order m) or alpha B op(L)
// (right, L of order n), read from the panel copy P (ldp rows) of B.
template <typename T>
KB_DEV T trmm_elem(int side_left, int trans, int unit, const T* L0, int ldl, int ord, const T* P, int ldp,
                   int r, int c)
{
    // trans 2: the conjugate of the triangle, read through this view
    struct View {
        const T* p;
        bool cj;
        KB_DEV T operator[](long i) const { return cj ? kb_conj(p[i]) : p[i]; }
    };
    const View L{L0, trans == 2};
    T s = T(0);
    if (side_left) {
        const T* pc = P + (long)c * ldp;
        if (!trans) {   // sum_{k<=r} L(r,k) B(k,c)
#pragma unroll 8
            for (int k = 0; k < r; ++k) s = fma(L[r + (long)k * ldl], pc[k], s);
            s = unit ? s + pc[r] : fma(L[r + (long)r * ldl], pc[r], s);
        } else {        // sum_{k>=r} L(k,r) B(k,c)
            const View lr{L0 + (long)r * ldl, trans == 2};
            s = unit ? pc[r] : lr[r] * pc[r];
#pragma unroll 8
            for (int k = r + 1; k < ord; ++k) s = fma(lr[k], pc[k], s);
        }
    } else {
        if (!trans) {   // sum_{k>=c} B(r,k) L(k,c)
            const View lc{L0 + (long)c * ldl, false};
            s = unit ? P[r + (long)c * ldp] : lc[c] * P[r + (long)c * ldp];
#pragma unroll 8
            for (int k = c + 1; k < ord; ++k) s = fma(P[r + (long)k * ldp], lc[k], s);
        } else {        // sum_{k<=c} B(r,k) L(c,k)
#pragma unroll 8
            for (int k = 0; k < c; ++k) s = fma(P[r + (long)k * ldp], L[c + (long)k * ldl], s);
            s = unit ? s + P[r + (long)c * ldp] : fma(P[r + (long)c * ldp], L[c + (long)c * ldl], s);
        }
    }
    return s;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_trmm(int side_left, int trans, int unit, int m, int n, T alpha, BatchArg<T> A, BatchArg<T> B, int stageA,
       int panel, int batch)
{
    extern __shared__ __align__(16) unsigned char kb_inv_lds[];
    const long b = blockIdx.x;
    if (b >= batch) return;
    const int t = threadIdx.x, nt = blockDim.x;
    const int ord = side_left ? m : n;
    const T* GA = A.get(b);
    T* GB = B.get(b);
    const int lda = A.ldim(b), ldb = B.ldim(b);
    T* SA = reinterpret_cast<T*>(kb_inv_lds);
    T* SP = SA + (stageA ? (long)ord * (ord + 1) : 0);
    const T* L = GA;
    int ldl = lda;
    // staging copies: TM_U elements per thread whose loads go out together
    // (unconditional, from opaque clamped addresses; a load under a branch of its
    // own would be waited for alone)
    constexpr int TM_U = 4;
    if (stageA) {
        const unsigned tot = (unsigned)(ord * ord);   // staged sizes fit 32 bits
        for (unsigned e0 = t; e0 < tot; e0 += TM_U * nt) {
            T v[TM_U];
#pragma unroll
            for (int u = 0; u < TM_U; ++u) {
                const unsigned e = e0 + u * nt;
                const int i = (int)(e % (unsigned)ord), j = (int)(e / (unsigned)ord);
                const bool in = e < tot && i >= j;
                const T* q = GA + (in ? i + (long)j * lda : 0);
                asm volatile("" : "+v"(q));
                if constexpr (__is_same(T, double) || __is_same(T, float)) v[u] = gld(q);
                else v[u] = *q;
            }
#pragma unroll
            for (int u = 0; u < TM_U; ++u) {
                const unsigned e = e0 + u * nt;
                const int i = (int)(e % (unsigned)ord), j = (int)(e / (unsigned)ord);
                if (e < tot && i >= j) SA[i + (long)j * (ord + 1)] = v[u];
            }
        }
        L = SA;
        ldl = ord + 1;
    }
    // Panel: columns [p0, p0 + w) of B (left) or rows [p0, p0 + w) (right).
    const int total = side_left ? n : m;
    for (int p0 = 0; p0 < total; p0 += panel) {
        const int w = total - p0 < panel ? total - p0 : panel;
        const int prow = side_left ? m : w, pcol = side_left ? w : n;
        __syncthreads();   // previous panel's reads done (and the triangle staged)
        const unsigned tot = (unsigned)(prow * pcol);   // staged sizes fit 32 bits
        for (unsigned e0 = t; e0 < tot; e0 += TM_U * nt) {
            T v[TM_U];
#pragma unroll
            for (int u = 0; u < TM_U; ++u) {
                const unsigned e = e0 + u * nt;
                const int i = (int)(e % (unsigned)prow), j = (int)(e / (unsigned)prow);
                const bool in = e < tot;
                const T* q = GB + (in ? (side_left ? i + (long)(p0 + j) * ldb : p0 + i + (long)j * ldb) : 0);
                asm volatile("" : "+v"(q));
                if constexpr (__is_same(T, double) || __is_same(T, float)) v[u] = gld(q);
                else v[u] = *q;
            }
#pragma unroll
            for (int u = 0; u < TM_U; ++u)
                if (e0 + u * nt < tot) SP[e0 + u * nt] = v[u];
        }
        __syncthreads();
        for (unsigned e = t; e < (unsigned)(prow * pcol); e += nt) {   // staged sizes fit 32 bits
            const int i = (int)(e % (unsigned)prow), j = (int)(e / (unsigned)prow);
            T v = T(0);
            if (alpha != T(0)) v = alpha * trmm_elem<T>(side_left, trans, unit, L, ldl, ord, SP, prow, i, j);
            if (side_left) GB[i + (long)(p0 + j) * ldb] = v;
            else GB[p0 + i + (long)j * ldb] = v;
        }
    }
}

template <typename T>
static int launch_inv(int op, int unit, int n, kb_mat_t A, int* info, const int* skip, int batch, hipStream_t s)
{
    if (n > KB_INV_MAX_N) return (int)hipErrorNotSupported;
    // register form: real types up to 64, single complex up to 16 (above, its
    // rows would not fit in registers), double complex never
    constexpr int REG_MAX = __is_same(T, kb_cd) ? 0 : __is_same(T, kb_cf) ? 16 : 64;
    if constexpr (REG_MAX >= 16) {
        if (n <= 16) {
            hipLaunchKernelGGL((k_inv_reg<T, 16>), dim3((batch + 3) / 4), dim3(64), 0, s, op, unit, n, to_arg<T>(A),
                               info, skip, batch);
            return (int)hipGetLastError();
        }
    }
    if constexpr (REG_MAX >= 64) {
        if (n <= 32) {
            hipLaunchKernelGGL((k_inv_reg<T, 32>), dim3((batch + 1) / 2), dim3(64), 0, s, op, unit, n, to_arg<T>(A),
                               info, skip, batch);
            return (int)hipGetLastError();
        }
        if (n <= 64) {
            hipLaunchKernelGGL((k_inv_reg<T, 64>), dim3(batch), dim3(64), 0, s, op, unit, n, to_arg<T>(A), info, skip,
                               batch);
            return (int)hipGetLastError();
        }
    }
    const size_t need = (size_t)n * (n + 1) * sizeof(T);
    const int staged = need <= KB_INV_LDS;
    const int threads = ((n + 63) / 64) * 64;
    hipLaunchKernelGGL(k_inv<T>, dim3(batch), dim3(threads), staged ? need : 0, s, op, unit, n, to_arg<T>(A),
                       info, skip, staged, batch);
    return (int)hipGetLastError();
}

template <typename T>
static int launch_trmm(int side_left, int trans, int unit, int m, int n, T alpha, kb_mat_t A, kb_mat_t B,
                       int batch, hipStream_t s)
{
    const int ord = side_left ? m : n;
    const size_t tri = (size_t)ord * (ord + 1) * sizeof(T);
    const int stageA = tri <= KB_TRMM_LDS * 5 / 8;
    const size_t left = KB_TRMM_LDS - (stageA ? tri : 0);
    const size_t line = (size_t)(side_left ? m : n) * sizeof(T);   // one column (left) / row (right) of B
    if (line > left) return (int)hipErrorNotSupported;
    const int total = side_left ? n : m;
    long panel = (long)(left / line);
    if (panel > total) panel = total;
    const size_t bytes = (stageA ? tri : 0) + (size_t)panel * line;
    hipLaunchKernelGGL(k_trmm<T>, dim3(batch), dim3(256), bytes, s, side_left, trans, unit, m, n, alpha,
                       to_arg<T>(A), to_arg<T>(B), stageA, (int)panel, batch);
    return (int)hipGetLastError();
}

extern "C" int kb_launch_inv(char prec, int op, int unit, int n, kb_mat_t A, int* info, const int* skip, int batch,
                             hipStream_t s)
{
    switch (prec) {
    case 'd': return launch_inv<double>(op, unit, n, A, info, skip, batch, s);
    case 's': return launch_inv<float>(op, unit, n, A, info, skip, batch, s);
    case 'z': return launch_inv<kb_cd>(op, unit, n, A, info, skip, batch, s);
    case 'c': return launch_inv<kb_cf>(op, unit, n, A, info, skip, batch, s);
    }
    return (int)hipErrorInvalidValue;
}

// alpha as a (real, imaginary) pair; trans 0 / 1 / 2 = none / transpose / conjugate transpose
extern "C" int kb_launch_trmm(char prec, int side_left, int trans, int unit, int m, int n, const double* alpha,
                              kb_mat_t A, kb_mat_t B, int batch, hipStream_t s)
{
    switch (prec) {
    case 'd': return launch_trmm<double>(side_left, trans, unit, m, n, alpha[0], A, B, batch, s);
    case 's': return launch_trmm<float>(side_left, trans, unit, m, n, (float)alpha[0], A, B, batch, s);
    case 'z': return launch_trmm<kb_cd>(side_left, trans, unit, m, n, kb_cd(alpha[0], alpha[1]), A, B, batch, s);
    case 'c': return launch_trmm<kb_cf>(side_left, trans, unit, m, n, kb_cf((float)alpha[0], (float)alpha[1]), A, B,
                                        batch, s);
    }
    return (int)hipErrorInvalidValue;
}
