// kb_blas3.hip -- batched GEMM, lower SYRK and lower TRSM (uniform, strided,
// pointer-array and non-uniform), the operations the reference's batched
// Cholesky recursion is built from:
//   GEMM  src/batch_triangular/Xgemm_batch_core.cuh:170-313,492-634 (cuBLAS/MAGMA there)
//   SYRK  src/batch_triangular/Xsyrk_batch_drivers.cuh:125-433
//   TRSM  src/batch_triangular/Xtrsm_batch_drivers.cuh:54-367
// Here they are native kernels with no host-side recursion and no pointer
// arrays to build; non-uniform batches read each matrix's sizes on the device
// (no host synchronisation, unlike the reference's MAGMA path,
// src/batch_triangular/Xtrsm_batch_drivers.cuh:315-320).
//
// The kernels are templates over the element type: the real entry points and
// the complex (C, Z) ones instantiate the same code (kb_cplx.h), the upper
// triangle is served by reading the matrix as its transpose, and a
// conjugate-transposed operand ('C') conjugates what it loads.
#include "kb_cplx.h"
#include "kb_device.h"
#include "kb_launch.h"

namespace {

constexpr int GT = 64;   // output tile
constexpr int GK = 32;   // k chunk

template <typename T> T scal(const double* v) { return T(v[0]); }
template <> kb_cf scal<kb_cf>(const double* v) { return kb_cf((float)v[0], (float)v[1]); }
template <> kb_cd scal<kb_cd>(const double* v) { return kb_cd(v[0], v[1]); }

template <typename T>
BatchArg<T> arg(kb_mat_t m)
{
    BatchArg<T> a;
    a.base = (T*)m.base; a.stride = m.stride; a.ptrs = (T* const*)m.ptrs; a.lds = m.lds; a.ld = m.ld;
    return a;
}

// C := alpha op(A) op(B) + beta C (op: 0 none, 1 transpose, 2 conjugate
// transpose); SYRK: the lower (or, with upper, the upper) part of
// C := alpha op(A) op(A)^T + beta C.
template <typename T, bool SYRK>
__global__ void __launch_bounds__(256)
k_gemm(int ta, int tb, int m0, int n0, int k0, T alpha, BatchArg<T> A, BatchArg<T> B, T beta,
       BatchArg<T> C, const int* vm, const int* vn, const int* vk, int batch, int upper)
{
    __shared__ T sA[GK][GT + 1];
    __shared__ T sB[GK][GT + 1];
    const long mat = blockIdx.x;
    if (mat >= batch) return;
    const int m = vm ? vm[mat] : m0;
    const int n = SYRK ? m : (vn ? vn[mat] : n0);
    const int k = vk ? vk[mat] : k0;
    if (m <= 0 || n <= 0) return;
    const T* Ap = A.get(mat);
    const T* Bp = SYRK ? Ap : B.get(mat);
    T* Cp = C.get(mat);
    const int lda = A.ldim(mat), ldb = SYRK ? lda : B.ldim(mat), ldc = C.ldim(mat);
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int tm = (m + GT - 1) / GT, tn = (n + GT - 1) / GT;
    const int total = SYRK ? tm * (tm + 1) / 2 : tm * tn;
    for (int t = blockIdx.y; t < total; t += gridDim.y) {
        int I, J;
        if (SYRK) {
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
        T acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
        if (alpha != T(0)) {
            for (int p0 = 0; p0 < k; p0 += GK) {
                for (int e = tid; e < GT * GK; e += 256) {
                    // op(A)(i, p): NoTrans -> A[i + p lda] (i fastest); Trans -> A[p + i lda] (p fastest)
                    int i, p;
                    if (ta == 0) { i = e % GT; p = e / GT; } else { p = e % GK; i = e / GK; }
                    const int gi = I * GT + i, gp = p0 + p;
                    T v = T(0);
                    if (gi < m && gp < k) v = ta == 0 ? Ap[gi + (long)gp * lda] : Ap[gp + (long)gi * lda];
                    sA[p][i] = ta == 2 ? kb_conj(v) : v;
                    // op(B)(p, j): NoTrans -> B[p + j ldb]; Trans -> B[j + p ldb].  SYRK uses op(A)^T.
                    const int tbb = SYRK ? (ta == 0 ? 1 : 0) : tb;
                    int j, q;
                    if (tbb == 0) { q = e % GK; j = e / GK; } else { j = e % GT; q = e / GT; }
                    const int gj = J * GT + j, gq = p0 + q;
                    T w = T(0);
                    if (gj < n && gq < k) w = tbb == 0 ? Bp[gq + (long)gj * ldb] : Bp[gj + (long)gq * ldb];
                    sB[q][j] = tbb == 2 ? kb_conj(w) : w;
                }
                __syncthreads();
                const int pe = min(GK, k - p0);
                for (int p = 0; p < pe; ++p) {
                    T va[4], vb[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) va[a] = sA[p][ty + 16 * a];
#pragma unroll
                    for (int b = 0; b < 4; ++b) vb[b] = sB[p][tx + 16 * b];
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) acc[a][b] = kb_fmat(va[a], vb[b], acc[a][b]);
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int gi = I * GT + ty + 16 * a, gj = J * GT + tx + 16 * b;
                if (gi < m && gj < n && (!SYRK || (upper ? gj >= gi : gj <= gi))) {
                    T* c = Cp + gi + (long)gj * ldc;
                    T r = alpha * acc[a][b];
                    if (beta != T(0)) r = kb_fmat(beta, *c, r);
                    *c = r;
                }
            }
        }
    }
}

// TRSM (unit diagonal or not), one thread per right-hand-side vector, on the lower view Lv
// of the triangle (Lv = L, or U^T read from the upper triangle; trans 2
// conjugates the elements read):
// forward:  x_j = (alpha b_j - sum_{p<j} Lv(j,p) x_p) / Lv(j,j)   (side R trans T, side L trans N on Lv)
// backward: x_j = (alpha b_j - sum_{p>j} Lv(p,j) x_p) / Lv(j,j)   (side R trans N, side L trans T on Lv)
constexpr int TB = 32;

template <typename T>
__global__ void __launch_bounds__(256)
k_trsm(int side_left, int trans, int m0, int n0, T alpha, BatchArg<T> A, BatchArg<T> B,
       const int* vm, const int* vn, int batch, int upper, int unit)
{
    __shared__ T sL[TB][TB + 1];
    __shared__ T sInv[TB];
    const long mat = blockIdx.x;
    if (mat >= batch) return;
    const int m = vm ? vm[mat] : m0;
    const int n = vn ? vn[mat] : n0;
    if (m <= 0 || n <= 0) return;
    const T* L = A.get(mat);
    T* Bp = B.get(mat);
    const int lda = A.ldim(mat), ldb = B.ldim(mat);
    const int len = side_left ? m : n;          // order of L
    const int nv = side_left ? n : m;           // number of vectors
    const long vs = side_left ? ldb : 1;        // vector stride
    const long es = side_left ? 1 : ldb;        // element stride
    // U op = Lv^T op: an upper triangle flips the transposition
    const int et = upper ? (trans == 0 ? 1 : 0) : (trans == 0 ? 0 : 1);
    const bool cj = trans == 2;
    auto lv = [&](int i, int j) {
        const T v = upper ? L[j + (long)i * lda] : L[i + (long)j * lda];
        return cj ? kb_conj(v) : v;
    };
    const bool fwd = side_left ? (et == 0) : (et != 0);
    const int nblk = (len + TB - 1) / TB;
    for (int v0 = blockIdx.y * 256; v0 < nv; v0 += gridDim.y * 256) {
        const int v = v0 + threadIdx.x;
        const bool act = v < nv;
        T* x = Bp + (long)(act ? v : 0) * vs;
        for (int bi = 0; bi < nblk; ++bi) {
            const int jb = fwd ? bi : nblk - 1 - bi;
            const int j0 = jb * TB, jk = min(TB, len - j0);
            T acc[TB];
#pragma unroll
            for (int jj = 0; jj < TB; ++jj) acc[jj] = (act && jj < jk) ? alpha * x[(long)(j0 + jj) * es] : T(0);
            // contributions of the blocks already solved
            for (int bp = 0; bp < bi; ++bp) {
                const int pb = fwd ? bp : nblk - 1 - bp;
                const int p0 = pb * TB, pk = min(TB, len - p0);
                __syncthreads();
                for (int e = threadIdx.x; e < TB * TB; e += 256) {
                    const int r = e % TB, c = e / TB;
                    // forward: sL[jj][pp] = L(j0+jj, p0+pp);  backward: sL[jj][pp] = L(p0+pp, j0+jj)
                    T val = T(0);
                    if (fwd) { if (r < jk && c < pk) val = lv(j0 + r, p0 + c); sL[r][c] = val; }
                    else     { if (r < pk && c < jk) val = lv(p0 + r, j0 + c); sL[c][r] = val; }
                }
                __syncthreads();
                for (int pp = 0; pp < pk; ++pp) {
                    const T xp = act ? x[(long)(p0 + pp) * es] : T(0);
#pragma unroll
                    for (int jj = 0; jj < TB; ++jj) acc[jj] = kb_fmat(-sL[jj][pp], xp, acc[jj]);
                }
            }
            // diagonal block
            __syncthreads();
            for (int e = threadIdx.x; e < TB * TB; e += 256) {
                const int r = e % TB, c = e / TB;
                sL[r][c] = (r < jk && c < jk && c <= r) ? lv(j0 + r, j0 + c) : T(0);
            }
            if (threadIdx.x < TB) {
                const int r = threadIdx.x;
                sInv[r] = (r < jk && !unit) ? T(1) / lv(j0 + r, j0 + r) : T(1);
            }
            __syncthreads();
            if (fwd) {
#pragma unroll
                for (int jj = 0; jj < TB; ++jj) {
                    acc[jj] = acc[jj] * sInv[jj];
#pragma unroll
                    for (int q = jj + 1; q < TB; ++q) acc[q] = kb_fmat(-sL[q][jj], acc[jj], acc[q]);
                }
            } else {
#pragma unroll
                for (int jj = TB - 1; jj >= 0; --jj) {
                    acc[jj] = acc[jj] * sInv[jj];
#pragma unroll
                    for (int q = 0; q < jj; ++q) acc[q] = kb_fmat(-sL[jj][q], acc[jj], acc[q]);
                }
            }
            if (act) {
#pragma unroll
                for (int jj = 0; jj < TB; ++jj)
                    if (jj < jk) x[(long)(j0 + jj) * es] = acc[jj];
            }
        }
    }
}

template <typename T>
int gemm_t(int ta, int tb, int m, int n, int k, T alpha, kb_mat_t A, kb_mat_t B, T beta,
           kb_mat_t C, const int* vm, const int* vn, const int* vk, int batch, hipStream_t s, bool syrk,
           int upper = 0)
{
    if constexpr (__is_same(T, double) || __is_same(T, float)) {
        // real batches: the matrix-core kernels of kb_gemm.hip, uniform or with the
        // non-uniform sizes bounded by m / n / k
        if ((!vm && !vn && !vk) || (vm && vk && (syrk || vn) && m > 0 && n > 0 && k > 0)) {
            const int e = kb_launch_gemm_fast(sizeof(T) == 8 ? 'd' : 's', syrk, upper, ta, syrk ? !ta : tb, m, n, k,
                                              (double)alpha, A, B, (double)beta, C, vm, syrk ? vm : vn, vk, batch, s);
            if (e != (int)hipErrorNotSupported) return e;
        }
    }
    int gy;
    if (vm) {
        gy = 16;
    } else {
        const int tm = (m + GT - 1) / GT, tn = (n + GT - 1) / GT;
        gy = syrk ? tm * (tm + 1) / 2 : tm * tn;
        if (gy > 64) gy = 64;
    }
    dim3 grid((unsigned)batch, (unsigned)gy);
    if (syrk)
        hipLaunchKernelGGL((k_gemm<T, true>), grid, dim3(256), 0, s, ta, tb, m, n, k, alpha, arg<T>(A),
                           arg<T>(A), beta, arg<T>(C), vm, vm, vk, batch, upper);
    else
        hipLaunchKernelGGL((k_gemm<T, false>), grid, dim3(256), 0, s, ta, tb, m, n, k, alpha, arg<T>(A),
                           arg<T>(B), beta, arg<T>(C), vm, vn, vk, batch, 0);
    return (int)hipGetLastError();
}

template <typename T>
int trsm_t(int side_left, int trans, int m, int n, T alpha, kb_mat_t A, kb_mat_t B,
           const int* vm, const int* vn, int batch, hipStream_t s, int upper = 0, int unit = 0)
{
    if constexpr (__is_same(T, double) || __is_same(T, float)) {
        // real batches: the register / matrix-core / blocked kernels of kb_trsm.hip,
        // uniform or with the non-uniform sizes bounded by m / n
        if ((!vm && !vn) || (vm && vn && m > 0 && n > 0)) {
            const int e = kb_launch_trsm_fast(sizeof(T) == 8 ? 'd' : 's', side_left, upper, trans, unit, m, n,
                                              (double)alpha, A, B, vm, vn, batch, s);
            if (e != (int)hipErrorNotSupported) return e;
        }
    }
    int gy;
    if (vm) {
        gy = 4;
    } else {
        const int nv = side_left ? n : m;
        gy = (nv + 255) / 256;
        if (gy > 64) gy = 64;
    }
    hipLaunchKernelGGL((k_trsm<T>), dim3((unsigned)batch, (unsigned)gy), dim3(256), 0, s, side_left, trans, m, n,
                       alpha, arg<T>(A), arg<T>(B), vm, vn, batch, upper, unit);
    return (int)hipGetLastError();
}

}  // namespace

extern "C" int kb_launch_gemm(char prec, int transA, int transB, int m, int n, int k, double alpha,
                              kb_mat_t A, kb_mat_t B, double beta, kb_mat_t C,
                              const int* vm, const int* vn, const int* vk, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    if (prec == 'd') return gemm_t<double>(transA, transB, m, n, k, alpha, A, B, beta, C, vm, vn, vk, batch, s, false);
    return gemm_t<float>(transA, transB, m, n, k, (float)alpha, A, B, (float)beta, C, vm, vn, vk, batch, s, false);
}

extern "C" int kb_launch_syrk(char prec, int trans, int m, int k, double alpha, kb_mat_t A,
                              double beta, kb_mat_t B, const int* vm, const int* vk, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    if (prec == 'd') return gemm_t<double>(trans, 0, m, m, k, alpha, A, A, beta, B, vm, vm, vk, batch, s, true);
    return gemm_t<float>(trans, 0, m, m, k, (float)alpha, A, A, (float)beta, B, vm, vm, vk, batch, s, true);
}

extern "C" int kb_launch_trsm(char prec, int side_left, int trans, int unit, int m, int n, double alpha,
                              kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    if (unit) return (int)hipErrorNotSupported;
    if (prec == 'd') return trsm_t<double>(side_left, trans, m, n, alpha, A, B, vm, vn, batch, s);
    return trsm_t<float>(side_left, trans, m, n, (float)alpha, A, B, vm, vn, batch, s);
}

// Any precision ('s', 'd', 'c', 'z'), scalars as (real, imaginary) pairs,
// op 0 / 1 / 2 = none / transpose / conjugate transpose, upper triangles.
extern "C" int kb_launch_gemm_x(char prec, int transA, int transB, int m, int n, int k, const double* alpha,
                                kb_mat_t A, kb_mat_t B, const double* beta, kb_mat_t C, const int* vm,
                                const int* vn, const int* vk, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    switch (prec) {
    case 'd': return gemm_t<double>(transA, transB, m, n, k, scal<double>(alpha), A, B, scal<double>(beta), C, vm, vn, vk, batch, s, false);
    case 's': return gemm_t<float>(transA, transB, m, n, k, scal<float>(alpha), A, B, scal<float>(beta), C, vm, vn, vk, batch, s, false);
    case 'z': return gemm_t<kb_cd>(transA, transB, m, n, k, scal<kb_cd>(alpha), A, B, scal<kb_cd>(beta), C, vm, vn, vk, batch, s, false);
    case 'c': return gemm_t<kb_cf>(transA, transB, m, n, k, scal<kb_cf>(alpha), A, B, scal<kb_cf>(beta), C, vm, vn, vk, batch, s, false);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_syrk_x(char prec, int upper, int trans, int m, int k, const double* alpha, kb_mat_t A,
                                const double* beta, kb_mat_t B, const int* vm, const int* vk, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    switch (prec) {
    case 'd': return gemm_t<double>(trans, 0, m, m, k, scal<double>(alpha), A, A, scal<double>(beta), B, vm, vm, vk, batch, s, true, upper);
    case 's': return gemm_t<float>(trans, 0, m, m, k, scal<float>(alpha), A, A, scal<float>(beta), B, vm, vm, vk, batch, s, true, upper);
    case 'z': return gemm_t<kb_cd>(trans, 0, m, m, k, scal<kb_cd>(alpha), A, A, scal<kb_cd>(beta), B, vm, vm, vk, batch, s, true, upper);
    case 'c': return gemm_t<kb_cf>(trans, 0, m, m, k, scal<kb_cf>(alpha), A, A, scal<kb_cf>(beta), B, vm, vm, vk, batch, s, true, upper);
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_trsm_x(char prec, int side_left, int upper, int trans, int unit, int m, int n,
                                const double* alpha, kb_mat_t A, kb_mat_t B, const int* vm, const int* vn, int batch,
                                hipStream_t s)
{
    if (batch < 1) return 0;
    switch (prec) {
    case 'd': return trsm_t<double>(side_left, trans, m, n, scal<double>(alpha), A, B, vm, vn, batch, s, upper, unit);
    case 's': return trsm_t<float>(side_left, trans, m, n, scal<float>(alpha), A, B, vm, vn, batch, s, upper, unit);
    case 'z': return trsm_t<kb_cd>(side_left, trans, m, n, scal<kb_cd>(alpha), A, B, vm, vn, batch, s, upper, unit);
    case 'c': return trsm_t<kb_cf>(side_left, trans, m, n, scal<kb_cf>(alpha), A, B, vm, vn, batch, s, upper, unit);
    }
    return (int)hipErrorInvalidValue;
}
