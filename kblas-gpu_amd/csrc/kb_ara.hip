// kb_ara.hip -- the SPD helpers of the randomized range finder (ARA):
// rank-revealing Cholesky of small Gram matrices, the FP32 -> FP64 Gram
// product G = A^T A, and the right triangular solve X R = B with the
// resulting upper factor (include/batch_ara.h:74-102,
// src/batch_svd/ara_util.cuh:163-275,345-408,560-808).
//
// k_ara_potrf: 32 lanes per matrix (eight per 256-thread workgroup), lane i
//   holds row i of the lower triangle of G in registers.  Step k publishes
//   column k through LDS (the diagonal entry comes from lane k), stops the
//   matrix when the updated G(k,k) <= max(0, max diag(G)) * eps, else
//   L(k,k) = sqrt, L(i,k) = G(i,k) / L(k,k) (true division), publishes the
//   scaled column and every lane updates its row, G(i,c) -= L(c,k) L(i,k) as
//   one FMA, c in (k, i].  The algorithm never reads the upper triangle, so
//   only the lower one is loaded.  R (dim x dim) = the transpose of the
//   factored lower part (past the rank: the unfactored Schur complement),
//   zero below the diagonal, written as FP32 or FP64; diag_R(k) *= L(k,k) for
//   k below the rank, 0 from the rank on (bs entries per matrix).
// k_ara_syrk: one wave per matrix, n <= 32: G = A^T A as 16x16x4 FP64
//   matrix-core products of the FP32 columns widened to FP64 (the products
//   are exact), 16 rows per step: tiles (0,0), (1,0), (1,1); the upper
//   off-diagonal tile is written as the mirror of (1,0), so G is exactly
//   symmetric.
// k_ara_trsm: one thread per row of B (256 rows per workgroup), the upper
//   triangle of R (cols <= bs <= 32) staged in LDS, the row substituted in
//   registers: x_j = (b_j - sum_{k<j} R(k,j) x_k) / R(j,j), the sum as FMAs
//   in k order, true division.
#include "kb_device.h"
#include "kb_launch.h"

constexpr double KB_ARA_EPS = 2.2204460492503131e-16;   // ara_util.cuh:221 (KBlasEpsilon<double>)

template <typename TO>
__global__ void __launch_bounds__(256)
k_ara_potrf(const int* __restrict__ dims, double* const* G, const int* __restrict__ ldg, TO* const* R,
            const int* __restrict__ ldr, double* diag, int bs, int* ranks, int num_ops)
{
    __shared__ double sc[8][2][32];   // per matrix: column k before and after scaling
    const int seg = threadIdx.x >> 5, i = threadIdx.x & 31;
    const long op = (long)blockIdx.x * 8 + seg;
    int dim = 0;
    if (op < num_ops) {
        dim = dims[op];
        dim = dim > bs ? bs : dim;
        if (dim <= 0 && i == 0) ranks[op] = 0;
    }
    double v[32];
    double dgv = 0.0, dl = 0.0;
    if (dim > 0) {
        const double* g = G[op];
        const int lg = ldg[op];
#pragma unroll
        for (int c = 0; c < 32; ++c) v[c] = (i < dim && c <= i) ? gld(g + i + (long)c * lg) : 0.0;
        if (i < dim) dgv = gld(g + i + (long)i * lg);
        if (i < bs) dl = gld(diag + (long)bs * op + i);
    } else {
#pragma unroll
        for (int c = 0; c < 32; ++c) v[c] = 0.0;
    }
    // tol = max(0, max_i G(i, i)) * eps over the matrix's 32 lanes
    double mx = dgv > 0.0 ? dgv : 0.0;
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
        const double o = __shfl_xor(mx, m, 32);
        mx = o > mx ? o : mx;
    }
    const double tol = mx * KB_ARA_EPS;
    bool alive = dim > 0;
    int rank = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        sc[seg][0][i] = v[k];   // G(i, k); lane k's is the diagonal entry
        if (!__syncthreads_or(alive && k < dim)) break;
        if (alive && k < dim) {
            const double akk = sc[seg][0][k];
            if (akk <= tol) {
                alive = false;
            } else {
                const double s = __builtin_sqrt(akk);
                if (i == k) {
                    v[k] = s;
                    dl = dl * s;
                } else if (i > k) {
                    v[k] = v[k] / s;
                }
                rank = k + 1;
            }
        }
        const bool upd = alive && k < dim;
        sc[seg][1][i] = v[k];   // L(i, k)
        __syncthreads();
        if (upd && i > k) {
#pragma unroll
            for (int c = k + 1; c < 32; ++c)
                if (c <= i) v[c] = __builtin_fma(-sc[seg][1][c], v[k], v[c]);
        }
    }
    if (dim > 0) {
        if (i == 0) ranks[op] = rank;
        if (i < bs) gst(diag + (long)bs * op + i, i >= rank ? 0.0 : dl);
        if (i < dim) {
            // lane i writes column i of R: R(r, i) = G(i, r) for r <= i, 0 below
            TO* rp = R[op] + (long)i * ldr[op];
#pragma unroll
            for (int r = 0; r < 32; ++r)
                if (r < dim) gst(rp + r, r <= i ? (TO)v[r] : TO(0));
        }
    }
}

__global__ void __launch_bounds__(256)
k_ara_syrk(const int* __restrict__ ms, const int* __restrict__ ns, const float* const* A, const int* __restrict__ lda,
           double* const* B, const int* __restrict__ ldb, int batch)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long op = (long)blockIdx.x * 4 + wave;
    if (op >= batch) return;   // no workgroup barriers below
    const int m = ms[op];
    int n = ns[op];
    if (m <= 0 || n <= 0) return;
    n = n > 32 ? 32 : n;
    const float* a = A[op];
    const int la = lda[op];
    const int li = lane & 15, lg = lane >> 4;
    const bool two = n > 16;
    const bool c0 = li < n, c1 = 16 + li < n;
    const float* p0 = a + (long)li * la;
    const float* p1 = a + (long)(16 + li) * la;
    kb_d4 t00 = {0.0, 0.0, 0.0, 0.0}, t10 = t00, t11 = t00;
    // lane (li, lg) feeds rows r0 + 4 lg + t of columns li and 16 + li to
    // product t: the K index of the product is the same row for both operands
    for (int r0 = 0; r0 < m; r0 += 16) {
        double x0[4], x1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int r = r0 + 4 * lg + t;
            const bool in = r < m;
            const bool i0 = in && c0, i1 = in && c1 && two;
            const float v0 = gld_opaque(i0 ? p0 + r : a), v1 = gld_opaque(i1 ? p1 + r : a);
            x0[t] = i0 ? (double)v0 : 0.0;
            x1[t] = i1 ? (double)v1 : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            t00 = mfma16x16x4(x0[t], x0[t], t00);
            if (two) {
                t10 = mfma16x16x4(x1[t], x0[t], t10);
                t11 = mfma16x16x4(x1[t], x1[t], t11);
            }
        }
    }
    // accumulator: lane (li, lg), register q -> tile row 4 q + lg, tile column li
    double* b = B[op];
    const int lb = ldb[op];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 4 * q + lg;
        if (r < n && li < n) gst(b + r + (long)li * lb, t00[q]);
        if (two && 16 + r < n && li < n) {
            gst(b + 16 + r + (long)li * lb, t10[q]);
            gst(b + li + (long)(16 + r) * lb, t10[q]);
        }
        if (two && 16 + r < n && 16 + li < n) gst(b + 16 + r + (long)(16 + li) * lb, t11[q]);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_ara_trsm(T* const* Bp, const int* __restrict__ ldb, T* const* Rp, const int* __restrict__ ldr,
           const int* __restrict__ rows_, const int* __restrict__ cols_, int bs, int nchunk, int num_ops)
{
    __shared__ T sR[32 * 33];   // R(k, j) at j * 33 + k
    const long op = blockIdx.x / nchunk;
    const int chunk = blockIdx.x % nchunk;
    if (op >= num_ops) return;
    const int rows = rows_[op];
    int cols = cols_[op];
    cols = cols > bs ? bs : cols;
    if (cols <= 0 || rows <= chunk * 256) return;   // uniform over the workgroup
    const T* rm = Rp[op];
    const int lr = ldr[op];
    for (int e = threadIdx.x; e < cols * cols; e += 256) {
        const int k = e % cols, j = e / cols;
        sR[j * 33 + k] = k <= j ? gld(rm + k + (long)j * lr) : T(0);
    }
    __syncthreads();
    const int r = chunk * 256 + threadIdx.x;
    if (r >= rows) return;
    T* b = Bp[op] + r;
    const int lb = ldb[op];
    T x[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) x[j] = j < cols ? gld(b + (long)j * lb) : T(0);
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        if (j < cols) {
            T s = x[j];
#pragma unroll
            for (int k = 0; k < j; ++k) s = __builtin_fma(-sR[j * 33 + k], x[k], s);
            x[j] = s / sR[j * 33 + j];
        }
    }
#pragma unroll
    for (int j = 0; j < 32; ++j)
        if (j < cols) gst(b + (long)j * lb, x[j]);
}

extern "C" int kb_launch_ara_potrf(char rprec, const int* dims, double* const* G, const int* ldg, void* const* R,
                                   const int* ldr, double* diag, int bs, int* ranks, int num_ops, hipStream_t s)
{
    if (num_ops <= 0) return 0;
    const unsigned grid = (unsigned)(((long)num_ops + 7) / 8);
    if (rprec == 'd')
        hipLaunchKernelGGL(k_ara_potrf<double>, dim3(grid), dim3(256), 0, s, dims, G, ldg, (double* const*)R, ldr,
                           diag, bs, ranks, num_ops);
    else
        hipLaunchKernelGGL(k_ara_potrf<float>, dim3(grid), dim3(256), 0, s, dims, G, ldg, (float* const*)R, ldr, diag,
                           bs, ranks, num_ops);
    return (int)hipGetLastError();
}

extern "C" int kb_launch_ara_syrk(const int* m, const int* n, const float* const* A, const int* lda,
                                  double* const* B, const int* ldb, int batch, hipStream_t s)
{
    if (batch <= 0) return 0;
    const unsigned grid = (unsigned)(((long)batch + 3) / 4);
    hipLaunchKernelGGL(k_ara_syrk, dim3(grid), dim3(256), 0, s, m, n, A, lda, B, ldb, batch);
    return (int)hipGetLastError();
}

extern "C" int kb_launch_ara_trsm(char prec, void* const* B, const int* ldb, void* const* R, const int* ldr,
                                  const int* rows, const int* cols, int bs, int max_rows, int num_ops, hipStream_t s)
{
    if (num_ops <= 0 || max_rows <= 0) return 0;
    const int nchunk = (max_rows + 255) / 256;
    const long grid = (long)num_ops * nchunk;
    if (grid > 0x7fffffffL) return (int)hipErrorInvalidValue;
    if (prec == 'd')
        hipLaunchKernelGGL(k_ara_trsm<double>, dim3((unsigned)grid), dim3(256), 0, s, (double* const*)B, ldb,
                           (double* const*)R, ldr, rows, cols, bs, nchunk, num_ops);
    else
        hipLaunchKernelGGL(k_ara_trsm<float>, dim3((unsigned)grid), dim3(256), 0, s, (float* const*)B, ldb,
                           (float* const*)R, ldr, rows, cols, bs, nchunk, num_ops);
    return (int)hipGetLastError();
}
