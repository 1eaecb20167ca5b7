// kb_gen.hip -- the generic batched Cholesky family: POTRF / POTRS / POSV for
// any element type (float, double, and the complex types of kb_cplx.h) and
// either triangle.  It serves what the specialised kernels do not: the upper
// triangle of the real entry points and every complex (C, Z) entry point.
//
// Arithmetic (src/batch_triangular/Xpotrf_batch_kernels.cuh:52-69, the
// reference's column step, instantiated there for C and Z as well): for
// j = 0..n-1: s = sqrt(a_jj), l_ij = a_ij / s (true division), then
// a_ik -= l_ij l_kj for j < k <= i.  For complex types this is the
// factorization A = L L^T of a complex symmetric matrix (no conjugation, the
// principal square root), as the reference's kernels compute it.  The upper
// triangle is read as the lower view Lv(i,j) = U(j,i), so A = U^T U with
// U = Lv^T.  A pivot that is not positive (real types) or is zero or not
// finite (complex) stops the factorization with info = j + 1 (LAPACK's
// convention); POSV then leaves that matrix's right-hand sides untouched.
//
// One 256-thread workgroup per matrix, worked on in place in memory: the
// pivot by thread 0, the column scaled by the threads over its rows, the
// trailing lower triangle one row per thread; the solves one right-hand side
// per thread (forward with Lv, backward with Lv^T).  This is the correctness
// path for these variants, not a tuned one.
#include "kb_cplx.h"
#include "kb_device.h"
#include "kb_launch.h"

namespace {

template <typename T>
struct LowerView {
    T* p;
    int ld;
    int upper;
    KB_DEV T& operator()(int i, int j) const { return upper ? p[j + (long)i * ld] : p[i + (long)j * ld]; }
};

template <typename T>
KB_DEV int gen_potrf(LowerView<T> L, int n, T* sh_s, int* sh_flag)
{
    const int tid = threadIdx.x;
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            const T d = L(j, j);
            if (kb_bad_pivot(d)) {
                *sh_flag = j + 1;
            } else {
                const T s = kb_sqrtt(d);
                L(j, j) = s;
                *sh_s = s;
            }
        }
        __syncthreads();
        const int f = *sh_flag;
        if (f) return f;
        const T s = *sh_s;
        for (int i = j + 1 + tid; i < n; i += blockDim.x) L(i, j) = L(i, j) / s;
        __syncthreads();
        for (int i = j + 1 + tid; i < n; i += blockDim.x) {
            const T lij = L(i, j);
            for (int k = j + 1; k <= i; ++k) L(i, k) = kb_fmat(-lij, L(k, j), L(i, k));
        }
        __syncthreads();
    }
    return 0;
}

// (Lv Lv^T) x = b for the right-hand sides c = tid, tid + blockDim.x, ...
template <typename T>
KB_DEV void gen_solve(LowerView<T> L, int n, T* B, int ldb, int side_left, int nrhs)
{
    for (int c = threadIdx.x; c < nrhs; c += blockDim.x) {
        for (int i = 0; i < n; ++i) {
            T v = *rhs_elem(B, ldb, side_left, i, c);
            for (int k = 0; k < i; ++k) v = kb_fmat(-L(i, k), *rhs_elem(B, ldb, side_left, k, c), v);
            *rhs_elem(B, ldb, side_left, i, c) = v / L(i, i);
        }
        for (int i = n - 1; i >= 0; --i) {
            T v = *rhs_elem(B, ldb, side_left, i, c);
            for (int k = i + 1; k < n; ++k) v = kb_fmat(-L(k, i), *rhs_elem(B, ldb, side_left, k, c), v);
            *rhs_elem(B, ldb, side_left, i, c) = v / L(i, i);
        }
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_gchol(int upper, int n, int nrhs, int side_left, BatchArg<T> A, BatchArg<T> B, int* info, int batch)
{
    __shared__ T sh_s;
    __shared__ int sh_flag;
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        if (threadIdx.x == 0) sh_flag = 0;
        __syncthreads();
        LowerView<T> L{A.get(mat), A.ldim(mat), upper};
        int inf = 0;
        if (MODE != KB_MODE_POTRS) inf = gen_potrf<T>(L, n, &sh_s, &sh_flag);
        if (MODE != KB_MODE_POTRF && inf == 0) gen_solve<T>(L, n, B.get(mat), B.ldim(mat), side_left, nrhs);
        if (MODE != KB_MODE_POTRS && threadIdx.x == 0 && info) info[mat] = inf;
        __syncthreads();
    }
}

template <typename T>
int gchol_t(int mode, int upper, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch,
            hipStream_t s)
{
    const unsigned grid = (unsigned)(batch < 4096 ? batch : 4096);
    const BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B);
    if (mode == KB_MODE_POTRF)
        hipLaunchKernelGGL((k_gchol<T, KB_MODE_POTRF>), dim3(grid), dim3(256), 0, s, upper, n, nrhs, side_left, a, b,
                           info, batch);
    else if (mode == KB_MODE_POSV)
        hipLaunchKernelGGL((k_gchol<T, KB_MODE_POSV>), dim3(grid), dim3(256), 0, s, upper, n, nrhs, side_left, a, b,
                           info, batch);
    else
        hipLaunchKernelGGL((k_gchol<T, KB_MODE_POTRS>), dim3(grid), dim3(256), 0, s, upper, n, nrhs, side_left, a, b,
                           info, batch);
    return (int)hipGetLastError();
}

}  // namespace

extern "C" int kb_launch_gchol(char prec, int mode, int upper, int n, int nrhs, int side_left, kb_mat_t A,
                               kb_mat_t B, int* info, int batch, hipStream_t s)
{
    if (batch < 1) return 0;
    switch (prec) {
    case 'd': return gchol_t<double>(mode, upper, n, nrhs, side_left, A, B, info, batch, s);
    case 's': return gchol_t<float>(mode, upper, n, nrhs, side_left, A, B, info, batch, s);
    case 'z': return gchol_t<kb_cd>(mode, upper, n, nrhs, side_left, A, B, info, batch, s);
    case 'c': return gchol_t<kb_cf>(mode, upper, n, nrhs, side_left, A, B, info, batch, s);
    }
    return (int)hipErrorInvalidValue;
}
