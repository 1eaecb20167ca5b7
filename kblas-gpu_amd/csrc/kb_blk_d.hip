// Instantiation unit: blocked (n > 64) Cholesky family, double.
#include "kb_blk_impl.h"

extern "C" int kb_blk_chol_d(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                             int batch, hipStream_t s)
{
    return launch_blk<double>(mode, n, nrhs, side_left, A, B, info, batch, s);
}

extern "C" int kb_blk_trsm_d(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                              kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s)
{
    return launch_blk_trsm<double>(fwd, side_left, unit, n, nrhs, (double)alpha, A, B, vm, vn, batch, s);
}

// Mixed-precision POSV entry for orders above the register range: LAPACK
// dsposv's full-precision path (ITER = -1, "fell back to full precision for
// implementation reasons"): X := B (and ITER := -1) here, then the FP64 POSV
// on (A, X), so A := L and X := (L L^T)^-1 B.
__global__ void __launch_bounds__(256)
k_sposv_full_prep(int n, int nrhs, int side_left, BatchArg<double> B, BatchArg<double> X, int* iter, int batch)
{
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        const double* Bp = B.get(mat);
        double* Xp = X.get(mat);
        const int ldb = B.ldim(mat), ldx = X.ldim(mat);
        copy_rhs<double>(Bp, ldb, Xp, ldx, side_left, n, nrhs, threadIdx.x, blockDim.x);
        if (threadIdx.x == 0 && iter) iter[mat] = -1;
    }
}

extern "C" int kb_launch_blk_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info,
                                   int* iter, int batch, hipStream_t s)
{
    const int grid = batch < 2048 ? batch : 2048;
    hipLaunchKernelGGL(k_sposv_full_prep, dim3(grid), dim3(256), 0, s, n, nrhs, side_left, to_arg<double>(B),
                       to_arg<double>(X), iter, batch);
    const int e = (int)hipGetLastError();
    if (e) return e;
    return launch_blk<double>(KB_MODE_POSV, n, nrhs, side_left, A, X, info, batch, s);
}
