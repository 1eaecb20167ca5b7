// Instantiation unit: blocked (n > 64) Cholesky family, double.
#include "kb_blk_impl.h"

extern "C" int kb_blk_chol_d(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                             int batch, hipStream_t s)
{
    return launch_blk<double>(mode, n, nrhs, side_left, A, B, info, batch, s);
}

// Mixed-precision POSV entry for orders above the register range: LAPACK
// dsposv's full-precision path (ITER = -1, "fell back to full precision for
// implementation reasons"): X := B, A := L in FP64, X := (L L^T)^-1 X.
__global__ void __launch_bounds__(KB_BLK_WG)
k_blk_sposv64(int n, int nrhs, int side_left, BatchArg<double> A, BatchArg<double> B, BatchArg<double> X,
              int* info, int* iter, int batch)
{
    __shared__ BlkSmem<double> S;
    for (long mat = blockIdx.x; mat < batch; mat += gridDim.x) {
        double* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        double* Bp = B.get(mat);
        double* Xp = X.get(mat);
        const int ldb = B.ldim(mat), ldx = X.ldim(mat);
        for (int e = threadIdx.x; e < n * nrhs; e += KB_BLK_WG)
            rhs_at(Xp, ldx, side_left, e % n, e / n) = rhs_at(Bp, ldb, side_left, e % n, e / n);
        __syncthreads();
        const int inf = blk_potrf<double>(Ap, lda, n, S);
        if (threadIdx.x == 0) {
            if (info) info[mat] = inf;
            if (iter) iter[mat] = -1;
        }
        if (inf == 0) blk_solve<double>(Ap, lda, n, Xp, ldx, side_left, nrhs, S);
        __syncthreads();
    }
}

extern "C" int kb_launch_blk_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info,
                                   int* iter, int batch, hipStream_t s)
{
    const int grid = batch < 2048 ? batch : 2048;
    hipLaunchKernelGGL(k_blk_sposv64, dim3(grid), dim3(KB_BLK_WG), 0, s, n, nrhs, side_left, to_arg<double>(A),
                       to_arg<double>(B), to_arg<double>(X), info, iter, batch);
    return (int)hipGetLastError();
}
