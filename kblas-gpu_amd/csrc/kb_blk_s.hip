// Instantiation unit: blocked (n > 64) Cholesky family, float.
#include "kb_blk_impl.h"

extern "C" int kb_blk_chol_s(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                             int batch, hipStream_t s)
{
    return launch_blk<float>(mode, n, nrhs, side_left, A, B, info, batch, s);
}

extern "C" int kb_blk_trsm_s(int fwd, int side_left, int unit, int n, int nrhs, double alpha, kb_mat_t A,
                              kb_mat_t B, const int* vm, const int* vn, int batch, hipStream_t s)
{
    return launch_blk_trsm<float>(fwd, side_left, unit, n, nrhs, (float)alpha, A, B, vm, vn, batch, s);
}
