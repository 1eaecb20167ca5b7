// Instantiation unit: DPP-row Cholesky family, float, padded order 16.
#include "kb_rg_launch.h"

extern "C" int kb_rg_chol_s16(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                               int batch, hipStream_t s)
{
    return rg_dispatch<float, 16>(mode, n, nrhs, side_left, A, B, info, batch, s);
}
