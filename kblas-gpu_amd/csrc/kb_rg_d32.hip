// Instantiation unit: DPP-row Cholesky family, double, padded order 32.
#include "kb_rg_launch.h"

extern "C" int kb_rg_chol_d32(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                               int batch, hipStream_t s)
{
    return rg_dispatch<double, 32>(mode, n, nrhs, side_left, A, B, info, batch, s);
}
