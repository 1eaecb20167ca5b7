// Instantiation unit: DPP-row Cholesky family, float, 32 < n <= 64.
#include "kb_rg4_impl.h"

extern "C" int kb_rg_chol_s64(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                               int batch, hipStream_t s)
{
    return rg4_dispatch<float>(mode, n, nrhs, side_left, A, B, info, batch, s);
}
