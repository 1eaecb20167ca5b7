// Instantiation unit: MFMA-tiled Cholesky family (32 < n <= 64), float.
#include "kb_tile_impl.h"

extern "C" int kb_tile_chol_s(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info,
                              int batch, hipStream_t s)
{
    return launch_tile<float>(mode, n, nrhs, side_left, A, B, info, batch, s);
}
