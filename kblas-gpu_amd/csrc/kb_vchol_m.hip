// Instantiation unit: non-uniform persistent Cholesky family (double, mixed).
#include "kb_vchol_impl.h"

extern "C" int kb_vchol_m(int mode, const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B,
                          kb_mat_t X, int* info, int* iter, int batch, void* ws, hipStream_t s)
{
    if (mode == KB_MODE_SPOSV) return launch_vchol_mixed<double, KB_MODE_SPOSV>(vm, vn, side_left, A, B, X, info, iter, batch, ws, s);
    return (int)hipErrorNotSupported;
}
