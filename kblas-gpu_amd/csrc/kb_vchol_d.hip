// Instantiation unit: non-uniform persistent Cholesky family (double).
#include "kb_vchol_impl.h"

extern "C" int kb_vchol_d(int mode, const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B,
                          kb_mat_t X, int* info, int* iter, int batch, void* ws, int split, int max_nrhs,
                          hipStream_t s)
{
    if (mode == KB_MODE_POTRF) return launch_vchol<double, KB_MODE_POTRF>(vm, vn, side_left, A, B, X, info, iter, batch, ws, split, max_nrhs, s);
    if (mode == KB_MODE_POSV) return launch_vchol<double, KB_MODE_POSV>(vm, vn, side_left, A, B, X, info, iter, batch, ws, split, max_nrhs, s);
    if (mode == KB_MODE_POTRS) return launch_vchol<double, KB_MODE_POTRS>(vm, vn, side_left, A, B, X, info, iter, batch, ws, split, max_nrhs, s);
    return (int)hipErrorNotSupported;
}
