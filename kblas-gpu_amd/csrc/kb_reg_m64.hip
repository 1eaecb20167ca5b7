// Instantiation unit: register-resident mixed-precision POSV, padded order 64.
#include "kb_reg_impl.h"

extern "C" int kb_reg_sposv_64(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                               int* info, int* iter, int batch, hipStream_t s)
{
    return sposv_rc<64>(n, nrhs, side_left, A, B, X, info, iter, batch, s);
}
