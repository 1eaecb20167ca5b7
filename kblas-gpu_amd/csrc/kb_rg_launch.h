// kb_rg_launch.h -- launchers for the DPP-row Cholesky family (kb_rg_impl.h).
#pragma once
#include "kb_rg_impl.h"

// one wavefront (four matrices) per workgroup: the most even spread of a
// batch over the CUs (the n = 16 POSV of 100,000 systems ran 4 % faster than
// with four wavefronts per workgroup, the n = 32 POTRF the same)
constexpr int KB_RG_WPB = 1;

template <typename T, int NB, int MODE, int RC>
static int rg_launch(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch,
                     hipStream_t s)
{
    constexpr int MPB = 4 * KB_RG_WPB;
    dim3 grid((unsigned)((batch + MPB - 1) / MPB)), block(64 * KB_RG_WPB);
    if (n == NB)
        hipLaunchKernelGGL((k_rg_chol<T, NB, MODE, RC, true, KB_RG_WPB>), grid, block, 0, s, n, nrhs, side_left,
                           to_arg<T>(A), to_arg<T>(B), info, batch);
    else
        hipLaunchKernelGGL((k_rg_chol<T, NB, MODE, RC, false, KB_RG_WPB>), grid, block, 0, s, n, nrhs, side_left,
                           to_arg<T>(A), to_arg<T>(B), info, batch);
    return (int)hipGetLastError();
}

template <typename T, int NB>
static int rg_dispatch(int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, int* info, int batch,
                       hipStream_t s)
{
    if (mode == KB_MODE_POTRF) return rg_launch<T, NB, KB_MODE_POTRF, 1>(n, nrhs, side_left, A, B, info, batch, s);
    if (mode == KB_MODE_POSV) {
        if (nrhs <= 1) return rg_launch<T, NB, KB_MODE_POSV, 1>(n, nrhs, side_left, A, B, info, batch, s);
        return rg_launch<T, NB, KB_MODE_POSV, 4>(n, nrhs, side_left, A, B, info, batch, s);
    }
    if (nrhs <= 1) return rg_launch<T, NB, KB_MODE_POTRS, 1>(n, nrhs, side_left, A, B, info, batch, s);
    return rg_launch<T, NB, KB_MODE_POTRS, 4>(n, nrhs, side_left, A, B, info, batch, s);
}
