// kb_reg.hip -- size dispatch for the register-resident kernels (n <= 64).
#include "kb_launch.h"

extern "C" int kb_rg_chol_d16(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_rg_chol_d32(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_rg_chol_s16(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_rg_chol_s32(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_rg_chol_d64(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_rg_chol_s64(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_reg_sposv_16(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);
extern "C" int kb_reg_sposv_32(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);
extern "C" int kb_tile_chol_d(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_tile_chol_s(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_tile_sposv(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);

// Four matrices per wavefront, one per 16-lane DPP row: n <= 16 and n <= 32
// whole in VGPRs; POTRF with 32 < n <= 48 in 16-column blocks.  Otherwise
// (32 < n <= 64) the MFMA-tiled kernel, one matrix per wavefront, which keeps
// the factor in registers for the solves (measured: the blocked DPP-row kernel
// re-reads its finished columns and loses once a solve follows or n > 48).
extern "C" int kb_launch_reg_chol(char prec, int mode, int n, int nrhs, int side_left,
                                  kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    if (prec == 'd') {
        if (n <= 16) return kb_rg_chol_d16(mode, n, nrhs, side_left, A, B, info, batch, s);
        if (n <= 32) return kb_rg_chol_d32(mode, n, nrhs, side_left, A, B, info, batch, s);
        if (n <= 48 && mode == KB_MODE_POTRF) return kb_rg_chol_d64(mode, n, nrhs, side_left, A, B, info, batch, s);
        return kb_tile_chol_d(mode, n, nrhs, side_left, A, B, info, batch, s);
    }
    if (n <= 16) return kb_rg_chol_s16(mode, n, nrhs, side_left, A, B, info, batch, s);
    if (n <= 32) return kb_rg_chol_s32(mode, n, nrhs, side_left, A, B, info, batch, s);
    if (n <= 48 && mode == KB_MODE_POTRF) return kb_rg_chol_s64(mode, n, nrhs, side_left, A, B, info, batch, s);
    return kb_tile_chol_s(mode, n, nrhs, side_left, A, B, info, batch, s);
}

extern "C" int kb_launch_reg_chol_v(char prec, int mode, int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B,
                                    int* info, int batch, int variant, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    const bool d = prec == 'd';
    switch (variant) {
    case KB_VAR_AUTO:
        return kb_launch_reg_chol(prec, mode, n, nrhs, side_left, A, B, info, batch, s);
    case KB_VAR_RG16:
        if (n > 16) break;
        return d ? kb_rg_chol_d16(mode, n, nrhs, side_left, A, B, info, batch, s)
                 : kb_rg_chol_s16(mode, n, nrhs, side_left, A, B, info, batch, s);
    case KB_VAR_RG32:
        if (n > 32) break;
        return d ? kb_rg_chol_d32(mode, n, nrhs, side_left, A, B, info, batch, s)
                 : kb_rg_chol_s32(mode, n, nrhs, side_left, A, B, info, batch, s);
    case KB_VAR_RG4:
        if (n <= 32 || mode != KB_MODE_POTRF) break;
        return d ? kb_rg_chol_d64(mode, n, nrhs, side_left, A, B, info, batch, s)
                 : kb_rg_chol_s64(mode, n, nrhs, side_left, A, B, info, batch, s);
    case KB_VAR_TILE:
        return d ? kb_tile_chol_d(mode, n, nrhs, side_left, A, B, info, batch, s)
                 : kb_tile_chol_s(mode, n, nrhs, side_left, A, B, info, batch, s);
    default:
        break;
    }
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_reg_sposv_v(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                                     int* info, int* iter, int batch, int variant, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    if (variant == KB_VAR_REG && n <= 32)
        return n <= 16 ? kb_reg_sposv_16(n, nrhs, side_left, A, B, X, info, iter, batch, s)
                       : kb_reg_sposv_32(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    if (variant == KB_VAR_TILE) return kb_tile_sposv(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    if (variant == KB_VAR_AUTO) return kb_launch_reg_sposv(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    return (int)hipErrorInvalidValue;
}

extern "C" int kb_launch_reg_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B,
                                   kb_mat_t X, int* info, int* iter, int batch, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    if (n <= 16) return kb_reg_sposv_16(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    if (n <= 32) return kb_reg_sposv_32(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    return kb_tile_sposv(n, nrhs, side_left, A, B, X, info, iter, batch, s);
}

extern "C" int kb_blk_chol_d(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
extern "C" int kb_blk_chol_s(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);

extern "C" int kb_launch_blk_chol(char prec, int mode, int n, int nrhs, int side_left,
                                  kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s)
{
    if (n < 1 || batch < 1) return (int)hipErrorInvalidValue;
    if (prec == 'd') return kb_blk_chol_d(mode, n, nrhs, side_left, A, B, info, batch, s);
    return kb_blk_chol_s(mode, n, nrhs, side_left, A, B, info, batch, s);
}

extern "C" int kb_vchol_d(int, const int*, const int*, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, void*, int, int,
                          hipStream_t);
extern "C" int kb_vchol_s(int, const int*, const int*, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, void*, int, int,
                          hipStream_t);
extern "C" int kb_vchol_m(int, const int*, const int*, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, void*, hipStream_t);

// hist + offsets + counters + the order array (see kb_vchol_impl.h)
extern "C" size_t kb_vchol_ws_bytes(int batch)
{
    return sizeof(int) * ((size_t)KB_VCHOL_WS_HDR_INTS + (size_t)(batch > 0 ? batch : 0));
}

extern "C" int kb_launch_vchol_split(char prec, int mode, const int* vm, const int* vn, int side_left,
                                     kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info, int* iter, int batch,
                                     int max_nrhs, void* ws, int split, hipStream_t s)
{
    if (batch < 1) return 0;
    if (!ws || split < 0 || split > KB_REG_MAX_N) return (int)hipErrorInvalidValue;
    if (prec == 'm') return kb_vchol_m(mode, vm, vn, side_left, A, B, X, info, iter, batch, ws, s);
    if (prec == 'd') return kb_vchol_d(mode, vm, vn, side_left, A, B, X, info, iter, batch, ws, split, max_nrhs, s);
    return kb_vchol_s(mode, vm, vn, side_left, A, B, X, info, iter, batch, ws, split, max_nrhs, s);
}

extern "C" int kb_launch_vchol(char prec, int mode, const int* vm, const int* vn, int side_left,
                               kb_mat_t A, kb_mat_t B, kb_mat_t X, int* info, int* iter, int batch,
                               int max_nrhs, void* ws, hipStream_t s)
{
    return kb_launch_vchol_split(prec, mode, vm, vn, side_left, A, B, X, info, iter, batch, max_nrhs, ws,
                                 KB_REG_MAX_N, s);
}
