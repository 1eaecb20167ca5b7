// kb_reg.hip -- size dispatch for the register-resident kernels (n <= 64).
#include "kb_launch.h"

#define KB_DECL_CHOL(p, nb) \
    extern "C" int kb_reg_chol_##p##nb(int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t);
KB_DECL_CHOL(d, 16) KB_DECL_CHOL(d, 32) KB_DECL_CHOL(d, 64)
KB_DECL_CHOL(s, 16) KB_DECL_CHOL(s, 32) KB_DECL_CHOL(s, 64)
extern "C" int kb_reg_sposv_16(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);
extern "C" int kb_reg_sposv_32(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);
extern "C" int kb_reg_sposv_64(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t);

extern "C" int kb_launch_reg_chol(char prec, int mode, int n, int nrhs, int side_left,
                                  kb_mat_t A, kb_mat_t B, int* info, int batch, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    if (prec == 'd') {
        if (n <= 16) return kb_reg_chol_d16(mode, n, nrhs, side_left, A, B, info, batch, s);
        if (n <= 32) return kb_reg_chol_d32(mode, n, nrhs, side_left, A, B, info, batch, s);
        return kb_reg_chol_d64(mode, n, nrhs, side_left, A, B, info, batch, s);
    }
    if (n <= 16) return kb_reg_chol_s16(mode, n, nrhs, side_left, A, B, info, batch, s);
    if (n <= 32) return kb_reg_chol_s32(mode, n, nrhs, side_left, A, B, info, batch, s);
    return kb_reg_chol_s64(mode, n, nrhs, side_left, A, B, info, batch, s);
}

extern "C" int kb_launch_reg_sposv(int n, int nrhs, int side_left, kb_mat_t A, kb_mat_t B,
                                   kb_mat_t X, int* info, int* iter, int batch, hipStream_t s)
{
    if (n < 1 || n > KB_REG_MAX_N || batch < 1) return (int)hipErrorInvalidValue;
    if (n <= 16) return kb_reg_sposv_16(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    if (n <= 32) return kb_reg_sposv_32(n, nrhs, side_left, A, B, X, info, iter, batch, s);
    return kb_reg_sposv_64(n, nrhs, side_left, A, B, X, info, iter, batch, s);
}
