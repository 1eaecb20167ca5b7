// kb_stubs.hip -- launchers whose kernels are not built yet report "not supported".
#include "kb_launch.h"

extern "C" int kb_launch_blk_chol(char, int, int, int, int, kb_mat_t, kb_mat_t, int*, int, hipStream_t)
{ return (int)hipErrorNotSupported; }
extern "C" int kb_launch_blk_sposv(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t)
{ return (int)hipErrorNotSupported; }
extern "C" int kb_launch_gemm(char, int, int, int, int, int, double, kb_mat_t, kb_mat_t, double, kb_mat_t,
                              const int*, const int*, const int*, int, hipStream_t)
{ return (int)hipErrorNotSupported; }
extern "C" int kb_launch_syrk(char, int, int, int, double, kb_mat_t, double, kb_mat_t, const int*, const int*,
                              int, hipStream_t)
{ return (int)hipErrorNotSupported; }
extern "C" int kb_launch_trsm(char, int, int, int, int, int, double, kb_mat_t, kb_mat_t, const int*, const int*,
                              int, hipStream_t)
{ return (int)hipErrorNotSupported; }
extern "C" size_t kb_vchol_ws_bytes(int batch) { return (size_t)batch * 4 + 4096; }
extern "C" int kb_launch_vchol(char, int, const int*, const int*, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*,
                               int, int, void*, hipStream_t)
{ return (int)hipErrorNotSupported; }
