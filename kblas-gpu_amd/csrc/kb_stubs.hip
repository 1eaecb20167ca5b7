// kb_stubs.hip -- launchers whose kernels are not built yet report "not supported".
#include "kb_launch.h"

extern "C" int kb_launch_blk_sposv(int, int, int, kb_mat_t, kb_mat_t, kb_mat_t, int*, int*, int, hipStream_t)
{ return (int)hipErrorNotSupported; }
