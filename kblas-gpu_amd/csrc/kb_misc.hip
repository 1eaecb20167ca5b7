// kb_misc.hip -- small utility kernels.
#include "kb_device.h"
#include "kb_launch.h"

__global__ void k_fill_int(int* p, int v, long count)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

extern "C" int kb_launch_fill_int(int* p, int v, long count, hipStream_t s)
{
    if (count <= 0) return 0;
    long blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(256), 0, s, p, v, count);
    return (int)hipGetLastError();
}
