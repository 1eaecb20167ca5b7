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

// Streaming copy of bytes (a multiple of 16) with 16-byte vector loads and
// stores, four in flight per thread per step: the achievable-bandwidth
// reference the benchmark sets its HBM fractions against.  Not part of the
// KBLAS interface.
__global__ void __launch_bounds__(256) k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     long n16)
{
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

extern "C" int kb_stream_copy(void* dst, const void* src, long bytes, int blocks, hipStream_t s)
{
    if (bytes <= 0 || (bytes & 15) || blocks < 1) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst,
                       bytes / 16);
    return (int)hipGetLastError();
}
