// kb_noalloc.hip -- a fixed device-resident scratch area for the non-uniform
// Cholesky queue, so that kblasDposv_vbatch_noalloc can run without a
// workspace allocation (callers under graph capture or in loops that must not
// call hipMalloc).
//
// The area is a file-scope __device__ array: every GPU the library runs on
// holds its own copy for the life of the process.  Host code never passes the
// array's name to a kernel (in host code it is only a stand-in symbol); the
// device address of each GPU's copy is looked up once per device with
// hipGetSymbolAddress and kept here.  All calls on one GPU share that GPU's
// copy, so calls that overlap on one GPU (two streams, two host threads) race;
// each call clears the queue header with a kernel before use.
#include <mutex>

#include <hip/hip_runtime.h>

#include "kb_launch.h"

#define KB_NOALLOC_MAX_DEVICES 64

// queue header + one int per matrix (kb_vchol_ws_bytes(KB_NOALLOC_MAX_BATCH))
static __device__ __align__(256) int scratch[KB_VCHOL_WS_HDR_INTS + KB_NOALLOC_MAX_BATCH];

static std::mutex g_mu;
static void* g_ptr[KB_NOALLOC_MAX_DEVICES];

// Look up this device's copy of the scratch array (the current device).
// Returns hipSuccess, or the lookup's error.
extern "C" int kb_noalloc_init(void)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= KB_NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev]) return (int)hipSuccess;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    g_ptr[dev] = p;
    return (int)hipSuccess;
}

// Device address of the current device's scratch copy (looked up on first use
// when kblasCreate ran on another device), or NULL with *err set.
extern "C" void* kb_noalloc_scratch(int* err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= KB_NOALLOC_MAX_DEVICES) {
        *err = (e != hipSuccess) ? (int)e : (int)hipErrorInvalidDevice;
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_ptr[dev]) {
            *err = 0;
            return g_ptr[dev];
        }
    }
    *err = kb_noalloc_init();
    if (*err) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[dev];
}
