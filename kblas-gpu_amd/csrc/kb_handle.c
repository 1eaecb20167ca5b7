/* kb_handle.c -- KBLAS handle, streams, timer, workspace, error strings.
 *
 * Behaviour follows the reference's runtime layer:
 *   kblasCreate / kblasDestroy        src/kblas_common.cu:35-51
 *   stream get/set, CreateStreams     src/kblas_struct.h:311-456
 *   workspace grow-only allocation    src/kblas_struct.h:206-269 (KBlasWorkspace::allocate)
 *   timer (two events, seconds)       src/kblas_gpu_timer.h:29-67
 *   error strings                     src/kblas_common.cu:170-202
 * All streams are created at the default priority.
 */
#include <stdlib.h>
#include <string.h>

#include "kb_internal.h"

int kb_hip_status(int hip_err) { return hip_err == 0 ? KBLAS_Success : KBLAS_UnknownError; }

int kblasCreate(kblasHandle_t* handle)
{
    if (!handle) return KBLAS_Error_WrongInput;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { *handle = NULL; return KBLAS_CUDA_Error; }
    struct KBlasHandle* h = (struct KBlasHandle*)calloc(1, sizeof(struct KBlasHandle));
    if (!h) return KBLAS_Error_Allocation;
    h->device = dev;
    h->stream = 0;
    /* this device's copy of the fixed scratch area of kblasDposv_vbatch_noalloc;
     * a failed lookup is reported by that entry point, which retries it */
    (void)kb_noalloc_init();
    kb_tune_init();
    *handle = h;
    return KBLAS_Success;
}

int kblasDestroy(kblasHandle_t* handle)
{
    if (!handle || !*handle) return KBLAS_Error_NotInitialized;
    struct KBlasHandle* h = *handle;
    kblasFreeWorkspace(h);
    if (h->timer_ready) {
        hipEventDestroy(h->ev_start);
        hipEventDestroy(h->ev_stop);
    }
    for (int i = 0; i < h->nstreams; ++i) hipStreamDestroy(h->streams[i]);
    free(h);
    *handle = NULL;
    return KBLAS_Success;
}

int kblasCreateStreams(kblasHandle_t h, int nStreams)
{
    KB_CHECK_HANDLE(h);
    if (nStreams < 0 || nStreams > KB_MAX_STREAMS) return KBLAS_Error_WrongInput;
    for (int i = h->nstreams; i < nStreams; ++i) {
        if (hipStreamCreateWithFlags(&h->streams[i], hipStreamNonBlocking) != hipSuccess)
            return KBLAS_CUDA_Error;
        h->nstreams = i + 1;
    }
    if (h->nstreams > 0) h->stream = h->streams[0];
    return KBLAS_Success;
}

hipStream_t kblasGetStream(kblasHandle_t h) { return h ? h->stream : 0; }
void kblasSetStream(kblasHandle_t h, hipStream_t stream) { if (h) h->stream = stream; }
void* kblasGetCublasHandle(kblasHandle_t h) { (void)h; return NULL; }
int kblasEnableMagma(kblasHandle_t h) { KB_CHECK_HANDLE(h); return KBLAS_Success; }

static int timer_init(kblasHandle_t h)
{
    if (h->timer_ready) return 1;
    if (hipEventCreate(&h->ev_start) != hipSuccess) return 0;
    if (hipEventCreate(&h->ev_stop) != hipSuccess) { hipEventDestroy(h->ev_start); return 0; }
    h->timer_ready = 1;
    return 1;
}

void kblasTimerTic(kblasHandle_t h)
{
    if (!h || !timer_init(h)) return;
    hipEventRecord(h->ev_start, h->stream);
}

void kblasTimerRecordEnd(kblasHandle_t h)
{
    if (!h || !timer_init(h)) return;
    hipEventRecord(h->ev_stop, h->stream);
}

double kblasTimerToc(kblasHandle_t h)
{
    if (!h || !timer_init(h)) return 0.0;
    hipEventRecord(h->ev_stop, h->stream);
    hipEventSynchronize(h->ev_stop);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, h->ev_start, h->ev_stop);
    return (double)ms / 1000.0;
}

const char* kblasGetErrorString(int error)
{
    switch (error) {
    case KBLAS_Success: return "success";
    case KBLAS_UnknownError: return "unknown error";
    case KBLAS_NotSupported: return "not supported";
    case KBLAS_NotImplemented: return "not implemented";
    case KBLAS_cuBLAS_Error: return "vendor BLAS error";
    case KBLAS_WrongConfig: return "wrong configuration";
    case KBLAS_CUDA_Error: return "device runtime (HIP) error";
    case KBLAS_InsufficientWorkspace: return "insufficient workspace supplied to function";
    case KBLAS_Error_Allocation: return "error allocating memory";
    case KBLAS_Error_Deallocation: return "error de-allocating memory";
    case KBLAS_Error_NotInitialized: return "handle not initialized";
    case KBLAS_Error_WrongInput: return "one of the input arguments is wrong";
    case KBLAS_MAGMA_Error: return "MAGMA error";
    case KBLAS_SVD_NoConvergence: return "SVD did not converge";
    default: return "unknown KBLAS error code";
    }
}

/* ---------------- workspace ---------------- */
static size_t zmax(size_t a, size_t b) { return a > b ? a : b; }

void kb_ws_request(kblasHandle_t h, const kb_ws_state* need)
{
    h->requested.h_data = zmax(h->requested.h_data, need->h_data);
    h->requested.h_ptrs = zmax(h->requested.h_ptrs, need->h_ptrs);
    h->requested.d_data = zmax(h->requested.d_data, need->d_data);
    h->requested.d_ptrs = zmax(h->requested.d_ptrs, need->d_ptrs);
}

int kb_ws_sufficient(const kb_ws_state* need, const kb_ws_state* have)
{
    return need->h_data <= have->h_data && need->h_ptrs <= have->h_ptrs &&
           need->d_data <= have->d_data && need->d_ptrs <= have->d_ptrs;
}

int kblasAllocateWorkspace(kblasHandle_t h)
{
    KB_CHECK_HANDLE(h);
    kb_ws_state* r = &h->requested;
    kb_ws_state* a = &h->allocated;
    if (r->h_data > a->h_data) {
        if (h->h_data) hipHostFree(h->h_data);
        h->h_data = NULL; a->h_data = 0;
        if (hipHostMalloc(&h->h_data, r->h_data, 0) != hipSuccess) return KBLAS_Error_Allocation;
        a->h_data = r->h_data;
    }
    if (r->h_ptrs > a->h_ptrs) {
        if (h->h_ptrs) hipHostFree(h->h_ptrs);
        h->h_ptrs = NULL; a->h_ptrs = 0;
        if (hipHostMalloc(&h->h_ptrs, r->h_ptrs, 0) != hipSuccess) return KBLAS_Error_Allocation;
        a->h_ptrs = r->h_ptrs;
    }
    if (r->d_data > a->d_data) {
        if (h->d_data) hipFree(h->d_data);
        h->d_data = NULL; a->d_data = 0;
        if (hipMalloc(&h->d_data, r->d_data) != hipSuccess) return KBLAS_Error_Allocation;
        a->d_data = r->d_data;
    }
    if (r->d_ptrs > a->d_ptrs) {
        if (h->d_ptrs) hipFree(h->d_ptrs);
        h->d_ptrs = NULL; a->d_ptrs = 0;
        if (hipMalloc(&h->d_ptrs, r->d_ptrs) != hipSuccess) return KBLAS_Error_Allocation;
        a->d_ptrs = r->d_ptrs;
    }
    return KBLAS_Success;
}

int kblasFreeWorkspace(kblasHandle_t h)
{
    KB_CHECK_HANDLE(h);
    int ok = 1;
    if (h->h_data && hipHostFree(h->h_data) != hipSuccess) ok = 0;
    if (h->h_ptrs && hipHostFree(h->h_ptrs) != hipSuccess) ok = 0;
    if (h->d_data && hipFree(h->d_data) != hipSuccess) ok = 0;
    if (h->d_ptrs && hipFree(h->d_ptrs) != hipSuccess) ok = 0;
    h->h_data = h->h_ptrs = h->d_data = h->d_ptrs = NULL;
    memset(&h->allocated, 0, sizeof(h->allocated));
    return ok ? KBLAS_Success : KBLAS_Error_Deallocation;
}
