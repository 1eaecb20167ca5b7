/*
 * kblas.h -- KBLAS handle and workspace API, MI355X (gfx950) build.
 *
 * Drop-in for the reference's include/kblas.h:35-108.  The only type change is
 * that streams are hipStream_t (the reference: cudaStream_t) and the vendor
 * BLAS handle is an opaque pointer (this library has no BLAS dependency).
 * Only the batched SPD path is provided (kblas_batch.h); the reference's
 * BLAS-2/3, SVD/QR and TLR headers are not part of this library.
 */
#ifndef KBLAS_H_
#define KBLAS_H_

#include <hip/hip_runtime_api.h>

#include "kblas_defs.h"

#ifdef __cplusplus
extern "C" {
#endif

struct KBlasHandle;
typedef struct KBlasWorkspace* kblasWorkspace_t;
typedef struct KBlasHandle* kblasHandle_t;

/* include/kblas.h:54 -- create a handle on the current device, stream 0. */
int kblasCreate(kblasHandle_t* handle);
/* include/kblas.h:59 -- free workspace, events and owned streams. */
int kblasDestroy(kblasHandle_t* handle);
/* include/kblas.h:60-62 -- event timer on the handle's stream; Toc returns seconds. */
void kblasTimerTic(kblasHandle_t handle);
void kblasTimerRecordEnd(kblasHandle_t handle);
double kblasTimerToc(kblasHandle_t handle);
/* include/kblas.h:67 -- create nStreams streams owned by the handle (default priority). */
int kblasCreateStreams(kblasHandle_t handle, int nStreams);
/* include/kblas.h:71-76 */
hipStream_t kblasGetStream(kblasHandle_t handle);
void kblasSetStream(kblasHandle_t handle, hipStream_t stream);
/* include/kblas.h:81 -- returns NULL: there is no vendor BLAS handle. */
void* kblasGetCublasHandle(kblasHandle_t handle);
/* include/kblas.h:86 -- accepted and ignored (non-uniform batches are native):
 * returns KBLAS_Success. */
int kblasEnableMagma(kblasHandle_t handle);
/* include/kblas.h:92 */
const char* kblasGetErrorString(int error);
/* include/kblas.h:102-108 -- grow-only allocation of what the *_wsquery calls asked for. */
int kblasAllocateWorkspace(kblasHandle_t handle);
int kblasFreeWorkspace(kblasHandle_t handle);

#ifdef __cplusplus
}
#endif

#include "kblas_batch.h"
#include "kblas_batch_ext.h"

#endif /* KBLAS_H_ */
