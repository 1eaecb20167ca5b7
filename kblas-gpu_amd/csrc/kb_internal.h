/* kb_internal.h -- handle and workspace internals of the host library (C). */
#ifndef KB_INTERNAL_H
#define KB_INTERNAL_H

#include <stddef.h>
#include <stdio.h>

#include "kblas.h"
#include "kb_launch.h"

#define KB_MAX_STREAMS 16

/* Workspace request / allocation (reference: KBlasWorkspaceState,
 * src/kblas_struct.h:43-91): host data, host pointers, device data, device
 * pointers, in bytes.  Requests accumulate as the max over *_wsquery calls. */
typedef struct {
    size_t h_data, h_ptrs, d_data, d_ptrs;
} kb_ws_state;

struct KBlasHandle {
    int device;
    hipStream_t stream;
    hipStream_t streams[KB_MAX_STREAMS];
    int nstreams;
    hipEvent_t ev_start, ev_stop;
    int timer_ready;
    kb_ws_state requested;
    kb_ws_state allocated;
    void* h_data;
    void* h_ptrs;
    void* d_data;
    void* d_ptrs;
};

/* sizes needed by each operation (0 for every uniform-size path) */
void kb_ws_need_vbatch(kb_ws_state* ws, int batchCount);
int  kb_ws_sufficient(const kb_ws_state* need, const kb_ws_state* have);
void kb_ws_request(kblasHandle_t h, const kb_ws_state* need);

/* per-size-bucket kernel choice (kb_tune.c); operation ids follow KB_MODE_* */
enum { KB_TUNE_POTRF = 0, KB_TUNE_POSV = 1, KB_TUNE_POTRS = 2, KB_TUNE_SPOSV = 3, KB_TUNE_VSPLIT = 4 };
void kb_tune_init(void);
int  kb_tune_variant(char prec, int op, int n);
int  kb_tune_legal(int op, int n, int variant);
int  kb_tune_op_id(const char* op);

/* status helpers */
int kb_hip_status(int hip_err);

#define KB_CHECK_HANDLE(h) do { if (!(h)) return KBLAS_Error_NotInitialized; } while (0)

#endif
