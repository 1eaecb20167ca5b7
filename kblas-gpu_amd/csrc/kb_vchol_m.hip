// Instantiation unit: non-uniform persistent Cholesky family (double, mixed).
#include "kb_vchol_impl.h"

// FP32 workspace slot of a workgroup of the mixed phase for orders above the
// register range: the factor (max_order^2 floats) and the right-hand sides
// (max_order * max_nrhs floats), rounded up to 256 bytes.
static long vsposv_slot_floats(int max_order, int max_nrhs)
{
    const long f = (long)max_order * max_order + (long)max_order * max_nrhs;
    return (f + 63) & ~63L;
}

static size_t vsposv_queue_bytes(int batch) { return (kb_vchol_ws_bytes(batch) + 255) & ~(size_t)255; }

static int clamp_order(int max_order) { return max_order < KB_SB_MAX_N ? max_order : KB_SB_MAX_N; }
static int clamp_rhs(int max_nrhs) { return max_nrhs < KB_SB_MAX_RHS ? max_nrhs : KB_SB_MAX_RHS; }

extern "C" size_t kb_vsposv_ws_bytes(int batch, int max_order, int max_nrhs)
{
    const size_t q = vsposv_queue_bytes(batch);
    if (batch < 1 || max_order <= KB_REG_MAX_N || max_nrhs <= 0) return q;
    const long sf = vsposv_slot_floats(clamp_order(max_order), clamp_rhs(max_nrhs));
    return q + (size_t)vsposv_grid(batch) * (size_t)sf * sizeof(float);
}

extern "C" int kb_launch_vsposv(const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                                int* info, int* iter, int batch, void* ws, size_t ws_bytes, int max_order,
                                int max_nrhs, hipStream_t s)
{
    if (batch < 1) return 0;
    if (!ws) return (int)hipErrorInvalidValue;
    float* slots = nullptr;
    long sf = 0;
    const int mo = clamp_order(max_order), mr = clamp_rhs(max_nrhs);
    if (max_order > KB_REG_MAX_N && max_nrhs > 0 && ws_bytes >= kb_vsposv_ws_bytes(batch, max_order, max_nrhs)) {
        slots = (float*)((char*)ws + vsposv_queue_bytes(batch));
        sf = vsposv_slot_floats(mo, mr);
    }
    return launch_vchol_mixed<double, KB_MODE_SPOSV>(vm, vn, side_left, A, B, X, info, iter, batch, ws, slots, sf,
                                                     mo, mr, s);
}

extern "C" int kb_vchol_m(int mode, const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B,
                          kb_mat_t X, int* info, int* iter, int batch, void* ws, hipStream_t s)
{
    if (mode == KB_MODE_SPOSV)
        return kb_launch_vsposv(vm, vn, side_left, A, B, X, info, iter, batch, ws, 0, 0, 0, s);
    return (int)hipErrorNotSupported;
}
