// kb_vchol_impl.h -- non-uniform batched Cholesky family over a size-sorted queue.
//
// The reference has no non-uniform POTRF/POSV; its non-uniform TRSM/SYRK go
// through MAGMA with a host round trip for the maximum sizes
// (src/batch_triangular/Xtrsm_batch_drivers.cuh:315-320).  Here:
//   1. a device counting sort orders the matrix ids by order n, largest
//      first (clear -> histogram -> scan -> scatter; no host synchronisation,
//      no allocation, fills done by kernels so the sequence can be captured in
//      a graph); the scan also records how many orders exceed the register
//      range;
//   2. a persistent kernel of 256-thread workgroups pulls the tickets of those
//      large matrices from the front of the queue (one workgroup per matrix,
//      blocked algorithm, largest first);
//   3. a second persistent kernel, queued behind it, gives every wavefront its
//      own tickets from the rest of the queue (register-resident routine, one
//      matrix per wavefront).  The two phases are separate kernels so each is
//      compiled and occupied for its own register and LDS needs.
#pragma once
#include "kb_blk_impl.h"
#include "kb_reg_impl.h"
#include "kb_sposv_big.h"

constexpr int KB_VQ_BINS = 4096;   // orders >= 4095 share the last bin
constexpr int KB_VQ_HDR = 2 * KB_VQ_BINS + 64;   // ints: hist, offsets, counters
static_assert(KB_VQ_HDR == KB_VCHOL_WS_HDR_INTS, "queue header size");
// counter[0]: next ticket of the workgroup phase; counter[1]: number of orders
// above the register range (tickets [0, counter[1])); counter[2]: next ticket
// of the wavefront phase (starts at counter[1])

struct VQueue {
    int* hist;
    int* offs;
    int* counter;
    int* order;
};

KB_DEV VQueue vq_view(void* ws, int batch)
{
    int* w = (int*)ws;
    VQueue q;
    q.hist = w;
    q.offs = w + KB_VQ_BINS;
    q.counter = w + 2 * KB_VQ_BINS;
    q.order = w + KB_VQ_HDR;
    (void)batch;
    return q;
}

// order of A[i]: side_left -> m[i]; right side (and POTRF, which passes vm = null) -> n[i]
KB_DEV int vq_order(const int* vm, const int* vn, int side_left, long i)
{
    return (side_left && vm) ? vm[i] : vn[i];
}
KB_DEV int vq_nrhs(const int* vm, const int* vn, int side_left, long i)
{
    if (!vm) return 0;
    return side_left ? vn[i] : vm[i];
}
KB_DEV int vq_bin(int n) { return n < 0 ? 0 : (n >= KB_VQ_BINS ? KB_VQ_BINS - 1 : n); }

namespace {

__global__ void k_vq_clear(void* ws)
{
    int* w = (int*)ws;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < KB_VQ_HDR; i += gridDim.x * blockDim.x) w[i] = 0;
}

__global__ void k_vq_hist(void* ws, const int* vm, const int* vn, int side_left, int batch)
{
    VQueue q = vq_view(ws, batch);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < batch; i += (long)gridDim.x * blockDim.x)
        atomicAdd(&q.hist[vq_bin(vq_order(vm, vn, side_left, i))], 1);
}

// exclusive scan in descending order of n: offs[b] = sum_{b' > b} hist[b'];
// counter[1] = counter[2] = the number of orders above split
__global__ void __launch_bounds__(1024) k_vq_scan(void* ws, int batch, int split)
{
    VQueue q = vq_view(ws, batch);
    __shared__ int part[1024];
    constexpr int PER = KB_VQ_BINS / 1024;
    const int t = threadIdx.x;
    // thread t owns bins (descending) [B-1-t*PER, ..., B-1-(t*PER+PER-1)]
    int s = 0;
    for (int j = 0; j < PER; ++j) s += q.hist[KB_VQ_BINS - 1 - (t * PER + j)];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;   // exclusive prefix of this thread's chunk
    for (int j = 0; j < PER; ++j) {
        const int b = KB_VQ_BINS - 1 - (t * PER + j);
        q.offs[b] = run;
        if (b == split) {   // run = number of orders above the split
            q.counter[1] = run;
            q.counter[2] = run;
        }
        run += q.hist[b];
    }
}

__global__ void k_vq_scatter(void* ws, const int* vm, const int* vn, int side_left, int batch)
{
    VQueue q = vq_view(ws, batch);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < batch; i += (long)gridDim.x * blockDim.x) {
        const int pos = atomicAdd(&q.offs[vq_bin(vq_order(vm, vn, side_left, i))], 1);
        q.order[pos] = (int)i;
    }
}

}  // namespace

template <typename T, int MODE>
KB_DEV void v_small(long mat, int n, int nrhs, int side_left, BatchArg<T>& A, BatchArg<T>& B,
                    BatchArg<T>& X, int* info, int* iter, float* lt, int lane)
{
    if (n <= 0) {
        if (lane == 0 && info) info[mat] = 0;
        if (lane == 0 && iter) iter[mat] = 0;
        return;
    }
    if constexpr (MODE == KB_MODE_SPOSV) {
        reg_sposv_one<64, 64, 4>(n, nrhs, side_left, (double*)A.get(mat), A.ldim(mat), (double*)B.get(mat),
                                 B.ldim(mat), (double*)X.get(mat), X.ldim(mat), info ? info + mat : nullptr,
                                 iter ? iter + mat : nullptr, lt, lane);
    } else if constexpr (MODE == KB_MODE_POTRF) {
        reg_chol_one<T, 64, 64, MODE, 1>(n, 0, 0, A.get(mat), A.ldim(mat), nullptr, 1,
                                         info ? info + mat : nullptr, nullptr, lane);
    } else {
        // FP64 factor reads its transpose back from A; FP32 keeps it in LDS
        T* Lt = (sizeof(T) == 4) ? (T*)lt : nullptr;
        reg_chol_one<T, 64, 64, MODE, 4>(n, nrhs, side_left, A.get(mat), A.ldim(mat), B.get(mat), B.ldim(mat),
                                         info ? info + mat : nullptr, Lt, lane);
    }
}

// Workgroup phase: the matrices above the register range, one workgroup each
// (CB: passes of 16 CB right-hand sides, as k_blk_chol).
template <typename T, int MODE, int CB>
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(CB == 2 ? 4 : 2)))
k_vchol_blk(const int* vm, const int* vn, int side_left, BatchArg<T> A, BatchArg<T> B, BatchArg<T> X,
            int* info, int* iter, int batch, void* ws)
{
    __shared__ BlkSmem<T> S;
    __shared__ int s_ticket;
    VQueue q = vq_view(ws, batch);
    const int nbig = q.counter[1];
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(q.counter, 1);
        __syncthreads();
        const int ticket = s_ticket;
        __syncthreads();
        if (ticket >= nbig) return;
        const long mat = q.order[ticket];
        const int n = vq_order(vm, vn, side_left, mat);
        const int nrhs = vq_nrhs(vm, vn, side_left, mat);
        if constexpr (MODE == KB_MODE_SPOSV) {
            // Orders above the register range are solved in FP64 directly
            // (LAPACK dsposv's ITER = -1: full precision for implementation
            // reasons); A receives L as in the dsposv fallback.
            T* Ap = A.get(mat);
            const int lda = A.ldim(mat);
            T* Bp = B.get(mat);
            T* Xp = X.get(mat);
            const int ldb = B.ldim(mat), ldx = X.ldim(mat);
            copy_rhs<T>(Bp, ldb, Xp, ldx, side_left, n, nrhs, threadIdx.x, KB_BLK_WG);
            __syncthreads();
            const int inf = blk_potrf<T, CB>(Ap, lda, n, S, Xp, ldx, side_left, nrhs);
            if (threadIdx.x == 0) { if (info) info[mat] = inf; if (iter) iter[mat] = -1; }
            if (inf == 0) blk_solve_bwd<T, CB>(Ap, lda, n, Xp, ldx, side_left, nrhs, S);
        } else {
            T* Ap = A.get(mat);
            const int lda = A.ldim(mat);
            if constexpr (MODE == KB_MODE_POTRF) {
                const int inf = blk_potrf<T, CB>(Ap, lda, n, S);
                if (threadIdx.x == 0 && info) info[mat] = inf;
            } else if constexpr (MODE == KB_MODE_POSV) {
                T* Bp = B.get(mat);
                const int ldb = B.ldim(mat);
                const int inf = blk_potrf<T, CB>(Ap, lda, n, S, Bp, ldb, side_left, nrhs);
                if (threadIdx.x == 0 && info) info[mat] = inf;
                if (inf == 0) blk_solve_bwd<T, CB>(Ap, lda, n, Bp, ldb, side_left, nrhs, S);
            } else {
                blk_solve<T, CB>(Ap, lda, n, B.get(mat), B.ldim(mat), side_left, nrhs, S);
            }
        }
        __syncthreads();
    }
}

// Wavefront phase: the rest of the queue, one matrix per wavefront.  Only the
// FP32 and mixed routines use the per-wavefront LDS area (their transposed
// factor); the FP64 ones read it back from memory.
template <typename T, int MODE>
__global__ void __launch_bounds__(KB_BLK_WG)
k_vchol_small(const int* vm, const int* vn, int side_left, BatchArg<T> A, BatchArg<T> B, BatchArg<T> X,
              int* info, int* iter, int batch, void* ws)
{
    constexpr bool LT = sizeof(T) == 4 || MODE == KB_MODE_SPOSV;
    __shared__ float lts[4][LT ? 64 * 65 : 1];
    VQueue q = vq_view(ws, batch);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* lt = lts[wave];
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(q.counter + 2, 1);
        const int ticket = __builtin_amdgcn_readfirstlane(t);
        if (ticket >= batch) return;
        const long mat = q.order[ticket];
        const int n = vq_order(vm, vn, side_left, mat);
        const int nrhs = vq_nrhs(vm, vn, side_left, mat);
        v_small<T, MODE>(mat, n, nrhs, side_left, A, B, X, info, iter, lt, lane);
    }
}

template <typename K>
static int vchol_grid(K kernel, int batch)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    // persistent grid: as many workgroups as can be resident at once
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, KB_BLK_WG, 0) != hipSuccess || per_cu < 1)
        per_cu = 2;
    const int grid = cus * per_cu;
    return grid > batch ? batch : grid;
}

template <typename T, int MODE>
static int launch_vchol(const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                        int* info, int* iter, int batch, void* ws, int split, int max_nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_vq_clear, dim3(8), dim3(1024), 0, s, ws);
    const int gb = (batch + 255) / 256 < 1024 ? (batch + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_vq_hist, dim3(gb), dim3(256), 0, s, ws, vm, vn, side_left, batch);
    hipLaunchKernelGGL(k_vq_scan, dim3(1), dim3(1024), 0, s, ws, batch, split);
    hipLaunchKernelGGL(k_vq_scatter, dim3(gb), dim3(256), 0, s, ws, vm, vn, side_left, batch);
    const BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B), x = to_arg<T>(X);
    // passes of 32 right-hand sides (four waves per SIMD) when the bound allows
    if (MODE == KB_MODE_POTRF || (max_nrhs > 0 && max_nrhs <= KB_BLK_NARROW))
        hipLaunchKernelGGL((k_vchol_blk<T, MODE, 2>), dim3(vchol_grid(k_vchol_blk<T, MODE, 2>, batch)),
                           dim3(KB_BLK_WG), 0, s, vm, vn, side_left, a, b, x, info, iter, batch, ws);
    else
        hipLaunchKernelGGL((k_vchol_blk<T, MODE, 4>), dim3(vchol_grid(k_vchol_blk<T, MODE, 4>, batch)),
                           dim3(KB_BLK_WG), 0, s, vm, vn, side_left, a, b, x, info, iter, batch, ws);
    hipLaunchKernelGGL((k_vchol_small<T, MODE>), dim3(vchol_grid(k_vchol_small<T, MODE>, (batch + 3) / 4)),
                       dim3(KB_BLK_WG), 0, s, vm, vn, side_left, a, b, x, info, iter, batch, ws);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed precision (SPOSV).  Orders above the register range run the mixed
// workgroup routine of kb_sposv_big.h in k_vsposv_big (FP32 factor and solves
// in a workspace slot per workgroup, FP64 residuals on the matrix cores;
// without a slot, dsposv's full-precision path, ITER = -1).  The rest of the
// queue runs the mixed register routine in k_vchol_mixed: a persistent kernel of
// 256-thread workgroups that can also take the large orders itself (workgroup
// phase first; once the queue reaches the register range every wavefront
// takes its own tickets); with small_only set it skips that phase and its
// wavefronts draw from the tickets after the large orders.  (Splitting this
// mode into k_vchol_blk<SPOSV> + k_vchol_small<SPOSV> hung on the GPU, cause
// not found, so neither is built for it.)
// ---------------------------------------------------------------------------
// The workgroup phase of the non-uniform mixed POSV: tickets [0, counter[1])
// of the size-sorted queue (the orders above the register range), one matrix
// per workgroup; workgroup b owns slot b of the FP32 workspace (slots == null:
// no workspace for the mixed path, every order here takes the FP64 path).
namespace {
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(2)))
k_vsposv_big(const int* vm, const int* vn, int side_left, BatchArg<double> A, BatchArg<double> B,
             BatchArg<double> X, int* info, int* iter, int batch, void* ws, float* slots, long slot_floats,
             int max_n, int max_nrhs)
{
    __shared__ SbSmem S;
    __shared__ int s_ticket;
    VQueue q = vq_view(ws, batch);
    const int nbig = q.counter[1];
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(q.counter, 1);
        __syncthreads();
        const int ticket = s_ticket;
        __syncthreads();
        if (ticket >= nbig) return;
        const long mat = q.order[ticket];
        const int n = vq_order(vm, vn, side_left, mat);
        const int nrhs = vq_nrhs(vm, vn, side_left, mat);
        double* Ap = A.get(mat);
        const int lda = A.ldim(mat);
        const double* Bp = B.get(mat);
        const int ldb = B.ldim(mat);
        double* Xp = X.get(mat);
        const int ldx = X.ldim(mat);
        if (nrhs <= 0) {
            // nothing to solve (dsposv's quick return)
            if (threadIdx.x == 0) { if (info) info[mat] = 0; if (iter) iter[mat] = 0; }
        } else if (slots && n <= max_n && nrhs <= max_nrhs && n <= KB_SB_MAX_N && nrhs <= KB_SB_MAX_RHS) {
            float* A32 = slots + (long)blockIdx.x * slot_floats;
            float* Z = A32 + (long)max_n * max_n;
            sb_sposv_one(mat, n, nrhs, side_left, Ap, lda, Bp, ldb, Xp, ldx, A32, Z, info, iter, S);
        } else {
            // dsposv's full-precision path (ITER = -1); A receives L
            copy_rhs<double>(Bp, ldb, Xp, ldx, side_left, n, nrhs, threadIdx.x, KB_BLK_WG);
            __syncthreads();
            const int inf = blk_potrf<double>(Ap, lda, n, S.d, Xp, ldx, side_left, nrhs);
            if (threadIdx.x == 0) { if (info) info[mat] = inf; if (iter) iter[mat] = -1; }
            if (inf == 0) blk_solve_bwd<double>(Ap, lda, n, Xp, ldx, side_left, nrhs, S.d);
        }
        __syncthreads();
    }
}
}  // namespace

// Shared memory of the persistent kernel: the blocked workgroup routine or,
// in the register phase, one float LDS area per wavefront (only the FP32 and
// mixed routines use it; the FP64 ones read their transposed factor back from
// memory, and leaving the area out lets more workgroups share a CU).
template <typename T, bool LT>
union VSmem {
    BlkSmem<T> blk;
    float lt[4][LT ? 64 * 65 : 1];
};

template <typename T, int MODE>
__global__ void __launch_bounds__(KB_BLK_WG) __attribute__((amdgpu_waves_per_eu(2)))
k_vchol_mixed(const int* vm, const int* vn, int side_left, BatchArg<T> A, BatchArg<T> B, BatchArg<T> X,
        int* info, int* iter, int batch, void* ws, int small_only)
{
    __shared__ VSmem<T, (sizeof(T) == 4 || MODE == KB_MODE_SPOSV)> S;
    __shared__ int s_ticket;
    VQueue q = vq_view(ws, batch);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // ---- workgroup phase: matrices above the register range ----
    int ticket = -1;
    while (!small_only) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(q.counter, 1);
        __syncthreads();
        ticket = s_ticket;
        __syncthreads();
        if (ticket >= batch) return;
        const long mat = q.order[ticket];
        const int n = vq_order(vm, vn, side_left, mat);
        if (n <= KB_REG_MAX_N) break;     // sorted: every later ticket is small too
        const int nrhs = vq_nrhs(vm, vn, side_left, mat);
        if constexpr (MODE == KB_MODE_SPOSV) {
            // Orders above the register range are solved in FP64 directly
            // (LAPACK dsposv's ITER = -1: full precision for implementation
            // reasons); A receives L as in the dsposv fallback.
            T* Ap = A.get(mat);
            const int lda = A.ldim(mat);
            T* Bp = B.get(mat);
            T* Xp = X.get(mat);
            const int ldb = B.ldim(mat), ldx = X.ldim(mat);
            copy_rhs<T>(Bp, ldb, Xp, ldx, side_left, n, nrhs, threadIdx.x, KB_BLK_WG);
            __syncthreads();
            const int inf = blk_potrf<T>(Ap, lda, n, S.blk, Xp, ldx, side_left, nrhs);
            if (threadIdx.x == 0) { if (info) info[mat] = inf; if (iter) iter[mat] = -1; }
            if (inf == 0) blk_solve_bwd<T>(Ap, lda, n, Xp, ldx, side_left, nrhs, S.blk);
        } else {
            T* Ap = A.get(mat);
            const int lda = A.ldim(mat);
            if constexpr (MODE == KB_MODE_POTRF) {
                const int inf = blk_potrf<T>(Ap, lda, n, S.blk);
                if (threadIdx.x == 0 && info) info[mat] = inf;
            } else if constexpr (MODE == KB_MODE_POSV) {
                T* Bp = B.get(mat);
                const int ldb = B.ldim(mat);
                const int inf = blk_potrf<T>(Ap, lda, n, S.blk, Bp, ldb, side_left, nrhs);
                if (threadIdx.x == 0 && info) info[mat] = inf;
                if (inf == 0) blk_solve_bwd<T>(Ap, lda, n, Bp, ldb, side_left, nrhs, S.blk);
            } else {
                blk_solve<T>(Ap, lda, n, B.get(mat), B.ldim(mat), side_left, nrhs, S.blk);
            }
        }
        __syncthreads();
    }
    // ---- wavefront phase: wave 0 keeps the ticket it holds, the others draw
    // (from counter[2], the tickets after the large orders, with small_only) ----
    float* lt = S.lt[wave];
    if (wave != 0) ticket = -1;
    int* ctr = q.counter + (small_only ? 2 : 0);
    for (;;) {
        if (ticket < 0) {
            int t = 0;
            if (lane == 0) t = atomicAdd(ctr, 1);
            ticket = __builtin_amdgcn_readfirstlane(t);
        }
        if (ticket >= batch) return;
        const long mat = q.order[ticket];
        const int n = vq_order(vm, vn, side_left, mat);
        const int nrhs = vq_nrhs(vm, vn, side_left, mat);
        int ln = lane;
        asm volatile("" : "+v"(ln));   // lane-dependent values formed per matrix, not kept live across phases
        v_small<T, MODE>(mat, n, nrhs, side_left, A, B, X, info, iter, lt, ln);
        ticket = -1;
    }
}

// Persistent grid of k_vsposv_big (also the number of FP32 workspace slots).
static int vsposv_grid(int batch) { return vchol_grid(k_vsposv_big, batch); }

template <typename T, int MODE>
static int launch_vchol_mixed(const int* vm, const int* vn, int side_left, kb_mat_t A, kb_mat_t B, kb_mat_t X,
                        int* info, int* iter, int batch, void* ws, float* slots, long slot_floats, int max_n,
                        int max_nrhs, hipStream_t s)
{
    hipLaunchKernelGGL(k_vq_clear, dim3(8), dim3(1024), 0, s, ws);
    const int gb = (batch + 255) / 256 < 1024 ? (batch + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_vq_hist, dim3(gb), dim3(256), 0, s, ws, vm, vn, side_left, batch);
    hipLaunchKernelGGL(k_vq_scan, dim3(1), dim3(1024), 0, s, ws, batch, KB_REG_MAX_N);
    hipLaunchKernelGGL(k_vq_scatter, dim3(gb), dim3(256), 0, s, ws, vm, vn, side_left, batch);
    const BatchArg<T> a = to_arg<T>(A), b = to_arg<T>(B), x = to_arg<T>(X);
    // orders above the register range: the mixed workgroup kernel (FP32 slots
    // in the workspace after the queue), FP64 for what does not fit
    const int gbig = vsposv_grid(batch);
    hipLaunchKernelGGL(k_vsposv_big, dim3(gbig), dim3(KB_BLK_WG), 0, s, vm, vn, side_left, a, b, x, info, iter,
                       batch, ws, slots, slot_floats, max_n, max_nrhs);
    // the rest: the mixed register routine, one matrix per wavefront
    hipLaunchKernelGGL((k_vchol_mixed<T, MODE>), dim3(vchol_grid(k_vchol_mixed<T, MODE>, batch)), dim3(KB_BLK_WG), 0,
                       s, vm, vn, side_left, a, b, x, info, iter, batch, ws, 1);
    return (int)hipGetLastError();
}
