// kb_device.h -- device-side building blocks for the gfx950 batched SPD kernels.
//
// Lane-group broadcasts replace the reference's sub-warp shuffles
// (include/kblas_operators.h:84-105, `shfl(x, lane, width)` built from two
// 32-bit __shfl_sync for FP64).  On CDNA4 a 16-lane group is one DPP row, so
// the broadcast is a single `v_mov_b64_dpp row_newbcast:J`; a 32-lane group
// uses ds_swizzle in bit-mask mode (no LDS storage, no bank conflicts); a
// 64-lane group has one matrix per wavefront, so the source lane is
// wave-uniform and v_readlane puts the value in an SGPR that every FMA can
// read directly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kb_launch.h"

#define KB_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// Batch operand: strided base, array of device pointers, or per-matrix leading
// dimensions.  One code path serves the reference's strided and pointer-array
// entry points and the non-uniform ones.
// ---------------------------------------------------------------------------
template <typename T>
struct BatchArg {
    T* base;             // strided mode: first matrix
    long stride;         // strided mode: elements between matrices
    T* const* ptrs;      // pointer-array mode (device array of device pointers)
    const int* lds;      // non-uniform leading dimensions (device), or nullptr
    int ld;              // uniform leading dimension
    KB_DEV T* get(long b) const
    {
        return ptrs ? ptrs[b] : base + b * stride;
    }
    KB_DEV int ldim(long b) const { return lds ? lds[b] : ld; }
};

// The same with an element offset into every matrix: sub-blocks of a batch,
// which a pointer array cannot take into its base.  Only the kernels that
// work on sub-blocks (matrix-core TRSM and GEMM) take this form, so the others
// keep two scalar registers per operand free.
template <typename T>
struct BatchArgO {
    T* base;
    long stride;
    T* const* ptrs;
    const int* lds;
    int ld;
    long off;
    KB_DEV T* get(long b) const
    {
        return (ptrs ? ptrs[b] : base + b * stride) + off;
    }
    KB_DEV int ldim(long b) const { return lds ? lds[b] : ld; }
};

// A strided operand takes its offset into the base; a pointer array with an
// offset needs to_arg_off.
template <typename T>
static BatchArg<T> to_arg(kb_mat_t m)
{
    BatchArg<T> a;
    a.base = (T*)m.base + (m.ptrs ? 0 : m.off);
    a.stride = m.stride;
    a.ptrs = (T* const*)m.ptrs;
    a.lds = m.lds;
    a.ld = m.ld;
    return a;
}
template <typename T>
static BatchArgO<T> to_arg_off(kb_mat_t m)
{
    BatchArgO<T> a;
    a.base = (T*)m.base;
    a.stride = m.stride;
    a.ptrs = (T* const*)m.ptrs;
    a.lds = m.lds;
    a.ld = m.ld;
    a.off = m.off;
    return a;
}

// Element (row, col) of right-hand side col: side_left, B is n x nrhs and the
// vector is column col; right side, B is nrhs x n and the vector is row col.
template <typename T>
KB_DEV T* rhs_elem(T* Bp, int ldb, int side_left, int row, int col)
{
    return side_left ? Bp + row + (long)col * ldb : Bp + col + (long)row * ldb;
}

// ---------------------------------------------------------------------------
// Global-memory accessors.  Batch operands always live in device (global)
// memory; going through an address_space(1) pointer lets the compiler emit
// global_load/global_store instead of flat accesses, which would also count
// against the LDS wait counter.
// ---------------------------------------------------------------------------
template <typename T>
KB_DEV T gld(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}

// gld for one of a group of loads meant to be in flight together: the caller
// clamps p to a valid address where the value goes unused and selects after
// the load; the address is kept opaque to the optimizer, which would
// otherwise see clamped loads repeat one another and put each load under a
// branch of its own (one wait per element).
template <typename T>
KB_DEV T gld_opaque(const T* p)
{
    asm volatile("" : "+v"(p));
    return gld(p);
}
template <typename T>
KB_DEV void gst(T* p, T v)
{
    *(__attribute__((address_space(1))) T*)p = v;
}

// X := B for an n-row block of nrhs right-hand sides, thread t of nt; the
// element order follows the storage (rows fastest for side_left, right-hand
// sides fastest otherwise) so consecutive threads touch consecutive addresses.
template <typename T>
KB_DEV void copy_rhs(const T* Bp, int ldb, T* Xp, int ldx, int side_left, int n, int nrhs, int t, int nt)
{
    const int fast = side_left ? n : nrhs;
    for (int e = t; e < n * nrhs; e += nt) {
        const int r = side_left ? e % fast : e / fast, c = side_left ? e / fast : e % fast;
        gst(rhs_elem(Xp, ldx, side_left, r, c), gld(rhs_elem(const_cast<T*>(Bp), ldb, side_left, r, c)));
    }
}

// Make this wavefront's global stores visible to its own later global loads
// (by other lanes of the same wavefront).  Workgroup scope suffices: the
// stores and loads go through the same CU's vector L1, so this is a wait for
// the stores to complete, not a device-scope fence, which on gfx950 writes
// back and invalidates the XCD's whole L2.
KB_DEV void wave_global_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Make this wavefront's LDS writes visible to its other lanes.
KB_DEV void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// bit casts
// ---------------------------------------------------------------------------
KB_DEV int f2i(float x) { return __builtin_bit_cast(int, x); }
KB_DEV float i2f(int x) { return __builtin_bit_cast(float, x); }
KB_DEV int2 d2i2(double x) { return __builtin_bit_cast(int2, x); }
KB_DEV double i22d(int2 x) { return __builtin_bit_cast(double, x); }

// ---------------------------------------------------------------------------
// Broadcast lane J of each G-lane group to the whole group (J compile-time).
// ---------------------------------------------------------------------------
template <int G, int J>
KB_DEV double bcast(double x)
{
    static_assert(J >= 0 && J < G, "lane out of group");
    if constexpr (G == 16) {
        return __builtin_amdgcn_update_dpp(0.0, x, 0x150 + J, 0xF, 0xF, false);
    } else if constexpr (G == 32) {
        int2 v = d2i2(x);
        v.x = __builtin_amdgcn_ds_swizzle(v.x, (J << 5));
        v.y = __builtin_amdgcn_ds_swizzle(v.y, (J << 5));
        return i22d(v);
    } else {
        int2 v = d2i2(x);
        v.x = __builtin_amdgcn_readlane(v.x, J);
        v.y = __builtin_amdgcn_readlane(v.y, J);
        return i22d(v);
    }
}

template <int G, int J>
KB_DEV float bcast(float x)
{
    static_assert(J >= 0 && J < G, "lane out of group");
    if constexpr (G == 16) {
        return __builtin_amdgcn_update_dpp(0.0f, x, 0x150 + J, 0xF, 0xF, false);
    } else if constexpr (G == 32) {
        return i2f(__builtin_amdgcn_ds_swizzle(f2i(x), (J << 5)));
    } else {
        return i2f(__builtin_amdgcn_readlane(f2i(x), J));
    }
}

// Broadcast with a lane index that becomes a compile-time constant once the
// enclosing loop is fully unrolled (the switch folds away).  G = 64 reads an
// SGPR-uniform lane directly.
#define KB_BC16(i) case i: return bcast<16, i>(x);
#define KB_BC32(i) case i: return bcast<32, i>(x);
#define KB_REP16(M, o) M(o+0) M(o+1) M(o+2) M(o+3) M(o+4) M(o+5) M(o+6) M(o+7) \
    M(o+8) M(o+9) M(o+10) M(o+11) M(o+12) M(o+13) M(o+14) M(o+15)

template <int G, typename T>
KB_DEV T bcastk(T x, int j)
{
    if constexpr (G == 16) {
        switch (j) { KB_REP16(KB_BC16, 0) default: return x; }
    } else if constexpr (G == 32) {
        switch (j) { KB_REP16(KB_BC32, 0) KB_REP16(KB_BC32, 16) default: return x; }
    } else {
        if constexpr (sizeof(T) == 8) {
            int2 v = d2i2(x);
            v.x = __builtin_amdgcn_readlane(v.x, j);
            v.y = __builtin_amdgcn_readlane(v.y, j);
            return i22d(v);
        } else {
            return i2f(__builtin_amdgcn_readlane(f2i(x), j));
        }
    }
}

// Runtime-lane broadcast within a G-lane group (ds_bpermute); used off the hot loops.
template <int G, typename T>
KB_DEV T bcast_rt(T x, int j)
{
    int lane = __lane_id();
    int src = (lane & ~(G - 1)) + j;
    if constexpr (sizeof(T) == 8) {
        int2 v = d2i2(x);
        v.x = __builtin_amdgcn_ds_bpermute(src << 2, v.x);
        v.y = __builtin_amdgcn_ds_bpermute(src << 2, v.y);
        return i22d(v);
    } else {
        return i2f(__builtin_amdgcn_ds_bpermute(src << 2, f2i(x)));
    }
}

// Group-wide max (all lanes of the G-lane group get the result).
template <int G, typename T>
KB_DEV T group_max(T x)
{
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        T y = __shfl_xor(x, off, 64);
        x = (y > x || y != y) ? y : x;   // NaN wins so a failed solve is never reported converged
    }
    return x;
}

template <int G>
KB_DEV int group_all(int pred)
{
#pragma unroll
    for (int off = 1; off < G; off <<= 1) pred &= __shfl_xor(pred, off, 64);
    return pred;
}

// 1/sqrt(x): hardware estimate refined by Newton-Raphson (FP64: v_rsq_f64 then
// twice r += r/2 (1 - x r^2), ~1 ulp); FP32 v_rsq_f32 is already ~1 ulp.
// Cholesky only needs the pivot's square root to a few ulp for backward
// stability; LAPACK's dpotf2 likewise scales the column by the reciprocal.
// ---------------------------------------------------------------------------
// fused row reads
// ---------------------------------------------------------------------------
// acc += src(lane K of this 16-lane row) * mul.  The s_nop supplies the two
// wait states a DPP read needs after a VALU write of its source register.
template <int K>
KB_DEV void fmac_row(double& acc, double src, double mul)
{
    asm("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(src), "v"(mul), "n"(K));
}
template <int K>
KB_DEV void fmac_row(float& acc, float src, float mul)
{
    asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(src), "v"(mul), "n"(K));
}

#define KB_FR(i) case i: fmac_row<i>(acc, src, mul); return;
// k becomes a constant once the caller's loops are unrolled; the switch folds.
template <typename T>
KB_DEV void fmac_rowk(T& acc, T src, T mul, int k)
{
    switch (k & 15) { KB_REP16(KB_FR, 0) }
}
#undef KB_FR

// value of lane K of this 16-lane row (compiler-visible DPP move, undefined old value)
#define KB_MR(i) case i: return __builtin_amdgcn_mov_dpp(x, 0x150 + i, 0xF, 0xF, false);
template <typename T>
KB_DEV T row_bcast(T x, int k)
{
    switch (k & 15) { KB_REP16(KB_MR, 0) default: return x; }
}
#undef KB_MR

// acc -= src(lane K of this 16-lane row) * mul
template <int K>
KB_DEV void fnmac_row(double& acc, double src, double mul)
{
    asm("s_nop 1\n\tv_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(src), "v"(mul), "n"(K));
}
template <int K>
KB_DEV void fnmac_row(float& acc, float src, float mul)
{
    asm("s_nop 1\n\tv_fmac_f32_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(src), "v"(mul), "n"(K));
}

#define KB_FNR(i) case i: fnmac_row<i>(acc, src, mul); return;
template <typename T>
KB_DEV void fnmac_rowk(T& acc, T src, T mul, int k)
{
    switch (k & 15) { KB_REP16(KB_FNR, 0) }
}
#undef KB_FNR

// ---------------------------------------------------------------------------
// 16x16x4 matrix-core products (FP64 and FP32 accumulators of 4 per lane)
// ---------------------------------------------------------------------------
typedef double kb_d4 __attribute__((ext_vector_type(4)));
typedef float kb_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct TileV;
template <> struct TileV<double> { typedef kb_d4 v4; };
template <> struct TileV<float> { typedef kb_f4 v4; };

KB_DEV kb_d4 mfma16x16x4(double a, double b, kb_d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
KB_DEV kb_f4 mfma16x16x4(float a, float b, kb_f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <typename T> KB_DEV T kb_rsqrt(T x);
template <> KB_DEV double kb_rsqrt<double>(double x)
{
    double r = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        double t = x * r;
        double e = __builtin_fma(-t, r, 1.0);
        r = __builtin_fma(0.5 * r, e, r);
    }
    return r;
}
template <> KB_DEV float kb_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// Reciprocal: hardware estimate and two Newton steps (FP64), within an ulp
// or so of 1/x, against the ~10-instruction IEEE division sequence.
template <typename T> KB_DEV T kb_rcp(T x);
template <> KB_DEV double kb_rcp<double>(double x)
{
    double r = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return r;
}
template <> KB_DEV float kb_rcp<float>(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

template <typename T> KB_DEV T kb_sqrt(T x);
template <> KB_DEV double kb_sqrt<double>(double x) { return __builtin_sqrt(x); }
template <> KB_DEV float kb_sqrt<float>(float x) { return __builtin_sqrtf(x); }
