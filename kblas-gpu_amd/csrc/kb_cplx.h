// kb_cplx.h -- the complex element types of the C / Z entry points and the
// scalar helpers the generic kernels (kb_blas3.hip, kb_gen.hip) use for every
// element type.  Storage is two packed reals (real part first), the layout of
// hipFloatComplex / hipDoubleComplex.
//
// Arithmetic follows the reference's complex operators
// (include/kblas_operators.h:150-260): products and sums without
// conjugation, a fused multiply-add as two FMA chains, quotient
// a conj(b) / |b|^2, and the principal square root from the modulus and the
// cosine of the argument (a non-negative real part, the imaginary part with
// the sign of the argument's).
#pragma once
#include "kb_device.h"

template <typename R>
struct KbCplx {
    R x, y;
    KbCplx() = default;
    __host__ __device__ constexpr KbCplx(R r) : x(r), y(R(0)) {}
    __host__ __device__ constexpr KbCplx(R r, R i) : x(r), y(i) {}
};

template <typename R> KB_DEV KbCplx<R> operator+(KbCplx<R> a, KbCplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> KB_DEV KbCplx<R> operator-(KbCplx<R> a, KbCplx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> KB_DEV KbCplx<R> operator-(KbCplx<R> a) { return {-a.x, -a.y}; }
template <typename R> KB_DEV KbCplx<R> operator*(KbCplx<R> a, KbCplx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y};
}
template <typename R> KB_DEV KbCplx<R> operator/(KbCplx<R> a, KbCplx<R> b)
{
    const R d = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
template <typename R> KB_DEV bool operator==(KbCplx<R> a, KbCplx<R> b) { return a.x == b.x && a.y == b.y; }
template <typename R> KB_DEV bool operator!=(KbCplx<R> a, KbCplx<R> b) { return !(a == b); }

// a * b + c
KB_DEV double kb_fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
KB_DEV float kb_fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename R> KB_DEV KbCplx<R> kb_fmat(KbCplx<R> a, KbCplx<R> b, KbCplx<R> c)
{
    return {kb_fmat(a.x, b.x, kb_fmat(-a.y, b.y, c.x)), kb_fmat(a.x, b.y, kb_fmat(a.y, b.x, c.y))};
}

// fma and the refined reciprocal under the names the real kernels call
template <typename R> KB_DEV KbCplx<R> fma(KbCplx<R> a, KbCplx<R> b, KbCplx<R> c) { return kb_fmat(a, b, c); }
template <> KB_DEV KbCplx<double> kb_rcp<KbCplx<double>>(KbCplx<double> x) { return KbCplx<double>(1.0) / x; }
template <> KB_DEV KbCplx<float> kb_rcp<KbCplx<float>>(KbCplx<float> x) { return KbCplx<float>(1.0f) / x; }

KB_DEV double kb_conj(double a) { return a; }
KB_DEV float kb_conj(float a) { return a; }
template <typename R> KB_DEV KbCplx<R> kb_conj(KbCplx<R> a) { return {a.x, -a.y}; }

KB_DEV double kb_sqrtt(double a) { return __builtin_sqrt(a); }
KB_DEV float kb_sqrtt(float a) { return __builtin_sqrtf(a); }
KB_DEV double kb_hypot(double a, double b) { return hypot(a, b); }
KB_DEV float kb_hypot(float a, float b) { return hypotf(a, b); }
template <typename R> KB_DEV KbCplx<R> kb_sqrtt(KbCplx<R> a)
{
    const R r = kb_hypot(a.x, a.y);
    if (r == R(0)) return {R(0), R(0)};
    const R c = a.x / r;
    R im = kb_sqrtt(r * (R(1) - c) / R(2));
    if (__builtin_signbit(a.y)) im = -im;
    return {kb_sqrtt(r * (c + R(1)) / R(2)), im};
}

// a pivot the factorization cannot continue with: not positive (real), zero or not finite (complex)
KB_DEV bool kb_bad_pivot(double a) { return !(a > 0.0); }
KB_DEV bool kb_bad_pivot(float a) { return !(a > 0.0f); }
template <typename R> KB_DEV bool kb_bad_pivot(KbCplx<R> a)
{
    const R m = a.x * a.x + a.y * a.y;
    return !(m > R(0)) || !(m < __builtin_inf());
}

typedef KbCplx<float> kb_cf;
typedef KbCplx<double> kb_cd;
