// Hyper-dual scalar for one Hessian entry per evaluation: x = v + a e1 + b e2
// + ab e1 e2 with e1^2 = e2^2 = 0.  Seeding variable i with a = 1 and
// variable j with b = 1 gives, for any composition f of the operations below,
// f.a = df/dx_i, f.b = df/dx_j and f.ab = d2f/dx_i dx_j (Fike and Alonso,
// AIAA 2011-886).  Four values per scalar instead of 1 + K + K(K+1)/2: the
// device-functor element kernels use it for K >= 9, where a full Dual<K>
// per intermediate no longer fits the register file.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>
#include "Dual.hpp"

namespace tad {

template <class T>
struct HyperDual
{
    T v, a, b, ab;

    HyperDual() = default;
    template <class U, std::enable_if_t<std::is_arithmetic<U>::value, int> = 0>
    __host__ __device__ HyperDual(U c) : v(T(c)), a(T(0)), b(T(0)), ab(T(0))
    {
    }
    __host__ __device__ HyperDual(T v_, T a_, T b_, T ab_) : v(v_), a(a_), b(b_), ab(ab_) {}
};

template <class U>
using if_scalar = std::enable_if_t<std::is_arithmetic<U>::value, int>;

template <class T>
__device__ __forceinline__ HyperDual<T> operator+(const HyperDual<T>& x, const HyperDual<T>& y)
{
    return HyperDual<T>(x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab);
}
template <class T>
__device__ __forceinline__ HyperDual<T> operator-(const HyperDual<T>& x, const HyperDual<T>& y)
{
    return HyperDual<T>(x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab);
}
template <class T>
__device__ __forceinline__ HyperDual<T> operator-(const HyperDual<T>& x)
{
    return HyperDual<T>(-x.v, -x.a, -x.b, -x.ab);
}
// product rule; the e1 e2 part collects both cross terms
template <class T>
__device__ __forceinline__ HyperDual<T> operator*(const HyperDual<T>& x, const HyperDual<T>& y)
{
    return HyperDual<T>(x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b,
                        x.ab * y.v + x.a * y.b + x.b * y.a + x.v * y.ab);
}

// f(x) with f(x.v) = f0, f'(x.v) = f1, f''(x.v) = f2
template <class T>
__device__ __forceinline__ HyperDual<T> chain(const HyperDual<T>& x, T f0, T f1, T f2)
{
    return HyperDual<T>(f0, f1 * x.a, f1 * x.b, f1 * x.ab + f2 * x.a * x.b);
}

template <class T>
__device__ __forceinline__ HyperDual<T> operator/(const HyperDual<T>& x, const HyperDual<T>& y)
{
    const T iy = T(1) / y.v;
    return x * chain(y, iy, -iy * iy, T(2) * iy * iy * iy);
}

template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator+(const HyperDual<T>& x, U s) { return HyperDual<T>(x.v + T(s), x.a, x.b, x.ab); }
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator+(U s, const HyperDual<T>& x) { return x + s; }
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator-(const HyperDual<T>& x, U s) { return HyperDual<T>(x.v - T(s), x.a, x.b, x.ab); }
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator-(U s, const HyperDual<T>& x) { return HyperDual<T>(T(s) - x.v, -x.a, -x.b, -x.ab); }
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator*(const HyperDual<T>& x, U s)
{
    const T c = T(s);
    return HyperDual<T>(x.v * c, x.a * c, x.b * c, x.ab * c);
}
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator*(U s, const HyperDual<T>& x) { return x * s; }
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator/(const HyperDual<T>& x, U s)
{
    const T c = T(s);
    return HyperDual<T>(x.v / c, x.a / c, x.b / c, x.ab / c);
}
template <class T, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDual<T> operator/(U s, const HyperDual<T>& x)
{
    const T iy = T(1) / x.v;
    return chain(x, T(s) * iy, -T(s) * iy * iy, T(2) * T(s) * iy * iy * iy);
}

template <class T, class R>
__device__ __forceinline__ HyperDual<T>& operator+=(HyperDual<T>& x, const R& y) { return x = x + y; }
template <class T, class R>
__device__ __forceinline__ HyperDual<T>& operator-=(HyperDual<T>& x, const R& y) { return x = x - y; }
template <class T, class R>
__device__ __forceinline__ HyperDual<T>& operator*=(HyperDual<T>& x, const R& y) { return x = x * y; }
template <class T, class R>
__device__ __forceinline__ HyperDual<T>& operator/=(HyperDual<T>& x, const R& y) { return x = x / y; }

#define TAD_HD_COMPARE(OP)                                                                                         \
    template <class T>                                                                                             \
    __device__ __forceinline__ bool operator OP(const HyperDual<T>& x, const HyperDual<T>& y) { return x.v OP y.v; } \
    template <class T, class U, if_scalar<U> = 0>                                                                  \
    __device__ __forceinline__ bool operator OP(const HyperDual<T>& x, U s) { return x.v OP T(s); }                 \
    template <class T, class U, if_scalar<U> = 0>                                                                  \
    __device__ __forceinline__ bool operator OP(U s, const HyperDual<T>& x) { return T(s) OP x.v; }
TAD_HD_COMPARE(<)
TAD_HD_COMPARE(<=)
TAD_HD_COMPARE(>)
TAD_HD_COMPARE(>=)
TAD_HD_COMPARE(==)
TAD_HD_COMPARE(!=)
#undef TAD_HD_COMPARE

template <class T>
__device__ __forceinline__ HyperDual<T> sqr(const HyperDual<T>& x) { return x * x; }
template <class T>
__device__ __forceinline__ HyperDual<T> sqrt(const HyperDual<T>& x)
{
    const T f0 = tad::sqrt(x.v);
    const T f1 = T(0.5) / f0;
    return chain(x, f0, f1, -f1 / (T(2) * x.v));
}
template <class T>
__device__ __forceinline__ HyperDual<T> exp(const HyperDual<T>& x)
{
    const T f0 = tad::exp(x.v);
    return chain(x, f0, f0, f0);
}
template <class T>
__device__ __forceinline__ HyperDual<T> log(const HyperDual<T>& x)
{
    const T ix = T(1) / x.v;
    return chain(x, tad::log(x.v), ix, -ix * ix);
}
template <class T>
__device__ __forceinline__ HyperDual<T> sin(const HyperDual<T>& x)
{
    const T s = tad::sin(x.v);
    return chain(x, s, tad::cos(x.v), -s);
}
template <class T>
__device__ __forceinline__ HyperDual<T> cos(const HyperDual<T>& x)
{
    const T c = tad::cos(x.v);
    return chain(x, c, -tad::sin(x.v), -c);
}
template <class T>
__device__ __forceinline__ HyperDual<T> pow(const HyperDual<T>& x, double e)
{
    const T f0 = tad::pow(x.v, e);
    const T ix = T(1) / x.v;
    const T f1 = T(e) * f0 * ix;
    return chain(x, f0, f1, T(e - 1.0) * f1 * ix);
}

} // namespace tad
