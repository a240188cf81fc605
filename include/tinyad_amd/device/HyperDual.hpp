// Hyper-dual scalar with several second directions, for Hessian entries of
// one row per evaluation: x = v + a e1 + sum_m (b_m e2m + ab_m e1 e2m), the
// e's nilpotent and mutually annihilating except the products e1 e2m.
// Seeding variable i with a = 1 and variables j_m with b_m = 1 gives, for
// any composition f of the operations below, f.a = df/dx_i and
// f.ab[m] = d2f/dx_i dx_jm (hyper-dual numbers, Fike and Alonso, AIAA
// 2011-886, with M second directions).  2 + 2M values per scalar instead of
// 1 + K + K(K+1)/2: the device-functor element kernels use it for K >= 7,
// where a full Dual<K> per intermediate no longer fits the register file.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>
#include "Dual.hpp"

namespace tad {

template <class U>
using if_scalar = std::enable_if_t<std::is_arithmetic<U>::value, int>;

// The value has its own type V (default T), as for Dual: FP32 derivative
// passes keep their values in FP64.
template <class T, int M, class V = T>
struct HyperDualM
{
    V v;
    T a, b[M], ab[M];

    HyperDualM() = default;
    template <class U, if_scalar<U> = 0>
    __host__ __device__ HyperDualM(U c) : v(V(c)), a(T(0))
    {
#pragma unroll
        for (int m = 0; m < M; ++m) b[m] = ab[m] = T(0);
    }
};

template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> operator+(const HyperDualM<T, M, V>& x, const HyperDualM<T, M, V>& y)
{
    HyperDualM<T, M, V> r;
    r.v = x.v + y.v;
    r.a = x.a + y.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] + y.b[m]; r.ab[m] = x.ab[m] + y.ab[m]; }
    return r;
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> operator-(const HyperDualM<T, M, V>& x, const HyperDualM<T, M, V>& y)
{
    HyperDualM<T, M, V> r;
    r.v = x.v - y.v;
    r.a = x.a - y.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] - y.b[m]; r.ab[m] = x.ab[m] - y.ab[m]; }
    return r;
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> operator-(const HyperDualM<T, M, V>& x)
{
    HyperDualM<T, M, V> r;
    r.v = -x.v;
    r.a = -x.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = -x.b[m]; r.ab[m] = -x.ab[m]; }
    return r;
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> operator*(const HyperDualM<T, M, V>& x, const HyperDualM<T, M, V>& y)
{
    HyperDualM<T, M, V> r;
    r.v = x.v * y.v;
    const T xv = T(x.v), yv = T(y.v);
    r.a = x.a * yv + xv * y.a;
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        r.b[m] = x.b[m] * yv + xv * y.b[m];
        r.ab[m] = x.ab[m] * yv + x.a * y.b[m] + x.b[m] * y.a + xv * y.ab[m];
    }
    return r;
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> chain(const HyperDualM<T, M, V>& x, V f0, T f1, T f2)
{
    HyperDualM<T, M, V> r;
    r.v = f0;
    r.a = f1 * x.a;
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        r.b[m] = f1 * x.b[m];
        r.ab[m] = f1 * x.ab[m] + f2 * x.a * x.b[m];
    }
    return r;
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> operator/(const HyperDualM<T, M, V>& x, const HyperDualM<T, M, V>& y)
{
    const V iy = V(1) / y.v;
    return x * chain(y, iy, T(-iy * iy), T(V(2) * iy * iy * iy));
}
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator+(const HyperDualM<T, M, V>& x, U s) { HyperDualM<T, M, V> r = x; r.v += V(s); return r; }
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator+(U s, const HyperDualM<T, M, V>& x) { return x + s; }
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator-(const HyperDualM<T, M, V>& x, U s) { HyperDualM<T, M, V> r = x; r.v -= V(s); return r; }
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator-(U s, const HyperDualM<T, M, V>& x) { HyperDualM<T, M, V> r = -x; r.v += V(s); return r; }
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator*(const HyperDualM<T, M, V>& x, U s)
{
    const T c = T(s);
    HyperDualM<T, M, V> r;
    r.v = x.v * V(s);
    r.a = x.a * c;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] * c; r.ab[m] = x.ab[m] * c; }
    return r;
}
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator*(U s, const HyperDualM<T, M, V>& x) { return x * s; }
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator/(const HyperDualM<T, M, V>& x, U s)
{
    const V is = V(1) / V(s);
    const T c = T(is);
    HyperDualM<T, M, V> r;
    r.v = x.v * is;
    r.a = x.a * c;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] * c; r.ab[m] = x.ab[m] * c; }
    return r;
}
template <class T, int M, class V, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M, V> operator/(U s, const HyperDualM<T, M, V>& x)
{
    const V iy = V(1) / x.v;
    const V sv = V(s);
    return chain(x, sv * iy, T(-sv * iy * iy), T(V(2) * sv * iy * iy * iy));
}
template <class T, int M, class V, class R>
__device__ __forceinline__ HyperDualM<T, M, V>& operator+=(HyperDualM<T, M, V>& x, const R& y) { return x = x + y; }
template <class T, int M, class V, class R>
__device__ __forceinline__ HyperDualM<T, M, V>& operator-=(HyperDualM<T, M, V>& x, const R& y) { return x = x - y; }
template <class T, int M, class V, class R>
__device__ __forceinline__ HyperDualM<T, M, V>& operator*=(HyperDualM<T, M, V>& x, const R& y) { return x = x * y; }
template <class T, int M, class V, class R>
__device__ __forceinline__ HyperDualM<T, M, V>& operator/=(HyperDualM<T, M, V>& x, const R& y) { return x = x / y; }
#define TAD_HDM_COMPARE(OP)                                                                                           \
    template <class T, int M, class V>                                                                                \
    __device__ __forceinline__ bool operator OP(const HyperDualM<T, M, V>& x, const HyperDualM<T, M, V>& y) { return x.v OP y.v; } \
    template <class T, int M, class V, class U, if_scalar<U> = 0>                                                     \
    __device__ __forceinline__ bool operator OP(const HyperDualM<T, M, V>& x, U s) { return x.v OP V(s); }            \
    template <class T, int M, class V, class U, if_scalar<U> = 0>                                                     \
    __device__ __forceinline__ bool operator OP(U s, const HyperDualM<T, M, V>& x) { return V(s) OP x.v; }
TAD_HDM_COMPARE(<)
TAD_HDM_COMPARE(<=)
TAD_HDM_COMPARE(>)
TAD_HDM_COMPARE(>=)
TAD_HDM_COMPARE(==)
TAD_HDM_COMPARE(!=)
#undef TAD_HDM_COMPARE
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> sqr(const HyperDualM<T, M, V>& x) { return x * x; }
// value and derivative factors in V, rounded to T once (as for Dual)
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> sqrt(const HyperDualM<T, M, V>& x)
{
    const V f0 = tad::sqrt(x.v);
    const V f1 = V(0.5) / f0;
    return chain(x, f0, T(f1), T(-f1 / (V(2) * x.v)));
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> exp(const HyperDualM<T, M, V>& x)
{
    const V f0 = tad::exp(x.v);
    return chain(x, f0, T(f0), T(f0));
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> log(const HyperDualM<T, M, V>& x)
{
    const V ix = V(1) / x.v;
    return chain(x, tad::log(x.v), T(ix), T(-ix * ix));
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> sin(const HyperDualM<T, M, V>& x)
{
    const V sv = tad::sin(x.v);
    return chain(x, sv, T(tad::cos(x.v)), T(-sv));
}
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> cos(const HyperDualM<T, M, V>& x)
{
    const V c = tad::cos(x.v);
    return chain(x, c, T(-tad::sin(x.v)), T(-c));
}
// a^e from p = a^(e-2) as p a and p a a (Scalar.hh:292-306): finite at 0 for e >= 2
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> pow(const HyperDualM<T, M, V>& x, double e)
{
    const V p2 = tad::pow(x.v, e - 2.0);
    const V p1 = p2 * x.v;
    return chain(x, p1 * x.v, T(V(e) * p1), T(V(e * (e - 1.0)) * p2));
}
// integer exponent: a^0 and a^1 exactly (Scalar.hh:266-290)
template <class T, int M, class V>
__device__ __forceinline__ HyperDualM<T, M, V> pow(const HyperDualM<T, M, V>& x, int e)
{
    if (e == 0) return chain(x, V(1), T(0), T(0));
    if (e == 1) return chain(x, x.v, T(1), T(0));
    const V p2 = tad::pow(x.v, double(e - 2));
    const V p1 = p2 * x.v;
    return chain(x, p1 * x.v, T(V(e) * p1), T(V(double(e) * double(e - 1)) * p2));
}

} // namespace tad
