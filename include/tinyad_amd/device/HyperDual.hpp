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

template <class T, int M>
struct HyperDualM
{
    T v, a, b[M], ab[M];

    HyperDualM() = default;
    template <class U, if_scalar<U> = 0>
    __host__ __device__ HyperDualM(U c) : v(T(c)), a(T(0))
    {
#pragma unroll
        for (int m = 0; m < M; ++m) b[m] = ab[m] = T(0);
    }
};

template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> operator+(const HyperDualM<T, M>& x, const HyperDualM<T, M>& y)
{
    HyperDualM<T, M> r;
    r.v = x.v + y.v;
    r.a = x.a + y.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] + y.b[m]; r.ab[m] = x.ab[m] + y.ab[m]; }
    return r;
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> operator-(const HyperDualM<T, M>& x, const HyperDualM<T, M>& y)
{
    HyperDualM<T, M> r;
    r.v = x.v - y.v;
    r.a = x.a - y.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] - y.b[m]; r.ab[m] = x.ab[m] - y.ab[m]; }
    return r;
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> operator-(const HyperDualM<T, M>& x)
{
    HyperDualM<T, M> r;
    r.v = -x.v;
    r.a = -x.a;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = -x.b[m]; r.ab[m] = -x.ab[m]; }
    return r;
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> operator*(const HyperDualM<T, M>& x, const HyperDualM<T, M>& y)
{
    HyperDualM<T, M> r;
    r.v = x.v * y.v;
    r.a = x.a * y.v + x.v * y.a;
#pragma unroll
    for (int m = 0; m < M; ++m)
    {
        r.b[m] = x.b[m] * y.v + x.v * y.b[m];
        r.ab[m] = x.ab[m] * y.v + x.a * y.b[m] + x.b[m] * y.a + x.v * y.ab[m];
    }
    return r;
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> chain(const HyperDualM<T, M>& x, T f0, T f1, T f2)
{
    HyperDualM<T, M> r;
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
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> operator/(const HyperDualM<T, M>& x, const HyperDualM<T, M>& y)
{
    const T iy = T(1) / y.v;
    return x * chain(y, iy, -iy * iy, T(2) * iy * iy * iy);
}
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator+(const HyperDualM<T, M>& x, U s) { HyperDualM<T, M> r = x; r.v += T(s); return r; }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator+(U s, const HyperDualM<T, M>& x) { return x + s; }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator-(const HyperDualM<T, M>& x, U s) { HyperDualM<T, M> r = x; r.v -= T(s); return r; }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator-(U s, const HyperDualM<T, M>& x) { HyperDualM<T, M> r = -x; r.v += T(s); return r; }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator*(const HyperDualM<T, M>& x, U s)
{
    const T c = T(s);
    HyperDualM<T, M> r;
    r.v = x.v * c;
    r.a = x.a * c;
#pragma unroll
    for (int m = 0; m < M; ++m) { r.b[m] = x.b[m] * c; r.ab[m] = x.ab[m] * c; }
    return r;
}
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator*(U s, const HyperDualM<T, M>& x) { return x * s; }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator/(const HyperDualM<T, M>& x, U s) { return x * (T(1) / T(s)); }
template <class T, int M, class U, if_scalar<U> = 0>
__device__ __forceinline__ HyperDualM<T, M> operator/(U s, const HyperDualM<T, M>& x)
{
    const T iy = T(1) / x.v;
    return chain(x, T(s) * iy, -T(s) * iy * iy, T(2) * T(s) * iy * iy * iy);
}
template <class T, int M, class R>
__device__ __forceinline__ HyperDualM<T, M>& operator+=(HyperDualM<T, M>& x, const R& y) { return x = x + y; }
template <class T, int M, class R>
__device__ __forceinline__ HyperDualM<T, M>& operator-=(HyperDualM<T, M>& x, const R& y) { return x = x - y; }
template <class T, int M, class R>
__device__ __forceinline__ HyperDualM<T, M>& operator*=(HyperDualM<T, M>& x, const R& y) { return x = x * y; }
template <class T, int M, class R>
__device__ __forceinline__ HyperDualM<T, M>& operator/=(HyperDualM<T, M>& x, const R& y) { return x = x / y; }
#define TAD_HDM_COMPARE(OP)                                                                                           \
    template <class T, int M>                                                                                         \
    __device__ __forceinline__ bool operator OP(const HyperDualM<T, M>& x, const HyperDualM<T, M>& y) { return x.v OP y.v; } \
    template <class T, int M, class U, if_scalar<U> = 0>                                                              \
    __device__ __forceinline__ bool operator OP(const HyperDualM<T, M>& x, U s) { return x.v OP T(s); }               \
    template <class T, int M, class U, if_scalar<U> = 0>                                                              \
    __device__ __forceinline__ bool operator OP(U s, const HyperDualM<T, M>& x) { return T(s) OP x.v; }
TAD_HDM_COMPARE(<)
TAD_HDM_COMPARE(<=)
TAD_HDM_COMPARE(>)
TAD_HDM_COMPARE(>=)
TAD_HDM_COMPARE(==)
TAD_HDM_COMPARE(!=)
#undef TAD_HDM_COMPARE
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> sqr(const HyperDualM<T, M>& x) { return x * x; }
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> sqrt(const HyperDualM<T, M>& x)
{
    const T f0 = tad::sqrt(x.v);
    const T f1 = T(0.5) / f0;
    return chain(x, f0, f1, -f1 / (T(2) * x.v));
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> exp(const HyperDualM<T, M>& x)
{
    const T f0 = tad::exp(x.v);
    return chain(x, f0, f0, f0);
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> log(const HyperDualM<T, M>& x)
{
    const T ix = T(1) / x.v;
    return chain(x, tad::log(x.v), ix, -ix * ix);
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> sin(const HyperDualM<T, M>& x)
{
    const T sv = tad::sin(x.v);
    return chain(x, sv, tad::cos(x.v), -sv);
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> cos(const HyperDualM<T, M>& x)
{
    const T c = tad::cos(x.v);
    return chain(x, c, -tad::sin(x.v), -c);
}
template <class T, int M>
__device__ __forceinline__ HyperDualM<T, M> pow(const HyperDualM<T, M>& x, double e)
{
    const T f0 = tad::pow(x.v, e);
    const T ix = T(1) / x.v;
    const T f1 = T(e) * f0 * ix;
    return chain(x, f0, f1, T(e - 1.0) * f1 * ix);
}

} // namespace tad
