// Device forward-mode second-order scalar Dual<K, T>: value, gradient (K) and
// symmetric Hessian stored packed (upper triangle, K(K+1)/2 entries).
//
// The operator formulas are the ones of the reference's Scalar<k, T>
// (include/TinyAD/Scalar.hh:220-234 unary minus, :249-264 sqr, :579-736 +/-,
// :738-821 *, :823-891 /).  The reference stores the full k x k Hessian and
// documents that it does not exploit symmetry (Scalar.hh:751-762); here only
// the upper triangle is stored and updated, which is the same arithmetic per
// stored entry and halves register use.  All loops have compile-time bounds so
// every entry lives in a register after unrolling.
#pragma once

#include <hip/hip_runtime.h>

namespace tad {

template <int K>
struct Sym
{
    static constexpr int n = K * (K + 1) / 2;
    // packed index of (i, j), i <= j
    __host__ __device__ static constexpr int idx(int i, int j)
    {
        return i * K - (i * (i - 1)) / 2 + (j - i);
    }
    __host__ __device__ static constexpr int at(int i, int j)
    {
        return i <= j ? idx(i, j) : idx(j, i);
    }
};

template <int K, class T>
struct Dual
{
    T v;
    T g[K];
    T h[Sym<K>::n];
};

template <int K, class T>
__device__ __forceinline__ Dual<K, T> dual_const(T v)
{
    Dual<K, T> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = T(0);
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = T(0);
    return r;
}

// Active variable with local index idx (Scalar.hh:70-78).
template <int K, class T>
__device__ __forceinline__ Dual<K, T> dual_var(T v, int idx)
{
    Dual<K, T> r = dual_const<K, T>(v);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = (i == idx) ? T(1) : T(0);
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator+(const Dual<K, T>& a, const Dual<K, T>& b)
{
    Dual<K, T> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] + b.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] + b.h[i];
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator-(const Dual<K, T>& a, const Dual<K, T>& b)
{
    Dual<K, T> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] - b.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] - b.h[i];
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator-(const Dual<K, T>& a)
{
    Dual<K, T> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = -a.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = -a.h[i];
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator+(const Dual<K, T>& a, T b)
{
    Dual<K, T> r = a;
    r.v += b;
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator+(T a, const Dual<K, T>& b) { return b + a; }

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator-(const Dual<K, T>& a, T b)
{
    Dual<K, T> r = a;
    r.v -= b;
    return r;
}

// passive - active (Scalar.hh:699-716)
template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator-(T a, const Dual<K, T>& b)
{
    Dual<K, T> r = -b;
    r.v = a - b.v;
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator*(const Dual<K, T>& a, T b)
{
    Dual<K, T> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] * b;
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] * b;
    return r;
}

template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator*(T a, const Dual<K, T>& b) { return b * a; }

// Hess(a b) = b Ha + ga gb^T + gb ga^T + a Hb
template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator*(const Dual<K, T>& a, const Dual<K, T>& b)
{
    Dual<K, T> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = b.v * a.g[i] + a.v * b.g[i];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = b.v * a.h[s] + a.g[i] * b.g[j] + b.g[i] * a.g[j] + a.v * b.h[s];
        }
    return r;
}

// Hess(a^2) = 2 (a Ha + ga ga^T)
template <int K, class T>
__device__ __forceinline__ Dual<K, T> sqr(const Dual<K, T>& a)
{
    Dual<K, T> r;
    r.v = a.v * a.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = T(2) * a.v * a.g[i];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = T(2) * (a.v * a.h[s] + a.g[i] * a.g[j]);
        }
    return r;
}

// a / b: grad (b ga - a gb) / b^2, Hess (Ha - g gb^T - gb g^T - (a/b) Hb) / b
template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator/(const Dual<K, T>& a, const Dual<K, T>& b)
{
    Dual<K, T> r;
    r.v = a.v / b.v;
    const T bb = b.v * b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = (b.v * a.g[i] - a.v * b.g[i]) / bb;
    const T ib = T(1) / b.v;
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = (a.h[s] - r.g[i] * b.g[j] - b.g[i] * r.g[j] - r.v * b.h[s]) * ib;
        }
    return r;
}

// passive / active (Scalar.hh:868-884)
template <int K, class T>
__device__ __forceinline__ Dual<K, T> operator/(T a, const Dual<K, T>& b)
{
    Dual<K, T> r;
    r.v = a / b.v;
    const T c = -a / (b.v * b.v);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = c * b.g[i];
    const T ib = T(1) / b.v;
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = (-r.g[i] * b.g[j] - b.g[i] * r.g[j] - r.v * b.h[s]) * ib;
        }
    return r;
}

} // namespace tad
