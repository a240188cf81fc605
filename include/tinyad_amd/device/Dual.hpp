// Device forward-mode second-order scalar Dual<K, T, V>: value (type V), gradient (K) and
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
#include <type_traits>

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

// The value has its own type V (default T): device-functor elements in mixed
// precision keep values in FP64 (V = double) and derivatives in FP32
// (T = float), so that differences of coordinates are formed before the
// switch to FP32, as the catalog's elements form their edge matrices in FP64.
// Derivative arithmetic always runs in T; with V = T every expression is the
// single-type one.
template <int K, class T, class V = T>
struct Dual
{
    V v;
    T g[K];
    T h[Sym<K>::n];

    Dual() = default;
    // passive constant: zero gradient and Hessian (Scalar.hh:59-68); implicit,
    // so generic energies can write S(0) or return INFINITY for S = Dual
    template <class U, std::enable_if_t<std::is_arithmetic<U>::value, int> = 0>
    __host__ __device__ Dual(U c) : v(V(c)), g{}, h{}
    {
    }
};

template <int K, class T, class V = T>
__device__ __forceinline__ Dual<K, T, V> dual_const(V v)
{
    Dual<K, T, V> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = T(0);
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = T(0);
    return r;
}

// Active variable with local index idx (Scalar.hh:70-78).
template <int K, class T, class V = T>
__device__ __forceinline__ Dual<K, T, V> dual_var(V v, int idx)
{
    Dual<K, T, V> r = dual_const<K, T, V>(v);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = (i == idx) ? T(1) : T(0);
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator+(const Dual<K, T, V>& a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] + b.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] + b.h[i];
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator-(const Dual<K, T, V>& a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] - b.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] - b.h[i];
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator-(const Dual<K, T, V>& a)
{
    Dual<K, T, V> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = -a.g[i];
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = -a.h[i];
    return r;
}

// Passive operands enter the value in V and the derivatives in T.
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator+(const Dual<K, T, V>& a, V b)
{
    Dual<K, T, V> r = a;
    r.v += b;
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator+(V a, const Dual<K, T, V>& b) { return b + a; }

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator-(const Dual<K, T, V>& a, V b)
{
    Dual<K, T, V> r = a;
    r.v -= b;
    return r;
}

// passive - active (Scalar.hh:699-716)
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator-(V a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r = -b;
    r.v = a - b.v;
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator*(const Dual<K, T, V>& a, V b)
{
    Dual<K, T, V> r;
    r.v = a.v * b;
    const T bt = T(b);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] * bt;
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] * bt;
    return r;
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator*(V a, const Dual<K, T, V>& b) { return b * a; }

// Hess(a b) = b Ha + ga gb^T + gb ga^T + a Hb
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator*(const Dual<K, T, V>& a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r;
    r.v = a.v * b.v;
    const T av = T(a.v), bv = T(b.v);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = bv * a.g[i] + av * b.g[i];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = bv * a.h[s] + a.g[i] * b.g[j] + b.g[i] * a.g[j] + av * b.h[s];
        }
    return r;
}

// Hess(a^2) = 2 (a Ha + ga ga^T)
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> sqr(const Dual<K, T, V>& a)
{
    Dual<K, T, V> r;
    r.v = a.v * a.v;
    const T av = T(a.v);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = T(2) * av * a.g[i];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = T(2) * (av * a.h[s] + a.g[i] * a.g[j]);
        }
    return r;
}

// a / b: grad (b ga - a gb) / b^2, Hess (Ha - g gb^T - gb g^T - (a/b) Hb) / b
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator/(const Dual<K, T, V>& a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r;
    r.v = a.v / b.v;
    const T av = T(a.v), bv = T(b.v), rv = T(r.v);
    const T bb = bv * bv;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = (bv * a.g[i] - av * b.g[i]) / bb;
    const T ib = T(1) / bv;
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = (a.h[s] - r.g[i] * b.g[j] - b.g[i] * r.g[j] - rv * b.h[s]) * ib;
        }
    return r;
}

// passive / active (Scalar.hh:868-884)
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator/(V a, const Dual<K, T, V>& b)
{
    Dual<K, T, V> r;
    r.v = a / b.v;
    const T c = T(-a / (b.v * b.v));
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = c * b.g[i];
    const T ib = T(V(1) / b.v), rv = T(r.v);
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = (-r.g[i] * b.g[j] - b.g[i] * r.g[j] - rv * b.h[s]) * ib;
        }
    return r;
}

// ---------------------------------------------------------------------------
// Operations for energies written once for a generic scalar S (S = double for
// the passive value, S = Dual<K, T, V> for derivatives); none of the library's
// built-in element kernels uses them.
// ---------------------------------------------------------------------------

// Active / passive (Scalar.hh:823-866): value a/b, gradient ga/b, Hessian Ha/b.
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> operator/(const Dual<K, T, V>& a, V b)
{
    Dual<K, T, V> r;
    r.v = a.v / b;
    const T bt = T(b);
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = a.g[i] / bt;
#pragma unroll
    for (int i = 0; i < Sym<K>::n; ++i) r.h[i] = a.h[i] / bt;
    return r;
}

// Passive operands of another arithmetic type (int literals, double constants
// next to FP32 values) are converted to the value type V first.
template <class V, class U>
using if_other_scalar = std::enable_if_t<std::is_arithmetic<U>::value && !std::is_same<V, U>::value, int>;

template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator+(const Dual<K, T, V>& a, U b) { return a + V(b); }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator+(U a, const Dual<K, T, V>& b) { return V(a) + b; }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator-(const Dual<K, T, V>& a, U b) { return a - V(b); }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator-(U a, const Dual<K, T, V>& b) { return V(a) - b; }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator*(const Dual<K, T, V>& a, U b) { return a * V(b); }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator*(U a, const Dual<K, T, V>& b) { return V(a) * b; }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator/(const Dual<K, T, V>& a, U b) { return a / V(b); }
template <int K, class T, class V, class U, if_other_scalar<V, U> = 0>
__device__ __forceinline__ Dual<K, T, V> operator/(U a, const Dual<K, T, V>& b) { return V(a) / b; }

template <int K, class T, class V, class R>
__device__ __forceinline__ Dual<K, T, V>& operator+=(Dual<K, T, V>& a, const R& b) { return a = a + b; }
template <int K, class T, class V, class R>
__device__ __forceinline__ Dual<K, T, V>& operator-=(Dual<K, T, V>& a, const R& b) { return a = a - b; }
template <int K, class T, class V, class R>
__device__ __forceinline__ Dual<K, T, V>& operator*=(Dual<K, T, V>& a, const R& b) { return a = a * b; }
template <int K, class T, class V, class R>
__device__ __forceinline__ Dual<K, T, V>& operator/=(Dual<K, T, V>& a, const R& b) { return a = a / b; }

// Comparisons look at values only (Scalar.hh:1045-1100).
template <class X> __host__ __device__ __forceinline__ auto dual_value(const X& x) { return x; }
template <int K, class T, class V> __host__ __device__ __forceinline__ V dual_value(const Dual<K, T, V>& x) { return x.v; }

#define TAD_DUAL_COMPARE(OP)                                                                                    \
    template <int K, class T, class V>                                                                          \
    __device__ __forceinline__ bool operator OP(const Dual<K, T, V>& a, const Dual<K, T, V>& b) { return a.v OP b.v; } \
    template <int K, class T, class V, class U, std::enable_if_t<std::is_arithmetic<U>::value, int> = 0>        \
    __device__ __forceinline__ bool operator OP(const Dual<K, T, V>& a, U b) { return a.v OP V(b); }             \
    template <int K, class T, class V, class U, std::enable_if_t<std::is_arithmetic<U>::value, int> = 0>        \
    __device__ __forceinline__ bool operator OP(U a, const Dual<K, T, V>& b) { return V(a) OP b.v; }
TAD_DUAL_COMPARE(<)
TAD_DUAL_COMPARE(<=)
TAD_DUAL_COMPARE(>)
TAD_DUAL_COMPARE(>=)
TAD_DUAL_COMPARE(==)
TAD_DUAL_COMPARE(!=)
#undef TAD_DUAL_COMPARE

// f(a) for a twice differentiable scalar function with f(a.v) = f0,
// f'(a.v) = f1, f''(a.v) = f2: gradient f1 ga, Hessian f1 Ha + f2 ga ga^T
// (the chain rule every unary function of Scalar.hh:266-577 applies).
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> chain(const Dual<K, T, V>& a, V f0, T f1, T f2)
{
    Dual<K, T, V> r;
    r.v = f0;
#pragma unroll
    for (int i = 0; i < K; ++i) r.g[i] = f1 * a.g[i];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j)
        {
            const int s = Sym<K>::idx(i, j);
            r.h[s] = f1 * a.h[s] + f2 * a.g[i] * a.g[j];
        }
    return r;
}

// Passive forms, so that tad::sqrt(x) etc. work for S = double and float too.
__host__ __device__ __forceinline__ double sqr(double a) { return a * a; }
__host__ __device__ __forceinline__ float sqr(float a) { return a * a; }
__device__ __forceinline__ double sqrt(double a) { return ::sqrt(a); }
__device__ __forceinline__ float sqrt(float a) { return ::sqrtf(a); }
__device__ __forceinline__ double exp(double a) { return ::exp(a); }
__device__ __forceinline__ float exp(float a) { return ::expf(a); }
__device__ __forceinline__ double log(double a) { return ::log(a); }
__device__ __forceinline__ float log(float a) { return ::logf(a); }
__device__ __forceinline__ double sin(double a) { return ::sin(a); }
__device__ __forceinline__ float sin(float a) { return ::sinf(a); }
__device__ __forceinline__ double cos(double a) { return ::cos(a); }
__device__ __forceinline__ float cos(float a) { return ::cosf(a); }
__device__ __forceinline__ double pow(double a, double e) { return ::pow(a, e); }
__device__ __forceinline__ float pow(float a, double e) { return ::powf(a, float(e)); }

// The unary functions compute value and derivative factors in V and round
// the factors to T once.

// sqrt: f' = 1 / (2 sqrt a), f'' = -f' / (2 a)   (Scalar.hh:266-286)
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> sqrt(const Dual<K, T, V>& a)
{
    const V f0 = tad::sqrt(a.v);
    const V f1 = V(0.5) / f0;
    return chain(a, f0, T(f1), T(-f1 / (V(2) * a.v)));
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> exp(const Dual<K, T, V>& a)
{
    const V f0 = tad::exp(a.v);
    return chain(a, f0, T(f0), T(f0));
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> log(const Dual<K, T, V>& a)
{
    const V ia = V(1) / a.v;
    return chain(a, tad::log(a.v), T(ia), T(-ia * ia));
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> sin(const Dual<K, T, V>& a)
{
    const V s = tad::sin(a.v);
    return chain(a, s, T(tad::cos(a.v)), T(-s));
}

template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> cos(const Dual<K, T, V>& a)
{
    const V c = tad::cos(a.v);
    return chain(a, c, T(-tad::sin(a.v)), T(-c));
}

// a^e for a passive exponent: f' = e a^(e-1), f'' = e (e-1) a^(e-2), built
// up from p = a^(e-2) as p a and p a a (Scalar.hh:292-306), so that a = 0
// stays finite for e >= 2.
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> pow(const Dual<K, T, V>& a, double e)
{
    const V p2 = tad::pow(a.v, e - 2.0);
    const V p1 = p2 * a.v;
    return chain(a, p1 * a.v, T(V(e) * p1), T(V(e * (e - 1.0)) * p2));
}

// Integer exponent: a^0 and a^1 exactly, the rest as above (Scalar.hh:266-290).
template <int K, class T, class V>
__device__ __forceinline__ Dual<K, T, V> pow(const Dual<K, T, V>& a, int e)
{
    if (e == 0) return chain(a, V(1), T(0), T(0));
    if (e == 1) return chain(a, a.v, T(1), T(0));
    const V p2 = tad::pow(a.v, double(e - 2));
    const V p1 = p2 * a.v;
    return chain(a, p1 * a.v, T(V(e) * p1), T(V(double(e) * double(e - 1)) * p2));
}

} // namespace tad
