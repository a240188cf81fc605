// Element and energy kernels over an element type E, one lane per element:
// the library instantiates them for its built-in energy catalog and
// DeviceElements.hh for energies written by the caller.  E provides N
// (valence), D (dimension), NP (parameters per element),
//   template <class T, int MODE, class Sink>
//   static double eval(const double* X, const double* p, T eps, const Sink&, bool& nonfinite)
// (value returned, derivatives handed to the sink) and
//   static double energy(const double* X, const double* p)  (passive FP64).
// Element data is SoA: conn[a * n + e], params[k * n + e]; x is AoS (d per handle).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include "Scatter.hpp"

namespace tad {

constexpr int kBlock = 256;   // lanes per workgroup of the plain element grid
constexpr int kMaxCand = 8;   // step sizes per energy launch

// E::padded (optional, default false): elements may end in padding entries
// (handle -1: coordinates read as 0, no slot), the elements of a
// dynamic-valence term.
template <class E, class = void>
struct padded_slots : std::false_type
{
};
template <class E>
struct padded_slots<E, std::void_t<decltype(E::padded)>> : std::bool_constant<E::padded>
{
};

// E::merges (optional, default false): an element may list a handle more than
// once; its entries are merged into the first occurrence before projection
// (dense records only) and only first occurrences get slots.
template <class E, class = void>
struct merged_slots : std::false_type
{
};
template <class E>
struct merged_slots<E, std::void_t<decltype(E::merges)>> : std::bool_constant<E::merges>
{
};

// ---------------------------------------------------------------------------
// Element kernel
// ---------------------------------------------------------------------------
// Scatter targets of one term: contribution arrays and the term's slot maps.
template <class T>
struct ElemOut
{
    T* G;
    T* Hc;
    const uint32_t* gpos;
    const uint32_t* hpos;
};

template <class E, class T, int MODE>
__device__ __forceinline__ void element_body(int64_t e, int64_t n, const int32_t* __restrict__ conn,
                                             const double* __restrict__ params, const double* __restrict__ x,
                                             double eps, double* __restrict__ elem_f, const ElemOut<T>& o,
                                             int* __restrict__ err)
{
    constexpr bool kMerge = merged_slots<E>::value && E::N > 1;
    double X[E::N * E::D];
    int n_active = 0;  // handles before the padding (padded terms)
    int32_t h[E::N];
#pragma unroll
    for (int a = 0; a < E::N; ++a)
    {
        const int64_t v = conn[a * n + e];
        h[a] = (int32_t)v;
        if constexpr (padded_slots<E>::value)
        {
            n_active += v >= 0 ? 1 : 0;
#pragma unroll
            for (int i = 0; i < E::D; ++i) X[a * E::D + i] = v >= 0 ? x[v * E::D + i] : 0.0;
        }
        else
        {
#pragma unroll
            for (int i = 0; i < E::D; ++i) X[a * E::D + i] = x[v * E::D + i];
        }
    }
    double p[E::NP > 0 ? E::NP : 1];
#pragma unroll
    for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
    bool nf = false;
    ScatterSink<E::N, E::D, T, padded_slots<E>::value || kMerge> out{o.G, o.Hc, o.gpos, o.hpos, e, n};
    if constexpr (kMerge)
    {
        // entry a repeats entry r < a (the first with its handle; padding
        // entries repeat nothing)
#pragma unroll
        for (int a = 1; a < E::N; ++a)
        {
            int r = a;
#pragma unroll
            for (int b = a - 1; b >= 0; --b) r = (h[b] == h[a] && h[a] >= 0) ? b : r;
            if (r != a)
            {
                out.dups |= 1u << a;
                out.reps |= (uint64_t)r << (4 * a);
            }
        }
    }
    if constexpr (padded_slots<E>::value)
        elem_f[e] = E::template eval<T, MODE>(X, p, T(eps), out, nf, n_active);
    else
        elem_f[e] = E::template eval<T, MODE>(X, p, T(eps), out, nf);
    if (nf) *err = 1;
}

// One lane per element, one tile of blockDim elements per workgroup.
template <class E, class T, int MODE>
__global__ __launch_bounds__(kBlock) void k_element(int64_t n, const int32_t* __restrict__ conn,
                                                    const double* __restrict__ params, const double* __restrict__ x,
                                                    double eps, double* __restrict__ elem_f, ElemOut<T> o,
                                                    int* __restrict__ err)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    element_body<E, T, MODE>(e, n, conn, params, x, eps, elem_f, o, err);
}

// ---------------------------------------------------------------------------
// Energy kernel (line search)
// ---------------------------------------------------------------------------
template <class E>
__global__ __launch_bounds__(kBlock) void k_energy(int64_t n, const int32_t* __restrict__ conn,
                                                   const double* __restrict__ params, const double* __restrict__ x,
                                                   const double* __restrict__ dir, const double* __restrict__ steps,
                                                   int ncand, double* __restrict__ out, int64_t stride)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    constexpr int NX = E::N * E::D;
    double X[NX], DX[NX];
#pragma unroll
    for (int a = 0; a < E::N; ++a)
    {
        const int64_t v = conn[a * n + e];
        if constexpr (padded_slots<E>::value)
        {
#pragma unroll
            for (int i = 0; i < E::D; ++i)
            {
                X[a * E::D + i] = v >= 0 ? x[v * E::D + i] : 0.0;
                DX[a * E::D + i] = (dir && v >= 0) ? dir[v * E::D + i] : 0.0;
            }
        }
        else
        {
#pragma unroll
            for (int i = 0; i < E::D; ++i)
            {
                X[a * E::D + i] = x[v * E::D + i];
                DX[a * E::D + i] = dir ? dir[v * E::D + i] : 0.0;
            }
        }
    }
    double p[E::NP > 0 ? E::NP : 1];
#pragma unroll
    for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
    if (!dir)
    {
        out[e] = E::energy(X, p);
        return;
    }
    for (int c = 0; c < ncand; ++c)
    {
        const double sc = steps[c];
        double Xc[NX];
#pragma unroll
        for (int i = 0; i < NX; ++i) Xc[i] = fma(sc, DX[i], X[i]);
        out[c * stride + e] = E::energy(Xc, p);
    }
}

} // namespace tad
