// Element type table shared by the element and energy kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_internal.hpp"
#include "tad_energies.hpp"
#include "tinyad_amd/device/Kernels.hpp"

namespace tad {

// ---------------------------------------------------------------------------
// Element types
// ---------------------------------------------------------------------------
template <int Dm>
struct SymdirElem
{
    static constexpr int N = Dm + 1, D = Dm, NP = Dm * Dm + 1;
    static constexpr bool tunable = true;  // has persistent launch variants
    template <class T, int MODE, class Sink>
    __device__ static double eval(const double* X, const double* p, T eps, const Sink& out, bool& nf)
    {
        return symdir_element<Dm, T, MODE>(X, p, eps, out, nf);
    }
    __device__ static double energy(const double* X, const double* p) { return energy_symdir<Dm>(X, p); }
};

template <int Dm>
struct PenaltyElem
{
    static constexpr int N = 1, D = Dm, NP = Dm + 1;
    static constexpr bool tunable = false;
    template <class T, int MODE, class Sink>
    __device__ static double eval(const double* X, const double* p, T eps, const Sink& out, bool& nf)
    {
        return penalty_element<Dm, T, MODE>(X, p, eps, out, nf);
    }
    __device__ static double energy(const double* X, const double* p) { return energy_penalty<Dm>(X, p); }
};

template <int NN, int Dm>
struct QuadElem
{
    static constexpr int N = NN, D = Dm, K = NN * Dm, NP = K * K + K + 1;
    static constexpr bool tunable = false;
    static constexpr bool merges = true;  // a handle may appear more than once
    template <class T, int MODE, class Sink>
    __device__ static double eval(const double* X, const double* p, T eps, const Sink& out, bool& nf)
    {
        return quadratic_element<NN, Dm, T, MODE>(X, p, eps, out, nf);
    }
    __device__ static double energy(const double* X, const double* p) { return energy_quadratic<K>(X, p); }
};

} // namespace tad
