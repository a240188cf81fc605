// Element kernel (gfx950): per-element value, gradient and (projected) Hessian,
// one lane per element.  Replaces the OpenMP element loop of
// ScalarObjectiveTerm.hh:242-254 and the per-element projection of
// HessianProjection.hh:48-101 (see tad_energies.hpp for the math).
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_elements.hpp"

namespace tad {

// ---------------------------------------------------------------------------
// Element kernel
// ---------------------------------------------------------------------------
template <class E, class T, int MODE>
__global__ __launch_bounds__(kBlock) void k_element(int64_t n, const int32_t* __restrict__ conn,
                                                    const double* __restrict__ params, const double* __restrict__ x,
                                                    double eps, double* __restrict__ elem_f, T* __restrict__ rec,
                                                    int* __restrict__ err)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    constexpr int RS = record_size(E::N, E::D);
    double X[E::N * E::D];
#pragma unroll
    for (int a = 0; a < E::N; ++a)
    {
        const int64_t v = conn[a * n + e];
#pragma unroll
        for (int i = 0; i < E::D; ++i) X[a * E::D + i] = x[v * E::D + i];
    }
    double p[E::NP];
#pragma unroll
    for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
    bool nf = false;
    elem_f[e] = E::template eval<T, MODE>(X, p, T(eps), rec + e * RS, nf);
    if (nf) *err = 1;
}

template <class E, class T>
static hipError_t launch_element_t(const TermDev& t, int mode, const double* x, double eps, double* elem_f,
                                   T* arena, int* err, hipStream_t s)
{
    const int64_t n = t.n_elem;
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    T* rec = arena + t.rec_base;
    double* ef = elem_f + t.f_base;
    switch (mode)
    {
        case MODE_GRAD: k_element<E, T, MODE_GRAD><<<grid, kBlock, 0, s>>>(n, t.conn, t.params, x, eps, ef, rec, err); break;
        case MODE_HESS: k_element<E, T, MODE_HESS><<<grid, kBlock, 0, s>>>(n, t.conn, t.params, x, eps, ef, rec, err); break;
        default: k_element<E, T, MODE_HESS_PROJ><<<grid, kBlock, 0, s>>>(n, t.conn, t.params, x, eps, ef, rec, err); break;
    }
    return hipGetLastError();
}

template <class T>
static hipError_t launch_element_p(const TermDev& t, int d, int mode, const double* x, double eps, double* elem_f,
                                   T* arena, int* err, hipStream_t s)
{
    switch (t.energy)
    {
        case E_SYMDIR_TRI: return launch_element_t<SymdirElem<2>, T>(t, mode, x, eps, elem_f, arena, err, s);
        case E_SYMDIR_TET: return launch_element_t<SymdirElem<3>, T>(t, mode, x, eps, elem_f, arena, err, s);
        case E_PENALTY:
            if (d == 1) return launch_element_t<PenaltyElem<1>, T>(t, mode, x, eps, elem_f, arena, err, s);
            if (d == 2) return launch_element_t<PenaltyElem<2>, T>(t, mode, x, eps, elem_f, arena, err, s);
            return launch_element_t<PenaltyElem<3>, T>(t, mode, x, eps, elem_f, arena, err, s);
        case E_QUADRATIC:
        {
#define TAD_QUAD(NN, DD) \
    if (t.N == NN && d == DD) return launch_element_t<QuadElem<NN, DD>, T>(t, mode, x, eps, elem_f, arena, err, s);
            TAD_QUAD(1, 1) TAD_QUAD(1, 2) TAD_QUAD(1, 3)
            TAD_QUAD(2, 1) TAD_QUAD(2, 2) TAD_QUAD(2, 3)
            TAD_QUAD(3, 1) TAD_QUAD(3, 2) TAD_QUAD(3, 3)
            TAD_QUAD(4, 1) TAD_QUAD(4, 2) TAD_QUAD(4, 3)
#undef TAD_QUAD
            return hipErrorInvalidValue;
        }
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_element(const TermDev& t, int d, int precision, int mode, const double* x, double eps,
                          double* elem_f, void* arena, int* err, hipStream_t s)
{
    if (precision == 0)
        return launch_element_p<double>(t, d, mode, x, eps, elem_f, (double*)arena, err, s);
    return launch_element_p<float>(t, d, mode, x, eps, elem_f, (float*)arena, err, s);
}

} // namespace tad
