// Energy kernel (gfx950): passive FP64 energies at x + s_c d for a batch of step
// sizes, the evaluation the Armijo line search performs per trial
// (ScalarObjectiveTerm.hh:162-186 as called from LineSearch.hh:49-52).
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_elements.hpp"

namespace tad {

// ---------------------------------------------------------------------------
// Energy kernel (line search)
// ---------------------------------------------------------------------------
constexpr int kMaxCand = 8;

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
#pragma unroll
        for (int i = 0; i < E::D; ++i)
        {
            X[a * E::D + i] = x[v * E::D + i];
            DX[a * E::D + i] = dir ? dir[v * E::D + i] : 0.0;
        }
    }
    double p[E::NP];
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

template <class E>
static hipError_t launch_energy_t(const TermDev& t, const double* x, const double* dir, const double* steps, int ncand,
                                  double* elem_f, int64_t stride, hipStream_t s)
{
    const int64_t n = t.n_elem;
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    k_energy<E><<<grid, kBlock, 0, s>>>(n, t.conn, t.params, x, dir, steps, ncand, elem_f + t.f_base, stride);
    return hipGetLastError();
}

hipError_t launch_energy(const TermDev& t, int d, const double* x, const double* dir, const double* steps, int ncand,
                         double* elem_f, int64_t stride, hipStream_t s)
{
    if (ncand < 1 || ncand > kMaxCand) return hipErrorInvalidValue;
    switch (t.energy)
    {
        case E_SYMDIR_TRI: return launch_energy_t<SymdirElem<2>>(t, x, dir, steps, ncand, elem_f, stride, s);
        case E_SYMDIR_TET: return launch_energy_t<SymdirElem<3>>(t, x, dir, steps, ncand, elem_f, stride, s);
        case E_PENALTY:
            if (d == 1) return launch_energy_t<PenaltyElem<1>>(t, x, dir, steps, ncand, elem_f, stride, s);
            if (d == 2) return launch_energy_t<PenaltyElem<2>>(t, x, dir, steps, ncand, elem_f, stride, s);
            return launch_energy_t<PenaltyElem<3>>(t, x, dir, steps, ncand, elem_f, stride, s);
        case E_QUADRATIC:
        {
#define TAD_QUAD(NN, DD) \
    if (t.N == NN && d == DD) return launch_energy_t<QuadElem<NN, DD>>(t, x, dir, steps, ncand, elem_f, stride, s);
            TAD_QUAD(1, 1) TAD_QUAD(1, 2) TAD_QUAD(1, 3)
            TAD_QUAD(2, 1) TAD_QUAD(2, 2) TAD_QUAD(2, 3)
            TAD_QUAD(3, 1) TAD_QUAD(3, 2)
            TAD_QUAD(4, 1) TAD_QUAD(4, 2)
#undef TAD_QUAD
            return hipErrorInvalidValue;
        }
        default: return hipErrorInvalidValue;
    }
}

} // namespace tad
