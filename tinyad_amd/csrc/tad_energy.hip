// Energy kernel (gfx950): passive FP64 energies at x + s_c d for a batch of step
// sizes, the evaluation the Armijo line search performs per trial
// (ScalarObjectiveTerm.hh:162-186 as called from LineSearch.hh:49-52).
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_elements.hpp"

namespace tad {

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
    if (t.dev_ops.energy)
    {
        if (t.n_elem == 0) return hipSuccess;
        tad_device_energy_args a{};
        a.n_elem = t.n_elem;
        a.conn = t.conn;
        a.params = t.params;
        a.x = x;
        a.dir = dir;
        a.steps = steps;
        a.ncand = ncand;
        a.out = elem_f + t.f_base;
        a.stride = stride;
        a.stream = (void*)s;
        return (hipError_t)t.dev_ops.energy(&a);
    }
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
