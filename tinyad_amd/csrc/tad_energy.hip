// Energy kernel (gfx950): passive FP64 energies at x + s_c d for a batch of step
// sizes, the evaluation the Armijo line search performs per trial
// (ScalarObjectiveTerm.hh:162-186 as called from LineSearch.hh:49-52).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
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

// ---------------------------------------------------------------------------
// Energy without library allocations (tad_fn_eval_noalloc)
// ---------------------------------------------------------------------------
// The per-workgroup partial sums live in this file-scope device array instead
// of memory the library allocates: one per workgroup of kBlock elements, so
// kNoAllocPartials * kBlock = 16,777,216 elements at most (512 KiB per
// device, held until the process ends).  Host code never passes the array's
// name to a kernel: each device's address comes from hipGetSymbolAddress
// (noalloc_scratch below).  Calls that overlap on one device share the array
// and race.
static __device__ __align__(256) double scratch[kNoAllocPartials];

hipError_t noalloc_scratch(int device, double** ptr)
{
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device])
    {
        // the caller has made `device` current: the symbol resolves to that
        // device's copy of the array
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double*>(p);
    }
    *ptr = addr[device];
    return hipSuccess;
}

// partials[blockIdx.x] = sum of the passive energies of the block's kBlock
// elements (0 past the end), a fixed-order tree sum in LDS.
template <class E>
__global__ __launch_bounds__(kBlock) void k_energy_partials(int64_t n, const int32_t* __restrict__ conn,
                                                            const double* __restrict__ params,
                                                            const double* __restrict__ x, double* __restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double v = 0.0;
    if (e < n)
    {
        double X[E::N * E::D];
#pragma unroll
        for (int a = 0; a < E::N; ++a)
        {
            const int64_t h = conn[a * n + e];
#pragma unroll
            for (int i = 0; i < E::D; ++i) X[a * E::D + i] = x[h * E::D + i];
        }
        double p[E::NP > 0 ? E::NP : 1];
#pragma unroll
        for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
        v = E::energy(X, p);
    }
    sh[threadIdx.x] = v;
#pragma unroll
    for (int w = kBlock / 2; w > 0; w /= 2)
    {
        __syncthreads();
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// out[0] = sum of partials[0 .. count), one workgroup, fixed order.
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partials, int64_t count,
                                                         double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += kBlock) acc += partials[i];
    sh[threadIdx.x] = acc;
#pragma unroll
    for (int w = kBlock / 2; w > 0; w /= 2)
    {
        __syncthreads();
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

template <class E>
static hipError_t launch_partials_t(const TermDev& t, const double* x, double* partials, hipStream_t s)
{
    const unsigned grid = (unsigned)((t.n_elem + kBlock - 1) / kBlock);
    k_energy_partials<E><<<grid, kBlock, 0, s>>>(t.n_elem, t.conn, t.params, x, partials);
    return hipGetLastError();
}

int64_t noalloc_partials(const TermDev* terms, int n_terms)
{
    int64_t c = 0;
    for (int i = 0; i < n_terms; ++i) c += (terms[i].n_elem + kBlock - 1) / kBlock;
    return c;
}

hipError_t launch_energy_noalloc(const TermDev* terms, int n_terms, int d, const double* x, double* scratch_ptr,
                                 double* f_out, hipStream_t s)
{
    const int64_t count = noalloc_partials(terms, n_terms);
    if (count > kNoAllocPartials) return hipErrorInvalidValue;
    double* part = scratch_ptr;
    for (int i = 0; i < n_terms; ++i)
    {
        const TermDev& t = terms[i];
        if (t.n_elem == 0) continue;
        hipError_t e = hipErrorInvalidValue;
        switch (t.energy)
        {
            case E_SYMDIR_TRI: e = launch_partials_t<SymdirElem<2>>(t, x, part, s); break;
            case E_SYMDIR_TET: e = launch_partials_t<SymdirElem<3>>(t, x, part, s); break;
            case E_PENALTY:
                if (d == 1) e = launch_partials_t<PenaltyElem<1>>(t, x, part, s);
                else if (d == 2) e = launch_partials_t<PenaltyElem<2>>(t, x, part, s);
                else e = launch_partials_t<PenaltyElem<3>>(t, x, part, s);
                break;
            case E_QUADRATIC:
            {
#define TAD_QUAD(NN, DD) \
    else if (t.N == NN && d == DD) e = launch_partials_t<QuadElem<NN, DD>>(t, x, part, s);
                if (false) {}
                TAD_QUAD(1, 1) TAD_QUAD(1, 2) TAD_QUAD(1, 3)
                TAD_QUAD(2, 1) TAD_QUAD(2, 2) TAD_QUAD(2, 3)
                TAD_QUAD(3, 1) TAD_QUAD(3, 2)
                TAD_QUAD(4, 1) TAD_QUAD(4, 2)
#undef TAD_QUAD
                break;
            }
            default: break;
        }
        if (e != hipSuccess) return e;
        part += (t.n_elem + kBlock - 1) / kBlock;
    }
    k_sum_partials<<<1, kBlock, 0, s>>>(scratch_ptr, count, f_out);
    return hipGetLastError();
}

} // namespace tad
