// Direction solves in a general linear subspace (gfx950):
// newton_direction_reduced_basis (NewtonDirection.hh:76-105) for a sparse
// n x m basis B that does not merely select variables.  The reference forms
// H_r = B^T H B + w I_m and factorizes it; here conjugate gradients run on
// y in R^m with the operator p -> B^T (H (B p)) + w p applied as three CSR
// products (B, the scalar CSR Hessian, B^T), preconditioned by diag(H_r).
// Every sum has a fixed order (kRowLanes lanes per row, fixed shuffle tree), so
// the solve is bitwise reproducible.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_internal.hpp"

namespace tad {

namespace {

constexpr int kThreads = 256;  // lanes per workgroup
constexpr int kRowLanes = 8;  // lanes per CSR row

// y[r] = sum_k a[k] x[c[k]] over row r (+ w add[r] when add is not null)
__global__ __launch_bounds__(kThreads) void k_csr_spmv(int64_t nrows, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci, const double* __restrict__ av,
                                                     const double* __restrict__ x, double w,
                                                     const double* __restrict__ add, double* __restrict__ y)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes;
    double acc = 0.0;
    if (r < nrows)
        for (int32_t k = rp[r] + lane; k < rp[r + 1]; k += kRowLanes) acc += av[k] * x[ci[k]];
#pragma unroll
    for (int off = kRowLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kRowLanes);
    if (r < nrows && lane == 0) y[r] = add ? acc + w * add[r] : acc;
}

// Entry (r, c) of a CSR matrix with sorted column indices (0 when absent).
__device__ __forceinline__ double csr_at(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                         const double* __restrict__ v, int64_t r, int32_t c)
{
    int32_t lo = rp[r], hi = rp[r + 1];
    const int32_t end = hi;
    while (lo < hi)
    {
        const int32_t mid = lo + (hi - lo) / 2;
        if (ci[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && ci[lo] == c) ? v[lo] : 0.0;
}

// dinv[j] = 1 / (w + sum over the entries (p, B_pj), (q, B_qj) of column j of
// B of B_pj B_qj H_pq), column j of B being row j of B^T; dinv[j] = 0 and
// flag[0] = 1 when that diagonal entry of B^T H B + w I is not positive.
__global__ __launch_bounds__(kThreads) void k_basis_diag(int64_t m, const int32_t* __restrict__ tp,
                                                       const int32_t* __restrict__ tc, const double* __restrict__ tv,
                                                       const int32_t* __restrict__ hrp,
                                                       const int32_t* __restrict__ hci, const double* __restrict__ H,
                                                       double w, double* __restrict__ dinv, int* __restrict__ flag)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    double s = 0.0;
    for (int32_t a = tp[j]; a < tp[j + 1]; ++a)
    {
        const int32_t p = tc[a];
        double t = 0.0;
        for (int32_t b = tp[j]; b < tp[j + 1]; ++b) t += tv[b] * csr_at(hrp, hci, H, p, tc[b]);
        s += tv[a] * t;
    }
    s += w;
    if (!(s > 0.0) || !(s < 1.0 / 0.0))
    {
        dinv[j] = 0.0;
        flag[0] = 1;
    }
    else
        dinv[j] = 1.0 / s;
}

__global__ void k_basis_scale(int64_t m, const double* __restrict__ r, const double* __restrict__ dinv,
                              double* __restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) z[i] = dinv[i] * r[i];
}

// y += alpha p, r -= alpha q
__global__ void k_basis_step(int64_t m, double alpha, const double* __restrict__ p, const double* __restrict__ q,
                             double* __restrict__ y, double* __restrict__ r)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    y[i] = fma(alpha, p[i], y[i]);
    r[i] = fma(-alpha, q[i], r[i]);
}

// p = z + beta p
__global__ void k_basis_direction(int64_t m, const double* __restrict__ z, double beta, double* __restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = fma(beta, p[i], z[i]);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

} // namespace

hipError_t launch_csr_spmv(int64_t nrows, const int32_t* row_ptr, const int32_t* col, const double* val,
                           const double* x, double w, const double* add, double* y, hipStream_t s)
{
    if (nrows == 0) return hipSuccess;
    k_csr_spmv<<<grid_of(nrows * kRowLanes), kThreads, 0, s>>>(nrows, row_ptr, col, val, x, w, add, y);
    return hipGetLastError();
}

hipError_t launch_basis_diag(int64_t m, const int32_t* bt_ptr, const int32_t* bt_col, const double* bt_val,
                             const int32_t* h_row_ptr, const int32_t* h_col, const double* H, double w, double* dinv,
                             int* flag, hipStream_t s)
{
    if (m == 0) return hipSuccess;
    k_basis_diag<<<grid_of(m), kThreads, 0, s>>>(m, bt_ptr, bt_col, bt_val, h_row_ptr, h_col, H, w, dinv, flag);
    return hipGetLastError();
}

hipError_t launch_basis_scale(int64_t m, const double* r, const double* dinv, double* z, hipStream_t s)
{
    if (m == 0) return hipSuccess;
    k_basis_scale<<<grid_of(m), kThreads, 0, s>>>(m, r, dinv, z);
    return hipGetLastError();
}

hipError_t launch_basis_step(int64_t m, double alpha, const double* p, const double* q, double* y, double* r,
                             hipStream_t s)
{
    if (m == 0) return hipSuccess;
    k_basis_step<<<grid_of(m), kThreads, 0, s>>>(m, alpha, p, q, y, r);
    return hipGetLastError();
}

hipError_t launch_basis_direction(int64_t m, const double* z, double beta, double* p, hipStream_t s)
{
    if (m == 0) return hipSuccess;
    k_basis_direction<<<grid_of(m), kThreads, 0, s>>>(m, z, beta, p);
    return hipGetLastError();
}

} // namespace tad
