// Dense symmetric positive definite inverse (gfx950) for the coarse matrix of
// the two-level preconditioner: A^-1 by blocked Gauss-Jordan sweeps (the
// sweep operator, Beaton 1964 / Goodnight 1979: sweeping every pivot of an
// SPD matrix turns it into -A^-1; without pivoting, every pivot block is a
// Schur complement and stays positive definite).  Per 64 x 64 pivot block K:
//   Q = M_KK^-1                     (one workgroup, scalar sweeps in LDS)
//   W = Q M_K*,  C = M_K*            (one workgroup per 64-column tile, MFMA)
//   M_IJ -= C_*I^T W_*J  (I, J != K), M_KJ = W_*J, M_IK = W_*I^T, M_KK = -Q
//                                    (one workgroup per tile, MFMA)
// The matrix is padded to a multiple of 64 with identity rows and columns
// (decoupled: their part of the result is -I and is dropped).  Products use
// v_mfma_f64_16x16x4_f64; every sum has a fixed order (deterministic).  The
// reference factorizes its whole Hessian with SimplicialLDLT instead
// (NewtonDirection.hh:34-41); this inverse only serves the preconditioner.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_internal.hpp"

namespace tad {

namespace {

constexpr int kTile = 64;
typedef double tad_v4d __attribute__((ext_vector_type(4)));

// S = A (n x n, column-major) padded to np x np with identity; clears info.
__global__ __launch_bounds__(256) void k_sweep_load(int64_t n, int64_t np, const double* __restrict__ A,
                                                    double* __restrict__ S, int* __restrict__ info)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id == 0) info[0] = 0;
    if (id >= np * np) return;
    const int64_t i = id % np, j = id / np;
    S[id] = (i < n && j < n) ? A[i + j * n] : (i == j ? 1.0 : 0.0);
}

// Q = M_KK^-1 by 64 scalar sweeps of the pivot block in LDS (256 threads:
// thread t owns column t % 64 and rows t / 64 + 4 m).  info = 1 when a pivot
// is not positive or not finite (M not positive definite at working precision).
__global__ __launch_bounds__(256) void k_sweep_pivot(int64_t np, int64_t k0, const double* __restrict__ S,
                                                     double* __restrict__ Q, int* __restrict__ info)
{
    __shared__ double L[kTile][kTile + 1];
    const int t = threadIdx.x;
    const int j = t % kTile, i0 = t / kTile;
#pragma unroll
    for (int m = 0; m < 16; ++m)
    {
        const int i = i0 + 4 * m;
        L[i][j] = S[(k0 + i) + (k0 + j) * np];
    }
    bool bad = false;
    for (int k = 0; k < kTile; ++k)
    {
        __syncthreads();
        const double d = L[k][k];
        const double akj = L[k][j];
        double aik[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) aik[m] = L[i0 + 4 * m][k];
        __syncthreads();
        bad = bad || !(d > 0.0) || !(d < 1.0 / 0.0);
        const double rd = 1.0 / d;
#pragma unroll
        for (int m = 0; m < 16; ++m)
        {
            const int i = i0 + 4 * m;
            double v;
            if (i == k && j == k) v = -rd;
            else if (i == k) v = akj * rd;
            else if (j == k) v = aik[m] * rd;
            else v = L[i][j] - aik[m] * (akj * rd);
            L[i][j] = v;
        }
    }
    __syncthreads();
    // after all sweeps L = -P^-1
#pragma unroll
    for (int m = 0; m < 16; ++m)
    {
        const int i = i0 + 4 * m;
        Q[i + j * kTile] = -L[i][j];
    }
    if (bad) info[0] = 1;
}

// acc[c] (16 x 16 block, rows 16 w .. of the tile, columns 16 c ..) =
// sum_k A(r, k) B(k, c) over k = 0..63, with A(r, k) = a[ra(r) + k * lda_k]
// style accessors given as functors.  MFMA 16x16x4 FP64 operand layout: lane l
// holds A(l % 16, l / 16) and B(l / 16, l % 16); the result register q of
// lane l is D(4 q + l / 16, l % 16) (rows interleave across the lane groups).
template <class FA, class FB>
__device__ __forceinline__ void tile_product(int w, FA a_at, FB b_at, tad_v4d acc[4])
{
    const int lane = threadIdx.x & 63;
    const int r = lane % 16, kk = lane / 16;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = tad_v4d{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 4
    for (int k0 = 0; k0 < kTile; k0 += 4)
    {
        const double av = a_at(16 * w + r, k0 + kk);
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const double bv = b_at(k0 + kk, 16 * c + r);
            acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[c], 0, 0, 0);
        }
    }
}

// Panel of pivot block K: C[:, J] = M_KJ, W[:, J] = Q M_KJ (both 64 x np,
// column-major with leading dimension 64); one workgroup per column tile J.
__global__ __launch_bounds__(256) void k_sweep_panel(int64_t np, int64_t k0, const double* __restrict__ S,
                                                     const double* __restrict__ Q, double* __restrict__ W,
                                                     double* __restrict__ C)
{
    const int64_t j0 = (int64_t)blockIdx.x * kTile;
    const int t = threadIdx.x;
    const int w = t / 64, lane = t & 63;
    for (int e = t; e < kTile * kTile; e += 256)
    {
        const int i = e % kTile, j = e / kTile;
        C[i + (j0 + j) * kTile] = S[(k0 + i) + (j0 + j) * np];
    }
    tad_v4d acc[4];
    tile_product(
        w, [&](int r, int k) { return Q[r + k * kTile]; },
        [&](int k, int c) { return S[(k0 + k) + (j0 + c) * np]; }, acc);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int row = 16 * w + 4 * q + lane / 16, col = 16 * c + lane % 16;
            W[row + (j0 + col) * kTile] = acc[c][q];
        }
}

// Trailing update of pivot block K, one workgroup per 64 x 64 tile (I, J):
// M_IJ -= C_*I^T W_*J off the pivot row and column, the pivot row and column
// from W, the pivot block -Q.
__global__ __launch_bounds__(256) void k_sweep_update(int64_t np, int64_t k0, double* __restrict__ S,
                                                      const double* __restrict__ Q, const double* __restrict__ W,
                                                      const double* __restrict__ C)
{
    const int64_t nt = np / kTile;
    const int64_t I = blockIdx.x % nt, J = blockIdx.x / nt;
    const int64_t i0 = I * kTile, j0 = J * kTile;
    const int t = threadIdx.x;
    const bool ri = i0 == k0, rj = j0 == k0;
    if (ri || rj)
    {
        for (int e = t; e < kTile * kTile; e += 256)
        {
            const int i = e % kTile, j = e / kTile;
            double v;
            if (ri && rj) v = -Q[i + j * kTile];
            else if (ri) v = W[i + (j0 + j) * kTile];
            else v = W[j + (i0 + i) * kTile];
            S[(i0 + i) + (j0 + j) * np] = v;
        }
        return;
    }
    const int w = t / 64, lane = t & 63;
    tad_v4d acc[4];
    tile_product(
        w, [&](int r, int k) { return C[k + (i0 + r) * kTile]; },
        [&](int k, int c) { return W[k + (j0 + c) * kTile]; }, acc);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int row = 16 * w + 4 * q + lane / 16, col = 16 * c + lane % 16;
            double* p = S + (i0 + row) + (j0 + col) * np;
            *p = *p - acc[c][q];
        }
}

// A = -(S + S^T) / 2 on the leading n x n part (the sweeps leave -A^-1).
__global__ __launch_bounds__(256) void k_sweep_store(int64_t n, int64_t np, const double* __restrict__ S,
                                                     double* __restrict__ A)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * n) return;
    const int64_t i = id % n, j = id / n;
    A[i + j * n] = -0.5 * (S[i + j * np] + S[j + i * np]);
}

} // namespace

int64_t spd_inverse_work(int64_t n)
{
    const int64_t np = (n + kTile - 1) / kTile * kTile;
    return np * np + 2 * kTile * np + kTile * kTile;
}

hipError_t launch_spd_inverse(int64_t n, double* A, double* work, int* info, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t np = (n + kTile - 1) / kTile * kTile;
    const int64_t nt = np / kTile;
    double* S = work;
    double* W = S + np * np;
    double* C = W + kTile * np;
    double* Q = C + kTile * np;
    k_sweep_load<<<(unsigned)((np * np + 255) / 256), 256, 0, s>>>(n, np, A, S, info);
    for (int64_t K = 0; K < nt; ++K)
    {
        const int64_t k0 = K * kTile;
        k_sweep_pivot<<<1, 256, 0, s>>>(np, k0, S, Q, info);
        k_sweep_panel<<<(unsigned)nt, 256, 0, s>>>(np, k0, S, Q, W, C);
        k_sweep_update<<<(unsigned)(nt * nt), 256, 0, s>>>(np, k0, S, Q, W, C);
    }
    k_sweep_store<<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(n, np, S, A);
    return hipGetLastError();
}

} // namespace tad
