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

// Q = M_KK^-1 by 64 scalar sweeps of the pivot block (256 threads: thread t
// holds column t % 64, rows 16 (t / 64) + m, m = 0..15, in registers; each
// step publishes the pivot row and column through double-buffered LDS, one
// barrier per step).  The block is loaded and stored through an LDS tile so
// that global accesses run down the columns.  info = 1 when a pivot is not
// positive or not finite (M not positive definite at working precision).
__global__ __launch_bounds__(256) void k_sweep_pivot(int64_t np, int64_t k0, const double* __restrict__ S,
                                                     double* __restrict__ Q, int* __restrict__ info)
{
    __shared__ double L[kTile][kTile + 1];
    __shared__ double prow[2][kTile];
    __shared__ double pcol[2][kTile];
    const int t = threadIdx.x;
    const int j = t % kTile, w = t / kTile;
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
        const int e = t + 256 * q;
        const int i = e % kTile, jj = e / kTile;
        L[i][jj] = S[(k0 + i) + (k0 + jj) * np];
    }
    __syncthreads();
    double r[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) r[m] = L[16 * w + m][j];
    bool bad = false;
    for (int qb = 0; qb < 4; ++qb)  // pivots 16 qb + s; rows 16 qb .. belong to wave qb
    {
#pragma unroll
        for (int s = 0; s < 16; ++s)
        {
            const int k = 16 * qb + s;
            const int buf = s & 1;
            if (w == qb) prow[buf][j] = r[s];
            if (j == k)
            {
#pragma unroll
                for (int m = 0; m < 16; ++m) pcol[buf][16 * w + m] = r[m];
            }
            __syncthreads();
            const double d = prow[buf][k];
            const double akj = prow[buf][j];
            double aik[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) aik[m] = pcol[buf][16 * w + m];
            bad = bad || !(d > 0.0) || !(d < 1.0 / 0.0);
            const double rd = 1.0 / d;
            const double arow = akj * rd;
            const bool is_col = j == k;
#pragma unroll
            for (int m = 0; m < 16; ++m)
            {
                const bool is_row = w == qb && m == s;
                double v = r[m] - aik[m] * arow;
                v = is_col ? aik[m] * rd : v;
                v = is_row ? (is_col ? -rd : arow) : v;
                r[m] = v;
            }
        }
    }
    // after all sweeps the block holds -P^-1
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) L[16 * w + m][j] = r[m];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
        const int e = t + 256 * q;
        const int i = e % kTile, jj = e / kTile;
        Q[i + jj * kTile] = -L[i][jj];
    }
    if (bad) info[0] = 1;
}

// Operand tiles in LDS, both stored [outer index][k] with one double of
// padding per row: A(r, k) = As[r][k], B(k, c) = Bs[c][k].
typedef double Tile[kTile][kTile + 1];

// acc[c] (16 x 16 block: rows 16 w .. 16 w + 15, columns 16 c ..) = sum over
// k = 0..63 of A(r, k) B(k, c), in k order.  MFMA 16x16x4 FP64 operand
// layout: lane l holds A(l % 16, l / 16) and B(l / 16, l % 16); the result
// register q of lane l is D(4 q + l / 16, l % 16) (rows interleave across the
// lane groups).
__device__ __forceinline__ void tile_product(int w, const Tile& As, const Tile& Bs, tad_v4d acc[4])
{
    const int lane = threadIdx.x & 63;
    const int r = lane % 16, kk = lane / 16;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = tad_v4d{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll 4
    for (int k0 = 0; k0 < kTile; k0 += 4)
    {
        const double av = As[16 * w + r][k0 + kk];
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const double bv = Bs[16 * c + r][k0 + kk];
            acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[c], 0, 0, 0);
        }
    }
}

// The product tile into R[row][col] (LDS).
__device__ __forceinline__ void tile_store(int w, const tad_v4d acc[4], Tile& R)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) R[16 * w + 4 * q + lane / 16][16 * c + lane % 16] = acc[c][q];
}

// Panel of pivot block K: C[:, J] = M_KJ, W[:, J] = Q M_KJ (both 64 x np,
// column-major with leading dimension 64); one workgroup per column tile J.
// Global accesses run down the columns; the operands pass through LDS.
__global__ __launch_bounds__(256) void k_sweep_panel(int64_t np, int64_t k0, const double* __restrict__ S,
                                                     const double* __restrict__ Q, double* __restrict__ W,
                                                     double* __restrict__ C)
{
    __shared__ Tile Qs, Ms;
    const int64_t j0 = (int64_t)blockIdx.x * kTile;
    const int t = threadIdx.x;
    const int w = t / 64;
#pragma unroll
    for (int p = 0; p < 16; ++p)
    {
        const int e = t + 256 * p;
        const int i = e % kTile, j = e / kTile;
        Qs[i][j] = Q[e];  // A(r = i, k = j) = Q(i, j)
        const double m = S[(k0 + i) + (j0 + j) * np];
        C[i + (j0 + j) * kTile] = m;
        Ms[j][i] = m;  // B(k = i, c = j)
    }
    __syncthreads();
    tad_v4d acc[4];
    tile_product(w, Qs, Ms, acc);
    __syncthreads();
    tile_store(w, acc, Qs);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 16; ++p)
    {
        const int e = t + 256 * p;
        const int i = e % kTile, j = e / kTile;
        W[i + (j0 + j) * kTile] = Qs[i][j];
    }
}

// Trailing update of pivot block K, one workgroup per 64 x 64 tile (I, J):
// M_IJ -= C_*I^T W_*J off the pivot row and column, the pivot row and column
// from W, the pivot block -Q.  C_*I and W_*J are contiguous 32 KB blocks.
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
    __shared__ Tile Cs, Ws;
    const int w = t / 64;
#pragma unroll
    for (int p = 0; p < 16; ++p)
    {
        const int e = t + 256 * p;
        const int k = e % kTile, o = e / kTile;
        Cs[o][k] = C[k + (i0 + o) * kTile];  // A(r = o, k) = C(k, i0 + o)
        Ws[o][k] = W[k + (j0 + o) * kTile];  // B(k, c = o) = W(k, j0 + o)
    }
    __syncthreads();
    tad_v4d acc[4];
    tile_product(w, Cs, Ws, acc);
    __syncthreads();
    tile_store(w, acc, Cs);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 16; ++p)
    {
        const int e = t + 256 * p;
        const int i = e % kTile, j = e / kTile;
        double* ptr = S + (i0 + i) + (j0 + j) * np;
        *ptr = *ptr - Cs[i][j];
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
