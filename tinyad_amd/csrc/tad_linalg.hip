// HIP kernels of the sparse second-order path for gfx950 (linear algebra part)
// and their launchers.
//
//   k_sum_*, k_dot_*   deterministic FP64 sums (fixed tree, no atomics)
//   k_assemble_*       element-order gather of gradient and CSR values
//                      through the precomputed maps (replaces the serial
//                      triplet loop ScalarObjectiveTerm.hh:257-277 and
//                      setFromTriplets, ScalarFunctionImpl.hh:398)
//   k_pcg_*, k_spmv    block-Jacobi PCG on the block-structured CSR
//                      (replaces SimplicialLDLT, NewtonDirection.hh:25-48)
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include "tad_internal.hpp"
#include "tad_energies.hpp"
#include "tad_elements.hpp"

namespace tad {

// ---------------------------------------------------------------------------
// Deterministic reductions
// ---------------------------------------------------------------------------
// Fixed-order block reduction: butterfly inside each 64-lane wave, then the
// four wave totals in order.  Same bits for the same inputs on every run.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) r += sh[w];
    return r;
}

// Same fixed-order reduction for a block of NT threads (NT a multiple of 64).
template <int NT>
__device__ __forceinline__ double block_sum_n(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r += sh[w];
    return r;
}

// acc over v[i], v[i + S], v[i + 2S], ... < hi in that order, four loads in flight
template <int S>
__device__ __forceinline__ double strided_sum(const double* __restrict__ v, int64_t i, int64_t hi)
{
    double acc = 0.0;
    for (; i + 3 * S < hi; i += 4 * S)
    {
        const double a0 = v[i], a1 = v[i + S], a2 = v[i + 2 * S], a3 = v[i + 3 * S];
        acc += a0;
        acc += a1;
        acc += a2;
        acc += a3;
    }
    for (; i < hi; i += S) acc += v[i];
    return acc;
}

// block b of row c sums the contiguous chunk [b*chunk, (b+1)*chunk)
__global__ __launch_bounds__(kBlock) void k_sum_stage1(const double* __restrict__ in, int64_t n, int64_t stride,
                                                       double* __restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int c = blockIdx.y;
    const int64_t chunk = (n + kReduceBlocks - 1) / kReduceBlocks;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    const double* row = in + c * stride;
    const double acc = strided_sum<kBlock>(row, lo + threadIdx.x, hi);
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[c * kReduceBlocks + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_stage2(const double* __restrict__ partials, double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < kReduceBlocks; i += kBlock) acc += partials[c * kReduceBlocks + i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[c] = s;
}

hipError_t launch_sum(const double* in, int64_t n, int ncand, int64_t stride, double* partials, double* out,
                      hipStream_t s)
{
    k_sum_stage1<<<dim3(kReduceBlocks, ncand), kBlock, 0, s>>>(in, n, stride, partials);
    k_sum_stage2<<<ncand, kBlock, 0, s>>>(partials, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_dot_stage1(const double* __restrict__ a, const double* __restrict__ b,
                                                       int64_t n, double* __restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int64_t chunk = (n + kReduceBlocks - 1) / kReduceBlocks;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    double acc = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kBlock) acc += a[i] * b[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

hipError_t launch_dot(const double* a, const double* b, int64_t n, double* partials, double* out, hipStream_t s)
{
    k_dot_stage1<<<kReduceBlocks, kBlock, 0, s>>>(a, b, n, partials);
    k_sum_stage2<<<1, kBlock, 0, s>>>(partials, out);
    return hipGetLastError();
}

__global__ void k_fill(double* p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_fill<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, n, v);
    return hipGetLastError();
}

__global__ void k_axpy_point(const double* __restrict__ x0, const double* __restrict__ dir,
                             const double* __restrict__ step, int64_t n, double* __restrict__ xo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xo[i] = fma(step[0], dir[i], x0[i]);
}

hipError_t launch_axpy_point(const double* x0, const double* dir, const double* step, int64_t n, double* x_out,
                             hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_axpy_point<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(x0, dir, step, n, x_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Assembly: contiguous sums over destination-sorted contributions
// ---------------------------------------------------------------------------
// The element kernels scatter every d x d block (and every per-vertex gradient
// part) to its slot; the slots of one destination are contiguous and in
// element order, so each output entry is a fixed-order sum over one run.
// sum_{c in [c0, c1)} src[c * S] in index order, four loads in flight
template <int S, class T>
__device__ __forceinline__ double run_sum(const T* __restrict__ src, int64_t c0, int64_t c1)
{
    double acc = 0.0;
    int64_t c = c0;
    for (; c + 4 <= c1; c += 4)
    {
        const double a0 = (double)src[c * S], a1 = (double)src[(c + 1) * S];
        const double a2 = (double)src[(c + 2) * S], a3 = (double)src[(c + 3) * S];
        acc += a0;
        acc += a1;
        acc += a2;
        acc += a3;
    }
    for (; c < c1; ++c) acc += (double)src[c * S];
    return acc;
}

// Runs exist for the upper blocks; a lower block (u, v), u > v, sums the run
// of (v, u) transposed.  D*D lanes share one block (lane m sums entry m over
// the run), so every contribution is read as one contiguous D*D-value segment
// by consecutive lanes; a 64-lane wave serves floor(64 / D^2) blocks.
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_assemble_hessian(int64_t nb, const int32_t* __restrict__ block_row,
                                                             const int32_t* __restrict__ brow_ptr,
                                                             const int64_t* __restrict__ hslot_ptr,
                                                             const int32_t* __restrict__ mirror,
                                                             const T* __restrict__ Hc,
                                                             const int32_t* __restrict__ row_ptr,
                                                             double* __restrict__ H)
{
    constexpr int DD = D * D;
    constexpr int BPW = 64 / DD;  // blocks per wave
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane >= BPW * DD) return;
    const int64_t b = wave * BPW + lane / DD;
    if (b >= nb) return;
    const int m = lane % DD;
    const int u = block_row[b];
    const int b0 = brow_ptr[u];
    const int k = (int)(b - b0);
    const int nbu = brow_ptr[u + 1] - b0;
    const int i = m / D, j = m % D;
    const int64_t mb = mirror[b];
    const int src = (mb == b) ? m : j * D + i;
    const int64_t c1 = hslot_ptr[mb + 1];
    const double acc = run_sum<DD>(Hc + src, hslot_ptr[mb], c1);
    H[(int64_t)row_ptr[D * u] + (int64_t)i * D * nbu + D * k + j] = acc;
}

// The same sums with one thread per block (all D*D entries of a contribution
// per thread, two contributions in flight); identical bits.  Which shape is
// faster depends on d, the precision and the run lengths (the autotuner times
// both).
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_tpb(int64_t nb, const int32_t* __restrict__ block_row,
                                                                 const int32_t* __restrict__ brow_ptr,
                                                                 const int64_t* __restrict__ hslot_ptr,
                                                                 const int32_t* __restrict__ mirror,
                                                                 const T* __restrict__ Hc,
                                                                 const int32_t* __restrict__ row_ptr,
                                                                 double* __restrict__ H)
{
    constexpr int DD = D * D;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int u = block_row[b];
    const int b0 = brow_ptr[u];
    const int k = (int)(b - b0);
    const int nbu = brow_ptr[u + 1] - b0;
    const int64_t mb = mirror[b];
    double acc[DD];
#pragma unroll
    for (int m = 0; m < DD; ++m) acc[m] = 0.0;
    int64_t c = hslot_ptr[mb];
    const int64_t c1 = hslot_ptr[mb + 1];
    for (; c + 2 <= c1; c += 2)
    {
        double v0[DD], v1[DD];
#pragma unroll
        for (int m = 0; m < DD; ++m)
        {
            v0[m] = (double)Hc[c * DD + m];
            v1[m] = (double)Hc[(c + 1) * DD + m];
        }
#pragma unroll
        for (int m = 0; m < DD; ++m)
        {
            acc[m] += v0[m];
            acc[m] += v1[m];
        }
    }
    for (; c < c1; ++c)
    {
#pragma unroll
        for (int m = 0; m < DD; ++m) acc[m] += (double)Hc[c * DD + m];
    }
    const bool tr = mb != b;
    const int64_t base = row_ptr[D * u];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            H[base + (int64_t)i * D * nbu + D * k + j] = tr ? acc[j * D + i] : acc[i * D + j];
}

// Upper-run shapes: each run is summed once and written to its upper block
// and, transposed, to the mirrored lower block, so every contribution is read
// once (shapes 0 / 1 read the off-diagonal runs twice).  D*D lanes per upper
// block, lane m = (i, j) writes H_upper(i, j) and H_lower(j, i).
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_upper(int64_t nu, const int64_t* __restrict__ uslot_ptr,
                                                                   const int4* __restrict__ uplace,
                                                                   const T* __restrict__ Hc, double* __restrict__ H)
{
    constexpr int DD = D * D;
    constexpr int BPW = 64 / DD;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane >= BPW * DD) return;
    const int64_t b = wave * BPW + lane / DD;
    if (b >= nu) return;
    const int m = lane % DD;
    const int i = m / D, j = m % D;
    const double acc = run_sum<DD>(Hc + m, uslot_ptr[b], uslot_ptr[b + 1]);
    const int4 pl = uplace[b];
    H[(int64_t)pl.x + (int64_t)i * pl.y + j] = acc;
    if (pl.z >= 0) H[(int64_t)pl.z + (int64_t)j * pl.w + i] = acc;
}

// One thread per upper block, two contributions in flight.
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_upper_tpb(int64_t nu,
                                                                       const int64_t* __restrict__ uslot_ptr,
                                                                       const int4* __restrict__ uplace,
                                                                       const T* __restrict__ Hc,
                                                                       double* __restrict__ H)
{
    constexpr int DD = D * D;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nu) return;
    double acc[DD];
#pragma unroll
    for (int m = 0; m < DD; ++m) acc[m] = 0.0;
    int64_t c = uslot_ptr[b];
    const int64_t c1 = uslot_ptr[b + 1];
    for (; c + 2 <= c1; c += 2)
    {
        double v0[DD], v1[DD];
#pragma unroll
        for (int m = 0; m < DD; ++m)
        {
            v0[m] = (double)Hc[c * DD + m];
            v1[m] = (double)Hc[(c + 1) * DD + m];
        }
#pragma unroll
        for (int m = 0; m < DD; ++m)
        {
            acc[m] += v0[m];
            acc[m] += v1[m];
        }
    }
    for (; c < c1; ++c)
    {
#pragma unroll
        for (int m = 0; m < DD; ++m) acc[m] += (double)Hc[c * DD + m];
    }
    const int4 pl = uplace[b];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) H[(int64_t)pl.x + (int64_t)i * pl.y + j] = acc[i * D + j];
    if (pl.z >= 0)
    {
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) H[(int64_t)pl.z + (int64_t)j * pl.w + i] = acc[i * D + j];
    }
}

// D lanes per handle, floor(64 / D) handles per wave.
template <class T, int D>
__global__ __launch_bounds__(kBlock) void k_assemble_gradient(int64_t nh, const int64_t* __restrict__ gslot_ptr,
                                                              const T* __restrict__ G, double* __restrict__ g)
{
    constexpr int HPW = 64 / D;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane >= HPW * D) return;
    const int64_t v = wave * HPW + lane / D;
    if (v >= nh) return;
    const int i = lane % D;
    g[v * D + i] = run_sum<D>(G + i, gslot_ptr[v], gslot_ptr[v + 1]);
}

template <class T, int D>
static void launch_assemble_hessian_td(int shape, const HessAsm& a, const T* Hc, double* H, hipStream_t s)
{
    constexpr int bpw = 64 / (D * D);
    const int64_t n = shape >= 2 ? a.n_upper : a.n_blocks;
    const unsigned gt = (unsigned)((n + kBlock - 1) / kBlock);
    const unsigned gl = (unsigned)((((n + bpw - 1) / bpw) * 64 + kBlock - 1) / kBlock);
    if (shape == 0)
        k_assemble_hessian<T, D><<<gl, kBlock, 0, s>>>(a.n_blocks, a.block_row, a.brow_ptr, a.hslot_ptr, a.mirror, Hc,
                                                       a.row_ptr, H);
    else if (shape == 1)
        k_assemble_hessian_tpb<T, D><<<gt, kBlock, 0, s>>>(a.n_blocks, a.block_row, a.brow_ptr, a.hslot_ptr, a.mirror,
                                                           Hc, a.row_ptr, H);
    else if (shape == 2)
        k_assemble_hessian_upper<T, D><<<gl, kBlock, 0, s>>>(a.n_upper, a.uslot_ptr, a.uplace, Hc, H);
    else
        k_assemble_hessian_upper_tpb<T, D><<<gt, kBlock, 0, s>>>(a.n_upper, a.uslot_ptr, a.uplace, Hc, H);
}

template <class T>
static hipError_t launch_assemble_hessian_t(int d, int shape, const HessAsm& a, const T* Hc, double* H, hipStream_t s)
{
    if (a.n_blocks == 0) return hipSuccess;
    if (shape < 0 || shape >= kAsmShapes) return hipErrorInvalidValue;
    if (d == 1) launch_assemble_hessian_td<T, 1>(shape, a, Hc, H, s);
    else if (d == 2) launch_assemble_hessian_td<T, 2>(shape, a, Hc, H, s);
    else launch_assemble_hessian_td<T, 3>(shape, a, Hc, H, s);
    return hipGetLastError();
}

hipError_t launch_assemble_hessian(int d, int precision, int shape, const HessAsm& a, const void* Hc, double* H,
                                   hipStream_t s)
{
    if (precision == 0) return launch_assemble_hessian_t<double>(d, shape, a, (const double*)Hc, H, s);
    return launch_assemble_hessian_t<float>(d, shape, a, (const float*)Hc, H, s);
}

template <class T>
static hipError_t launch_assemble_gradient_t(int d, int64_t nh, const int64_t* gslot_ptr, const T* G, double* g,
                                             hipStream_t s)
{
    if (nh == 0) return hipSuccess;
    const int hpw = 64 / d;
    const int64_t waves = (nh + hpw - 1) / hpw;
    const unsigned grid = (unsigned)((waves * 64 + kBlock - 1) / kBlock);
    if (d == 1) k_assemble_gradient<T, 1><<<grid, kBlock, 0, s>>>(nh, gslot_ptr, G, g);
    else if (d == 2) k_assemble_gradient<T, 2><<<grid, kBlock, 0, s>>>(nh, gslot_ptr, G, g);
    else k_assemble_gradient<T, 3><<<grid, kBlock, 0, s>>>(nh, gslot_ptr, G, g);
    return hipGetLastError();
}

hipError_t launch_assemble_gradient(int d, int precision, int64_t nh, const int64_t* gslot_ptr, const void* G,
                                    double* g, hipStream_t s)
{
    if (precision == 0) return launch_assemble_gradient_t<double>(d, nh, gslot_ptr, (const double*)G, g, s);
    return launch_assemble_gradient_t<float>(d, nh, gslot_ptr, (const float*)G, g, s);
}

// ---------------------------------------------------------------------------
// PCG
// ---------------------------------------------------------------------------
// Reductions without atomics or fences: every kernel writes one partial per
// block; the kernel that consumes a scalar re-sums the producer's partials in
// index order in every block (identical bits everywhere, a few KB from L2).
// Scalars that change per iteration live in double-buffered partial slots
// indexed by the iteration number, so no block of a kernel ever reads a value
// another block of the same kernel writes.  Block 0 records bookkeeping
// (iteration count, residual, termination) for later kernels and the host.
//
// Partial buffer layout (doubles): [pq: Gs][rz: 2 x Gv][rr: 2 x Gv][rr0: Gv]

// Block-uniform read of the termination flag: block 0 of the same kernel may
// set it while other blocks start, so one thread reads it for the whole block
// and every thread then takes the same branch around the barriers below.
__device__ __forceinline__ bool block_done(const PcgScalars* sc)
{
    __shared__ int flag;
    if (threadIdx.x == 0) flag = sc->done;
    __syncthreads();
    return flag != 0;
}

// Every block re-sums all n partials: thread t adds parts[t], parts[t + NT],
// ... in index order (four loads in flight), then the fixed block tree.
template <int NT = kBlock>
__device__ __forceinline__ double sum_parts(const double* __restrict__ parts, int n, double* sh)
{
    double acc = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * NT < n; i += 4 * NT)
    {
        const double a0 = parts[i], a1 = parts[i + NT], a2 = parts[i + 2 * NT], a3 = parts[i + 3 * NT];
        acc += a0;
        acc += a1;
        acc += a2;
        acc += a3;
    }
    for (; i < n; i += NT) acc += parts[i];
    return block_sum_n<NT>(acc, sh);
}

template <int D>
__device__ __forceinline__ void invert_block(const double* B, double* o)
{
    if constexpr (D == 1)
        o[0] = 1.0 / B[0];
    else if constexpr (D == 2)
    {
        const double det = B[0] * B[3] - B[1] * B[2];
        o[0] = B[3] / det; o[1] = -B[1] / det; o[2] = -B[2] / det; o[3] = B[0] / det;
    }
    else
    {
        const double c00 = B[4] * B[8] - B[5] * B[7];
        const double c01 = B[5] * B[6] - B[3] * B[8];
        const double c02 = B[3] * B[7] - B[4] * B[6];
        const double det = B[0] * c00 + B[1] * c01 + B[2] * c02;
        const double id = 1.0 / det;
        o[0] = c00 * id; o[3] = c01 * id; o[6] = c02 * id;
        o[1] = (B[2] * B[7] - B[1] * B[8]) * id;
        o[4] = (B[0] * B[8] - B[2] * B[6]) * id;
        o[7] = (B[1] * B[6] - B[0] * B[7]) * id;
        o[2] = (B[1] * B[5] - B[2] * B[4]) * id;
        o[5] = (B[2] * B[3] - B[0] * B[5]) * id;
        o[8] = (B[0] * B[4] - B[1] * B[3]) * id;
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_pcg_setup(int64_t nh, const int32_t* __restrict__ brow_ptr,
                                                      const int32_t* __restrict__ diag_block,
                                                      const int32_t* __restrict__ row_ptr,
                                                      const double* __restrict__ H, double w,
                                                      double* __restrict__ Dinv)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nh) return;
    const int nbu = brow_ptr[u + 1] - brow_ptr[u];
    const int k = diag_block[u];
    const int64_t base = row_ptr[D * u];
    double B[D * D], o[D * D];
    if (k < 0)
    {
        // handle used by no element: its rows are empty, (w I + H) is w I there
#pragma unroll
        for (int m = 0; m < D * D; ++m) Dinv[u * D * D + m] = (m % (D + 1) == 0) ? (w > 0.0 ? 1.0 / w : 1.0) : 0.0;
        return;
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            B[i * D + j] = H[base + (int64_t)i * D * nbu + D * k + j] + (i == j ? w : 0.0);
    invert_block<D>(B, o);
#pragma unroll
    for (int m = 0; m < D * D; ++m) Dinv[u * D * D + m] = o[m];
}

// x = 0, r = b, z = M^-1 r; partials rho_0 = r.z (slot 1) and ||b||^2 (rr0)
template <int D>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int64_t nh, const double* __restrict__ b,
                                                     const double* __restrict__ Dinv, double* __restrict__ x,
                                                     double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ parts, int Gs, int Gv, PcgScalars* sc)
{
    __shared__ double sh[kBlock];
    double rz = 0.0, rr = 0.0;
    for (int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x; u < nh; u += (int64_t)gridDim.x * kBlock)
    {
        double rv[D];
#pragma unroll
        for (int i = 0; i < D; ++i)
        {
            rv[i] = b[u * D + i];
            x[u * D + i] = 0.0;
            r[u * D + i] = rv[i];
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
        {
            double zi = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) zi += Dinv[u * D * D + i * D + j] * rv[j];
            z[u * D + i] = zi;
            rz += rv[i] * zi;
            rr += rv[i] * rv[i];
        }
    }
    const double srz = block_sum(rz, sh);
    const double srr = block_sum(rr, sh);
    if (threadIdx.x == 0)
    {
        parts[Gs + Gv + blockIdx.x] = srz;      // rho_0 in slot 1
        parts[Gs + 4 * Gv + blockIdx.x] = srr;  // rr0
        if (blockIdx.x == 0)
        {
            sc->iters = 0;
            sc->done = 0;
            sc->breakdown = 0;
        }
    }
}

// Chronopoulos-Gear form of preconditioned CG (the iterates of standard PCG
// in exact arithmetic, one reduction point per iteration):
//   s_k = A z_k,  delta_k = z_k.s_k,  rho_k = r_k.z_k
//   k = 0: beta = 0, alpha = rho_0 / delta_0,            p = z, q = s
//   k > 0: beta = rho_k / rho_{k-1}, alpha = rho_k / (delta_k - beta rho_k / alpha_{k-1}),
//          p = z + beta p,  q = s + beta q
//   x += alpha p,  r -= alpha q,  z = M^-1 r
// so the SpMV gathers only z, which is complete when it runs.

// CG iteration k, first half: termination test on rr_k = ||r_k||^2 (k = 0:
// ||b||^2, recorded), then s = (w I + H) z and partial z.s.
template <int D, int LPB>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t nh, const int32_t* __restrict__ brow_ptr,
                                                 const int32_t* __restrict__ bcol, const int32_t* __restrict__ row_ptr,
                                                 const double* __restrict__ H, double w, const double* __restrict__ z,
                                                 double* __restrict__ sv, double* __restrict__ parts, int Gs, int Gv,
                                                 int k, PcgScalars* sc)
{
    if (block_done(sc)) return;
    __shared__ double sh[kBlock];
    double rr, bb;
    if (k == 0)
    {
        bb = sum_parts(parts + Gs + 4 * Gv, Gv, sh);
        rr = bb;
    }
    else
    {
        rr = sum_parts(parts + Gs + 2 * Gv + ((k + 1) & 1) * Gv, Gv, sh);
        bb = sc->bb;
    }
    const bool stop = rr <= sc->tol2 * bb || k >= sc->max_iters;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        if (k == 0) sc->bb = bb;
        sc->rr = rr;
        sc->iters = k;
        if (stop) sc->done = 1;
    }
    if (stop) return;
    const int lane = threadIdx.x % LPB;
    constexpr int rows_per_block = kBlock / LPB;
    double zs = 0.0;
    for (int64_t u0 = (int64_t)blockIdx.x * rows_per_block; u0 < nh; u0 += (int64_t)gridDim.x * rows_per_block)
    {
        const int64_t u = u0 + threadIdx.x / LPB;
        double acc[D];
#pragma unroll
        for (int i = 0; i < D; ++i) acc[i] = 0.0;
        if (u < nh)
        {
            const int b0 = brow_ptr[u];
            const int nbu = brow_ptr[u + 1] - b0;
            const int64_t base = row_ptr[D * u];
            for (int kk = lane; kk < nbu; kk += LPB)
            {
                const int64_t v = bcol[b0 + kk];
                double zv[D];
#pragma unroll
                for (int j = 0; j < D; ++j) zv[j] = z[v * D + j];
#pragma unroll
                for (int i = 0; i < D; ++i)
#pragma unroll
                    for (int j = 0; j < D; ++j) acc[i] += H[base + (int64_t)i * D * nbu + D * kk + j] * zv[j];
            }
        }
#pragma unroll
        for (int off = LPB / 2; off > 0; off >>= 1)
#pragma unroll
            for (int i = 0; i < D; ++i) acc[i] += __shfl_xor(acc[i], off, LPB);
        if (u < nh && lane == 0)
        {
#pragma unroll
            for (int i = 0; i < D; ++i)
            {
                const double zi = z[u * D + i];
                const double si = acc[i] + w * zi;
                sv[u * D + i] = si;
                zs += zi * si;
            }
        }
    }
    const double tot = block_sum(zs, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

// CG iteration k, second half: scalars, direction recurrences, x / r / z
// update; partials of rho_{k+1} = r.z and ||r||^2 into slot k & 1.  alpha and
// rho of iteration k are kept (slot k & 1 of sc) for iteration k + 1.
// A block of 64 D threads owns 64 consecutive vertices per step: thread t
// updates variable t of the chunk (every vector access is a contiguous run),
// and the D x D preconditioner blocks and the new residual pass through LDS
// for z = M^-1 r.
constexpr int kUpdVerts = 64;

template <int D>
__global__ __launch_bounds__(kUpdVerts * D) void k_pcg_update(int64_t nh, const double* __restrict__ Dinv,
                                                              double* __restrict__ x, double* __restrict__ r,
                                                              double* __restrict__ z, double* __restrict__ p,
                                                              double* __restrict__ q,
                                                              const double* __restrict__ sv,
                                                              double* __restrict__ parts, int Gs, int Gv, int k,
                                                              PcgScalars* sc)
{
    constexpr int NT = kUpdVerts * D;
    if (block_done(sc)) return;
    __shared__ double sh[NT];
    __shared__ double sD[kUpdVerts * D * D];
    __shared__ double sR[NT];
    const int64_t n = nh * D;
    const int64_t dn = nh * D * D;
    const int t = threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * kUpdVerts;
    // operands of one chunk: this thread's variable and its third of the
    // chunk's D x D preconditioner blocks
    double zc = 0.0, pc = 0.0, qc = 0.0, sc_ = 0.0, xc = 0.0, rc = 0.0, dc[D];
    auto load = [&](int64_t c0) {
        const int64_t iv = c0 * D + t;
        if (iv < n)
        {
            zc = z[iv];
            pc = p[iv];
            qc = q[iv];
            sc_ = sv[iv];
            xc = x[iv];
            rc = r[iv];
        }
#pragma unroll
        for (int m = 0; m < D; ++m)
        {
            const int64_t id = c0 * D * D + t + (int64_t)m * NT;
            dc[m] = id < dn ? Dinv[id] : 0.0;
        }
    };
    int64_t c0 = (int64_t)blockIdx.x * kUpdVerts;
    if (c0 < nh) load(c0);  // in flight during the scalar re-sums
    const double delta = sum_parts<NT>(parts, Gs, sh);
    const double rho = sum_parts<NT>(parts + Gs + ((k + 1) & 1) * Gv, Gv, sh);
    double beta = 0.0, den = delta;
    if (k > 0)
    {
        const double rho_prev = sc->rho_hist[(k + 1) & 1];
        const double alpha_prev = sc->alpha_hist[(k + 1) & 1];
        beta = rho / rho_prev;
        den = delta - beta * rho / alpha_prev;
        // p.Ap from the recurrence can cancel to <= 0 near convergence even
        // for a positive definite matrix: restart along z (beta = 0, p = z,
        // p.Ap = z.Az) instead of reporting a breakdown
        if (!(den > 0.0))
        {
            beta = 0.0;
            den = delta;
        }
    }
    if (!(den > 0.0))
    {
        // not positive definite along z: the same decision in every block
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            sc->breakdown = 1;
            sc->done = 1;
            sc->iters = k;
        }
        return;
    }
    const double alpha = rho / den;
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        sc->rho_hist[k & 1] = rho;
        sc->alpha_hist[k & 1] = alpha;
    }
    double rz = 0.0, rr = 0.0;
    for (; c0 < nh; c0 += stride)
    {
        const int64_t iv = c0 * D + t;
        double rv = 0.0;
        if (iv < n)
        {
            const double pi = k > 0 ? fma(beta, pc, zc) : zc;
            const double qi = k > 0 ? fma(beta, qc, sc_) : sc_;
            p[iv] = pi;
            q[iv] = qi;
            x[iv] = xc + alpha * pi;
            rv = rc - alpha * qi;
            r[iv] = rv;
        }
        __syncthreads();  // the previous chunk's LDS reads are complete
        sR[t] = rv;
#pragma unroll
        for (int m = 0; m < D; ++m) sD[t + m * NT] = dc[m];
        __syncthreads();
        if (c0 + stride < nh) load(c0 + stride);  // next chunk in flight during this one's z
        if (iv < n)
        {
            const int vl = t / D, i = t - vl * D;
            double zi = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) zi += sD[vl * D * D + i * D + j] * sR[vl * D + j];
            z[iv] = zi;
            rz += rv * zi;
            rr += rv * rv;
        }
    }
    const double srz = block_sum_n<NT>(rz, sh);
    const double srr = block_sum_n<NT>(rr, sh);
    if (threadIdx.x == 0)
    {
        parts[Gs + (k & 1) * Gv + blockIdx.x] = srz;
        parts[Gs + 2 * Gv + (k & 1) * Gv + blockIdx.x] = srr;
    }
}

void pcg_grids(int d, int64_t nh, int* Gs, int* Gv)
{
    const int lpb = d == 3 ? 16 : 8;
    const int64_t rows_per_block = kBlock / lpb;
    const int64_t gs = (nh + rows_per_block - 1) / rows_per_block;
    const int64_t gv = (nh + kUpdVerts - 1) / kUpdVerts;
    *Gs = (int)std::max<int64_t>(1, std::min<int64_t>(gs, 2048));
    *Gv = (int)std::max<int64_t>(1, std::min<int64_t>(gv, 1024));
}

int64_t pcg_parts_size(int d, int64_t nh)
{
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    return Gs + 5 * (int64_t)Gv;
}

hipError_t launch_pcg_setup(int d, int64_t nh, const int32_t* brow_ptr, const int32_t* diag_block,
                            const int32_t* row_ptr, const double* H, double w, double* Dinv, hipStream_t s)
{
    const unsigned grid = (unsigned)((nh + kBlock - 1) / kBlock);
    if (d == 1) k_pcg_setup<1><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, Dinv);
    else if (d == 2) k_pcg_setup<2><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, Dinv);
    else k_pcg_setup<3><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, Dinv);
    return hipGetLastError();
}

hipError_t launch_pcg_init(int d, int64_t nh, const double* b, const double* Dinv, double* x, double* r, double* z,
                           double* parts, PcgScalars* sc, hipStream_t s)
{
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    if (d == 1) k_pcg_init<1><<<Gv, kBlock, 0, s>>>(nh, b, Dinv, x, r, z, parts, Gs, Gv, sc);
    else if (d == 2) k_pcg_init<2><<<Gv, kBlock, 0, s>>>(nh, b, Dinv, x, r, z, parts, Gs, Gv, sc);
    else k_pcg_init<3><<<Gv, kBlock, 0, s>>>(nh, b, Dinv, x, r, z, parts, Gs, Gv, sc);
    return hipGetLastError();
}

hipError_t launch_pcg_iteration(int d, int64_t nh, const int32_t* brow_ptr, const int32_t* bcol,
                                const int32_t* row_ptr, const double* H, double w, const double* Dinv, double* x,
                                double* r, double* z, double* pbuf, double* q, double* parts, PcgScalars* sc, int k,
                                bool with_update, hipStream_t s, Timer* timer)
{
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    double* p = pbuf;
    double* sv = pbuf + nh * d;
    if (timer) timer->begin("spmv", s);
    if (d == 1) k_spmv<1, 8><<<Gs, kBlock, 0, s>>>(nh, brow_ptr, bcol, row_ptr, H, w, z, sv, parts, Gs, Gv, k, sc);
    else if (d == 2) k_spmv<2, 8><<<Gs, kBlock, 0, s>>>(nh, brow_ptr, bcol, row_ptr, H, w, z, sv, parts, Gs, Gv, k, sc);
    else k_spmv<3, 16><<<Gs, kBlock, 0, s>>>(nh, brow_ptr, bcol, row_ptr, H, w, z, sv, parts, Gs, Gv, k, sc);
    if (timer) timer->end("spmv", s);
    if (!with_update) return hipGetLastError();
    if (timer) timer->begin("pcg_update", s);
    if (d == 1) k_pcg_update<1><<<Gv, kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, sc);
    else if (d == 2) k_pcg_update<2><<<Gv, 2 * kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, sc);
    else k_pcg_update<3><<<Gv, 3 * kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, sc);
    if (timer) timer->end("pcg_update", s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-segment energy sums (fixed order: terms in order, fixed thread stride
// within each range, fixed block tree)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_segment_sum(const double* __restrict__ elem_f,
                                                        const int64_t* __restrict__ ranges, int n_terms,
                                                        int64_t n_seg, double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    const int64_t sgm = blockIdx.x;
    double acc = 0.0;
    for (int t = 0; t < n_terms; ++t)
    {
        const int64_t a = ranges[(t * n_seg + sgm) * 2], b = ranges[(t * n_seg + sgm) * 2 + 1];
        int64_t i = a + threadIdx.x;
        for (; i + 3 * kBlock < b; i += 4 * kBlock)  // four loads in flight, same order
        {
            const double v0 = elem_f[i], v1 = elem_f[i + kBlock], v2 = elem_f[i + 2 * kBlock],
                         v3 = elem_f[i + 3 * kBlock];
            acc += v0;
            acc += v1;
            acc += v2;
            acc += v3;
        }
        for (; i < b; i += kBlock) acc += elem_f[i];
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[sgm] = s;
}

hipError_t launch_segment_sum(const double* elem_f, const int64_t* ranges, int n_terms, int64_t n_seg, double* out,
                              hipStream_t s)
{
    if (n_seg == 0) return hipSuccess;
    k_segment_sum<<<(unsigned)n_seg, kBlock, 0, s>>>(elem_f, ranges, n_terms, n_seg, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dual self test (reference operator known answers)
// ---------------------------------------------------------------------------
__global__ void k_selftest_dual(int which, double xv, double yv, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Dual<2, double> x = dual_var<2, double>(xv, 0), y = dual_var<2, double>(yv, 1);
    Dual<2, double> f;
    switch (which)
    {
        case 0: f = sqr(x) / y; break;
        case 1:
        {
            const Dual<2, double> a = 0.5 * sqr(x) - sqr(y) + 2.0 * x - y;
            const Dual<2, double> b = -sqr(x - 2.0) - sqr(y - 3.0) + 1.0;
            f = a / b;
            break;
        }
        case 2: f = (sqr(x) + x) * (sqr(y) - y) / (y - 1.0); break;
        default: f = -(x * y) + x - y + 2.0 * x * 3.0 - 1.0; break;
    }
    out[0] = f.v;
    out[1] = f.g[0];
    out[2] = f.g[1];
    out[3] = f.h[Sym<2>::idx(0, 0)];
    out[4] = f.h[Sym<2>::idx(0, 1)];
    out[5] = f.h[Sym<2>::idx(0, 1)];
    out[6] = f.h[Sym<2>::idx(1, 1)];
}

hipError_t launch_selftest_dual(int which, double x, double y, double* out_dev, hipStream_t s)
{
    k_selftest_dual<<<1, 64, 0, s>>>(which, x, y, out_dev);
    return hipGetLastError();
}

} // namespace tad
