// HIP kernels of the sparse second-order path for gfx950 (linear algebra part)
// and their launchers.
//
//   k_sum_*, k_dot_*   deterministic FP64 sums (fixed tree, no atomics)
//   k_assemble_*       element-order gather of gradient and CSR values
//                      through the precomputed maps (replaces the serial
//                      triplet loop ScalarObjectiveTerm.hh:257-277 and
//                      setFromTriplets, ScalarFunctionImpl.hh:398)
//   k_pcg_*, k_spmv_sym  block-Jacobi PCG on the block-structured CSR
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
__global__ __launch_bounds__(kBlock) void k_assemble_hessian(int64_t b_begin, int64_t nb, const int32_t* __restrict__ block_row,
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
    const int64_t b = b_begin + wave * BPW + lane / DD;
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
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_tpb(int64_t b_begin, int64_t nb,
                                                                 const int32_t* __restrict__ block_row,
                                                                 const int32_t* __restrict__ brow_ptr,
                                                                 const int64_t* __restrict__ hslot_ptr,
                                                                 const int32_t* __restrict__ mirror,
                                                                 const T* __restrict__ Hc,
                                                                 const int32_t* __restrict__ row_ptr,
                                                                 double* __restrict__ H)
{
    constexpr int DD = D * D;
    const int64_t b = b_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_upper(int64_t b_begin, int64_t nu,
                                                                   const int64_t* __restrict__ uslot_ptr,
                                                                   const int4* __restrict__ uplace,
                                                                   const T* __restrict__ Hc, double* __restrict__ H)
{
    constexpr int DD = D * D;
    constexpr int BPW = 64 / DD;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane >= BPW * DD) return;
    const int64_t b = b_begin + wave * BPW + lane / DD;
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
__global__ __launch_bounds__(kBlock) void k_assemble_hessian_upper_tpb(int64_t b_begin, int64_t nu,
                                                                       const int64_t* __restrict__ uslot_ptr,
                                                                       const int4* __restrict__ uplace,
                                                                       const T* __restrict__ Hc,
                                                                       double* __restrict__ H)
{
    constexpr int DD = D * D;
    const int64_t b = b_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ __launch_bounds__(kBlock) void k_assemble_gradient(int64_t v_begin, int64_t nh,
                                                              const int64_t* __restrict__ gslot_ptr,
                                                              const T* __restrict__ G, double* __restrict__ g)
{
    constexpr int HPW = 64 / D;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (lane >= HPW * D) return;
    const int64_t v = v_begin + wave * HPW + lane / D;
    if (v >= nh) return;
    const int i = lane % D;
    g[v * D + i] = run_sum<D>(G + i, gslot_ptr[v], gslot_ptr[v + 1]);
}

// Blocks [b0, b1) of the shape's index space: CSR blocks (shapes 0, 1) or
// upper blocks in CSR order (shapes 2, 3).
template <class T, int D>
static void launch_assemble_hessian_td(int shape, const HessAsm& a, int64_t b0, int64_t b1, const T* Hc, double* H,
                                       hipStream_t s)
{
    constexpr int bpw = 64 / (D * D);
    const int64_t n = b1 - b0;
    if (n <= 0) return;
    const unsigned gt = (unsigned)((n + kBlock - 1) / kBlock);
    const unsigned gl = (unsigned)((((n + bpw - 1) / bpw) * 64 + kBlock - 1) / kBlock);
    if (shape == 0)
        k_assemble_hessian<T, D><<<gl, kBlock, 0, s>>>(b0, b1, a.block_row, a.brow_ptr, a.hslot_ptr, a.mirror, Hc,
                                                       a.row_ptr, H);
    else if (shape == 1)
        k_assemble_hessian_tpb<T, D><<<gt, kBlock, 0, s>>>(b0, b1, a.block_row, a.brow_ptr, a.hslot_ptr, a.mirror,
                                                           Hc, a.row_ptr, H);
    else if (shape == 2)
        k_assemble_hessian_upper<T, D><<<gl, kBlock, 0, s>>>(b0, b1, a.uslot_ptr, a.uplace, Hc, H);
    else
        k_assemble_hessian_upper_tpb<T, D><<<gt, kBlock, 0, s>>>(b0, b1, a.uslot_ptr, a.uplace, Hc, H);
}

template <class T>
static hipError_t launch_assemble_hessian_t(int d, int shape, const HessAsm& a, int64_t r0, int64_t r1, const T* Hc,
                                            double* H, hipStream_t s)
{
    if (a.n_blocks == 0) return hipSuccess;
    if (shape < 0 || shape >= kAsmShapes) return hipErrorInvalidValue;
    int64_t b0, b1;
    if (shape >= 2)
    {
        b0 = r0 < 0 ? 0 : a.urow_ptr_h[r0];
        b1 = r0 < 0 ? a.n_upper : a.urow_ptr_h[r1];
    }
    else
    {
        b0 = r0 < 0 ? 0 : a.brow_ptr_h[r0];
        b1 = r0 < 0 ? a.n_blocks : a.brow_ptr_h[r1];
    }
    if (d == 1) launch_assemble_hessian_td<T, 1>(shape, a, b0, b1, Hc, H, s);
    else if (d == 2) launch_assemble_hessian_td<T, 2>(shape, a, b0, b1, Hc, H, s);
    else launch_assemble_hessian_td<T, 3>(shape, a, b0, b1, Hc, H, s);
    return hipGetLastError();
}

hipError_t launch_assemble_hessian(int d, int precision, int shape, const HessAsm& a, const void* Hc, double* H,
                                   hipStream_t s, int64_t r0, int64_t r1)
{
    if (precision == 0) return launch_assemble_hessian_t<double>(d, shape, a, r0, r1, (const double*)Hc, H, s);
    return launch_assemble_hessian_t<float>(d, shape, a, r0, r1, (const float*)Hc, H, s);
}

template <class T>
static hipError_t launch_assemble_gradient_t(int d, int64_t v0, int64_t v1, const int64_t* gslot_ptr, const T* G,
                                             double* g, hipStream_t s)
{
    const int64_t nh = v1 - v0;
    if (nh <= 0) return hipSuccess;
    const int hpw = 64 / d;
    const int64_t waves = (nh + hpw - 1) / hpw;
    const unsigned grid = (unsigned)((waves * 64 + kBlock - 1) / kBlock);
    if (d == 1) k_assemble_gradient<T, 1><<<grid, kBlock, 0, s>>>(v0, v1, gslot_ptr, G, g);
    else if (d == 2) k_assemble_gradient<T, 2><<<grid, kBlock, 0, s>>>(v0, v1, gslot_ptr, G, g);
    else k_assemble_gradient<T, 3><<<grid, kBlock, 0, s>>>(v0, v1, gslot_ptr, G, g);
    return hipGetLastError();
}

hipError_t launch_assemble_gradient(int d, int precision, int64_t nh, const int64_t* gslot_ptr, const void* G,
                                    double* g, hipStream_t s, int64_t v0, int64_t v1)
{
    if (v0 < 0)
    {
        v0 = 0;
        v1 = nh;
    }
    if (precision == 0) return launch_assemble_gradient_t<double>(d, v0, v1, gslot_ptr, (const double*)G, g, s);
    return launch_assemble_gradient_t<float>(d, v0, v1, gslot_ptr, (const float*)G, g, s);
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

// The block-Jacobi inverse keeps the upper triangle of each symmetric D x D
// block, row by row: (0,0) (0,1) .. (0,D-1) (1,1) .. (D-1,D-1).
template <int D>
struct PackedBlock
{
    static constexpr int n = D * (D + 1) / 2;
    __device__ static constexpr int at(int i, int j)
    {
        return i <= j ? i * D - (i * (i - 1)) / 2 + (j - i) : j * D - (j * (j - 1)) / 2 + (i - j);
    }
};

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
                                                      const unsigned char* __restrict__ free_vars,
                                                      double* __restrict__ Dinv)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nh) return;
    const int nbu = brow_ptr[u + 1] - brow_ptr[u];
    const int k = diag_block[u];
    const int64_t base = row_ptr[D * u];
    // reduced basis: a fixed component takes a unit row / column here and a
    // zero row / column in the inverse (the free sub-block's inverse)
    bool fr[D];
#pragma unroll
    for (int i = 0; i < D; ++i) fr[i] = free_vars ? free_vars[u * D + i] != 0 : true;
    double B[D * D], o[D * D];
    if (k < 0)
    {
        // handle used by no element: its rows are empty, (w I + H) is w I there
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = i; j < D; ++j)
                Dinv[u * PackedBlock<D>::n + PackedBlock<D>::at(i, j)] =
                    (i == j && fr[i]) ? (w > 0.0 ? 1.0 / w : 1.0) : 0.0;
        return;
    }
    // the block's lower entries (the solve uses H's scalar lower triangle)
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
        {
            const int il = i > j ? i : j, jl = i > j ? j : i;
            B[i * D + j] = (fr[i] && fr[j]) ? H[base + (int64_t)il * D * nbu + D * k + jl] + (i == j ? w : 0.0)
                                            : (i == j ? 1.0 : 0.0);
        }
    invert_block<D>(B, o);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j)
            if (!(fr[i] && fr[j])) o[i * D + j] = 0.0;
    // the upper triangle: the inverse of a symmetric block is symmetric (the
    // cofactor products above commute), and the stored preconditioner is
    // symmetric exactly even where H's block is so only up to rounding
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = i; j < D; ++j) Dinv[u * PackedBlock<D>::n + PackedBlock<D>::at(i, j)] = o[i * D + j];
}

// x = 0, r = b, z = M^-1 r; partials rho_0 = r.z (slot 1) and ||b||^2 (rr0)
template <int D>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int64_t nh, const double* __restrict__ b,
                                                     const unsigned char* __restrict__ free_vars,
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
            rv[i] = (free_vars && !free_vars[u * D + i]) ? 0.0 : b[u * D + i];
            x[u * D + i] = 0.0;
            r[u * D + i] = rv[i];
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
        {
            double zi = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) zi += Dinv[u * PackedBlock<D>::n + PackedBlock<D>::at(i, j)] * rv[j];
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

// The solver's copy of H (the lower triangle's blocks, diagonal included, in
// CSR order; D*D contiguous row-major values each), one thread per value.
// A diagonal block takes its upper entries from its lower ones, so the copy
// is the symmetric matrix of H's scalar lower triangle.
template <int D>
__global__ __launch_bounds__(kBlock) void k_sym_pack(int64_t n_sym, const int32_t* __restrict__ sym_full,
                                                     const int32_t* __restrict__ block_row,
                                                     const int32_t* __restrict__ brow_ptr,
                                                     const int32_t* __restrict__ bcol,
                                                     const int32_t* __restrict__ row_ptr,
                                                     const double* __restrict__ H, double* __restrict__ Hs)
{
    constexpr int DD = D * D;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_sym * DD) return;
    const int64_t l = id / DD;
    const int m = (int)(id - l * DD);
    int i = m / D, j = m - i * D;
    const int32_t b = sym_full[l];
    const int32_t u = block_row[b];
    if (bcol[b] == u && i < j)
    {
        const int t = i;
        i = j;
        j = t;
    }
    const int32_t b0 = brow_ptr[u];
    const int64_t nbu = brow_ptr[u + 1] - b0;
    Hs[id] = H[(int64_t)row_ptr[D * u] + i * D * nbu + (int64_t)D * (b - b0) + j];
}

// CG iteration k, first half: termination test on rr_k = ||r_k||^2 (k = 0:
// ||b||^2, recorded), then s = (w I + H) z and partial z.s.  The termination
// flag, the residual partials and the first row group's block data are all
// requested before the test, so their memory latencies overlap.  LPB lanes
// per block row, lane k takes blocks k, k + LPB, ... of the row, then a
// shuffle tree.  H comes from the solver's copy: block b of row u is stored
// block src[b], or the transpose of stored block ~src[b] (the lower block
// (v, u) it mirrors), so every stored off-diagonal block is read twice per
// product.  XCD: the workgroups of one XCD (blockIdx % 8, the dispatch order
// across the eight XCDs) take a contiguous eighth of the row groups in
// order, so the block an upper block reads belongs to a row that the same
// XCD streams shortly after, and the second read can hit in that XCD's L2.
// Requires gridDim.x % 8 == 0 for XCD.
template <int D, int LPB, bool XCD>
__global__ __launch_bounds__(kBlock) void k_spmv_sym(int64_t nh, const int32_t* __restrict__ brow_ptr,
                                                     const int32_t* __restrict__ bcol,
                                                     const int32_t* __restrict__ src, const double* __restrict__ Hs,
                                                     double w, const double* __restrict__ z,
                                                     double* __restrict__ sv, double* __restrict__ parts, int Gs,
                                                     int Gv, int k, const int* __restrict__ kdev, PcgScalars* sc)
{
    static_assert(kBlock * 4 >= 1024, "the residual partials (at most 1024) are read four per thread");
    constexpr int DD = D * D;
    if (kdev) k = *kdev;
    __shared__ double sh[kBlock];
    __shared__ int flag;
    const int done_v = threadIdx.x == 0 ? sc->done : 0;
    const double* rparts = k == 0 ? parts + Gs + 4 * Gv : parts + Gs + 2 * Gv + ((k + 1) & 1) * Gv;
    double pa[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const int i = threadIdx.x + q * kBlock;
        pa[q] = i < Gv ? rparts[i] : 0.0;
    }
    const int lane = threadIdx.x % LPB;
    constexpr int rows_per_block = kBlock / LPB;
    const int64_t n_groups = (nh + rows_per_block - 1) / rows_per_block;
    int64_t g0, gstep, gend;
    if constexpr (XCD)
    {
        const int64_t per = (n_groups + 7) / 8;
        const int64_t x = blockIdx.x & 7;
        g0 = x * per + (blockIdx.x >> 3);
        gstep = gridDim.x >> 3;
        gend = min(n_groups, (x + 1) * per);
    }
    else
    {
        g0 = blockIdx.x;
        gstep = gridDim.x;
        gend = n_groups;
    }
    auto load_block = [&](int64_t b, double* h) {
        const int32_t si = src[b];
        const double* __restrict__ hb = Hs + (int64_t)(si < 0 ? ~si : si) * DD;
        double t[DD];
#pragma unroll
        for (int m = 0; m < DD; ++m) t[m] = hb[m];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) h[i * D + j] = si < 0 ? t[j * D + i] : t[i * D + j];
    };
    int f_nbu = 0, f_b0 = 0;
    int64_t f_v = 0;
    double f_h[DD];
    if (g0 < gend)
    {
        const int64_t u = g0 * rows_per_block + threadIdx.x / LPB;
        if (u < nh)
        {
            f_b0 = brow_ptr[u];
            f_nbu = brow_ptr[u + 1] - f_b0;
            if (lane < f_nbu)
            {
                f_v = bcol[f_b0 + lane];
                load_block(f_b0 + lane, f_h);
            }
        }
    }
    if (threadIdx.x == 0) flag = done_v;
    __syncthreads();
    if (flag) return;
    double racc = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if ((int)threadIdx.x + q * kBlock < Gv) racc += pa[q];
    const double rsum = block_sum_n<kBlock>(racc, sh);
    double rr, bb;
    if (k == 0)
    {
        bb = rsum;
        rr = bb;
    }
    else
    {
        rr = rsum;
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
    double zs = 0.0;
    bool first = true;
    for (int64_t g = g0; g < gend; g += gstep)
    {
        const int64_t u = g * rows_per_block + threadIdx.x / LPB;
        double acc[D];
#pragma unroll
        for (int i = 0; i < D; ++i) acc[i] = 0.0;
        if (u < nh)
        {
            int b0, nbu, kk0;
            if (first)
            {
                b0 = f_b0;
                nbu = f_nbu;
                if (lane < nbu)
                {
                    double zv[D];
#pragma unroll
                    for (int j = 0; j < D; ++j) zv[j] = z[f_v * D + j];
#pragma unroll
                    for (int i = 0; i < D; ++i)
#pragma unroll
                        for (int j = 0; j < D; ++j) acc[i] += f_h[i * D + j] * zv[j];
                }
                kk0 = lane + LPB;
            }
            else
            {
                b0 = brow_ptr[u];
                nbu = brow_ptr[u + 1] - b0;
                kk0 = lane;
            }
            for (int kk = kk0; kk < nbu; kk += LPB)
            {
                const int64_t v = bcol[b0 + kk];
                double h[DD];
                load_block(b0 + kk, h);
                double zv[D];
#pragma unroll
                for (int j = 0; j < D; ++j) zv[j] = z[v * D + j];
#pragma unroll
                for (int i = 0; i < D; ++i)
#pragma unroll
                    for (int j = 0; j < D; ++j) acc[i] += h[i * D + j] * zv[j];
            }
        }
        first = false;
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
                                                              const int* __restrict__ kdev,
                                                              const double* __restrict__ cparts, int Gc,
                                                              const unsigned char* __restrict__ free_vars,
                                                              PcgScalars* sc)
{
    if (kdev) k = *kdev;
    constexpr int NT = kUpdVerts * D;
    constexpr int PB = PackedBlock<D>::n;
    constexpr int ND = (kUpdVerts * PB + NT - 1) / NT;  // packed block entries per thread and chunk
    if (block_done(sc)) return;
    __shared__ double sh[NT];
    __shared__ double sD[ND * NT];
    __shared__ double sR[NT];
    const int64_t n = nh * D;
    const int64_t dn = nh * PB;
    const int t = threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * kUpdVerts;
    // operands of one chunk: this thread's variable and its share of the
    // chunk's packed preconditioner blocks
    double zc = 0.0, pc = 0.0, qc = 0.0, sc_ = 0.0, xc = 0.0, rc = 0.0, dc[ND];
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
        for (int m = 0; m < ND; ++m)
        {
            const int64_t id = c0 * PB + t + (int64_t)m * NT;
            dc[m] = (t + m * NT < kUpdVerts * PB && id < dn) ? Dinv[id] : 0.0;
        }
    };
    int64_t c0 = (int64_t)blockIdx.x * kUpdVerts;
    if (c0 < nh) load(c0);  // in flight during the scalar re-sums
    const double delta = sum_parts<NT>(parts, Gs, sh);
    double rho = sum_parts<NT>(parts + Gs + ((k + 1) & 1) * Gv, Gv, sh);
    // r.z with the coarse correction: r.(D^-1 r) + (P^T r).(A_c^-1 P^T r)
    if (Gc > 0) rho += sum_parts<NT>(cparts + ((k + 1) & 1) * Gc, Gc, sh);
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
            if (free_vars && !free_vars[iv]) rv = 0.0;  // reduced basis: no residual on fixed variables
            r[iv] = rv;
        }
        __syncthreads();  // the previous chunk's LDS reads are complete
        sR[t] = rv;
#pragma unroll
        for (int m = 0; m < ND; ++m) sD[t + m * NT] = dc[m];
        __syncthreads();
        if (c0 + stride < nh) load(c0 + stride);  // next chunk in flight during this one's z
        if (iv < n)
        {
            const int vl = t / D, i = t - vl * D;
            double zi = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) zi += sD[vl * PB + PackedBlock<D>::at(i, j)] * sR[vl * D + j];
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

// ---------------------------------------------------------------------------
// Two-level preconditioner: additive coarse correction on handle aggregates
// ---------------------------------------------------------------------------
// z = D^-1 r + P A_c^-1 P^T r with A_c = P^T (w I + H) P.  P has KC columns per
// aggregate; its rows for handle v are R_v (D x KC): the identity on the
// translations and, for KC > D, the rotations omega about the aggregate's
// first handle, delta x_v = omega x dv (D = 3) or omega (-dv_1, dv_0) (D = 2).
// Every sum below runs in a fixed order (fixed lane assignment and a fixed
// shuffle tree), so the correction is bitwise reproducible.
template <int D, int KC>
__device__ __forceinline__ void coarse_rows(const double* __restrict__ dv, double R[D][KC])
{
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int c = 0; c < KC; ++c) R[i][c] = (c == i) ? 1.0 : 0.0;
    if constexpr (KC > D)
    {
        if constexpr (D == 3)
        {
            R[0][4] = dv[2];
            R[0][5] = -dv[1];
            R[1][3] = -dv[2];
            R[1][5] = dv[0];
            R[2][3] = dv[1];
            R[2][4] = -dv[0];
        }
        else if constexpr (D == 2)
        {
            R[0][2] = -dv[1];
            R[1][2] = dv[0];
        }
    }
}

// sum over the 64 lanes of a wave, fixed xor tree (every lane gets the total)
template <int L>
__device__ __forceinline__ void wave_sum(double* v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int m = 0; m < L; ++m) v[m] += __shfl_xor(v[m], off, 64);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_coarse_dvec(int64_t nh, const double* __restrict__ x,
                                                        const int32_t* __restrict__ agg_of,
                                                        const int32_t* __restrict__ agg_ptr,
                                                        const int32_t* __restrict__ agg_vert, double* __restrict__ dvec)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nh) return;
    const int64_t root = agg_vert[agg_ptr[agg_of[v]]];
#pragma unroll
    for (int i = 0; i < D; ++i) dvec[v * D + i] = x[v * D + i] - x[root * D + i];
}

// one wave per chunk of a coarse block pair's fine blocks (u, v): sum of
// R_u^T (H_uv + w I [u == v]) R_v over the chunk
template <int D, int KC>
__global__ __launch_bounds__(kBlock) void k_coarse_galerkin_chunk(int64_t n_gchunks,
                                                                  const int64_t* __restrict__ gchunk_ptr,
                                                                  const int32_t* __restrict__ pair_blocks,
                                                                  const int32_t* __restrict__ block_row,
                                                                  const int32_t* __restrict__ brow_ptr,
                                                                  const int32_t* __restrict__ bcol,
                                                                  const int32_t* __restrict__ row_ptr,
                                                                  const double* __restrict__ H, double w,
                                                                  const double* __restrict__ dvec,
                                                                  const unsigned char* __restrict__ free_vars,
                                                                  double* __restrict__ gpart)
{
    const int lane = threadIdx.x & 63;
    const int64_t ch = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (ch >= n_gchunks) return;
    double acc[KC * KC];
#pragma unroll
    for (int m = 0; m < KC * KC; ++m) acc[m] = 0.0;
    const int64_t c1 = gchunk_ptr[ch + 1];
    for (int64_t c = gchunk_ptr[ch] + lane; c < c1; c += 64)
    {
        const int32_t b = pair_blocks[c];
        const int u = block_row[b];
        const int v = bcol[b];
        const int b0 = brow_ptr[u];
        const int nbu = brow_ptr[u + 1] - b0;
        const int64_t base = (int64_t)row_ptr[D * u] + (int64_t)D * (b - b0);
        double Hb[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j)
                Hb[i][j] = H[base + (int64_t)i * D * nbu + j] + ((u == v && i == j) ? w : 0.0);
        double Ru[D][KC], Rv[D][KC];
        coarse_rows<D, KC>(dvec + (int64_t)u * D, Ru);
        coarse_rows<D, KC>(dvec + (int64_t)v * D, Rv);
        if (free_vars)  // reduced basis: P restricted to the free variables
        {
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int c2 = 0; c2 < KC; ++c2)
                {
                    if (!free_vars[(int64_t)u * D + i]) Ru[i][c2] = 0.0;
                    if (!free_vars[(int64_t)v * D + i]) Rv[i][c2] = 0.0;
                }
        }
        double T[D][KC];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int c2 = 0; c2 < KC; ++c2)
            {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < D; ++j) t += Hb[i][j] * Rv[j][c2];
                T[i][c2] = t;
            }
#pragma unroll
        for (int ca = 0; ca < KC; ++ca)
#pragma unroll
            for (int c2 = 0; c2 < KC; ++c2)
            {
                double t = 0.0;
#pragma unroll
                for (int i = 0; i < D; ++i) t += Ru[i][ca] * T[i][c2];
                acc[ca * KC + c2] += t;
            }
    }
    wave_sum<KC * KC>(acc);
#pragma unroll
    for (int m = 0; m < KC * KC; ++m)
        if (lane == m) gpart[ch * (KC * KC) + m] = acc[m];
}

// one thread per (pair, entry): chunk partials in order into A_c's (a, b) block
template <int KC>
__global__ __launch_bounds__(kBlock) void k_coarse_galerkin_pairs(int64_t n_pairs, const int32_t* __restrict__ pair_chunk,
                                                                  const int32_t* __restrict__ pair_ab,
                                                                  const double* __restrict__ gpart, int64_t nc,
                                                                  double* __restrict__ Ac)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_pairs * KC * KC) return;
    const int64_t pr = id / (KC * KC);
    const int m = (int)(id % (KC * KC));
    double t = 0.0;
    for (int32_t ch = pair_chunk[pr]; ch < pair_chunk[pr + 1]; ++ch) t += gpart[(int64_t)ch * (KC * KC) + m];
    const int64_t a = pair_ab[2 * pr], bb = pair_ab[2 * pr + 1];
    Ac[(a * KC + m / KC) + (bb * KC + m % KC) * nc] = t;
}

// A_jj <- A_jj (1 + rel), and 1 where A_jj = 0: a positive semidefinite
// Galerkin matrix of a reduced basis (coarse directions with no free
// variable) becomes positive definite; the added directions prolong to zero.
__global__ __launch_bounds__(kBlock) void k_shift_diagonal(int64_t n, double* __restrict__ A, double rel)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double a = A[j + j * n];
    A[j + j * n] = a == 0.0 ? 1.0 : a * (1.0 + rel);
}

__global__ __launch_bounds__(kBlock) void k_symmetrize_lower(int64_t n, double* __restrict__ A)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * n) return;
    const int64_t i = id % n, j = id / n;  // entry (i, j), column-major
    if (i < j) A[i + j * n] = A[j + i * n];
}

// r_c = P^T r: one wave per chunk of an aggregate's handles, then one
// thread per coarse unknown sums its aggregate's chunk partials in order
template <int D, int KC>
__global__ __launch_bounds__(kBlock) void k_coarse_restrict_chunk(int64_t n_chunks,
                                                                  const int32_t* __restrict__ chunk_ptr,
                                                                  const int32_t* __restrict__ agg_vert,
                                                                  const double* __restrict__ dvec,
                                                                  const double* __restrict__ r,
                                                                  double* __restrict__ part, PcgScalars* sc)
{
    if (block_done(sc)) return;
    const int lane = threadIdx.x & 63;
    const int64_t ch = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (ch >= n_chunks) return;
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;
    const int64_t m1 = chunk_ptr[ch + 1];
    for (int64_t m = chunk_ptr[ch] + lane; m < m1; m += 64)
    {
        const int64_t v = agg_vert[m];
        double R[D][KC];
        coarse_rows<D, KC>(dvec + v * D, R);
#pragma unroll
        for (int i = 0; i < D; ++i)
        {
            const double ri = r[v * D + i];
#pragma unroll
            for (int c = 0; c < KC; ++c) acc[c] += R[i][c] * ri;
        }
    }
    wave_sum<KC>(acc);
#pragma unroll
    for (int c = 0; c < KC; ++c)
        if (lane == c) part[ch * KC + c] = acc[c];
}

template <int KC>
__global__ __launch_bounds__(kBlock) void k_coarse_restrict_sum(int64_t n_agg, const int32_t* __restrict__ agg_chunk,
                                                                const double* __restrict__ part,
                                                                double* __restrict__ rc, PcgScalars* sc)
{
    if (block_done(sc)) return;
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_agg * KC) return;
    const int64_t a = id / KC;
    const int c = (int)(id % KC);
    double t = 0.0;
    for (int32_t ch = agg_chunk[a]; ch < agg_chunk[a + 1]; ++ch) t += part[(int64_t)ch * KC + c];
    rc[id] = t;
}

// Both restriction stages in one launch: one workgroup of kCoarseBlock
// threads per aggregate, its waves take the aggregate's chunks in turn (each
// chunk reduced exactly as by k_coarse_restrict_chunk), the chunk partials go
// to LDS and are summed in chunk order as by k_coarse_restrict_sum, so r_c
// has the same bits.  For aggregates of at most kRestrictMaxChunks chunks.
constexpr int kRestrictMaxChunks = 64;
// lanes per aggregate in the restriction and the coarse product (60^3 tet
// cube, 459 aggregates: 512 and 256 22.4 us per CG iteration for both
// kernels, 1,024 24.2 us)
constexpr int kCoarseBlock = 512;

template <int D, int KC>
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_restrict(const int32_t* __restrict__ agg_chunk,
                                                            const int32_t* __restrict__ chunk_ptr,
                                                            const int32_t* __restrict__ agg_vert,
                                                            const double* __restrict__ dvec,
                                                            const double* __restrict__ r, double* __restrict__ rc,
                                                            PcgScalars* sc)
{
    if (block_done(sc)) return;
    __shared__ double cpart[kRestrictMaxChunks * KC];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t a = blockIdx.x;
    const int32_t c0 = agg_chunk[a], c1 = agg_chunk[a + 1];
    for (int32_t ch = c0 + wv; ch < c1; ch += kCoarseBlock / 64)
    {
        double acc[KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[c] = 0.0;
        const int64_t m1 = chunk_ptr[ch + 1];
        for (int64_t m = chunk_ptr[ch] + lane; m < m1; m += 64)
        {
            const int64_t v = agg_vert[m];
            double R[D][KC];
            coarse_rows<D, KC>(dvec + v * D, R);
#pragma unroll
            for (int i = 0; i < D; ++i)
            {
                const double ri = r[v * D + i];
#pragma unroll
                for (int c = 0; c < KC; ++c) acc[c] += R[i][c] * ri;
            }
        }
        wave_sum<KC>(acc);
#pragma unroll
        for (int c = 0; c < KC; ++c)
            if (lane == c) cpart[(ch - c0) * KC + c] = acc[c];
    }
    __syncthreads();
    if (threadIdx.x < KC)
    {
        double t = 0.0;
        for (int32_t k = 0; k < c1 - c0; ++k) t += cpart[k * KC + threadIdx.x];
        rc[a * KC + threadIdx.x] = t;
    }
}

// Coarse product and prolongation in one launch, one workgroup of
// kCoarseBlock threads per aggregate a: y_a = rows a KC .. a KC + KC - 1 of
// A_c^-1 r_c (row i of the symmetric inverse is its contiguous column i;
// thread t takes the entries t, t + kCoarseBlock, ..., then a fixed wave tree
// and the waves' partials in order), one partial r_c[a] . y_a, then
// z_v += R_v y_a for every member v of the aggregate.
template <int D, int KC, class TA>
__global__ __launch_bounds__(kCoarseBlock) void k_coarse_apply_prolong(int64_t nc, const TA* __restrict__ Ainv,
                                                                       const double* __restrict__ rc,
                                                                       const int32_t* __restrict__ agg_ptr,
                                                                       const int32_t* __restrict__ agg_vert,
                                                                       const double* __restrict__ dvec,
                                                                       const unsigned char* __restrict__ free_vars,
                                                                       double* __restrict__ z,
                                                                       double* __restrict__ cparts_all, int slot,
                                                                       int Gc, const int* __restrict__ kdev,
                                                                       PcgScalars* sc)
{
    if (block_done(sc)) return;
    double* __restrict__ cparts = cparts_all + (int64_t)(kdev ? (*kdev & 1) : slot) * Gc;
    constexpr int NW = kCoarseBlock / 64;
    __shared__ double wpart[NW][KC];
    __shared__ double y[KC];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t a = blockIdx.x;
    const TA* __restrict__ col = Ainv + a * KC * nc;
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;
    for (int64_t j = threadIdx.x; j < nc; j += kCoarseBlock)
    {
        const double rj = rc[j];
#pragma unroll
        for (int c = 0; c < KC; ++c) acc[c] += (double)col[c * nc + j] * rj;
    }
    wave_sum<KC>(acc);
#pragma unroll
    for (int c = 0; c < KC; ++c)
        if (lane == c) wpart[wv][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < KC)
    {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) t += wpart[k][threadIdx.x];
        y[threadIdx.x] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < KC; ++c) t += rc[a * KC + c] * y[c];
        cparts[a] = t;
    }
    double ya[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) ya[c] = y[c];
    const int32_t m1 = agg_ptr[a + 1];
    for (int32_t m = agg_ptr[a] + threadIdx.x; m < m1; m += kCoarseBlock)
    {
        const int64_t v = agg_vert[m];
        double R[D][KC];
        coarse_rows<D, KC>(dvec + v * D, R);
#pragma unroll
        for (int i = 0; i < D; ++i)
        {
            if (free_vars && !free_vars[v * D + i]) continue;
            double t = z[v * D + i];
#pragma unroll
            for (int c = 0; c < KC; ++c) t += R[i][c] * ya[c];
            z[v * D + i] = t;
        }
    }
}

int coarse_apply_groups(int64_t n_agg) { return (int)std::max<int64_t>(1, n_agg); }

hipError_t launch_coarse_dvec(int d, int64_t nh, const double* x, const CoarseDev& c, double* dvec, hipStream_t s)
{
    if (nh == 0) return hipSuccess;
    const unsigned g = (unsigned)((nh + kBlock - 1) / kBlock);
    if (d == 1) k_coarse_dvec<1><<<g, kBlock, 0, s>>>(nh, x, c.agg_of, c.agg_ptr, c.agg_vert, dvec);
    else if (d == 2) k_coarse_dvec<2><<<g, kBlock, 0, s>>>(nh, x, c.agg_of, c.agg_ptr, c.agg_vert, dvec);
    else k_coarse_dvec<3><<<g, kBlock, 0, s>>>(nh, x, c.agg_of, c.agg_ptr, c.agg_vert, dvec);
    return hipGetLastError();
}

template <int D, int KC>
static void galerkin_t(int64_t n_pairs, int64_t n_gchunks, const int64_t* gchunk_ptr, const int32_t* pair_chunk,
                       const int32_t* pair_blocks, const int32_t* pair_ab, const int32_t* block_row,
                       const int32_t* brow_ptr, const int32_t* bcol, const int32_t* row_ptr, const double* H, double w,
                       const CoarseDev& c, double* gpart, double* Ac, hipStream_t s)
{
    const unsigned g1 = (unsigned)((n_gchunks * 64 + kBlock - 1) / kBlock);
    k_coarse_galerkin_chunk<D, KC><<<g1, kBlock, 0, s>>>(n_gchunks, gchunk_ptr, pair_blocks, block_row, brow_ptr,
                                                          bcol, row_ptr, H, w, c.dvec, c.free_vars, gpart);
    const unsigned g2 = (unsigned)((n_pairs * KC * KC + kBlock - 1) / kBlock);
    k_coarse_galerkin_pairs<KC><<<g2, kBlock, 0, s>>>(n_pairs, pair_chunk, pair_ab, gpart, c.nc, Ac);
}

hipError_t launch_coarse_galerkin(int d, int64_t n_pairs, int64_t n_gchunks, const int64_t* gchunk_ptr,
                                  const int32_t* pair_chunk, const int32_t* pair_blocks, const int32_t* pair_ab,
                                  const int32_t* block_row, const int32_t* brow_ptr, const int32_t* bcol,
                                  const int32_t* row_ptr, const double* H, double w, const CoarseDev& c,
                                  double* gpart, double* Ac, hipStream_t s)
{
    if (n_pairs == 0) return hipSuccess;
#define TAD_GALERKIN(DD, KK) galerkin_t<DD, KK>(n_pairs, n_gchunks, gchunk_ptr, pair_chunk, pair_blocks, pair_ab, \
                                                 block_row, brow_ptr, bcol, row_ptr, H, w, c, gpart, Ac, s)
    if (d == 3 && c.kc == 6) TAD_GALERKIN(3, 6);
    else if (d == 3 && c.kc == 3) TAD_GALERKIN(3, 3);
    else if (d == 2 && c.kc == 3) TAD_GALERKIN(2, 3);
    else if (d == 2 && c.kc == 2) TAD_GALERKIN(2, 2);
    else if (d == 1 && c.kc == 1) TAD_GALERKIN(1, 1);
    else return hipErrorInvalidValue;
#undef TAD_GALERKIN
    return hipGetLastError();
}

hipError_t launch_shift_diagonal(int64_t n, double* A, double rel, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_shift_diagonal<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, A, rel);
    return hipGetLastError();
}

// |A| row sums, one wave per row: lane-strided partial sums, then a fixed
// wave tree; k_rowmax then takes their maximum (||A||_inf).
__global__ __launch_bounds__(kBlock) void k_abs_rowsum(int64_t n, const double* __restrict__ A,
                                                       double* __restrict__ rowsum)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= n) return;
    double t = 0.0;
    for (int64_t j = lane; j < n; j += 64) t += fabs(A[i * n + j]);
    wave_sum<1>(&t);
    if (lane == 0) rowsum[i] = t;
}

__global__ __launch_bounds__(kBlock) void k_rowmax(int64_t n, const double* __restrict__ rowsum,
                                                   double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) m = fmax(m, rowsum[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ __launch_bounds__(kBlock) void k_coarse_to32(int64_t n, const double* __restrict__ A, double delta,
                                                        float* __restrict__ A32)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * n) return;
    const int64_t i = id / n, j = id - i * n;
    A32[id] = (float)(A[id] + (i == j ? delta : 0.0));
}

hipError_t launch_abs_rowmax(int64_t n, const double* A, double* rowsum, double* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_abs_rowsum<<<(unsigned)((n * 64 + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, A, rowsum);
    k_rowmax<<<1, kBlock, 0, s>>>(n, rowsum, out);
    return hipGetLastError();
}

hipError_t launch_coarse_to32(int64_t n, const double* A, double delta, float* A32, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_coarse_to32<<<(unsigned)((n * n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, A, delta, A32);
    return hipGetLastError();
}

hipError_t launch_symmetrize_lower(int64_t n, double* A, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_symmetrize_lower<<<(unsigned)((n * n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, A);
    return hipGetLastError();
}

template <int D, int KC>
static void coarse_correction_t(int64_t nh, const CoarseDev& c, const double* r, double* z, int slot, PcgScalars* sc,
                                hipStream_t s, const int* kdev)
{
    if (c.max_agg_chunks <= kRestrictMaxChunks)
    {
        k_coarse_restrict<D, KC><<<(unsigned)c.n_agg, kCoarseBlock, 0, s>>>(c.agg_chunk, c.chunk_ptr, c.agg_vert,
                                                                            c.dvec, r, c.rc, sc);
    }
    else
    {
        const unsigned gch = (unsigned)((c.n_chunks * 64 + kBlock - 1) / kBlock);
        k_coarse_restrict_chunk<D, KC><<<gch, kBlock, 0, s>>>(c.n_chunks, c.chunk_ptr, c.agg_vert, c.dvec, r,
                                                               c.chunk_part, sc);
        const unsigned gs = (unsigned)((c.n_agg * KC + kBlock - 1) / kBlock);
        k_coarse_restrict_sum<KC><<<gs, kBlock, 0, s>>>(c.n_agg, c.agg_chunk, c.chunk_part, c.rc, sc);
    }
    if (c.Ainv32)
        k_coarse_apply_prolong<D, KC, float><<<(unsigned)c.n_agg, kCoarseBlock, 0, s>>>(
            c.nc, c.Ainv32, c.rc, c.agg_ptr, c.agg_vert, c.dvec, c.free_vars, z, c.cparts, slot, c.Gc, kdev, sc);
    else
        k_coarse_apply_prolong<D, KC, double><<<(unsigned)c.n_agg, kCoarseBlock, 0, s>>>(
            c.nc, c.Ainv, c.rc, c.agg_ptr, c.agg_vert, c.dvec, c.free_vars, z, c.cparts, slot, c.Gc, kdev, sc);
}

hipError_t launch_coarse_correction(int d, int64_t nh, const CoarseDev& c, const double* r, double* z, int slot,
                                    PcgScalars* sc, hipStream_t s, Timer* timer, const int* kdev)
{
    if (c.kc == 0 || c.n_agg == 0) return hipSuccess;
    if (timer) timer->begin("pcg_coarse", s);
    if (d == 3 && c.kc == 6) coarse_correction_t<3, 6>(nh, c, r, z, slot, sc, s, kdev);
    else if (d == 3 && c.kc == 3) coarse_correction_t<3, 3>(nh, c, r, z, slot, sc, s, kdev);
    else if (d == 2 && c.kc == 3) coarse_correction_t<2, 3>(nh, c, r, z, slot, sc, s, kdev);
    else if (d == 2 && c.kc == 2) coarse_correction_t<2, 2>(nh, c, r, z, slot, sc, s, kdev);
    else if (d == 1 && c.kc == 1) coarse_correction_t<1, 1>(nh, c, r, z, slot, sc, s, kdev);
    else return hipErrorInvalidValue;
    if (timer) timer->end("pcg_coarse", s);
    return hipGetLastError();
}

constexpr int kSpmvLanes = 8;

void pcg_grids(int d, int64_t nh, int* Gs, int* Gv)
{
    // SpMV: kSpmvLanes lanes per block row, at most 1,024 workgroups (60^3
    // tet cube, 227K block rows: 8 lanes and 1,024 workgroups 40.7 us per
    // product against 42.1 with 768, 49.4 with 512, 43.1 with 1,536 and 44.8
    // with 2,048; 16 lanes with 2,048 workgroups 47.3, 4 lanes with 1,536 50.1)
    (void)d;
    const int64_t rows_per_block = kBlock / kSpmvLanes;
    const int64_t gs = (nh + rows_per_block - 1) / rows_per_block;
    const int64_t gv = (nh + kUpdVerts - 1) / kUpdVerts;
    *Gs = (int)std::max<int64_t>(1, std::min<int64_t>(gs, 1024));
    *Gv = (int)std::max<int64_t>(1, std::min<int64_t>(gv, 1024));
}

int64_t pcg_parts_size(int d, int64_t nh)
{
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    return Gs + 5 * (int64_t)Gv;
}

hipError_t launch_pcg_setup(int d, int64_t nh, const int32_t* brow_ptr, const int32_t* diag_block,
                            const int32_t* row_ptr, const double* H, double w, const unsigned char* free_vars,
                            double* Dinv, hipStream_t s)
{
    const unsigned grid = (unsigned)((nh + kBlock - 1) / kBlock);
    if (d == 1) k_pcg_setup<1><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, free_vars, Dinv);
    else if (d == 2) k_pcg_setup<2><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, free_vars, Dinv);
    else k_pcg_setup<3><<<grid, kBlock, 0, s>>>(nh, brow_ptr, diag_block, row_ptr, H, w, free_vars, Dinv);
    return hipGetLastError();
}

hipError_t launch_pcg_init(int d, int64_t nh, const double* b, const unsigned char* free_vars, const double* Dinv,
                           double* x, double* r, double* z, double* parts, PcgScalars* sc, const CoarseDev* co,
                           hipStream_t s)
{
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    if (d == 1) k_pcg_init<1><<<Gv, kBlock, 0, s>>>(nh, b, free_vars, Dinv, x, r, z, parts, Gs, Gv, sc);
    else if (d == 2) k_pcg_init<2><<<Gv, kBlock, 0, s>>>(nh, b, free_vars, Dinv, x, r, z, parts, Gs, Gv, sc);
    else k_pcg_init<3><<<Gv, kBlock, 0, s>>>(nh, b, free_vars, Dinv, x, r, z, parts, Gs, Gv, sc);
    // z_0 gets the coarse correction; its r.z partials join rho_0 in slot 1
    if (co && co->kc > 0 && co->n_agg > 0)
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_coarse_correction(d, nh, *co, r, z, 1, sc, s, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_sym_pack(int d, int64_t n_sym, const int32_t* sym_full, const int32_t* block_row,
                           const int32_t* brow_ptr, const int32_t* bcol, const int32_t* row_ptr, const double* H,
                           double* sym_H, hipStream_t s)
{
    if (n_sym == 0) return hipSuccess;
    const unsigned grid = (unsigned)((n_sym * d * d + kBlock - 1) / kBlock);
    if (d == 1)
        k_sym_pack<1><<<grid, kBlock, 0, s>>>(n_sym, sym_full, block_row, brow_ptr, bcol, row_ptr, H, sym_H);
    else if (d == 2)
        k_sym_pack<2><<<grid, kBlock, 0, s>>>(n_sym, sym_full, block_row, brow_ptr, bcol, row_ptr, H, sym_H);
    else
        k_sym_pack<3><<<grid, kBlock, 0, s>>>(n_sym, sym_full, block_row, brow_ptr, bcol, row_ptr, H, sym_H);
    return hipGetLastError();
}

template <int D, int LPB>
static void spmv_sym_t(int64_t nh, const int32_t* brow_ptr, const int32_t* bcol, const int32_t* src,
                       const double* Hs, double w, const double* z, double* sv, double* parts, int Gs, int Gv, int k,
                       const int* kdev, PcgScalars* sc, hipStream_t s)
{
    if (Gs % 8 == 0)
        k_spmv_sym<D, LPB, true><<<Gs, kBlock, 0, s>>>(nh, brow_ptr, bcol, src, Hs, w, z, sv, parts, Gs, Gv, k,
                                                        kdev, sc);
    else
        k_spmv_sym<D, LPB, false><<<Gs, kBlock, 0, s>>>(nh, brow_ptr, bcol, src, Hs, w, z, sv, parts, Gs, Gv, k,
                                                         kdev, sc);
}

hipError_t launch_pcg_iteration(int d, int64_t nh, const int32_t* brow_ptr, const int32_t* bcol,
                                const int32_t* sym_src, const double* sym_H, double w, const double* Dinv, double* x,
                                double* r, double* z, double* pbuf, double* q, double* parts, PcgScalars* sc, int k,
                                bool with_update, const unsigned char* free_vars, const CoarseDev* co, hipStream_t s,
                                Timer* timer, bool read_k)
{
    const int* kdev = read_k ? &sc->k : nullptr;
    const bool coarse = co && co->kc > 0 && co->n_agg > 0;
    const double* cparts = coarse ? co->cparts : nullptr;
    const int Gc = coarse ? co->Gc : 0;
    int Gs, Gv;
    pcg_grids(d, nh, &Gs, &Gv);
    double* p = pbuf;
    double* sv = pbuf + nh * d;
    if (!sym_H || !sym_src) return hipErrorInvalidValue;
    if (timer) timer->begin("spmv", s);
    if (d == 1)
        spmv_sym_t<1, kSpmvLanes>(nh, brow_ptr, bcol, sym_src, sym_H, w, z, sv, parts, Gs, Gv, k, kdev, sc, s);
    else if (d == 2)
        spmv_sym_t<2, kSpmvLanes>(nh, brow_ptr, bcol, sym_src, sym_H, w, z, sv, parts, Gs, Gv, k, kdev, sc, s);
    else
        spmv_sym_t<3, kSpmvLanes>(nh, brow_ptr, bcol, sym_src, sym_H, w, z, sv, parts, Gs, Gv, k, kdev, sc, s);
    if (timer) timer->end("spmv", s);
    if (!with_update) return hipGetLastError();
    if (timer) timer->begin("pcg_update", s);
    if (d == 1)
        k_pcg_update<1><<<Gv, kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, kdev, cparts, Gc,
                                                  free_vars, sc);
    else if (d == 2)
        k_pcg_update<2><<<Gv, 2 * kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, kdev, cparts, Gc,
                                                      free_vars, sc);
    else
        k_pcg_update<3><<<Gv, 3 * kUpdVerts, 0, s>>>(nh, Dinv, x, r, z, p, q, sv, parts, Gs, Gv, k, kdev, cparts, Gc,
                                                      free_vars, sc);
    if (timer) timer->end("pcg_update", s);
    // z of iteration k + 1 gets the coarse correction; its r.z partials go to
    // slot k & 1 next to the update's
    if (coarse)
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_coarse_correction(d, nh, *co, r, z, k & 1, sc, s, timer, kdev);
    }
    return hipGetLastError();
}

__global__ void k_pcg_advance(PcgScalars* sc)
{
    if (threadIdx.x == 0) sc->k += 1;
}

hipError_t launch_pcg_advance(PcgScalars* sc, hipStream_t s)
{
    k_pcg_advance<<<1, 64, 0, s>>>(sc);
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
