// Element kernel (gfx950): per-element value, gradient and (projected) Hessian,
// one lane per element.  Replaces the OpenMP element loop of
// ScalarObjectiveTerm.hh:242-254 and the per-element projection of
// HessianProjection.hh:48-101 (see tad_energies.hpp for the math).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include "tad_elements.hpp"

namespace tad {

// Waves per SIMD the element kernels are compiled for, by contribution type.
// FP64: 2 (the projected tet kernel then keeps 124 bytes per lane in scratch
// memory; at the 256 + 31 registers it needs unconstrained it runs one wave
// per SIMD, and the second wave hides enough latency to make it 11 % faster,
// with identical results).  FP32: 1 (no constraint; 150 registers already
// give three waves).
template <class T>
struct ElemWaves
{
    static constexpr int value = sizeof(T) == 8 ? 2 : 1;
};

// ---------------------------------------------------------------------------
// Grouped terms: Hessian blocks summed per wave in LDS
// ---------------------------------------------------------------------------
// The kElemGroup = 64 lanes of a wave evaluate 64 consecutive elements (one
// group, aligned to 64, whatever the launch shape).  Each element adds its
// blocks into the group's LDS accumulators (one per distinct upper block of
// the group, zeroed first): block by block, and for one block in the passes
// the plan gives, so that no two lanes add to one accumulator at once.  Then
// the wave writes the group's sums to their contribution slots.  The order of
// the additions is fixed by the plan and the element data, not by timing, so
// the sums are reproducible bit for bit.
struct GroupOut
{
    const uint8_t* gpass;
    const uint32_t* gout_ptr;
    const uint32_t* gout;
    int group_max;
};

// Orders the LDS accesses of the lanes of one wave (program order across the
// wave; no hardware barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int Dm, class T>
struct GroupSink
{
    T* __restrict__ G;
    const uint32_t* __restrict__ gpos;
    const uint32_t* __restrict__ hcode;
    const uint8_t* __restrict__ npass;  // this group's passes per block
    T* acc;                             // this wave's accumulators (LDS)
    int64_t e, n;

    __device__ __forceinline__ void grad(int v, const T* g) const { store_run<Dm>(G + (int64_t)gpos[v * n + e] * Dm, g); }
    // block (v, u), v <= u: added to its group accumulator, transposed to
    // the orientation of the upper CSR block when the code says so
    __device__ __forceinline__ void block(int v, int u, const T* b) const
    {
        const int bi = block_index(N, v, u);
        const uint32_t c = hcode[bi * n + e];
        const int gs = (int)(c & 0xffffu);
        const int pass = (int)((c >> 16) & 0xffu);
        const bool tr = (c >> 31) != 0;
        T o[Dm * Dm];
#pragma unroll
        for (int i = 0; i < Dm; ++i)
#pragma unroll
            for (int j = 0; j < Dm; ++j) o[i * Dm + j] = tr ? b[j * Dm + i] : b[i * Dm + j];
        T* a = acc + gs * (Dm * Dm);
        const int np = npass[bi];
        for (int p = 0; p < np; ++p)
        {
            if (p == pass)
            {
#pragma unroll
                for (int m = 0; m < Dm * Dm; ++m) a[m] += o[m];
            }
            wave_sync();
        }
    }
};

// One group: zero its accumulators, evaluate its elements (those below e1),
// write its block sums.  Called by all 64 lanes of the wave; `first` is the
// group's first element.
template <class E, class T, int MODE>
__device__ __forceinline__ void group_body(int64_t first, int64_t e1, int64_t n, const int32_t* __restrict__ conn,
                                           const double* __restrict__ params, const double* __restrict__ x,
                                           double eps, double* __restrict__ elem_f, const ElemOut<T>& o,
                                           const GroupOut& go, T* acc, int* __restrict__ err)
{
    constexpr int DD = E::D * E::D;
    constexpr int NB = n_blocks(E::N);
    const int lane = threadIdx.x & 63;
    const int64_t q = first / kElemGroup;
    const uint32_t g0 = go.gout_ptr[q];
    const int G = (int)(go.gout_ptr[q + 1] - g0);
    for (int i = lane; i < G * DD; i += 64) acc[i] = T(0);
    wave_sync();
    const int64_t e = first + lane;
    if (e < e1)
    {
        double X[E::N * E::D];
#pragma unroll
        for (int a = 0; a < E::N; ++a)
        {
            const int64_t v = conn[a * n + e];
#pragma unroll
            for (int i = 0; i < E::D; ++i) X[a * E::D + i] = x[v * E::D + i];
        }
        double p[E::NP > 0 ? E::NP : 1];
#pragma unroll
        for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
        bool nf = false;
        const GroupSink<E::N, E::D, T> out{ o.G, o.gpos, o.hpos, go.gpass + q * NB, acc, e, n };
        elem_f[e] = E::template eval<T, MODE>(X, p, T(eps), out, nf);
        if (nf) *err = 1;
    }
    wave_sync();
    // value i of the group's sums: slot block i / DD, entry i % DD (a wave
    // writes whole blocks with consecutive lanes)
    for (int i = lane; i < G * DD; i += 64) o.Hc[(int64_t)go.gout[g0 + i / DD] * DD + i % DD] = acc[i];
    wave_sync();
}

// Dynamic LDS of the grouped kernels: kBlock/64 (or BLOCK/64) waves, each
// with group_max accumulators of d x d values.
extern __shared__ double tad_group_lds[];

// Waves per SIMD the grouped kernels are compiled for: FP64 as the plain
// kernels (2); FP32 3, which holds the pass loops of the grouped sums in 168
// registers (unconstrained they take 191 and two waves: FP32 tets 0.73 ms
// instead of 0.52 ms at the 60^3 cube).
template <class T>
struct GroupWaves
{
    static constexpr int value = sizeof(T) == 8 ? 2 : 3;
};

template <class E, class T, int MODE>
__global__ __launch_bounds__(kBlock, GroupWaves<T>::value) void k_element_range_grouped(
    int64_t e0, int64_t e1, int64_t n, const int32_t* __restrict__ conn, const double* __restrict__ params,
    const double* __restrict__ x, double eps, double* __restrict__ elem_f, ElemOut<T> o, GroupOut go,
    int* __restrict__ err)
{
    T* acc = reinterpret_cast<T*>(tad_group_lds) + (threadIdx.x / 64) * (go.group_max * E::D * E::D);
    const int64_t first = e0 + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    if (first >= e1) return;  // whole wave past the end
    group_body<E, T, MODE>(first, e1, n, conn, params, x, eps, elem_f, o, go, acc, err);
}

template <class E, class T, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK, GroupWaves<T>::value) void k_element_persistent_grouped(
    int64_t e0, int64_t e1, int64_t n, const int32_t* __restrict__ conn, const double* __restrict__ params,
    const double* __restrict__ x, double eps, double* __restrict__ elem_f, ElemOut<T> o, GroupOut go,
    int* __restrict__ err, int map_xcd)
{
    T* acc = reinterpret_cast<T*>(tad_group_lds) + (threadIdx.x / 64) * (go.group_max * E::D * E::D);
    const int64_t n_tiles = (e1 - e0 + BLOCK - 1) / BLOCK;
    int64_t t1, stride, first_t;
    if (map_xcd)
    {
        const int xcd = blockIdx.x % 8;
        const int64_t local = blockIdx.x / 8;
        stride = gridDim.x / 8;
        const int64_t t0 = n_tiles * xcd / 8;
        t1 = n_tiles * (xcd + 1) / 8;
        first_t = t0 + local;
    }
    else
    {
        t1 = n_tiles;
        stride = gridDim.x;
        first_t = blockIdx.x;
    }
    for (int64_t t = first_t; t < t1; t += stride)
    {
        const int64_t first = e0 + t * BLOCK + (threadIdx.x & ~63);
        if (first < e1) group_body<E, T, MODE>(first, e1, n, conn, params, x, eps, elem_f, o, go, acc, err);
    }
}

// Persistent variant: a fixed grid walks the tiles.  map_xcd = 1 gives each of
// the 8 XCDs (workgroups are dispatched to XCDs round-robin by id) one
// contiguous eighth of the tiles, so the vertex gathers of neighbouring
// elements hit the same XCD's L2; map_xcd = 0 strides tiles across the grid.
// Requires gridDim.x to be a multiple of 8 when map_xcd = 1.
// Elements [e0, e1) of a term of n elements (n: the SoA stride).
template <class E, class T, int MODE, int BLOCK>
__global__ __launch_bounds__(BLOCK, ElemWaves<T>::value) void k_element_persistent(int64_t e0, int64_t e1, int64_t n,
                                                              const int32_t* __restrict__ conn,
                                                              const double* __restrict__ params,
                                                              const double* __restrict__ x, double eps,
                                                              double* __restrict__ elem_f, ElemOut<T> o,
                                                              int* __restrict__ err, int map_xcd)
{
    const int64_t n_tiles = (e1 - e0 + BLOCK - 1) / BLOCK;
    int64_t t0, t1, stride, first;
    if (map_xcd)
    {
        const int xcd = blockIdx.x % 8;
        const int64_t local = blockIdx.x / 8;
        stride = gridDim.x / 8;
        t0 = n_tiles * xcd / 8;
        t1 = n_tiles * (xcd + 1) / 8;
        first = t0 + local;
    }
    else
    {
        t1 = n_tiles;
        stride = gridDim.x;
        first = blockIdx.x;
    }
    for (int64_t t = first; t < t1; t += stride)
    {
        const int64_t e = e0 + t * BLOCK + threadIdx.x;
        if (e < e1) element_body<E, T, MODE>(e, n, conn, params, x, eps, elem_f, o, err);
    }
}

// Plain grid over the elements [e0, e1) of a term of n elements.
template <class E, class T, int MODE>
__global__ __launch_bounds__(kBlock, ElemWaves<T>::value) void k_element_range(int64_t e0, int64_t e1, int64_t n,
                                                          const int32_t* __restrict__ conn,
                                                          const double* __restrict__ params,
                                                          const double* __restrict__ x, double eps,
                                                          double* __restrict__ elem_f, ElemOut<T> o,
                                                          int* __restrict__ err)
{
    const int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    element_body<E, T, MODE>(e, n, conn, params, x, eps, elem_f, o, err);
}

// Persistent grid over the elements [e0, e1); the grid shrinks to one
// workgroup per tile for short ranges (a multiple of 8 with the per-XCD map).
template <class E, class T, int MODE, int BLOCK>
static void launch_persistent(const TermDev& t, const double* x, double eps, double* ef, const ElemOut<T>& o,
                              int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    const int64_t tiles = (e1 - e0 + BLOCK - 1) / BLOCK;
    int64_t grid = std::min<int64_t>(t.tune_grid, tiles);
    if (t.tune_map == 1) grid = (grid + 7) / 8 * 8;
    k_element_persistent<E, T, MODE, BLOCK><<<(unsigned)grid, BLOCK, 0, s>>>(e0, e1, t.n_elem, t.conn, t.params, x,
                                                                            eps, ef, o, err, t.tune_map);
}

// Grouped launches (second-order modes of grouped terms).
template <class E, class T, int MODE, int BLOCK>
static void launch_persistent_grouped(const TermDev& t, const double* x, double eps, double* ef, const ElemOut<T>& o,
                                      const GroupOut& go, int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    const int64_t tiles = (e1 - e0 + BLOCK - 1) / BLOCK;
    int64_t grid = std::min<int64_t>(t.tune_grid, tiles);
    if (t.tune_map == 1) grid = (grid + 7) / 8 * 8;
    const size_t lds = (size_t)(BLOCK / 64) * go.group_max * E::D * E::D * sizeof(T);
    k_element_persistent_grouped<E, T, MODE, BLOCK><<<(unsigned)grid, BLOCK, lds, s>>>(
        e0, e1, t.n_elem, t.conn, t.params, x, eps, ef, o, go, err, t.tune_map);
}

template <class E, class T, int MODE>
static hipError_t launch_grouped(const TermDev& t, const double* x, double eps, double* ef, const ElemOut<T>& o,
                                 int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    const GroupOut go{ t.gpass, t.gout_ptr, t.gout, t.group_max };
    if constexpr (E::tunable)
    {
        if (MODE == MODE_HESS_PROJ && t.tune_block > 0)
        {
            if (t.tune_map == 1 && t.tune_grid % 8 != 0) return hipErrorInvalidValue;
            if (t.tune_block == 64) launch_persistent_grouped<E, T, MODE, 64>(t, x, eps, ef, o, go, err, s, e0, e1);
            else if (t.tune_block == 128)
                launch_persistent_grouped<E, T, MODE, 128>(t, x, eps, ef, o, go, err, s, e0, e1);
            else launch_persistent_grouped<E, T, MODE, 256>(t, x, eps, ef, o, go, err, s, e0, e1);
            return hipGetLastError();
        }
    }
    const dim3 grid((unsigned)((e1 - e0 + kBlock - 1) / kBlock));
    const size_t lds = (size_t)(kBlock / 64) * go.group_max * E::D * E::D * sizeof(T);
    k_element_range_grouped<E, T, MODE><<<grid, kBlock, lds, s>>>(e0, e1, t.n_elem, t.conn, t.params, x, eps, ef, o,
                                                                   go, err);
    return hipGetLastError();
}

template <class E, class T>
static hipError_t launch_element_t(const TermDev& t, int mode, const double* x, double eps, double* elem_f,
                                   T* G, T* Hc, int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    const int64_t n = t.n_elem;
    if (e1 <= e0) return hipSuccess;
    const ElemOut<T> out{G, Hc, t.gpos, t.hpos};
    double* ef = elem_f + t.f_base;
    // the element range of a grouped term starts on a group boundary (pipeline
    // chunks are whole 256-element tiles)
    if (t.grouped && mode != MODE_GRAD)
    {
        if (e0 % kElemGroup != 0) return hipErrorInvalidValue;
        if (mode == MODE_HESS) return launch_grouped<E, T, MODE_HESS>(t, x, eps, ef, out, err, s, e0, e1);
        return launch_grouped<E, T, MODE_HESS_PROJ>(t, x, eps, ef, out, err, s, e0, e1);
    }
    if constexpr (E::tunable)
    {
        // the launch shape chosen by tad_fn_autotune (projected Hessian path)
        if (mode == MODE_HESS_PROJ && t.tune_block > 0)
        {
            if (t.tune_map == 1 && t.tune_grid % 8 != 0) return hipErrorInvalidValue;
            if (t.tune_block == 64) launch_persistent<E, T, MODE_HESS_PROJ, 64>(t, x, eps, ef, out, err, s, e0, e1);
            else if (t.tune_block == 128)
                launch_persistent<E, T, MODE_HESS_PROJ, 128>(t, x, eps, ef, out, err, s, e0, e1);
            else launch_persistent<E, T, MODE_HESS_PROJ, 256>(t, x, eps, ef, out, err, s, e0, e1);
            return hipGetLastError();
        }
    }
    const dim3 grid((unsigned)((e1 - e0 + kBlock - 1) / kBlock));
    switch (mode)
    {
        case MODE_GRAD:
            k_element_range<E, T, MODE_GRAD><<<grid, kBlock, 0, s>>>(e0, e1, n, t.conn, t.params, x, eps, ef, out, err);
            break;
        case MODE_HESS:
            k_element_range<E, T, MODE_HESS><<<grid, kBlock, 0, s>>>(e0, e1, n, t.conn, t.params, x, eps, ef, out, err);
            break;
        default:
            k_element_range<E, T, MODE_HESS_PROJ><<<grid, kBlock, 0, s>>>(e0, e1, n, t.conn, t.params, x, eps, ef, out,
                                                                          err);
            break;
    }
    return hipGetLastError();
}

template <class T>
static hipError_t launch_element_p(const TermDev& t, int d, int mode, const double* x, double eps, double* elem_f,
                                   T* G, T* Hc, int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    switch (t.energy)
    {
        case E_SYMDIR_TRI: return launch_element_t<SymdirElem<2>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
        case E_SYMDIR_TET: return launch_element_t<SymdirElem<3>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
        case E_PENALTY:
            if (d == 1) return launch_element_t<PenaltyElem<1>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
            if (d == 2) return launch_element_t<PenaltyElem<2>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
            return launch_element_t<PenaltyElem<3>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
        case E_QUADRATIC:
        {
#define TAD_QUAD(NN, DD) \
    if (t.N == NN && d == DD)  \
        return launch_element_t<QuadElem<NN, DD>, T>(t, mode, x, eps, elem_f, G, Hc, err, s, e0, e1);
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

hipError_t launch_element_range(const TermDev& t, int d, int precision, int mode, const double* x, double eps,
                                double* elem_f, void* G, void* Hc, int* err, hipStream_t s, int64_t e0, int64_t e1)
{
    if (e0 < 0 || e1 > t.n_elem || e0 > e1) return hipErrorInvalidValue;
    if (t.dev_ops.element)
    {
        if (e0 != 0 || e1 != t.n_elem) return hipErrorInvalidValue;  // device energies launch whole
        return launch_element(t, d, precision, mode, x, eps, elem_f, G, Hc, err, s);
    }
    if (precision == 0)
        return launch_element_p<double>(t, d, mode, x, eps, elem_f, (double*)G, (double*)Hc, err, s, e0, e1);
    return launch_element_p<float>(t, d, mode, x, eps, elem_f, (float*)G, (float*)Hc, err, s, e0, e1);
}

hipError_t launch_element(const TermDev& t, int d, int precision, int mode, const double* x, double eps,
                          double* elem_f, void* G, void* Hc, int* err, hipStream_t s)
{
    if (t.dev_ops.element)
    {
        if (t.n_elem == 0) return hipSuccess;
        tad_device_element_args a{};
        a.n_elem = t.n_elem;
        a.conn = t.conn;
        a.params = t.params;
        a.x = x;
        a.eps = eps;
        a.elem_f = elem_f + t.f_base;
        a.G = G;
        a.Hc = Hc;
        a.gpos = t.gpos;
        a.hpos = t.hpos;
        a.err = err;
        a.mode = mode;
        a.precision = precision;
        a.stream = (void*)s;
        return (hipError_t)t.dev_ops.element(&a);
    }
    if (precision == 0)
        return launch_element_p<double>(t, d, mode, x, eps, elem_f, (double*)G, (double*)Hc, err, s, 0, t.n_elem);
    return launch_element_p<float>(t, d, mode, x, eps, elem_f, (float*)G, (float*)Hc, err, s, 0, t.n_elem);
}

} // namespace tad
