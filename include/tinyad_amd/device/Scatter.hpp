// Scatter of per-element derivative records into the library's destination-
// sorted contribution slots, shared by the library's built-in element kernels
// and by element kernels compiled from user energies (DeviceElements.hh).
// The slot maps (gpos, hpos) are built by the library for every term; their
// layout is fixed by tad_device_ops version TAD_DEVICE_OPS_VERSION.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include "Dual.hpp"
#include "Modes.hpp"
#include "Projection.hpp"

namespace tad {

__host__ __device__ constexpr int n_blocks(int N) { return N * (N + 1) / 2; }
__host__ __device__ constexpr int block_index(int N, int v, int u) { return v * N - (v * (v - 1)) / 2 + (u - v); }

// Contiguous run of L values.  FP32 slots are only 4-byte aligned;
// gfx9-family compute queues run in unaligned mode, where dword-aligned
// vector stores are legal, so a 3 x 3 block goes out as two 16-byte stores
// and one 4-byte store instead of nine (the scattered per-lane writes of the
// element kernel are bound by L2 write requests: the FP32 tet kernel runs
// 1.35x faster).  FP64 slots stay with 8-byte stores: 16-byte stores at
// 8-byte-aligned slots measured slower there, in every evaluation mode.
typedef float tad_f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float tad_f2u __attribute__((ext_vector_type(2), aligned(4)));

template <int L>
__device__ __forceinline__ void store_run(double* __restrict__ dst, const double* v)
{
#pragma unroll
    for (int m = 0; m < L; ++m) dst[m] = v[m];
}

template <int L>
__device__ __forceinline__ void store_run(float* __restrict__ dst, const float* v)
{
#pragma unroll
    for (int m = 0; m + 4 <= L; m += 4)
    {
        const tad_f4u w = { v[m], v[m + 1], v[m + 2], v[m + 3] };
        *reinterpret_cast<tad_f4u*>(dst + m) = w;
    }
    constexpr int r = L % 4;
    if constexpr (r >= 2)
    {
        const tad_f2u w = { v[L - r], v[L - r + 1] };
        *reinterpret_cast<tad_f2u*>(dst + L - r) = w;
    }
    if constexpr (r % 2 == 1) dst[L - 1] = v[L - 1];
}

// Scatter of one element's derivative parts through the precomputed maps:
// gpos[a n + e] is the slot of vertex a's gradient part in G (d values per
// slot); hpos[block_index(a, b) n + e] (a <= b) the slot of block (a, b) in Hc
// (d*d values per slot), top bit set when the slot belongs to the transposed
// (upper) CSR block.  Slots are sorted by destination (vertex, upper CSR
// block) and, within a destination, by term and element, so the assembly sums
// contiguous runs in the reference's element order.
// Slot of a padding entry (handle -1) of an element of a dynamic-valence
// term, and of every entry of a handle the element lists again after its
// first occurrence: nothing is written for it.
constexpr uint32_t kNoSlot = 0xffffffffu;

// dups: bit a set when entry a of the element repeats the handle of an
// earlier entry; reps: 4 bits per entry, the first entry with that handle.
// write_dense_record merges such entries into their first occurrence before
// projecting (the reference gives a handle one set of local variables however
// often the element asks for it, Element.hh:208-243).
template <int N, int Dm, class T, bool Padded = false>
struct ScatterSink
{
    T* __restrict__ G;
    T* __restrict__ Hc;
    const uint32_t* __restrict__ gpos;
    const uint32_t* __restrict__ hpos;
    int64_t e, n;
    uint32_t dups = 0;
    uint64_t reps = 0;

    __device__ __forceinline__ void grad(int v, const T* g) const
    {
        const uint32_t pos = gpos[v * n + e];
        if constexpr (Padded)
            if (pos == kNoSlot) return;
        store_run<Dm>(G + (int64_t)pos * Dm, g);
    }
    // block (v, u), v <= u, in the orientation of its slot
    __device__ __forceinline__ void block(int v, int u, const T* b) const
    {
        const uint32_t pos = hpos[block_index(N, v, u) * n + e];
        if constexpr (Padded)
            if (pos == kNoSlot) return;
        T* dst = Hc + (int64_t)(pos & 0x7fffffffu) * (Dm * Dm);
        if constexpr (sizeof(T) == 8)
        {
            if (pos >> 31)
            {
#pragma unroll
                for (int i = 0; i < Dm; ++i)
#pragma unroll
                    for (int j = 0; j < Dm; ++j) dst[j * Dm + i] = b[i * Dm + j];
            }
            else
            {
#pragma unroll
                for (int m = 0; m < Dm * Dm; ++m) dst[m] = b[m];
            }
        }
        else
        {
            // the transpose as a register select: every lane stores the same
            // contiguous pattern of wide stores
            const bool tr = (pos >> 31) != 0;
            T o[Dm * Dm];
#pragma unroll
            for (int i = 0; i < Dm; ++i)
#pragma unroll
                for (int j = 0; j < Dm; ++j) o[i * Dm + j] = tr ? b[j * Dm + i] : b[i * Dm + j];
            store_run<Dm * Dm>(dst, o);
        }
    }
};

__device__ __forceinline__ double tad_inf() { return __longlong_as_double(0x7ff0000000000000LL); }

template <class T>
__device__ __forceinline__ bool finite_val(T x) { return isfinite(x); }

// positive_diagonally_dominant (HessianProjection.hh:23-42) on a packed
// symmetric K x K matrix: H_ii >= sum_{j != i} |H_ij| + eps for every row.
template <int K, class T>
__device__ __forceinline__ bool positive_diagonally_dominant(const T* Hp, T eps)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        T off = T(0);
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j != i) off += fabs(Hp[Sym<K>::at(i, j)]);
        ok = ok && !(Hp[Sym<K>::idx(i, i)] < off + eps);
    }
    return ok;
}

// Writes the N(N+1)/2 upper d x d blocks of a packed (N d) x (N d) matrix.
template <int N, int Dm, class T, class Sink>
__device__ __forceinline__ void write_blocks(const T* Hp, const Sink& out, bool& nonfinite)
{
    constexpr int K = N * Dm;
#pragma unroll
    for (int v = 0; v < N; ++v)
#pragma unroll
        for (int u = v; u < N; ++u)
        {
            T blk[Dm * Dm];
#pragma unroll
            for (int i = 0; i < Dm; ++i)
#pragma unroll
                for (int ip = 0; ip < Dm; ++ip)
                {
                    const T val = Hp[Sym<K>::at(v * Dm + i, u * Dm + ip)];
                    blk[i * Dm + ip] = val;
                    nonfinite = nonfinite || !finite_val(val);
                }
            out.block(v, u, blk);
        }
}

template <class S, class = void>
struct sink_merges : std::false_type
{
};
template <class S>
struct sink_merges<S, std::void_t<decltype(S::dups)>> : std::true_type
{
};

// Merges entry a into entry b (b < a, same handle) in place: g_b += g_a and,
// with HESS, the congruence with the matrix that adds variable block a to
// block b (rows and columns of b gain those of a, the diagonal block gains
// H_ba + H_ab + H_aa); entry a is left with zero gradient, rows and columns.
template <int N, int Dm, class T, bool HESS, int a, int b>
__device__ __forceinline__ void merge_entry(T* g, T* Hp)
{
    constexpr int K = N * Dm;
#pragma unroll
    for (int i = 0; i < Dm; ++i)
    {
        g[b * Dm + i] += g[a * Dm + i];
        g[a * Dm + i] = T(0);
    }
    if constexpr (HESS)
    {
#pragma unroll
        for (int i = 0; i < Dm; ++i)
#pragma unroll
            for (int k = i; k < Dm; ++k)
                Hp[Sym<K>::idx(b * Dm + i, b * Dm + k)] += Hp[Sym<K>::at(b * Dm + i, a * Dm + k)] +
                                                           Hp[Sym<K>::at(a * Dm + i, b * Dm + k)] +
                                                           Hp[Sym<K>::at(a * Dm + i, a * Dm + k)];
#pragma unroll
        for (int j = 0; j < K; ++j)
        {
            if (j / Dm == a || j / Dm == b) continue;
#pragma unroll
            for (int i = 0; i < Dm; ++i) Hp[Sym<K>::at(b * Dm + i, j)] += Hp[Sym<K>::at(a * Dm + i, j)];
        }
#pragma unroll
        for (int i = 0; i < Dm; ++i)
#pragma unroll
            for (int j = 0; j < K; ++j) Hp[Sym<K>::at(a * Dm + i, j)] = T(0);
    }
}

template <int N, int Dm, class T, bool HESS, int a, int b>
__device__ __forceinline__ void merge_pairs(T* g, T* Hp, uint32_t dups, uint64_t reps)
{
    if constexpr (a < N)
    {
        if constexpr (b < a)
        {
            if (((dups >> a) & 1u) && (int)((reps >> (4 * a)) & 15u) == b) merge_entry<N, Dm, T, HESS, a, b>(g, Hp);
            merge_pairs<N, Dm, T, HESS, a, b + 1>(g, Hp, dups, reps);
        }
        else
            merge_pairs<N, Dm, T, HESS, a + 1, 0>(g, Hp, dups, reps);
    }
}

// Repeated entries folded into their first occurrence, in entry order, so the
// record is the reference's element record up to the order of its local
// variables (the repeated entries' rows and columns stay as zeros, as the
// reference's unused trailing local variables do).
template <int N, int Dm, class T, bool HESS>
__device__ __forceinline__ void merge_repeated(T* g, T* Hp, uint32_t dups, uint64_t reps)
{
    merge_pairs<N, Dm, T, HESS, 1, 0>(g, Hp, dups, reps);
}

// Entry v of column a (0 <= a < N-1) of the Helmert basis of
// {w in R^N : sum w = 0}: 1/sqrt((a+1)(a+2)) for v <= a, -(a+1)/sqrt((a+1)(a+2))
// for v = a+1, 0 below (orthonormal columns; every index is a compile-time
// constant where it is used, so the values fold to constants).
__device__ __forceinline__ double helmert(int v, int a)
{
    const double s = 1.0 / __builtin_sqrt((double)((a + 1) * (a + 2)));
    return v <= a ? s : (v == a + 1 ? -(double)(a + 1) * s : 0.0);
}

// Translating every vertex of the element leaves its energy unchanged up to
// rounding: for every row (u,i) and coordinate j, sum_v H[(u,i),(v,j)] is
// below 64 ulp of the largest |H| entry.
template <int N, int Dm, class T>
__device__ __forceinline__ bool translation_invariant(const T* Hp)
{
    constexpr int K = N * Dm;
    T hmax = T(0);
#pragma unroll
    for (int q = 0; q < Sym<K>::n; ++q) hmax = fmax(hmax, fabs(Hp[q]));
    const T tol = T(64) * QlTraits<T>::eps * hmax;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
        for (int j = 0; j < Dm; ++j)
        {
            T sum = T(0);
#pragma unroll
            for (int v = 0; v < N; ++v) sum += Hp[Sym<K>::at(r, v * Dm + j)];
            ok = ok && (fabs(sum) <= tol);
        }
    return ok;
}

// Projection of a translation-invariant element Hessian on its (N-1) d
// dimensional part, as the catalog's simplices do: with Q = U (x) I_d (U the
// Helmert basis) and K = Q^T H Q,
//   P_eps(H) = Q P_eps(K) Q^T + max(eps, 0) (11^T / N) (x) I_d,
// the reference's projection of H (HessianProjection.hh:48-101) whose d
// translations are null vectors (clamped to eps, or left at 0 for eps < 0).
// Writes the blocks of the result; Hp is consumed.
template <int N, int Dm, class T, class Sink>
__device__ __forceinline__ void write_deflated_projection(const T* Hp, T eps, const Sink& out, bool& nonfinite)
{
    constexpr int K = N * Dm;
    constexpr int KD = (N - 1) * Dm;
    T Km[Sym<KD>::n];
    // one column (b, j') of K at a time: T1[(u,i)] = sum_v H[(u,i),(v,j')] U_vb
#pragma unroll
    for (int b = 0; b + 1 < N; ++b)
#pragma unroll
        for (int jp = 0; jp < Dm; ++jp)
        {
            const int c2 = b * Dm + jp;
            T T1[K];
#pragma unroll
            for (int r = 0; r < K; ++r)
            {
                T acc = T(0);
#pragma unroll
                for (int v = 0; v < N; ++v)
                    if (helmert(v, b) != 0.0) acc += T(helmert(v, b)) * Hp[Sym<K>::at(r, v * Dm + jp)];
                T1[r] = acc;
            }
#pragma unroll
            for (int a = 0; a + 1 < N; ++a)
#pragma unroll
                for (int i = 0; i < Dm; ++i)
                {
                    const int c1 = a * Dm + i;
                    if (c1 > c2) continue;
                    T acc = T(0);
#pragma unroll
                    for (int u = 0; u < N; ++u)
                        if (helmert(u, a) != 0.0) acc += T(helmert(u, a)) * T1[u * Dm + i];
                    Km[Sym<KD>::idx(c1, c2)] = acc;
                }
        }
    project_psd<KD, T, false>(Km, eps);
    const T ct = eps > T(0) ? eps / T(N) : T(0);
    // H[(v,i),(u,i')] = sum_{a,b} U_va U_ub P[(a,i),(b,i')] + ct delta_ii', one
    // vertex row at a time
#pragma unroll
    for (int v = 0; v < N; ++v)
    {
        T S1[Dm][KD];
#pragma unroll
        for (int i = 0; i < Dm; ++i)
#pragma unroll
            for (int c2 = 0; c2 < KD; ++c2)
            {
                T acc = T(0);
#pragma unroll
                for (int a = 0; a + 1 < N; ++a)
                    if (helmert(v, a) != 0.0) acc += T(helmert(v, a)) * Km[Sym<KD>::at(a * Dm + i, c2)];
                S1[i][c2] = acc;
            }
#pragma unroll
        for (int u = v; u < N; ++u)
        {
            T blk[Dm * Dm];
#pragma unroll
            for (int i = 0; i < Dm; ++i)
#pragma unroll
                for (int ip = 0; ip < Dm; ++ip)
                {
                    T acc = (i == ip) ? ct : T(0);
#pragma unroll
                    for (int b = 0; b + 1 < N; ++b)
                        if (helmert(u, b) != 0.0) acc += T(helmert(u, b)) * S1[i][b * Dm + ip];
                    blk[i * Dm + ip] = acc;
                    nonfinite = nonfinite || !finite_val(acc);
                }
            out.block(v, u, blk);
        }
    }
}

// Generic small dense terms (penalty, quadratic, device functors): full
// K x K Hessian in packed form.  Deflate: the energy is translation invariant
// (a function of vertex position differences, as geometry energies are), so
// its projected Hessian goes through the (K - d) x (K - d) eigensolve; an
// element whose Hessian rows do not sum to zero is reported as non-finite.
// Otherwise the Hessian is projected whole.  (The two paths are separate
// instantiations: one kernel holding both needs far more registers than
// either.)
template <int N, int Dm, class T, int MODE, bool Deflate = false, class Sink>
__device__ __forceinline__ void write_dense_record(T* g, T* Hp, T eps, const Sink& out, bool& nonfinite)
{
    constexpr int K = N * Dm;
    if constexpr (N > 1 && sink_merges<Sink>::value)
        if (out.dups) merge_repeated<N, Dm, T, (MODE >= MODE_HESS)>(g, Hp, out.dups, out.reps);
#pragma unroll
    for (int i = 0; i < K; ++i) nonfinite = nonfinite || !finite_val(g[i]);
#pragma unroll
    for (int v = 0; v < N; ++v) out.grad(v, g + v * Dm);
    if constexpr (MODE >= MODE_HESS)
    {
        if constexpr (MODE == MODE_HESS_PROJ)
        {
            if (!positive_diagonally_dominant<K, T>(Hp, eps))
            {
                if constexpr (Deflate && N > 1)
                {
                    nonfinite = nonfinite || !translation_invariant<N, Dm, T>(Hp);
                    write_deflated_projection<N, Dm, T>(Hp, eps, out, nonfinite);
                    return;
                }
                else
                    project_psd<K, T>(Hp, eps);
            }
        }
        write_blocks<N, Dm, T>(Hp, out, nonfinite);
    }
}

} // namespace tad
