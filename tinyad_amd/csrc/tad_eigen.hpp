// In-register symmetric eigensolve and positive-definite projection for small
// packed symmetric matrices (n <= 12), one matrix per lane.
//
// Semantics follow include/TinyAD/Utils/HessianProjection.hh:48-101:
//   eps >= 0: eigenvalues below eps are raised to eps;
//   eps <  0: negative eigenvalues are replaced by their absolute value;
//   if no eigenvalue changed, the matrix is returned unchanged.
// The reference first tries a positive-diagonal-dominance early out
// (HessianProjection.hh:23-42) and otherwise runs Eigen's dense
// SelfAdjointEigenSolver.  Here the early out is a Cholesky factorization of
// A - max(eps, 0) I: it succeeds exactly when every eigenvalue exceeds the
// clamp level, which covers every matrix the diagonal-dominance test accepts
// (Gershgorin) and more, and in that case the reference also returns the
// matrix unchanged.  Otherwise a Jacobi eigensolve (deterministic, robust on
// clustered spectra) in round-robin pair order runs and the clamped spectrum
// is re-assembled.
#pragma once

#include <hip/hip_runtime.h>
#include "tad_dual.hpp"

namespace tad {

template <class T> struct EigTraits;
template <> struct EigTraits<double>
{
    static constexpr double tol = 1e-15;
    static constexpr int max_sweeps = 14;
};
template <> struct EigTraits<float>
{
    static constexpr float tol = 2e-7f;
    static constexpr int max_sweeps = 10;
};

// Reciprocal square root and reciprocal from the hardware approximations
// refined by Newton steps (two steps take the ~2^-23 relative error of the
// FP64 instructions below 1 ulp).  Arguments are positive and finite where used.
__device__ __forceinline__ double tad_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}
__device__ __forceinline__ float tad_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ double tad_rcp(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    y = y * (2.0 - x * y);
    y = y * (2.0 - x * y);
    return y;
}
__device__ __forceinline__ float tad_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Round-robin ("tournament") order of the n(n-1)/2 index pairs: every round
// holds floor(n/2) disjoint pairs, so the rotations of one round commute and
// their angle computations are independent of each other.
template <int n>
struct Tournament
{
    static constexpr int m = n + (n & 1);        // players, with a dummy when n is odd
    static constexpr int rounds = m - 1;
    static constexpr int per_round = n / 2;      // real pairs per round
    int p[rounds][per_round > 0 ? per_round : 1];
    int q[rounds][per_round > 0 ? per_round : 1];
    constexpr Tournament() : p(), q()
    {
        for (int r = 0; r < rounds; ++r)
        {
            int k = 0;
            for (int i = 0; i < m / 2; ++i)
            {
                int a = (i == 0) ? m - 1 : (r + i) % (m - 1);
                int b = (r - i + (m - 1)) % (m - 1);
                if (i == 0) b = r;
                if (a >= n || b >= n) continue;  // pair with the dummy
                p[r][k] = a < b ? a : b;
                q[r][k] = a < b ? b : a;
                ++k;
            }
        }
    }
};

// True when A - shift I is positive definite (Cholesky succeeds).
template <int n, class T>
__device__ __forceinline__ bool cholesky_pd(const T* A, T shift)
{
    T L[Sym<n>::n];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < n; ++j)
    {
        T djj = A[Sym<n>::idx(j, j)] - shift;
#pragma unroll
        for (int k = 0; k < j; ++k) djj -= L[Sym<n>::idx(k, j)] * L[Sym<n>::idx(k, j)];
        ok = ok && (djj > T(0));
        const T ljj = sqrt(djj > T(0) ? djj : T(1));
        const T il = T(1) / ljj;
        L[Sym<n>::idx(j, j)] = ljj;
#pragma unroll
        for (int i = j + 1; i < n; ++i)
        {
            T s = A[Sym<n>::idx(j, i)];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[Sym<n>::idx(k, i)] * L[Sym<n>::idx(k, j)];
            L[Sym<n>::idx(j, i)] = s * il;  // L stored transposed: entry (j, i) = L_ij
        }
    }
    return ok;
}

// One Jacobi rotation of rows/columns p, q (angle from the current entries).
template <int n, class T>
__device__ __forceinline__ void jacobi_angle(const T* A, int p, int q, T& c, T& s, T& t)
{
    const T apq = A[Sym<n>::idx(p, q)];
    const T dlt = A[Sym<n>::idx(q, q)] - A[Sym<n>::idx(p, p)];
    const T y = dlt * dlt + T(4) * apq * apq;
    // tangent of the rotation angle in the stable form sign(d) 2 a_pq / (|d| + sqrt(d^2 + 4 a_pq^2))
    if constexpr (sizeof(T) == 4)
    {
        const T rt = y > T(0) ? y * tad_rsqrt(y > T(0) ? y : T(1)) : T(0);
        const T den = fabs(dlt) + rt;
        t = (dlt >= T(0) ? T(2) : T(-2)) * apq * tad_rcp(den > T(0) ? den : T(1));
        t = (den > T(0)) ? t : T(0);
        c = tad_rsqrt(T(1) + t * t);
    }
    else
    {
        // FP64 keeps the library sqrt / divide: with the refined hardware
        // approximations the register allocator spills A and V to scratch
        const T den = fabs(dlt) + sqrt(y);
        t = (dlt >= T(0) ? T(2) : T(-2)) * apq / (den > T(0) ? den : T(1));
        t = (den > T(0)) ? t : T(0);
        c = T(1) / sqrt(T(1) + t * t);
    }
    s = t * c;
}

template <int n, class T>
__device__ __forceinline__ void jacobi_apply(T* A, T* V, int p, int q, T c, T s, T t)
{
    const T apq = A[Sym<n>::idx(p, q)];
    A[Sym<n>::idx(p, p)] -= t * apq;
    A[Sym<n>::idx(q, q)] += t * apq;
    A[Sym<n>::idx(p, q)] = T(0);
#pragma unroll
    for (int rr = 0; rr < n; ++rr)
    {
        if (rr == p || rr == q) continue;
        const int irp = Sym<n>::at(rr, p), irq = Sym<n>::at(rr, q);
        const T arp = A[irp], arq = A[irq];
        A[irp] = c * arp - s * arq;
        A[irq] = s * arp + c * arq;
    }
#pragma unroll
    for (int rr = 0; rr < n; ++rr)
    {
        const T vrp = V[rr * n + p], vrq = V[rr * n + q];
        V[rr * n + p] = c * vrp - s * vrq;
        V[rr * n + q] = s * vrp + c * vrq;
    }
}

// Jacobi eigensolve: on exit A holds (approximately) the eigenvalues on its
// diagonal and V (row-major n x n) the eigenvectors as columns.  Sweeps stop
// when the off-diagonal mass is below tol^2 of the total.
// FP32 sweeps visit the pairs in round-robin order: each round's rotations
// are disjoint, commute, and their angle computations (the long dependent
// rsq/rcp chains) are independent, which gives the scheduler parallel work.
// FP64 sweeps use the cyclic row order: A and V already take 252 of the 256
// VGPRs, and in-flight rotations of a round would spill.
template <int n, class T>
__device__ __forceinline__ void jacobi_eigen(T* A, T* V)
{
#pragma unroll
    for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? T(1) : T(0);

    for (int sweep = 0; sweep < EigTraits<T>::max_sweeps; ++sweep)
    {
        T off = T(0), dia = T(0);
#pragma unroll
        for (int i = 0; i < n; ++i)
        {
            dia += A[Sym<n>::idx(i, i)] * A[Sym<n>::idx(i, i)];
#pragma unroll
            for (int j = i + 1; j < n; ++j) off += A[Sym<n>::idx(i, j)] * A[Sym<n>::idx(i, j)];
        }
        const T tol = EigTraits<T>::tol;
        if (!(off > tol * tol * (dia + T(2) * off)))
            break;
        if constexpr (sizeof(T) == 4)
        {
            constexpr Tournament<n> tour{};
            constexpr int NR = Tournament<n>::rounds;
            constexpr int NP = Tournament<n>::per_round > 0 ? Tournament<n>::per_round : 1;
#pragma unroll
            for (int r = 0; r < NR; ++r)
            {
                T cs[NP], sn[NP], tn[NP];
#pragma unroll
                for (int k = 0; k < Tournament<n>::per_round; ++k)
                    jacobi_angle<n, T>(A, tour.p[r][k], tour.q[r][k], cs[k], sn[k], tn[k]);
#pragma unroll
                for (int k = 0; k < Tournament<n>::per_round; ++k)
                    jacobi_apply<n, T>(A, V, tour.p[r][k], tour.q[r][k], cs[k], sn[k], tn[k]);
            }
        }
        else
        {
#pragma unroll
            for (int p = 0; p < n - 1; ++p)
#pragma unroll
                for (int q = p + 1; q < n; ++q)
                {
                    T c, s, t;
                    jacobi_angle<n, T>(A, p, q, c, s, t);
                    jacobi_apply<n, T>(A, V, p, q, c, s, t);
                }
        }
    }
}

// Projects packed symmetric A in place.  Returns true when A was modified.
// KeepUnchanged: when no eigenvalue needs clamping, A is left bit-identical
// (the reference returns its input untouched then); this keeps a copy of A
// alive through the eigensolve.  Without it the spectrum is re-assembled
// whenever the Cholesky test failed (equal to A up to rounding when nothing
// was clamped), which saves n(n+1)/2 registers during the eigensolve.
template <int n, class T, bool KeepUnchanged = true>
__device__ __forceinline__ bool project_psd(T* A, T eps)
{
    const T shift = eps > T(0) ? eps : T(0);
    if (cholesky_pd<n, T>(A, shift))
        return false;
    T Wbuf[KeepUnchanged ? Sym<n>::n : 1];
    T* W = A;
    if constexpr (KeepUnchanged)
    {
#pragma unroll
        for (int i = 0; i < Sym<n>::n; ++i) Wbuf[i] = A[i];
        W = Wbuf;
    }
    T V[n * n];
    jacobi_eigen<n, T>(W, V);
    T lam[n];
    bool changed = false;
#pragma unroll
    for (int i = 0; i < n; ++i)
    {
        T l = W[Sym<n>::idx(i, i)];
        if (eps < T(0))
        {
            changed = changed || (l < T(0));
            l = l < T(0) ? -l : l;
        }
        else
        {
            changed = changed || (l < eps);
            l = l < eps ? eps : l;
        }
        lam[i] = l;
    }
    if constexpr (KeepUnchanged)
        if (!changed)
            return false;
#pragma unroll
    for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = i; j < n; ++j)
        {
            T acc = T(0);
#pragma unroll
            for (int m = 0; m < n; ++m) acc += V[i * n + m] * lam[m] * V[j * n + m];
            A[Sym<n>::idx(i, j)] = acc;
        }
    return true;
}

} // namespace tad
