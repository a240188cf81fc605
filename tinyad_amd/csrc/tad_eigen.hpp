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
// matrix unchanged.  Otherwise a cyclic Jacobi eigensolve (deterministic,
// robust on clustered spectra) runs and the clamped spectrum is re-assembled.
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

__device__ __forceinline__ double tad_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float tad_rsqrt(float x) { return rsqrtf(x); }

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

// Cyclic Jacobi: on exit A holds (approximately) the eigenvalues on its
// diagonal and V (row-major n x n) the eigenvectors as columns.
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
#pragma unroll
        for (int p = 0; p < n - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < n; ++q)
            {
                const T apq = A[Sym<n>::idx(p, q)];
                const T app = A[Sym<n>::idx(p, p)];
                const T aqq = A[Sym<n>::idx(q, q)];
                // tan of the rotation angle, stable form without dividing by apq
                const T dlt = aqq - app;
                const T den = fabs(dlt) + sqrt(dlt * dlt + T(4) * apq * apq);
                T t = (dlt >= T(0) ? T(2) : T(-2)) * apq / (den > T(0) ? den : T(1));
                t = (den > T(0)) ? t : T(0);
                const T c = tad_rsqrt(T(1) + t * t);
                const T s = t * c;
                A[Sym<n>::idx(p, p)] = app - t * apq;
                A[Sym<n>::idx(q, q)] = aqq + t * apq;
                A[Sym<n>::idx(p, q)] = T(0);
#pragma unroll
                for (int r = 0; r < n; ++r)
                {
                    if (r == p || r == q) continue;
                    const int irp = Sym<n>::at(r, p), irq = Sym<n>::at(r, q);
                    const T arp = A[irp], arq = A[irq];
                    A[irp] = c * arp - s * arq;
                    A[irq] = s * arp + c * arq;
                }
#pragma unroll
                for (int r = 0; r < n; ++r)
                {
                    const T vrp = V[r * n + p], vrq = V[r * n + q];
                    V[r * n + p] = c * vrp - s * vrq;
                    V[r * n + q] = s * vrp + c * vrq;
                }
            }
    }
}

// Projects packed symmetric A in place.  Returns true when A was modified.
// keep_unchanged: when no eigenvalue needed clamping, A is left bit-identical
// (the reference returns the input untouched in that case).
template <int n, class T>
__device__ __forceinline__ bool project_psd(T* A, T eps)
{
    const T shift = eps > T(0) ? eps : T(0);
    if (cholesky_pd<n, T>(A, shift))
        return false;
    T W[Sym<n>::n];
#pragma unroll
    for (int i = 0; i < Sym<n>::n; ++i) W[i] = A[i];
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
