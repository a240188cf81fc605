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
// matrix unchanged.  Otherwise a Householder tridiagonalization and implicit
// QL iteration (the two stages SelfAdjointEigenSolver also uses) give the
// spectrum, and the clamped spectrum is re-assembled.
#pragma once

#include <hip/hip_runtime.h>
#include "Dual.hpp"

namespace tad {

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

// 1/sqrt(x), x > 0 finite: the hardware approximation refined by Newton
// steps (two in FP64: ~2^-22 -> below 1 ulp; one in FP32)
__device__ __forceinline__ double ql_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}
__device__ __forceinline__ float ql_rsqrt(float x)
{
    float y = __builtin_amdgcn_rsqf(x);
    y = y * (1.5f - 0.5f * x * y * y);
    return y;
}

template <class T> struct QlTraits;
template <> struct QlTraits<double>
{
    static constexpr double eps = 2.220446049250313e-16;  // 2^-52
    static constexpr int max_iters = 30;
};
template <> struct QlTraits<float>
{
    static constexpr float eps = 1.1920929e-07f;  // 2^-23
    static constexpr int max_iters = 30;
};

// Symmetric eigensolve by Householder reduction to tridiagonal form followed
// by the implicit QL iteration with a Wilkinson-type shift (Bowdler, Martin,
// Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306, the method of EISPACK
// tql2; Eigen's SelfAdjointEigenSolver follows the same two stages).
// A (packed, n x n) is consumed; on exit Z (row-major) holds orthonormal
// eigenvectors as columns and lam the eigenvalues.  Every register index is a
// compile-time constant: data-dependent loop bounds of the QL sweep are
// applied as per-step activity masks, so nothing goes to scratch.
// Entry of e holding coupling i of the reversed tridiagonal matrix (e[n-1]
// is the zero sentinel).
__host__ __device__ constexpr int ql_rev_e(int n, int i) { return i <= n - 2 ? n - 2 - i : n - 1; }

template <int n, class T>
__device__ __forceinline__ void tridiag_ql_eigen(T* A, T* Z, T* lam)
{
    T e[n];
    T tau[n > 2 ? n - 2 : 1];
    // 1. Householder: column k below the diagonal -> (e_k, 0, ..., 0); the
    //    reflector vector v overwrites that column (v_0 in row k + 1).
#pragma unroll
    for (int k = 0; k + 2 < n; ++k)
    {
        const int L = n - k - 1;
        T sigma = T(0);
#pragma unroll
        for (int i = 1; i < L; ++i) sigma += A[Sym<n>::idx(k, k + 1 + i)] * A[Sym<n>::idx(k, k + 1 + i)];
        const T x0 = A[Sym<n>::idx(k, k + 1)];
        const bool has = sigma > T(0);
        const T mu = sqrt(x0 * x0 + sigma);
        const T alpha = x0 > T(0) ? -mu : mu;
        const T v0 = x0 - alpha;
        const T vtv = v0 * v0 + sigma;
        const T tk = has ? T(2) / (has ? vtv : T(1)) : T(0);
        e[k] = has ? alpha : x0;
        tau[k] = tk;
        A[Sym<n>::idx(k, k + 1)] = v0;
        // p = tau B v, w = p - (tau/2)(v.p) v, B -= v w^T + w v^T
        T pv[n], vv[n];
#pragma unroll
        for (int i = 0; i < L; ++i) vv[i] = A[Sym<n>::idx(k, k + 1 + i)];
        T vp = T(0);
#pragma unroll
        for (int i = 0; i < L; ++i)
        {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < L; ++j) acc += A[Sym<n>::at(k + 1 + i, k + 1 + j)] * vv[j];
            pv[i] = tk * acc;
            vp += vv[i] * pv[i];
        }
        const T K = T(0.5) * tk * vp;
#pragma unroll
        for (int i = 0; i < L; ++i) pv[i] -= K * vv[i];
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int j = i; j < L; ++j) A[Sym<n>::idx(k + 1 + i, k + 1 + j)] -= vv[i] * pv[j] + pv[i] * vv[j];
    }
#pragma unroll
    for (int i = 0; i < n; ++i) lam[i] = A[Sym<n>::idx(i, i)];
    if (n >= 2) e[n - 2] = A[Sym<n>::idx(n - 2, n - 1)];
    e[n - 1] = T(0);
    // 2. Z = P_0 P_1 ... P_{n-3}, accumulated backwards (P_k acts on rows and
    //    columns k+1..n-1, where the partial product is non-trivial).
#pragma unroll
    for (int i = 0; i < n; ++i)
#pragma unroll
        for (int j = 0; j < n; ++j) Z[i * n + j] = (i == j) ? T(1) : T(0);
#pragma unroll
    for (int k = n - 3; k >= 0; --k)
    {
        const int L = n - k - 1;
#pragma unroll
        for (int j = k + 1; j < n; ++j)
        {
            T s = T(0);
#pragma unroll
            for (int i = 0; i < L; ++i) s += A[Sym<n>::idx(k, k + 1 + i)] * Z[(k + 1 + i) * n + j];
            s *= tau[k];
#pragma unroll
            for (int i = 0; i < L; ++i) Z[(k + 1 + i) * n + j] -= s * A[Sym<n>::idx(k, k + 1 + i)];
        }
    }
    // 3. implicit QL on the tridiagonal matrix in reverse order, J T J with
    //    J the reversal (lr[i] = lam[n-1-i], er[i] couples lr[i] and lr[i+1]):
    //    the reduction above runs from the first row down, and deflating the
    //    reversed matrix from the top needs fewer sweeps (element Hessians of
    //    the 60^3 twisted cube: 16.6 instead of 19.0 QL iterations per 9 x 9
    //    matrix, 18 instead of 22 for the slowest lane of a wave).  Column i
    //    of the reversed problem is column n-1-i of Z.
    // (index maps only: lr(i) is lam[n-1-i], er(i) is e[n-2-i] and er(n-1)
    // the zero sentinel e[n-1], so the sweep works on the same registers)
    const T eps = QlTraits<T>::eps;
#pragma unroll
    for (int l = 0; l + 1 < n; ++l)
    {
        for (int it = 0; it < QlTraits<T>::max_iters; ++it)
        {
            int m = n - 1;
#pragma unroll
            for (int j = n - 2; j >= l; --j)
                if (fabs(e[ql_rev_e(n, j)]) <= eps * (fabs(lam[n - 1 - (j)]) + fabs(lam[n - 1 - (j + 1)]))) m = j;
            if (m == l) break;
            T dm = lam[n - 1 - (n - 1)];
#pragma unroll
            for (int j = l; j + 1 < n; ++j)
                if (j == m) dm = lam[n - 1 - (j)];
            T g = (lam[n - 1 - (l + 1)] - lam[n - 1 - (l)]) / (T(2) * e[ql_rev_e(n, l)]);
            T r = sqrt(g * g + T(1));
            g = dm - lam[n - 1 - (l)] + e[ql_rev_e(n, l)] / (g + (g >= T(0) ? r : -r));
            T s = T(1), c = T(1), p = T(0);
            bool alive = true;  // false once a rotation underflowed (r == 0)
#pragma unroll
            for (int i = n - 2; i >= l; --i)
            {
                if (alive && i < m)
                {
                    const T f = s * e[ql_rev_e(n, i)];
                    const T b = c * e[ql_rev_e(n, i)];
                    // r = |(f, g)| and 1/r from one refined reciprocal square root
                    const T qq = f * f + g * g;
                    const T ir = qq > T(0) ? ql_rsqrt(qq) : T(0);
                    r = qq * ir;
                    e[ql_rev_e(n, i + 1)] = r;
                    if (r == T(0))
                    {
                        lam[n - 1 - (i + 1)] -= p;
#pragma unroll
                        for (int j = l; j + 1 < n; ++j)
                            if (j == m) e[ql_rev_e(n, j)] = T(0);
                        alive = false;
                    }
                    else
                    {
                        s = f * ir;
                        c = g * ir;
                        g = lam[n - 1 - (i + 1)] - p;
                        r = (lam[n - 1 - (i)] - g) * s + T(2) * c * b;
                        p = s * r;
                        lam[n - 1 - (i + 1)] = g + p;
                        g = c * r - b;
#pragma unroll
                        for (int k = 0; k < n; ++k)
                        {
                            const T zi = Z[k * n + (n - 1 - i)], zi1 = Z[k * n + (n - 2 - i)];
                            Z[k * n + (n - 2 - i)] = s * zi + c * zi1;
                            Z[k * n + (n - 1 - i)] = c * zi - s * zi1;
                        }
                    }
                }
            }
            if (alive)
            {
                lam[n - 1 - (l)] -= p;
                e[ql_rev_e(n, l)] = g;
#pragma unroll
                for (int j = l; j + 1 < n; ++j)
                    if (j == m) e[ql_rev_e(n, j)] = T(0);
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
    T lam[n];
    tridiag_ql_eigen<n, T>(W, V, lam);
    bool changed = false;
#pragma unroll
    for (int i = 0; i < n; ++i)
    {
        T l = lam[i];
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
