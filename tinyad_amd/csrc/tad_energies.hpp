// Device energy catalog: passive energies (FP64) and second-order element
// evaluations (FP64 or FP32) for the elements of one objective term.
//
// Symmetric Dirichlet simplices (triangle in the plane, tetrahedron in space)
// follow the reference test lambdas (tests/NewtonTest.cc:28-44,
// tests/ScalarTestHessianBlock.cc:50-64): with M = col_mat(b - a, c - a[, d - a]),
//   E = +inf                                if det(M) <= 0,
//   E = w (||J||_F^2 + ||J^-1||_F^2)        with J = M Mr^-1 otherwise.
//
// Derivatives.  The reference propagates Double<k> (k = 6 / 12) through the
// energy.  Every intermediate of this energy is a function of J, and J is a
// constant linear map of the element's k vertex coordinates:
//   dJ_ij / dX_vi' = delta_ii' D_vj,  D = [-1^T Mr^-1 ; Mr^-1]  (N x d).
// So the Double<k> value, gradient and Hessian are carried in factored form:
// forward mode over the d*d entries of J (Dual<d*d>) and, at the end, the
// exact chain rule through D.  The operators applied to the seeds are written
// out with their structural zeros removed (e.g. a product of two seeds has two
// Hessian entries); the arithmetic on the remaining entries is the forward-mode
// arithmetic of the reference.
//
// Projection.  Translating all vertices leaves E unchanged, so the element
// Hessian H has the d translations in its kernel and every other eigenvector
// in their orthogonal complement Q = U (x) I_d (U: a fixed orthonormal basis of
// {w in R^N : sum w = 0}).  With K = Q^T H Q (d*d x d*d):
//   P_eps(H) = Q P_eps(K) Q^T + max(eps, 0) (11^T / N) (x) I_d,
// i.e. the reference's per-element eigen-projection
// (include/TinyAD/Utils/HessianProjection.hh:48-101) computed on the 9 x 9
// (tet) or 4 x 4 (triangle) block that can carry non-zero eigenvalues.
//
// Output: every element hands its per-vertex gradient parts (d values) and
// the N(N+1)/2 upper d x d Hessian blocks (v <= u, row-major) to a sink; the
// library's sink scatters them (and the transposes of the off-diagonal
// blocks) to their slots in the destination-sorted contribution arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "tinyad_amd/device/Scatter.hpp"
#include "tad_internal.hpp"

namespace tad {

enum EnergyId
{
    E_SYMDIR_TRI = 1,
    E_SYMDIR_TET = 2,
    E_PENALTY = 3,
    E_QUADRATIC = 4,
};

// ---------------------------------------------------------------------------
// Passive energies (FP64), shared by the line-search energy kernel and by the
// value output of the derivative kernels so both give the same bits.
// ---------------------------------------------------------------------------

// X: N*d AoS element coordinates; p: [Mr^-1 (d x d row-major), w]
template <int Dm>
__device__ __forceinline__ double energy_symdir(const double* X, const double* p)
{
    double M[Dm][Dm];
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int k = 0; k < Dm; ++k) M[i][k] = X[(k + 1) * Dm + i] - X[i];
    double J[Dm][Dm];
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int j = 0; j < Dm; ++j)
        {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < Dm; ++k) s += M[i][k] * p[k * Dm + j];
            J[i][j] = s;
        }
    double nJ = 0.0;
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int j = 0; j < Dm; ++j) nJ += J[i][j] * J[i][j];
    if constexpr (Dm == 2)
    {
        const double detM = M[0][0] * M[1][1] - M[0][1] * M[1][0];
        if (!(detM > 0.0)) return tad_inf();
        const double detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        return p[4] * (nJ + nJ / (detJ * detJ));
    }
    else
    {
        const double detM = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                          - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                          + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        if (!(detM > 0.0)) return tad_inf();
        const double C[9] = {
            J[1][1] * J[2][2] - J[1][2] * J[2][1], J[1][2] * J[2][0] - J[1][0] * J[2][2], J[1][0] * J[2][1] - J[1][1] * J[2][0],
            J[0][2] * J[2][1] - J[0][1] * J[2][2], J[0][0] * J[2][2] - J[0][2] * J[2][0], J[0][1] * J[2][0] - J[0][0] * J[2][1],
            J[0][1] * J[1][2] - J[0][2] * J[1][1], J[0][2] * J[1][0] - J[0][0] * J[1][2], J[0][0] * J[1][1] - J[0][1] * J[1][0] };
        const double detJ = J[0][0] * C[0] + J[0][1] * C[1] + J[0][2] * C[2];
        double nC = 0.0;
#pragma unroll
        for (int t = 0; t < 9; ++t) nC += C[t] * C[t];
        return p[9] * (nJ + nC / (detJ * detJ));
    }
}

template <int D>
__device__ __forceinline__ double energy_penalty(const double* X, const double* p)
{
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) acc += (p[i] - X[i]) * (p[i] - X[i]);
    return p[D] * acc;
}

// 0.5 x^T A x + b^T x + c over the element's K local variables
template <int K>
__device__ __forceinline__ double energy_quadratic(const double* X, const double* p)
{
    double acc = p[K * K + K];
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        double r = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) r += p[i * K + j] * X[j];
        acc += (0.5 * r + p[K * K + i]) * X[i];
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Forward mode over J for symmetric Dirichlet.
// ---------------------------------------------------------------------------

// Accumulate sqr(C) into A and (optionally) s * C into det, where
// C = J_a J_b - J_c J_d is a product-difference of four distinct seeds.
template <int KJ, class T, int a, int b, int c, int d>
__device__ __forceinline__ T cofactor_accumulate(const T* J, Dual<KJ, T>& A)
{
    using S = Sym<KJ>;
    const T Cv = J[a] * J[b] - J[c] * J[d];
    // gradient entries of C at a, b, c, d
    const int ix[4] = { a, b, c, d };
    const T gC[4] = { J[b], J[a], -J[d], -J[c] };
    A.v += Cv * Cv;
#pragma unroll
    for (int m = 0; m < 4; ++m) A.g[ix[m]] += T(2) * Cv * gC[m];
    // Hess(C^2) = 2 (C Hess(C) + gC gC^T); Hess(C) has +1 at (a,b), -1 at (c,d)
    A.h[S::at(a, b)] += T(2) * Cv;
    A.h[S::at(c, d)] -= T(2) * Cv;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = m; n < 4; ++n)
            A.h[S::at(ix[m], ix[n])] += T(2) * gC[m] * gC[n];
    return Cv;
}

// det += J_s * C, with J_s a seed (index s) and C = J_a J_b - J_c J_d.
template <int KJ, class T, int s, int a, int b, int c, int d>
__device__ __forceinline__ void det_accumulate(const T* J, T Cv, Dual<KJ, T>& det)
{
    using S = Sym<KJ>;
    det.v += J[s] * Cv;
    det.g[s] += Cv;
    det.g[a] += J[s] * J[b];
    det.g[b] += J[s] * J[a];
    det.g[c] -= J[s] * J[d];
    det.g[d] -= J[s] * J[c];
    det.h[S::at(s, a)] += J[b];
    det.h[S::at(s, b)] += J[a];
    det.h[S::at(s, c)] -= J[d];
    det.h[S::at(s, d)] -= J[c];
    det.h[S::at(a, b)] += J[s];
    det.h[S::at(c, d)] -= J[s];
}

template <int KJ, class T>
__device__ __forceinline__ Dual<KJ, T> squared_norm_seeds(const T* J)
{
    Dual<KJ, T> r = dual_const<KJ, T>(T(0));
#pragma unroll
    for (int p = 0; p < KJ; ++p)
    {
        r.v += J[p] * J[p];
        r.g[p] = T(2) * J[p];
        r.h[Sym<KJ>::idx(p, p)] = T(2);
    }
    return r;
}

// E(J) = ||J||^2 + ||J^-1||^2 (unweighted), as a Dual over the entries of J.
template <int Dm, class T>
__device__ __forceinline__ Dual<Dm * Dm, T> symdir_dual(const T* J)
{
    constexpr int KJ = Dm * Dm;
    const Dual<KJ, T> nJ = squared_norm_seeds<KJ, T>(J);
    if constexpr (Dm == 2)
    {
        // det = J00 J11 - J01 J10; ||J^-1||^2 = ||J||^2 / det^2
        Dual<KJ, T> det = dual_const<KJ, T>(T(0));
        det.v = J[0] * J[3] - J[1] * J[2];
        det.g[0] = J[3]; det.g[3] = J[0]; det.g[1] = -J[2]; det.g[2] = -J[1];
        det.h[Sym<KJ>::at(0, 3)] = T(1);
        det.h[Sym<KJ>::at(1, 2)] = T(-1);
        const Dual<KJ, T> q = nJ / sqr(det);
        return nJ + q;
    }
    else
    {
        // ||J^-1||^2 = ||cof J||^2 / det^2, det = sum_k J_0k C_0k
        Dual<KJ, T> A = dual_const<KJ, T>(T(0));
        Dual<KJ, T> det = dual_const<KJ, T>(T(0));
        T c;
        c = cofactor_accumulate<KJ, T, 4, 8, 5, 7>(J, A); det_accumulate<KJ, T, 0, 4, 8, 5, 7>(J, c, det);
        c = cofactor_accumulate<KJ, T, 5, 6, 3, 8>(J, A); det_accumulate<KJ, T, 1, 5, 6, 3, 8>(J, c, det);
        c = cofactor_accumulate<KJ, T, 3, 7, 4, 6>(J, A); det_accumulate<KJ, T, 2, 3, 7, 4, 6>(J, c, det);
        cofactor_accumulate<KJ, T, 2, 7, 1, 8>(J, A);
        cofactor_accumulate<KJ, T, 0, 8, 2, 6>(J, A);
        cofactor_accumulate<KJ, T, 1, 6, 0, 7>(J, A);
        cofactor_accumulate<KJ, T, 1, 5, 2, 4>(J, A);
        cofactor_accumulate<KJ, T, 2, 3, 0, 5>(J, A);
        cofactor_accumulate<KJ, T, 0, 4, 1, 3>(J, A);
        // E = nJ + A r with r = det^-2, entry by entry (the product rule of
        // Scalar.hh:738-821 with r's derivatives from the chain rule):
        //   r' = -2 det^-3 det',  r'' = 6 det^-4 det' det'^T - 2 det^-3 det''
        // so only A, det and the result are live (det'' keeps its structural
        // zeros).
        const T id = T(1) / det.v;
        const T rv = id * id;
        const T c3 = T(-2) * rv * id;
        const T c4 = T(6) * rv * rv;
        T rg[KJ];
#pragma unroll
        for (int i = 0; i < KJ; ++i) rg[i] = c3 * det.g[i];
        Dual<KJ, T> E;
        E.v = nJ.v + A.v * rv;
#pragma unroll
        for (int i = 0; i < KJ; ++i) E.g[i] = nJ.g[i] + (A.g[i] * rv + A.v * rg[i]);
#pragma unroll
        for (int i = 0; i < KJ; ++i)
#pragma unroll
            for (int j = i; j < KJ; ++j)
            {
                const int k = Sym<KJ>::idx(i, j);
                const T rh = c4 * det.g[i] * det.g[j] + c3 * det.h[k];
                E.h[k] = nJ.h[k] + (A.v * rh + rv * A.h[k] + A.g[i] * rg[j] + rg[i] * A.g[j]);
            }
        return E;
    }
}

// Helmert basis of {w in R^N : sum w = 0}, column-wise (N x (N-1)).
template <int N> struct Helmert;
template <> struct Helmert<3>
{
    __device__ static constexpr double u(int v, int a)
    {
        constexpr double t[3][2] = { { 0.70710678118654752440, 0.40824829046386301637 },
                                     { -0.70710678118654752440, 0.40824829046386301637 },
                                     { 0.0, -0.81649658092772603273 } };
        return t[v][a];
    }
};
template <> struct Helmert<4>
{
    __device__ static constexpr double u(int v, int a)
    {
        constexpr double t[4][3] = { { 0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225 },
                                     { -0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225 },
                                     { 0.0, -0.81649658092772603273, 0.28867513459481288225 },
                                     { 0.0, 0.0, -0.86602540378443864676 } };
        return t[v][a];
    }
};

// Entry of the unprojected element Hessian in x coordinates:
// H_X[(v,i),(u,i')] = w sum_{j,j'} D_vj D_uj' H_J[(i,j),(i',j')]
template <int Dm, class T>
__device__ __forceinline__ T symdir_hx_entry(const Dual<Dm * Dm, T>& E, const T (*Dmat)[Dm], T w, int v, int i,
                                             int u, int ip)
{
    T s = T(0);
#pragma unroll
    for (int j = 0; j < Dm; ++j)
    {
        T t = T(0);
#pragma unroll
        for (int jp = 0; jp < Dm; ++jp) t += E.h[Sym<Dm * Dm>::at(i * Dm + j, ip * Dm + jp)] * Dmat[u][jp];
        s += Dmat[v][j] * t;
    }
    return w * s;
}

// Writes the unprojected element Hessian as N(N+1)/2 upper d x d blocks.
template <int Dm, class T, class Sink>
__device__ __forceinline__ void symdir_write_hx(const Dual<Dm * Dm, T>& E, const T (*Dmat)[Dm], T w, const Sink& out,
                                                bool& nonfinite)
{
    constexpr int N = Dm + 1;
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
                    const T val = symdir_hx_entry<Dm, T>(E, Dmat, w, v, i, u, ip);
                    blk[i * Dm + ip] = val;
                    nonfinite = nonfinite || !finite_val(val);
                }
            out.block(v, u, blk);
        }
}


// positive_diagonally_dominant (HessianProjection.hh:23-42) of H = Q K Q^T,
// entries H[(v,i),(u,i')] = sum_{a,b} U_va U_ub K[(a,i),(b,i')] formed on the fly.
template <int N, int Dm, class T>
__device__ __forceinline__ bool deflated_pdd(const T* Km, T eps)
{
    constexpr int KJ = Dm * Dm;
    bool ok = true;
#pragma unroll
    for (int v = 0; v < N; ++v)
#pragma unroll
        for (int i = 0; i < Dm; ++i)
        {
            T off = T(0), dia = T(0);
#pragma unroll
            for (int u = 0; u < N; ++u)
#pragma unroll
                for (int ip = 0; ip < Dm; ++ip)
                {
                    T h = T(0);
#pragma unroll
                    for (int a = 0; a < Dm; ++a)
#pragma unroll
                        for (int b = 0; b < Dm; ++b)
                            if (Helmert<N>::u(v, a) != 0.0 && Helmert<N>::u(u, b) != 0.0)
                                h += T(Helmert<N>::u(v, a) * Helmert<N>::u(u, b)) *
                                     Km[Sym<KJ>::at(a * Dm + i, b * Dm + ip)];
                    if (u == v && ip == i) dia = h;
                    else off += fabs(h);
                }
            ok = ok && !(dia < off + eps);
        }
    return ok;
}

// Second-order evaluation of one symmetric-Dirichlet simplex.
// X: element coordinates (FP64, AoS N x Dm); p: [Mr^-1, w] (FP64).
// Writes the record (gradient + Hessian blocks) in type T; returns the FP64 energy.
template <int Dm, class T, int MODE, class Sink>
__device__ __forceinline__ double symdir_element(const double* X, const double* p, T eps, const Sink& out,
                                                 bool& nonfinite)
{
    constexpr int N = Dm + 1;
    constexpr int KJ = Dm * Dm;
    const double f = energy_symdir<Dm>(X, p);

    if (!(f < tad_inf()))
    {
        // inverted element: +inf energy, zero derivatives; the projection of
        // the zero Hessian is eps I (clamp mode) or 0 (absolute-value mode)
        const T zero[Dm] = {};
#pragma unroll
        for (int v = 0; v < N; ++v) out.grad(v, zero);
        if constexpr (MODE >= MODE_HESS)
        {
            const T dg = (MODE == MODE_HESS_PROJ && eps > T(0)) ? eps : T(0);
#pragma unroll
            for (int v = 0; v < N; ++v)
#pragma unroll
                for (int u = v; u < N; ++u)
                {
                    T blk[Dm * Dm];
#pragma unroll
                    for (int i = 0; i < Dm; ++i)
#pragma unroll
                        for (int j = 0; j < Dm; ++j) blk[i * Dm + j] = (v == u && i == j) ? dg : T(0);
                    out.block(v, u, blk);
                }
        }
        return f;
    }

    // J = M Mr^-1 in FP64, then the derivative work in T
    double Md[Dm][Dm];
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int k = 0; k < Dm; ++k) Md[i][k] = X[(k + 1) * Dm + i] - X[i];
    T J[KJ];
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int j = 0; j < Dm; ++j)
        {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < Dm; ++k) s += Md[i][k] * p[k * Dm + j];
            J[i * Dm + j] = T(s);
        }
    const T w = T(p[KJ]);
    // D (N x Dm): dJ_ij/dX_vi = D_vj
    T Dmat[N][Dm];
#pragma unroll
    for (int j = 0; j < Dm; ++j)
    {
        T s = T(0);
#pragma unroll
        for (int k = 0; k < Dm; ++k)
        {
            Dmat[k + 1][j] = T(p[k * Dm + j]);
            s += Dmat[k + 1][j];
        }
        Dmat[0][j] = -s;
    }

    Dual<KJ, T> E = symdir_dual<Dm, T>(J);

    // gradient: g_X[v, i] = w sum_j D_vj dE/dJ_ij
#pragma unroll
    for (int v = 0; v < N; ++v)
    {
        T gv[Dm];
#pragma unroll
        for (int i = 0; i < Dm; ++i)
        {
            T s = T(0);
#pragma unroll
            for (int j = 0; j < Dm; ++j) s += Dmat[v][j] * E.g[i * Dm + j];
            gv[i] = w * s;
            nonfinite = nonfinite || !finite_val(gv[i]);
        }
        out.grad(v, gv);
    }
    if constexpr (MODE == MODE_GRAD)
        return f;

    if constexpr (MODE == MODE_HESS)
    {
        symdir_write_hx<Dm, T>(E, Dmat, w, out, nonfinite);
        return f;
    }
    else
    {
        // Diagonal-dominance early out of the reference.  With eps > 0 it
        // cannot fire here: the translations are in the kernel of H, so
        // H_ii = -sum_{u != v} H_(v,i),(u,i) up to rounding and H_ii >=
        // sum |H_ij| + eps is impossible unless eps is below rounding level.
        // E_aj = sum_v U_va D_vj ((N-1) x Dm, square)
        T Em[Dm][Dm];
#pragma unroll
        for (int a = 0; a < Dm; ++a)
#pragma unroll
            for (int j = 0; j < Dm; ++j)
            {
                T s = T(0);
#pragma unroll
                for (int v = 0; v < N; ++v)
                    if (Helmert<N>::u(v, a) != 0.0) s += T(Helmert<N>::u(v, a)) * Dmat[v][j];
                Em[a][j] = s;
            }
        // K[(a,i),(b,i')] = w sum_{j,j'} E_aj E_bj' H_J[(i,j),(i',j')], one
        // column (b, i') at a time: T1[(i,j)] = sum_j' H_J[(i,j),(i',j')] E_bj'
        T Km[Sym<KJ>::n];
#pragma unroll
        for (int b = 0; b < Dm; ++b)
#pragma unroll
            for (int ip = 0; ip < Dm; ++ip)
            {
                const int r2 = b * Dm + ip;
                T T1[KJ];
#pragma unroll
                for (int r = 0; r < KJ; ++r)
                {
                    T s = T(0);
#pragma unroll
                    for (int jp = 0; jp < Dm; ++jp) s += E.h[Sym<KJ>::at(r, ip * Dm + jp)] * Em[b][jp];
                    T1[r] = s;
                }
#pragma unroll
                for (int a = 0; a < Dm; ++a)
#pragma unroll
                    for (int i = 0; i < Dm; ++i)
                    {
                        const int r1 = a * Dm + i;
                        if (r2 < r1) continue;
                        T s = T(0);
#pragma unroll
                        for (int j = 0; j < Dm; ++j) s += Em[a][j] * T1[i * Dm + j];
                        Km[Sym<KJ>::idx(r1, r2)] = w * s;
                    }
            }
        // Diagonal-dominance early out of the reference on H = Q K Q^T (exact
        // when the translations are in the kernel of H); it leaves H
        // unprojected.  Only evaluated for eps <= 0, see above.
        const bool pdd = (eps <= T(0)) && deflated_pdd<N, Dm, T>(Km, eps);
        if (!pdd) project_psd<KJ, T, false>(Km, eps);
        // H[(v,i),(u,i')] = sum_{a,b} U_va U_ub P[(a,i),(b,i')] + max(eps,0)/N delta_ii'
        const T ct = eps > T(0) ? eps / T(N) : T(0);
        // one vertex row at a time: S1[i][(b,i')] = sum_a U_va P[(a,i),(b,i')]
        // is live for the blocks (v, u >= v) only
#pragma unroll
        for (int v = 0; v < N; ++v)
        {
            // form this vertex row's products here: scheduled ahead of the
            // previous row's block writes (the grouped sums' pass loops) they
            // were spilled to scratch memory (FP64 plain grid: 328 -> 220
            // bytes per lane, same time)
#pragma unroll
            for (int q = 0; q < Sym<KJ>::n; ++q) asm volatile("" : "+v"(Km[q]));
            T S1[Dm][KJ];
#pragma unroll
            for (int i = 0; i < Dm; ++i)
#pragma unroll
                for (int c2 = 0; c2 < KJ; ++c2)
                {
                    T s = T(0);
#pragma unroll
                    for (int a = 0; a < Dm; ++a)
                        if (Helmert<N>::u(v, a) != 0.0)
                            s += T(Helmert<N>::u(v, a)) * Km[Sym<KJ>::at(a * Dm + i, c2)];
                    S1[i][c2] = s;
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
                        T s = (i == ip) ? ct : T(0);
#pragma unroll
                        for (int b = 0; b < Dm; ++b)
                            if (Helmert<N>::u(u, b) != 0.0)
                                s += T(Helmert<N>::u(u, b)) * S1[i][b * Dm + ip];
                        blk[i * Dm + ip] = s;
                        nonfinite = nonfinite || !finite_val(s);
                    }
                out.block(v, u, blk);
            }
        }
        return f;
    }
}


template <int Dm, class T, int MODE, class Sink>
__device__ __forceinline__ double penalty_element(const double* X, const double* p, T eps, const Sink& out,
                                                  bool& nonfinite)
{
    const double f = energy_penalty<Dm>(X, p);
    T g[Dm], Hp[Sym<Dm>::n];
    const T w = T(p[Dm]);
#pragma unroll
    for (int i = 0; i < Dm; ++i) g[i] = T(-2) * w * T(p[i] - X[i]);
#pragma unroll
    for (int i = 0; i < Dm; ++i)
#pragma unroll
        for (int j = i; j < Dm; ++j) Hp[Sym<Dm>::idx(i, j)] = (i == j) ? T(2) * w : T(0);
    write_dense_record<1, Dm, T, MODE>(g, Hp, eps, out, nonfinite);
    return f;
}

template <int N, int Dm, class T, int MODE, class Sink>
__device__ __forceinline__ double quadratic_element(const double* X, const double* p, T eps, const Sink& out,
                                                    bool& nonfinite)
{
    constexpr int K = N * Dm;
    const double f = energy_quadratic<K>(X, p);
    T g[K], Hp[Sym<K>::n];
#pragma unroll
    for (int i = 0; i < K; ++i)
    {
        double s = p[K * K + i];
#pragma unroll
        for (int j = 0; j < K; ++j) s += 0.5 * (p[i * K + j] + p[j * K + i]) * X[j];
        g[i] = T(s);
#pragma unroll
        for (int j = i; j < K; ++j) Hp[Sym<K>::idx(i, j)] = T(0.5 * (p[i * K + j] + p[j * K + i]));
    }
    write_dense_record<N, Dm, T, MODE>(g, Hp, eps, out, nonfinite);
    return f;
}

} // namespace tad
