// Mixed second-derivative products for implicit differentiation through the
// optimum x*(theta) of E(x, theta), theta = per-vertex rest positions.
//
// For one symmetric Dirichlet tet with rest edges Mr(theta), R = Mr^-1 and
// weight w = w_scale det(Mr) / 6:
//   E = w Psi(J),  Psi(J) = ||J||^2 + ||J^-1||^2,  J = M(x) R.
// The kernel returns per element v = (d2E / dtheta dx)^T lambda, the
// theta-gradient of Phi = grad_x E . lambda = w <G, A>, where
//   A = Lam R (Lam: edge matrix of lambda), G = dPsi/dJ = 2 J - 2 K K^T K,
//   K = J^-T.
// Perturbing Mr by dMr: dR = -R dMr R, dJ = -J dMr R, dw = w tr(R dMr), so
//   dPhi/dMr = w (<G, A> I - J^T P - A^T G) R^T,   P = d2Psi/dJ2 [A]
//            = 2 A - 2 (dK K^T K + K dK^T K + K K^T dK),  dK = -(J^-1 A J^-1)^T,
// and the edge map Mr[:, k] = theta_{k+1} - theta_0 sends column k to vertex
// k + 1 and minus the column sum to vertex 0.  This is the product the
// reference would form from the truncated mixed Hessian block
// Scalar<24, T, true, 0, 12, 12, 12> (Scalar.hh:24-47, :188-196) contracted
// with lambda, in a dozen 3 x 3 products instead of a 24-variable number.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "tad_internal.hpp"
#include "tad_energies.hpp"

namespace tad {

struct M3
{
    double a[3][3];
};

__device__ __forceinline__ M3 mul(const M3& X, const M3& Y)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i][j] = X.a[i][0] * Y.a[0][j] + X.a[i][1] * Y.a[1][j] + X.a[i][2] * Y.a[2][j];
    return r;
}

__device__ __forceinline__ M3 transpose(const M3& X)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i][j] = X.a[j][i];
    return r;
}

__device__ __forceinline__ double det(const M3& X)
{
    return X.a[0][0] * (X.a[1][1] * X.a[2][2] - X.a[1][2] * X.a[2][1])
         - X.a[0][1] * (X.a[1][0] * X.a[2][2] - X.a[1][2] * X.a[2][0])
         + X.a[0][2] * (X.a[1][0] * X.a[2][1] - X.a[1][1] * X.a[2][0]);
}

// X^-1 from the adjugate (det passed in)
__device__ __forceinline__ M3 inverse(const M3& X, double dt)
{
    const double id = 1.0 / dt;
    M3 r;
    r.a[0][0] = (X.a[1][1] * X.a[2][2] - X.a[1][2] * X.a[2][1]) * id;
    r.a[0][1] = (X.a[0][2] * X.a[2][1] - X.a[0][1] * X.a[2][2]) * id;
    r.a[0][2] = (X.a[0][1] * X.a[1][2] - X.a[0][2] * X.a[1][1]) * id;
    r.a[1][0] = (X.a[1][2] * X.a[2][0] - X.a[1][0] * X.a[2][2]) * id;
    r.a[1][1] = (X.a[0][0] * X.a[2][2] - X.a[0][2] * X.a[2][0]) * id;
    r.a[1][2] = (X.a[0][2] * X.a[1][0] - X.a[0][0] * X.a[1][2]) * id;
    r.a[2][0] = (X.a[1][0] * X.a[2][1] - X.a[1][1] * X.a[2][0]) * id;
    r.a[2][1] = (X.a[0][1] * X.a[2][0] - X.a[0][0] * X.a[2][1]) * id;
    r.a[2][2] = (X.a[0][0] * X.a[1][1] - X.a[0][1] * X.a[1][0]) * id;
    return r;
}

// edge matrix col_mat(v1 - v0, v2 - v0, v3 - v0) of 4 packed 3-vectors
__device__ __forceinline__ M3 edges(const double* v)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) r.a[i][k] = v[3 * (k + 1) + i] - v[i];
    return r;
}

// out: the element's 4 x 3 values, written to its gradient slots
// (gpos[a n + e]: slot of vertex a), summed per vertex by the gradient assembly.
__global__ __launch_bounds__(256) void k_tet_rest_vjp(int64_t n, const int32_t* __restrict__ conn,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ rest,
                                                      const double* __restrict__ lam, double w_scale,
                                                      const uint32_t* __restrict__ gpos, double* __restrict__ gslots)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double Xv[12], Lv[12], Rv[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
        const int64_t v = conn[a * n + e];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            Xv[3 * a + i] = x[3 * v + i];
            Lv[3 * a + i] = lam[3 * v + i];
            Rv[3 * a + i] = rest[3 * v + i];
        }
    }
    double* dst[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) dst[a] = gslots + (int64_t)gpos[a * n + e] * 3;
    const M3 M = edges(Xv);
    const M3 Mr = edges(Rv);
    const double detM = det(M);
    const double detR = det(Mr);
    // inverted element (energy +inf) or degenerate rest shape: no contribution
    if (!(detM > 0.0) || detR == 0.0)
    {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int i = 0; i < 3; ++i) dst[a][i] = 0.0;
        return;
    }
    const M3 R = inverse(Mr, detR);
    const M3 J = mul(M, R);
    const M3 A = mul(edges(Lv), R);
    const M3 Ji = inverse(J, det(J));
    const M3 K = transpose(Ji);
    const M3 KtK = mul(Ji, K);      // K^T K
    const M3 KKt = mul(K, Ji);      // K K^T
    const M3 dK = transpose(mul(mul(Ji, A), Ji));  // -dK
    const M3 T0 = mul(K, KtK);      // K K^T K
    const M3 T1 = mul(dK, KtK);     // -(dK K^T K)
    const M3 T2 = mul(mul(K, transpose(dK)), K);   // -(K dK^T K)
    const M3 T3 = mul(KKt, dK);     // -(K K^T dK)
    M3 G, P;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            G.a[i][j] = 2.0 * J.a[i][j] - 2.0 * T0.a[i][j];
            P.a[i][j] = 2.0 * A.a[i][j] + 2.0 * (T1.a[i][j] + T2.a[i][j] + T3.a[i][j]);
            s += G.a[i][j] * A.a[i][j];
        }
    const M3 JtP = mul(transpose(J), P);
    const M3 AtG = mul(transpose(A), G);
    M3 Q;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Q.a[i][j] = (i == j ? s : 0.0) - JtP.a[i][j] - AtG.a[i][j];
    const double w = w_scale * detR / 6.0;
    const M3 Gm = mul(Q, transpose(R));
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            dst[k + 1][i] = w * Gm.a[i][k];
            sum += Gm.a[i][k];
        }
        dst[0][i] = -w * sum;
    }
}

hipError_t launch_tet_rest_vjp(const TermDev& t, const double* x, const double* rest, const double* lam,
                               double w_scale, double* G, hipStream_t s)
{
    if (t.n_elem == 0) return hipSuccess;
    const unsigned grid = (unsigned)((t.n_elem + 255) / 256);
    k_tet_rest_vjp<<<grid, 256, 0, s>>>(t.n_elem, t.conn, x, rest, lam, w_scale, t.gpos, G);
    return hipGetLastError();
}

} // namespace tad
