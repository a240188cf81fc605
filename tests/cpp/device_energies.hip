// Energies written as generic device functors, compiled into
// tests/cpp/bin/libtad_test_device_energies.so and handed to the library
// through tad_fn_add_elements_device.  The parity tests compare them with the
// library's built-in catalog (itself pinned to the oracle), with the oracle
// directly, and (for the transcendental one) with finite differences.
#include <tinyad_amd/DeviceElements.hh>

#include <cmath>

namespace {

// Symmetric Dirichlet on a triangle as the reference's test lambda states it
// (tests/NewtonTest.cc:28-44): M = col_mat(b - a, c - a), +inf if det M <= 0,
// else w (||M Mr^-1||^2 + ||Mr M^-1||^2).  p = [Mr^-1 (2 x 2 row-major), w].
// For 2 x 2 matrices ||J^-1||_F = ||J||_F / |det J|.
struct SymdirTri
{
    template <class S>
    __device__ S operator()(const S* x, const double* p) const
    {
        const S m00 = x[2] - x[0], m10 = x[3] - x[1];
        const S m01 = x[4] - x[0], m11 = x[5] - x[1];
        if (m00 * m11 - m01 * m10 <= 0.0) return S(INFINITY);
        const S j00 = m00 * p[0] + m01 * p[2], j01 = m00 * p[1] + m01 * p[3];
        const S j10 = m10 * p[0] + m11 * p[2], j11 = m10 * p[1] + m11 * p[3];
        const S n2 = tad::sqr(j00) + tad::sqr(j01) + tad::sqr(j10) + tad::sqr(j11);
        return p[4] * (n2 + n2 / tad::sqr(j00 * j11 - j01 * j10));
    }
};

// The same energy on a tetrahedron; ||J^-1||_F = ||cof J||_F / |det J|.
// p = [Mr^-1 (3 x 3 row-major), w].
struct SymdirTet
{
    template <class S>
    __device__ S operator()(const S* x, const double* p) const
    {
        S m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) m[i][c] = x[3 * (c + 1) + i] - x[i];
        const S detm = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (detm <= 0.0) return S(INFINITY);
        S j[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) j[i][c] = m[i][0] * p[0 * 3 + c] + m[i][1] * p[1 * 3 + c] + m[i][2] * p[2 * 3 + c];
        S nj = S(0.0), nc = S(0.0);
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c)
            {
                nj += tad::sqr(j[i][c]);
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
                nc += tad::sqr(j[i1][c1] * j[i2][c2] - j[i1][c2] * j[i2][c1]);
            }
        const S detj = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                       j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                       j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
        return p[9] * (nj + nc / tad::sqr(detj));
    }
};

// The same two energies declaring that they depend on coordinate differences
// only (projected through the eigensolve without the translations).
struct SymdirTriInv : SymdirTri
{
    static constexpr bool translation_invariant = true;
};
struct SymdirTetInv : SymdirTet
{
    static constexpr bool translation_invariant = true;
};

// |x_0|^2 + |x_1 - x_0|^2 over two 2D handles, wrongly declared translation
// invariant (the library must report it rather than project it).
struct FalseInvariant
{
    static constexpr bool translation_invariant = true;
    template <class S>
    __device__ S operator()(const S* x, const double*) const
    {
        return tad::sqr(x[0]) + tad::sqr(x[1]) + tad::sqr(x[2] - x[0]) + tad::sqr(x[3] - x[1]);
    }
};

// 0.5 x^T A x + b^T x + c over two 2D handles (K = 4): p = [A (4 x 4), b (4), c],
// the parameter row of the catalog's quadratic element.
struct Quadratic4
{
    template <class S>
    __device__ S operator()(const S* x, const double* p) const
    {
        S e = S(p[20]);
        for (int i = 0; i < 4; ++i)
        {
            e += p[16 + i] * x[i];
            for (int k = 0; k < 4; ++k) e += 0.5 * p[4 * i + k] * (x[i] * x[k]);
        }
        return e;
    }
};

// A smooth energy that uses every transcendental helper, on two 2D handles:
// p0 exp(0.3 x0 x3) + log(1 + x1^2) + sqrt(1 + x0^2 + x2^2)
//   + sin(x1) cos(x2) + p1 (1 + x3^2)^1.5 + 1 / (2 + x0^2).
struct Smooth4
{
    template <class S>
    __device__ S operator()(const S* x, const double* p) const
    {
        return p[0] * tad::exp(0.3 * x[0] * x[3]) + tad::log(1 + x[1] * x[1]) +
               tad::sqrt(1.0 + x[0] * x[0] + x[2] * x[2]) + tad::sin(x[1]) * tad::cos(x[2]) +
               p[1] * tad::pow(1.0 + x[3] * x[3], 1.5) + 1.0 / (2 + x[0] * x[0]);
    }
};

// The same helpers on four 2D handles (K = 8, hyper-dual passes), with
// active / active division and an integer power:
// p0 exp(0.3 x0 x7) + log(1 + x1^2) + sqrt(1 + x2^2 + x4^2) + sin(x3) cos(x5)
//   + p1 (1 + x6^2)^1.5 + 1 / (2 + x0^2) + x5 / (1.5 + x7^2) + (x2 - x6)^3.
struct Smooth8
{
    template <class S>
    __device__ S operator()(const S* x, const double* p) const
    {
        return p[0] * tad::exp(0.3 * x[0] * x[7]) + tad::log(1 + x[1] * x[1]) +
               tad::sqrt(1.0 + x[2] * x[2] + x[4] * x[4]) + tad::sin(x[3]) * tad::cos(x[5]) +
               p[1] * tad::pow(1.0 + x[6] * x[6], 1.5) + 1.0 / (2 + x[0] * x[0]) + x[5] / (1.5 + x[7] * x[7]) +
               tad::pow(x[2] - x[6], 3);
    }
};

// Powers with the base at 0 allowed (Scalar.hh:266-306): x0^2 and x1^3 with
// double exponents, x2^1, x3^2, x4^3 and x5^0 with int exponents, and x_k^2
// (double exponent) for the coordinates of handles past the third; N 2D
// handles, K = 2N (N = 3: one Dual pass, N = 4: hyper-dual passes).
template <int N>
struct Powers
{
    template <class S>
    __device__ S operator()(const S* x, const double*) const
    {
        S e = tad::pow(x[0], 2.0) + tad::pow(x[1], 3.0) + tad::pow(x[2], 1) + tad::pow(x[3], 2) +
              tad::pow(x[4], 3) + tad::pow(x[5], 0);
        for (int k = 6; k < 2 * N; ++k) e += tad::pow(x[k], 2.0);
        return e;
    }
};

// ||sum of the element's 2D handle values||^2, DynamicElementsTest.cc:9-33
// (padding slots read as 0, so the sum runs over the real handles only).
struct SumSquares2
{
    template <int N>
    struct At
    {
        template <class S>
        __device__ S operator()(const S* x, const double*) const
        {
            S sx = S(0.0), sy = S(0.0);
            for (int a = 0; a < N; ++a)
            {
                sx += x[2 * a];
                sy += x[2 * a + 1];
            }
            return tad::sqr(sx) + tad::sqr(sy);
        }
    };
};

// One-ring Dirichlet energy of a vertex (DynamicElementsTest.cc:93-131):
// x[0] the vertex, x[1 .. count) its neighbours, p[0] = count (1D values).
struct OneRing
{
    template <int N>
    struct At
    {
        template <class S>
        __device__ S operator()(const S* x, const double* p) const
        {
            S e = S(0.0);
            for (int i = 1; i < N; ++i)
                if (i < p[0]) e += 0.25 * tad::sqr(x[0] - x[i]);
            return e;
        }
    };
};

// Residuals of the reference's Gauss-Newton test (tests/GaussNewtonTest.cc:30-60):
// s (J00, J01, J10, J11, Ji00, Ji01, Ji10, Ji11) with J = M Mr^-1 and
// Ji = Mr M^-1 (Mr = (Mr^-1)^-1 from the parameters), every residual +inf if
// det M <= 0.  p = [Mr^-1 (2 x 2 row-major), s].
struct SymdirTriResidual
{
    template <class S>
    __device__ void operator()(const S* x, const double* p, S* r) const
    {
        const S m00 = x[2] - x[0], m10 = x[3] - x[1];
        const S m01 = x[4] - x[0], m11 = x[5] - x[1];
        const S det = m00 * m11 - m01 * m10;
        if (det <= 0.0)
        {
            for (int o = 0; o < 8; ++o) r[o] = S(INFINITY);
            return;
        }
        const double dr = p[0] * p[3] - p[1] * p[2];
        const double q00 = p[3] / dr, q01 = -p[1] / dr, q10 = -p[2] / dr, q11 = p[0] / dr;
        const S inv = 1.0 / det;
        const S n00 = m11 * inv, n01 = -m01 * inv, n10 = -m10 * inv, n11 = m00 * inv;
        r[0] = p[4] * (m00 * p[0] + m01 * p[2]);
        r[1] = p[4] * (m00 * p[1] + m01 * p[3]);
        r[2] = p[4] * (m10 * p[0] + m11 * p[2]);
        r[3] = p[4] * (m10 * p[1] + m11 * p[3]);
        r[4] = p[4] * (q00 * n00 + q01 * n10);
        r[5] = p[4] * (q00 * n01 + q01 * n11);
        r[6] = p[4] * (q10 * n00 + q11 * n10);
        r[7] = p[4] * (q10 * n01 + q11 * n11);
    }
};

// Pin residuals s (target - x) of one handle (GaussNewtonTest.cc:62-69 with
// weight s).  p = [target (D), s].
template <int D>
struct PinResidual
{
    template <class S>
    __device__ void operator()(const S* x, const double* p, S* r) const
    {
        for (int i = 0; i < D; ++i) r[i] = p[D] * (p[i] - x[i]);
    }
};

// Symmetric Dirichlet tet residuals: s (J (9 entries), J^-1 (9 entries)),
// J^-1 = cof(J)^T / det J; +inf if det M <= 0.  p = [Mr^-1 (3 x 3 row-major), s].
struct SymdirTetResidual
{
    template <class S>
    __device__ void operator()(const S* x, const double* p, S* r) const
    {
        S m[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) m[i][c] = x[3 * (c + 1) + i] - x[i];
        const S detm = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (detm <= 0.0)
        {
            for (int o = 0; o < 18; ++o) r[o] = S(INFINITY);
            return;
        }
        S j[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) j[i][c] = m[i][0] * p[0 * 3 + c] + m[i][1] * p[1 * 3 + c] + m[i][2] * p[2 * 3 + c];
        const S detj = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) -
                       j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                       j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
        const S inv = 1.0 / detj;
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c)
            {
                r[3 * i + c] = p[9] * j[i][c];
                // (J^-1)_{ic} = cof(J)_{ci} / det J
                const int c1 = (c + 1) % 3, c2 = (c + 2) % 3, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                r[9 + 3 * i + c] = p[9] * ((j[c1][i1] * j[c2][i2] - j[c1][i2] * j[c2][i1]) * inv);
            }
    }
};

const tad_device_residual_ops g_res_tri = tinyad_amd::device_residual<SymdirTriResidual, 3, 2, 8, 5>();
const tad_device_residual_ops g_res_pin2 = tinyad_amd::device_residual<PinResidual<2>, 1, 2, 2, 3>();
const tad_device_residual_ops g_res_pin3 = tinyad_amd::device_residual<PinResidual<3>, 1, 3, 3, 4>();
const tad_device_residual_ops g_res_tet = tinyad_amd::device_residual<SymdirTetResidual, 4, 3, 18, 10>();

const tad_device_ops g_tri = tinyad_amd::device_element<SymdirTri, 3, 2, 5>();
const tad_device_ops g_tet = tinyad_amd::device_element<SymdirTet, 4, 3, 10>();
const tad_device_ops g_tri_inv = tinyad_amd::device_element<SymdirTriInv, 3, 2, 5>();
const tad_device_ops g_tet_inv = tinyad_amd::device_element<SymdirTetInv, 4, 3, 10>();
const tad_device_ops g_false_inv = tinyad_amd::device_element<FalseInvariant, 2, 2, 0>();
const tad_device_ops g_quad = tinyad_amd::device_element<Quadratic4, 2, 2, 21>();
const tad_device_ops g_smooth = tinyad_amd::device_element<Smooth4, 2, 2, 2>();
const tad_device_ops g_smooth8 = tinyad_amd::device_element<Smooth8, 4, 2, 2>();
const tad_device_ops g_powers[2] = { tinyad_amd::device_element<Powers<3>, 3, 2, 0>(),
                                     tinyad_amd::device_element<Powers<4>, 4, 2, 0>() };
const tad_device_ops g_sum_sq[2] = { tinyad_amd::device_element<SumSquares2::At<3>, 3, 2, 0>(),
                                     tinyad_amd::device_element<SumSquares2::At<1>, 1, 2, 0>() };
const tad_device_ops g_one_ring[4] = { tinyad_amd::device_element<OneRing::At<6>, 6, 1, 1>(),
                                       tinyad_amd::device_element<OneRing::At<7>, 7, 1, 1>(),
                                       tinyad_amd::device_element<OneRing::At<8>, 8, 1, 1>(),
                                       tinyad_amd::device_element<OneRing::At<10>, 10, 1, 1>() };

} // namespace

extern "C" {
const tad_device_ops* tad_test_symdir_tri_ops() { return &g_tri; }
const tad_device_ops* tad_test_symdir_tet_ops() { return &g_tet; }
const tad_device_ops* tad_test_symdir_tri_inv_ops() { return &g_tri_inv; }
const tad_device_ops* tad_test_symdir_tet_inv_ops() { return &g_tet_inv; }
const tad_device_ops* tad_test_false_invariant_ops() { return &g_false_inv; }
const tad_device_ops* tad_test_quadratic4_ops() { return &g_quad; }
const tad_device_ops* tad_test_smooth4_ops() { return &g_smooth; }
const tad_device_ops* tad_test_smooth8_ops() { return &g_smooth8; }
// 3 and 4 handles (K = 6 and 8)
const tad_device_ops* tad_test_powers_ops(int i) { return (i >= 0 && i < 2) ? &g_powers[i] : nullptr; }
// valences 3 and 1, in that order (add_elements_dynamic<3, 1>)
const tad_device_ops* tad_test_sum_squares_ops(int i) { return (i >= 0 && i < 2) ? &g_sum_sq[i] : nullptr; }
// valences 6, 7, 8, 10 (add_elements_dynamic<6, 7, 8, 10>)
const tad_device_ops* tad_test_one_ring_ops(int i) { return (i >= 0 && i < 4) ? &g_one_ring[i] : nullptr; }
// residual elements of vector functions
const tad_device_residual_ops* tad_test_res_symdir_tri_ops() { return &g_res_tri; }
const tad_device_residual_ops* tad_test_res_pin_ops(int d) { return d == 2 ? &g_res_pin2 : d == 3 ? &g_res_pin3 : nullptr; }
const tad_device_residual_ops* tad_test_res_symdir_tet_ops() { return &g_res_tet; }
}
