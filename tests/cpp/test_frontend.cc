// C++ front-end checks, written like the reference's gtest cases
// (tests/NewtonTest.cc, tests/ScalarFunctionTest.cc, tests/ExceptionTest.cc)
// but against include/tinyad_amd/ScalarFunction.hh.  Prints one line per
// case; exit status 0 when every case passes.  Needs a GPU.
#include <cstdint>

// HandleTypeTest.cc:8-45: custom handle types made known by idx_from_handle()
// overloads declared before the header is included
struct CustomVariableHandle
{
    CustomVariableHandle() = default;
    CustomVariableHandle(int i) : idx(i) {}
    int idx = -1;
};
struct CustomElementHandle
{
    CustomElementHandle() = default;
    CustomElementHandle(int i) : idx(i) {}
    int idx = -1;
};
namespace tinyad_amd {
inline int64_t idx_from_handle(const CustomVariableHandle& h) { return h.idx; }
inline int64_t idx_from_handle(const CustomElementHandle& h) { return h.idx; }
} // namespace tinyad_amd

#include <tinyad_amd/ScalarFunction.hh>

#include <array>
#include <cmath>
#include <memory>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static int g_failed = 0;
#define EXPECT(cond)                                                                   \
    do                                                                                 \
    {                                                                                  \
        if (!(cond))                                                                   \
        {                                                                              \
            std::printf("  check failed at line %d: %s\n", __LINE__, #cond);        \
            ok = false;                                                                \
        }                                                                              \
    } while (0)

using tinyad_amd::VectorXd;

// symmetric Dirichlet on triangles as a device functor (tests/cpp/device_energies.hip)
extern "C" const tad_device_ops* tad_test_symdir_tri_ops();
// ||sum of the element's handle values||^2 compiled at valences 3 (i = 0) and 1 (i = 1)
extern "C" const tad_device_ops* tad_test_sum_squares_ops(int i);
// residuals (tests/cpp/device_energies.hip): symmetric Dirichlet triangles (8), pins (d)
extern "C" const tad_device_residual_ops* tad_test_res_symdir_tri_ops();
extern "C" const tad_device_residual_ops* tad_test_res_pin_ops(int d);

static std::vector<double> quad_params(const std::vector<double>& A, const std::vector<double>& b, double c)
{
    std::vector<double> p = A;
    p.insert(p.end(), b.begin(), b.end());
    p.push_back(c);
    return p;
}

static bool test_basic()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(1));
    func.add_elements<1>(tinyad_amd::range(1), TAD_E_QUADRATIC, [](int) { return std::array<int, 1>{ 0 }; },
                         [](int) { return quad_params({ 2, 0, 0, 2 }, { 0, 0 }, 0); });
    VectorXd x = { 1.0, 2.0 };
    EXPECT(func.eval(x) == 5.0);
    auto [f, g] = func.eval_with_gradient(x);
    EXPECT(f == 5.0 && g[0] == 2.0 && g[1] == 4.0);
    auto [fh, gh, H] = func.eval_with_derivatives(x);
    EXPECT(fh == 5.0 && H.coeff(0, 0) == 2.0 && H.coeff(1, 1) == 2.0);
    // move construction / assignment keep the function usable (ScalarFunctionTest.cc:181-252)
    auto func2 = std::move(func);
    EXPECT(func2.eval(x) == 5.0);
    tinyad_amd::ScalarFunction<2> func3;
    func3 = std::move(func2);
    EXPECT(func3(x) == 5.0);
    bool threw = false;
    try { func2.eval(x); } catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw);
    return ok;
}

static bool test_2d_nonconvex()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(1));
    func.add_elements<1>(tinyad_amd::range(1), TAD_E_QUADRATIC, [](int) { return std::array<int, 1>{ 0 }; },
                         [](int) { return quad_params({ -4, -2, -2, -2 }, { -1, 0 }, -1); });
    VectorXd x = { 1.0, 2.0 };
    auto [f, g, H] = func.eval_with_derivatives(x);
    auto [f2, g2, Hp] = func.eval_with_hessian_proj(x);
    EXPECT(f == -12 && f2 == -12 && g[0] == -9 && g[1] == -6);
    EXPECT(H.coeff(0, 0) == -4 && H.coeff(0, 1) == -2 && H.coeff(1, 1) == -2);
    const double a = Hp.coeff(0, 0), b = Hp.coeff(0, 1), c = Hp.coeff(1, 1);
    EXPECT(a > 0 && a * c - b * b > 0);  // positive definite
    return ok;
}

// ScalarFunctionTest.cc:153-178: an element that asks for the same handle twice
// gets one set of variables for it.  Here a valence-2 quadratic element lists
// handle 0 twice: E = (|x_0|^2 + |x_0|^2) / 2, so f = |x_0|^2, g_0 = 2 x_0,
// H_00 = 2 I, and handle 1 stays out of the pattern.
static bool test_element_variables()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(2));
    func.add_elements<2>(tinyad_amd::range(1), TAD_E_QUADRATIC, [](int) { return std::array<int, 2>{ 0, 0 }; },
                         [](int) {
                             return quad_params({ 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 }, { 0, 0, 0, 0 }, 0);
                         });
    VectorXd x = { 1.0, 2.0, 0.0, 0.0 };
    auto [f, g, H] = func.eval_with_hessian_proj(x);
    EXPECT(f == 5.0 && g[0] == 2.0 && g[1] == 4.0 && g[2] == 0.0 && g[3] == 0.0);
    EXPECT(H.nonZeros() == 4 && H.coeff(0, 0) == 2.0 && H.coeff(1, 1) == 2.0 && H.coeff(0, 1) == 0.0);
    EXPECT(func.eval(x) == 5.0);
    return ok;
}

// tests/NewtonTest.cc:12-95 on the 6-vertex planar fixture (tests/Meshes.hh:12-42)
// NewtonTest.cc:14-80; the triangle energy from the catalog, or with ops
// given, from a device functor written like the reference's lambda
static bool test_newton(VectorXd* x_out, const tad_device_ops* ops = nullptr)
{
    bool ok = true;
    const double V_rest[6][2] = { { 0, 0 }, { 1, 0 }, { 0, 1 }, { 1, 1 }, { 0, 2 }, { 1, 2 } };
    const int F[4][3] = { { 0, 1, 2 }, { 1, 3, 2 }, { 2, 3, 5 }, { 2, 5, 4 } };
    const int b[2] = { 0, 4 };
    const double bc[2][2] = { { 0.0, 0.0 }, { -2.0, 0.0 } };
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(6));
    auto conn_of = [&](int f) { return std::array<int, 3>{ F[f][0], F[f][1], F[f][2] }; };
    auto params_of = [&](int f) {
        const double m00 = V_rest[F[f][1]][0] - V_rest[F[f][0]][0], m10 = V_rest[F[f][1]][1] - V_rest[F[f][0]][1];
        const double m01 = V_rest[F[f][2]][0] - V_rest[F[f][0]][0], m11 = V_rest[F[f][2]][1] - V_rest[F[f][0]][1];
        const double det = m00 * m11 - m01 * m10;
        return std::vector<double>{ m11 / det, -m01 / det, -m10 / det, m00 / det, 1.0 / 4.0 };
    };
    if (ops) func.add_elements<3>(tinyad_amd::range(4), *ops, conn_of, params_of);
    else func.add_elements<3>(tinyad_amd::range(4), TAD_E_SYMDIR_TRI, conn_of, params_of);
    func.add_elements<1>(tinyad_amd::range(2), TAD_E_PENALTY, [&](int i) { return std::array<int, 1>{ b[i] }; },
                         [&](int i) { return std::vector<double>{ bc[i][0], bc[i][1], 1.0 }; });
    VectorXd x = func.x_from_data([&](int v) { return std::array<double, 2>{ 0.5 * V_rest[v][0], 0.25 * V_rest[v][1] }; });
    {
        auto [f, g, H_proj] = func.eval_with_hessian_proj(x);
        EXPECT(H_proj.nonZeros() == 4 * 6 + 8 * (6 + 4 - 1));
    }
    tinyad_amd::LinearSolver solver;
    solver.opts.tol = 1e-14;
    for (int i = 0; i < 10; ++i)
    {
        auto [f, g, H_proj] = func.eval_with_hessian_proj(x);
        VectorXd d = tinyad_amd::newton_direction(func, g, H_proj, solver);
        x = tinyad_amd::line_search(x, d, f, g, func);
    }
    auto [f, g, H_proj] = func.eval_with_hessian_proj(x);
    double gmax = 0;
    for (double gi : g) gmax = std::fmax(gmax, std::fabs(gi));
    EXPECT(func.eval(x) == f);
    EXPECT(std::fabs(f - 4.0) <= 1e-12);
    EXPECT(gmax <= 1e-10);
    std::vector<std::array<double, 2>> V(6);
    func.x_to_data(x, [&](int v, const std::array<double, 2>& p) { V[v] = p; });
    EXPECT(V[4][0] == x[8] && V[4][1] == x[9]);
    if (x_out) *x_out = x;
    return ok;
}

// tests/GaussNewtonTest.cc:13-162: the Newton test's energy as residuals
// (8 per triangle scaled by 1/sqrt(F.rows()), 2 per pin), compared with the
// scalar form at every iterate (f and g to 1e-12), 20 Gauss-Newton iterations
// with w = 1e-12 and a line search on eval_sum_of_squares; f -> 4 +- 0.1.
static bool test_gauss_newton()
{
    bool ok = true;
    const double V_rest[6][2] = { { 0, 0 }, { 1, 0 }, { 0, 1 }, { 1, 1 }, { 0, 2 }, { 1, 2 } };
    const int F[4][3] = { { 0, 1, 2 }, { 1, 3, 2 }, { 2, 3, 5 }, { 2, 5, 4 } };
    const int b[2] = { 0, 4 };
    const double bc[2][2] = { { 0.0, 0.0 }, { -2.0, 0.0 } };
    auto conn_of = [&](int f) { return std::array<int, 3>{ F[f][0], F[f][1], F[f][2] }; };
    auto rest_params = [&](int f, double w) {
        const double m00 = V_rest[F[f][1]][0] - V_rest[F[f][0]][0], m10 = V_rest[F[f][1]][1] - V_rest[F[f][0]][1];
        const double m01 = V_rest[F[f][2]][0] - V_rest[F[f][0]][0], m11 = V_rest[F[f][2]][1] - V_rest[F[f][0]][1];
        const double det = m00 * m11 - m01 * m10;
        return std::vector<double>{ m11 / det, -m01 / det, -m10 / det, m00 / det, w };
    };
    auto pin_of = [&](int i) { return std::array<int, 1>{ b[i] }; };
    auto pin_params = [&](int i) { return std::vector<double>{ bc[i][0], bc[i][1], 1.0 }; };
    auto func_sos = tinyad_amd::vector_function<2>(tinyad_amd::range(6));
    func_sos.add_elements<3, 8>(tinyad_amd::range(4), *tad_test_res_symdir_tri_ops(), conn_of,
                                [&](int f) { return rest_params(f, 1.0 / std::sqrt(4.0)); });
    func_sos.add_elements<1, 2>(tinyad_amd::range(2), *tad_test_res_pin_ops(2), pin_of, pin_params);
    auto func_ref = tinyad_amd::scalar_function<2>(tinyad_amd::range(6));
    func_ref.add_elements<3>(tinyad_amd::range(4), TAD_E_SYMDIR_TRI, conn_of,
                             [&](int f) { return rest_params(f, 1.0 / 4.0); });
    func_ref.add_elements<1>(tinyad_amd::range(2), TAD_E_PENALTY, pin_of, pin_params);
    EXPECT(func_sos.n_outputs == 36);
    VectorXd x = func_sos.x_from_data([&](int v) { return std::array<double, 2>{ 0.5 * V_rest[v][0], 0.25 * V_rest[v][1] }; });
    tinyad_amd::LinearSolver solver;
    solver.opts.tol = 1e-13;
    for (int i = 0; i < 20; ++i)
    {
        auto [f_ref, g_ref, H_ref] = func_ref.eval_with_hessian_proj(x);
        EXPECT(std::fabs(func_sos.eval_sum_of_squares(x) - f_ref) <= 1e-12);
        auto [f, g, r, J] = func_sos.eval_sum_of_squares_with_derivatives(x);
        EXPECT(std::fabs(f - f_ref) <= 1e-12);
        double dg = 0.0;
        for (size_t k = 0; k < g.size(); ++k) dg = std::fmax(dg, std::fabs(g[k] - g_ref[k]));
        EXPECT(dg <= 1e-12);
        EXPECT(J.rows() == 36 && J.cols() == 12 && J.nonZeros() == 4 * 8 * 6 + 2 * 2 * 2);
        // g = 2 J^T r from the returned J and r
        VectorXd jtr(12, 0.0);
        for (int64_t row = 0; row < J.rows(); ++row)
            for (int32_t k = J.outer[row]; k < J.outer[row + 1]; ++k) jtr[J.inner[k]] += J.values[k] * r[row];
        double dj = 0.0;
        for (size_t k = 0; k < g.size(); ++k) dj = std::fmax(dj, std::fabs(2.0 * jtr[k] - g[k]));
        EXPECT(dj <= 1e-12);
        const VectorXd d = tinyad_amd::gauss_newton_direction(func_sos, x, solver, 1e-12);
        x = tinyad_amd::line_search(x, d, f, g, [&](const VectorXd& z) { return func_sos.eval_sum_of_squares(z); });
    }
    auto [f, g, r, J] = func_sos.eval_sum_of_squares_with_derivatives(x);
    double gmax = 0.0;
    for (double gi : g) gmax = std::fmax(gmax, std::fabs(gi));
    EXPECT(std::fabs(f - 4.0) <= 0.1 && gmax <= 0.1);
    const VectorXd rr = func_sos(x);
    double drr = 0.0;
    for (size_t k = 0; k < rr.size(); ++k) drr = std::fmax(drr, std::fabs(rr[k] - r[k]));
    EXPECT(rr.size() == 36 && drr <= 1e-14);
    bool threw = false;
    try { func_sos.add_elements<3, 8>(tinyad_amd::range(1), *tad_test_res_pin_ops(2), conn_of, pin_params); }
    catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw);  // valence / output count differ from the ops
    return ok;
}

static bool test_newton_deterministic()
{
    bool ok = true;
    VectorXd x0, x1;
    ok = test_newton(&x0) && ok;
    ok = test_newton(&x1) && ok;
    EXPECT(std::memcmp(x0.data(), x1.data(), x0.size() * sizeof(double)) == 0);
    return ok;
}

// ExceptionTest.cc:8-54 semantics: a failing call throws, the object stays usable
static bool test_errors_and_recovery()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<1>(tinyad_amd::range(1));
    func.add_elements<1>(tinyad_amd::range(10), TAD_E_QUADRATIC, [](int) { return std::array<int, 1>{ 0 }; },
                         [](int) { return quad_params({ 0 }, { 0 }, 1.0); });
    int caught = 0;
    for (int i = 0; i < 2; ++i)
    {
        try { func.eval_with_hessian_proj(VectorXd(3, 0.0)); }
        catch (const std::runtime_error&) { ++caught; }
    }
    EXPECT(caught == 2);
    auto [f, g, H] = func.eval_with_hessian_proj(VectorXd(1, 0.0));
    EXPECT(std::fabs(f - 10.0) <= 1e-12);
    bool threw = false;
    try
    {
        func.add_elements<1>(tinyad_amd::range(1), TAD_E_QUADRATIC, [](int) { return std::array<int, 1>{ 5 }; },
                             [](int) { return quad_params({ 0 }, { 0 }, 1.0); });
    }
    catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw);
    return ok;
}

// Two 2-triangle meshes as segments of one function: per-mesh energies equal
// separate functions, autotuning keeps every bit, set_params equals a function
// built with the new rows.
static bool test_segments_autotune_params()
{
    bool ok = true;
    const std::array<int, 3> tris[2] = { { 0, 1, 2 }, { 1, 3, 2 } };
    auto params = [](double w) { return std::vector<double>{ 1, 0, 0, 1, w }; };  // Mr = I
    auto build = [&](int meshes, double w) {
        auto f = tinyad_amd::scalar_function<2>(tinyad_amd::range(4 * meshes));
        f.add_elements<3>(tinyad_amd::range(2 * meshes), TAD_E_SYMDIR_TRI,
                          [&](int e) { auto t = tris[e % 2]; for (int& v : t) v += 4 * (e / 2); return t; },
                          [&](int) { return params(w); });
        return f;
    };
    VectorXd x = { 0, 0, 1.2, 0, 0, 0.9, 1.1, 1.0, 0, 0, 1.5, 0.1, 0.2, 1.1, 1.3, 1.2 };
    auto both = build(2, 0.5);
    both.set_segments({ 0, 4, 8 });
    auto [f, g, H] = both.eval_with_hessian_proj(x);
    std::vector<double> fs = both.segment_energies();
    EXPECT(fs.size() == 2);
    for (int m = 0; m < 2; ++m)
    {
        auto one = build(1, 0.5);
        VectorXd xm(x.begin() + 8 * m, x.begin() + 8 * m + 8);
        EXPECT(std::fabs(one.eval(xm) - fs[m]) <= 1e-12 * std::fabs(fs[m]));
    }
    EXPECT(std::fabs(fs[0] + fs[1] - f) <= 1e-12 * std::fabs(f));
    const std::string rep = both.autotune(x);
    EXPECT(rep.find("best") != std::string::npos);
    auto [f2, g2, H2] = both.eval_with_hessian_proj(x);
    EXPECT(f2 == f && g2 == g && H2.values == H.values);
    std::vector<double> rows;
    for (int e = 0; e < 4; ++e) { auto r = params(2.0); rows.insert(rows.end(), r.begin(), r.end()); }
    both.set_params(0, rows);
    auto fresh = build(2, 2.0);
    EXPECT(both.eval(x) == fresh.eval(x));
    // launch shapes: store, change, restore; results keep their bits
    const auto tuned = both.launch_shape(0);
    const int tuned_asm = both.assembly_shape();
    auto [f3, g3, H3] = both.eval_with_hessian_proj(x);
    both.set_assembly_shape((tuned_asm + 1) % 4);
    both.set_launch_shape(0, { 64, 8, 1 });
    auto [f4, g4, H4] = both.eval_with_hessian_proj(x);
    EXPECT(f4 == f3 && g4 == g3 && H4.values == H3.values);
    both.set_launch_shape(0, tuned);
    both.set_assembly_shape(tuned_asm);
    EXPECT(both.launch_shape(0).block == tuned.block && both.assembly_shape() == tuned_asm);
    return ok;
}

// Newton direction with and without the two-level preconditioner on a pinned
// 24 x 24 grid of triangles: same direction, the coarse space reported.
static bool test_two_level_solve()
{
    bool ok = true;
    const int n = 24;
    auto vid = [&](int i, int j) { return i * n + j; };
    std::vector<std::array<int, 3>> F;
    for (int i = 0; i + 1 < n; ++i)
        for (int j = 0; j + 1 < n; ++j)
        {
            F.push_back({ vid(i, j), vid(i + 1, j), vid(i + 1, j + 1) });
            F.push_back({ vid(i, j), vid(i + 1, j + 1), vid(i, j + 1) });
        }
    const double h = 1.0 / (n - 1);
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(n * n));
    func.add_elements<3>(tinyad_amd::range((int)F.size()), TAD_E_SYMDIR_TRI, [&](int f) { return F[f]; },
                         [&](int) { return std::vector<double>{ 1 / h, 0, 0, 1 / h, 1e-3 }; });
    const int pins[3] = { vid(0, 0), vid(n - 1, 0), vid(0, n - 1) };
    func.add_elements<1>(tinyad_amd::range(3), TAD_E_PENALTY, [&](int k) { return std::array<int, 1>{ pins[k] }; },
                         [&](int k) {
                             const int v = pins[k];
                             return std::vector<double>{ (v / n) * h, (v % n) * h, 1.0 };
                         });
    VectorXd x = func.x_from_data([&](int v) {
        const double a = (v / n) * h, b = (v % n) * h;
        return std::array<double, 2>{ 1.1 * a + 0.001 * std::sin(7.0 * v), 0.9 * b + 0.001 * std::cos(5.0 * v) };
    });
    auto [f, g, H] = func.eval_with_hessian_proj(x);
    tinyad_amd::LinearSolver plain, two;
    plain.opts.coarse_aggregate = 0;
    plain.opts.tol = two.opts.tol = 1e-12;
    two.opts.coarse_aggregate = 32;
    VectorXd d0 = tinyad_amd::newton_direction(func, g, H, plain);
    {
        // the reference's Newton loop with a solver that refreshes the coarse
        // space on every 3rd direction
        tinyad_amd::LinearSolver lag;
        lag.coarse_refresh = 3;
        lag.opts.coarse_aggregate = 32;
        VectorXd xl = x;
        std::vector<int> reused;
        for (int it = 0; it < 5; ++it)
        {
            auto [fl, gl, Hl] = func.eval_with_hessian_proj(xl);
            VectorXd dl = tinyad_amd::newton_direction(func, gl, Hl, lag);
            reused.push_back(lag.last_stats.coarse_reused);
            xl = tinyad_amd::line_search(xl, dl, fl, gl, func);
        }
        EXPECT((reused == std::vector<int>{ 0, 1, 1, 0, 1 }));
        // a fresh solver on the same function builds its own coarse space on
        // its first call, although a compatible one is cached on the function
        tinyad_amd::LinearSolver fresh;
        fresh.coarse_refresh = 3;
        fresh.opts.coarse_aggregate = 32;
        auto [ff, gf, Hf] = func.eval_with_hessian_proj(xl);
        tinyad_amd::newton_direction(func, gf, Hf, fresh);
        EXPECT(fresh.last_stats.coarse_reused == 0 && fresh.last_stats.coarse_unknowns > 0);
        tinyad_amd::newton_direction(func, gf, Hf, fresh);
        EXPECT(fresh.last_stats.coarse_reused == 1);
        // a Hessian of the wrong size is rejected before the solve
        bool threw = false;
        try
        {
            tinyad_amd::SparseMatrix Hbad = Hf;
            Hbad.values.pop_back();
            tinyad_amd::newton_direction(func, gf, Hbad, fresh);
        }
        catch (const std::runtime_error&)
        {
            threw = true;
        }
        EXPECT(threw);
    }
    VectorXd d1 = tinyad_amd::newton_direction(func, g, H, two);
    double dmax = 0, diff = 0;
    for (size_t i = 0; i < d0.size(); ++i)
    {
        dmax = std::fmax(dmax, std::fabs(d0[i]));
        diff = std::fmax(diff, std::fabs(d0[i] - d1[i]));
    }
    EXPECT(plain.last_stats.coarse_unknowns == 0 && two.last_stats.coarse_unknowns % 3 == 0 &&
           two.last_stats.coarse_unknowns >= 30);
    EXPECT(two.last_stats.converged && two.last_stats.iters < plain.last_stats.iters);
    EXPECT(diff <= 1e-8 * dmax);
    return ok;
}

// SwitchSolverTest (SwitchSolverTest.cc:16-47): x^2 + y^2 at (1, 1), the Newton
// direction and the reduced-basis one with B = I are (-1, -1); a basis that
// keeps only x gives (-1, 0) whatever its scale; the basis (1, 1)^T gives
// (-1, -1), and with w = 1 (B^T H B + w = 5, B^T g = 4) (-0.8, -0.8); a basis
// with an empty column (B^T H B + w I singular) throws.
static bool test_reduced_basis()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(1));
    func.add_elements<1>(tinyad_amd::range(1), TAD_E_QUADRATIC, [](int) { return std::array<int, 1>{ 0 }; },
                         [](int) { return std::vector<double>{ 2, 0, 0, 2, 0, 0, 0 }; });
    VectorXd x = { 1.0, 1.0 };
    auto [f, g, H] = func.eval_with_hessian_proj(x);
    EXPECT(g[0] == 2.0 && g[1] == 2.0);
    tinyad_amd::LinearSolver solver;
    VectorXd d = tinyad_amd::newton_direction(func, g, H, solver);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && std::fabs(d[1] + 1.0) <= 1e-12);
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, tinyad_amd::identity(2), solver);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && std::fabs(d[1] + 1.0) <= 1e-12);
    tinyad_amd::SparseMatrix Bx;
    Bx.n_rows = 2;
    Bx.n_cols = 1;
    Bx.outer = { 0, 1, 1 };
    Bx.inner = { 0 };
    Bx.values = { 1.0 };
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, Bx, solver);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && d[1] == 0.0);
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, Bx);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && d[1] == 0.0);
    Bx.values = { 2.0 };
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, Bx, solver);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && d[1] == 0.0);
    tinyad_amd::SparseMatrix B1;
    B1.n_rows = 2;
    B1.n_cols = 1;
    B1.outer = { 0, 1, 2 };
    B1.inner = { 0, 0 };
    B1.values = { 1.0, 1.0 };
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, B1, solver);
    EXPECT(std::fabs(d[0] + 1.0) <= 1e-12 && std::fabs(d[1] + 1.0) <= 1e-12);
    d = tinyad_amd::newton_direction_reduced_basis(func, g, H, B1, solver, 1.0);
    EXPECT(std::fabs(d[0] + 0.8) <= 1e-12 && std::fabs(d[1] + 0.8) <= 1e-12);
    tinyad_amd::SparseMatrix B2 = B1;
    B2.n_cols = 2;  // column 1 has no entry
    bool threw = false;
    try { tinyad_amd::newton_direction_reduced_basis(func, g, H, B2, solver); }
    catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw);
    return ok;
}

// DynamicElementsTest.cc:9-33: element e accesses handles 0 .. e-1; valences
// <3, 1> give two terms of two elements each
static bool test_dynamic_elements()
{
    bool ok = true;
    auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(4));
    func.add_elements_dynamic(tinyad_amd::range(4), { *tad_test_sum_squares_ops(0), *tad_test_sum_squares_ops(1) },
                              [](int e) { return tinyad_amd::range(e); }, [](int) { return std::vector<double>{}; });
    EXPECT(func.n_terms() == 2);
    EXPECT(func.term_elements(0) == 2 && func.term_elements(1) == 2);
    VectorXd x(8, 1.0);
    auto [f, g, H_proj] = func.eval_with_hessian_proj(x);
    EXPECT(f == 28.0);
    EXPECT(g[0] == 12.0 && g[2] == 10.0 && g[4] == 6.0 && g[6] == 0.0);
    bool threw = false;
    try
    {
        func.add_elements_dynamic(tinyad_amd::range(1), { *tad_test_sum_squares_ops(0) },
                                  [](int) { return tinyad_amd::range(4); }, [](int) { return std::vector<double>{}; });
    }
    catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw && func.n_terms() == 2);
    return ok;
}

// Handle types of mesh libraries (Support/OpenMesh.hh, PMP.hh:
// h.idx(); GeometryCentral.hh: h.getIndex(); Polymesh.hh: h.idx.value),
// mocked here without the libraries
struct OmVertex { int i; int idx() const { return i; } };
struct GcVertex { size_t k; size_t getIndex() const { return k; } };
struct PmVertex { struct { int value; } idx; };

template <class H, class MakeH>
static bool check_handles(MakeH make)
{
    bool ok = true;
    const double V_rest[6][2] = { { 0, 0 }, { 1, 0 }, { 0, 1 }, { 1, 1 }, { 0, 2 }, { 1, 2 } };
    const int F[4][3] = { { 0, 1, 2 }, { 1, 3, 2 }, { 2, 3, 5 }, { 2, 5, 4 } };
    auto quad = [](int) { return quad_params({ 2, 0.5, 0.5, 1 }, { 0.1, -0.2 }, 0.0); };
    // the same energy over integer handles and over handle objects listed in reverse
    auto fi = tinyad_amd::scalar_function<2>(tinyad_amd::range(6));
    std::vector<H> hs;
    for (int v = 5; v >= 0; --v) hs.push_back(make(v));
    auto fh = tinyad_amd::scalar_function<2>(hs);
    fi.add_elements<1>(tinyad_amd::range(6), TAD_E_QUADRATIC, [](int v) { return std::array<int, 1>{ v }; }, quad);
    fh.template add_elements<1>(tinyad_amd::range(6), TAD_E_QUADRATIC, [&](int v) { return std::array<H, 1>{ make(v) }; },
                       quad);
    auto rest = [&](int f) {
        const double m00 = V_rest[F[f][1]][0] - V_rest[F[f][0]][0], m10 = V_rest[F[f][1]][1] - V_rest[F[f][0]][1];
        const double m01 = V_rest[F[f][2]][0] - V_rest[F[f][0]][0], m11 = V_rest[F[f][2]][1] - V_rest[F[f][0]][1];
        const double det = m00 * m11 - m01 * m10;
        return std::vector<double>{ m11 / det, -m01 / det, -m10 / det, m00 / det, 0.25 };
    };
    fi.add_elements<3>(tinyad_amd::range(4), TAD_E_SYMDIR_TRI,
                       [&](int f) { return std::array<int, 3>{ F[f][0], F[f][1], F[f][2] }; }, rest);
    fh.template add_elements<3>(tinyad_amd::range(4), TAD_E_SYMDIR_TRI,
                       [&](int f) { return std::array<H, 3>{ make(F[f][0]), make(F[f][1]), make(F[f][2]) }; }, rest);
    auto pos = [&](int v) { return std::array<double, 2>{ 0.9 * V_rest[v][0] + 0.01 * v, 1.1 * V_rest[v][1] }; };
    const VectorXd xi = fi.x_from_data([&](int v) { return pos(v); });
    const VectorXd xh = fh.x_from_data([&](H h) { return pos((int)tinyad_amd::idx_from_handle(h)); });
    EXPECT(xi == xh);
    auto [f1, g1, H1] = fi.eval_with_hessian_proj(xi);
    auto [f2, g2, H2] = fh.eval_with_hessian_proj(xh);
    EXPECT(f1 == f2 && g1 == g2 && H1.values == H2.values);
    int written = 0;
    fh.x_to_data(xh, [&](H h, const std::array<double, 2>& p) {
        const int v = (int)tinyad_amd::idx_from_handle(h);
        if (p[0] == xh[2 * v] && p[1] == xh[2 * v + 1]) ++written;
    });
    EXPECT(written == 6);
    auto d = tinyad_amd::newton_direction(fh, g2, H2, *std::make_unique<tinyad_amd::LinearSolver>());
    EXPECT(d.size() == 12);
    return ok;
}

// HandleTypeTest.cc:47-71: 5 variables, 10 elements, element e reads variable e / 2
template <class VH, class EH>
static bool check_reference_handle_test()
{
    bool ok = true;
    std::vector<VH> vhs;
    for (int i = 0; i < 5; ++i) vhs.push_back(VH(i));
    std::vector<EH> ehs;
    for (int i = 0; i < 10; ++i) ehs.push_back(EH(i));
    auto func = tinyad_amd::scalar_function<1>(vhs);
    func.template add_elements<1>(ehs, TAD_E_QUADRATIC,
        [](const EH& e) { return std::array<VH, 1>{ VH((int)tinyad_amd::idx_from_handle(e) / 2) }; },
        [](const EH&) { return quad_params({ 0.0 }, { 1.0 }, 0.0); });
    auto [f, g, H] = func.eval_with_hessian_proj(VectorXd(func.n_vars, 0.0));
    EXPECT(f == 0.0);
    for (double gi : g) EXPECT(gi == 2.0);
    return ok;
}

static bool test_handle_types()
{
    bool ok = true;
    ok = check_reference_handle_test<int, int>() && ok;
    ok = check_reference_handle_test<long int, long int>() && ok;
    ok = check_reference_handle_test<CustomVariableHandle, CustomElementHandle>() && ok;
    ok = check_handles<OmVertex>([](int v) { return OmVertex{ v }; }) && ok;
    ok = check_handles<GcVertex>([](int v) { return GcVertex{ (size_t)v }; }) && ok;
    ok = check_handles<PmVertex>([](int v) { PmVertex h; h.idx.value = v; return h; }) && ok;
    bool threw = false;
    try { tinyad_amd::scalar_function<2>(std::vector<OmVertex>{ { 0 }, { 0 } }); }
    catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw);  // not a permutation of 0..n-1
    return ok;
}

int main()
{
    struct Case { const char* name; bool (*fn)(); };
    const Case cases[] = {
        { "ScalarFunctionTest.Basic+Move", test_basic },
        { "ScalarFunctionTest.2DNonConvex", test_2d_nonconvex },
        { "ScalarFunctionTest.ElementVariables", test_element_variables },
        { "NewtonTest.2DDeformation", []() { return test_newton(nullptr); } },
        { "NewtonTest.Deterministic", test_newton_deterministic },
        { "NewtonTest.2DDeformation.DeviceFunctor", []() { return test_newton(nullptr, tad_test_symdir_tri_ops()); } },
        { "GaussNewtonTest.2DDeformationGaussNewton", test_gauss_newton },
        { "ExceptionTest", test_errors_and_recovery },
        { "Batching.SegmentsAutotuneParams", test_segments_autotune_params },
        { "Solver.TwoLevel", test_two_level_solve },
        { "SwitchSolverTest.ReducedBasis", test_reduced_basis },
        { "DynamicElementsTest.DynamicElementsTest", test_dynamic_elements },
        { "HandleTypeTest.MeshLibraryHandles", test_handle_types },
    };
    for (const Case& c : cases)
    {
        bool ok = false;
        try { ok = c.fn(); }
        catch (const std::exception& ex) { std::printf("  exception: %s\n", ex.what()); }
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", c.name);
        if (!ok) ++g_failed;
    }
    std::printf("%d failed\n", g_failed);
    return g_failed == 0 ? 0 : 1;
}
