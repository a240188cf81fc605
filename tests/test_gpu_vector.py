"""Vector functions on the GPU (VectorFunction.hh:19-192): residual elements as
device functors (tests/cpp/device_energies.hip), their residuals r, Jacobian
J (VectorObjectiveTerm.hh:162-246), sum of squares f = r . r with
g = 2 J^T r (VectorFunctionImpl.hh:228-275) and the Gauss-Newton direction
(w I + J^T J) d = -J^T r (GaussNewtonDirection.hh:24-47), against the CPU
oracle's restatement (oracle.VectorFunction, itself pinned to torch Jacobians
and the scalar energies in tests/test_oracle.py):
  * the reference's Gauss-Newton test (GaussNewtonTest.cc:113-162): 20
    iterations on the planar fixture, f and g equal to the scalar form's
    (1e-12), r, J and the direction equal to the oracle's, f -> 4 +- 0.1;
  * a 40 x 40 grid: r and J (identical pattern) to 1e-12, f, g and
    2 J^T J (the scalar function's Hessian pattern) to 1e-10 in FP64 and
    2e-5 with FP32 derivatives;
  * tet residuals (18 per tet, K = 12): the sum of squares equals the
    symmetric Dirichlet catalog energy and g its gradient, J matches central
    differences of r, and Gauss-Newton steps decrease f;
  * inverted elements, wrong element kinds and non-finite gradients.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
import tinyad_amd as tad
from tinyad_amd import configs

pytestmark = pytest.mark.gpu

PLUGIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "libtad_test_device_energies.so")


def nrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300))


@pytest.fixture(scope="module")
def res_ops():
    if not os.path.exists(PLUGIN):
        raise RuntimeError(f"{PLUGIN} is missing: run python -m tinyad_amd.build first")
    L = ctypes.CDLL(PLUGIN)
    for name in ["tad_test_res_symdir_tri_ops", "tad_test_res_symdir_tet_ops", "tad_test_res_pin_ops"]:
        getattr(L, name).restype = ctypes.c_void_p
    L.tad_test_res_pin_ops.argtypes = [ctypes.c_int]
    return {configs.R_SYMDIR_TRI: L.tad_test_res_symdir_tri_ops(), "tet": L.tad_test_res_symdir_tet_ops(),
            (configs.R_PIN, 2): L.tad_test_res_pin_ops(2), (configs.R_PIN, 3): L.tad_test_res_pin_ops(3), "_lib": L}


def gpu_vector_function(P, res_ops, ctx, precision="fp64"):
    fv = tad.VectorFunction(P.d, P.n_handles, ctx, precision)
    for t in P.terms:
        key = t.energy if t.energy == configs.R_SYMDIR_TRI else (t.energy, P.d)
        fv.add_elements_device(res_ops[key], t.conn, t.params)
    return fv


def test_gauss_newton_test_reproduction(ctx, res_ops):
    """tests/GaussNewtonTest.cc:113-162 on the GPU: per iteration the sum of
    squares and g = 2 J^T r equal the scalar form's (1e-12 absolute, the
    reference's bar for double), r and J equal the oracle's, the direction
    with w = 1e-12 the oracle's direct solve; line search on the sum of
    squares; after 20 iterations f = 4 +- 0.1 and |g| <= 0.1."""
    import torch
    Pv = configs.gauss_newton_test_problem()
    fv = gpu_vector_function(Pv, res_ops, ctx)
    fs = tad.ScalarFunction.from_problem(configs.newton_test_problem(), ctx)
    fo = oracle.VectorFunction.from_problem(Pv)
    assert fv.n_outputs == fo.n_outputs == 36
    x = Pv.x0.copy()
    dev = torch.device("cuda", 0)
    for it in range(20):
        f_ref, g_ref = fs.eval_with_gradient(x)
        assert abs(fv.eval_sum_of_squares(x) - f_ref) <= 1e-12
        f, g, r, J = fv.eval_sum_of_squares_with_derivatives(x)
        assert abs(f - f_ref) <= 1e-12 and np.abs(g - g_ref).max() <= 1e-12, it
        ro, Jo = fo.eval_with_jacobian(x)
        assert nrel(r, ro) <= 1e-14 and nrel(fv.eval(x), ro) <= 1e-14
        assert np.array_equal(J.indptr, Jo.indptr) and np.array_equal(J.indices, Jo.indices)
        assert nrel(J.data, Jo.data) <= 1e-13
        d = fv.gauss_newton_direction(x, 1e-12, tol=1e-13)
        do = oracle.gauss_newton_direction(ro, Jo, 1e-12)
        assert nrel(d, do) <= 1e-8, (it, nrel(d, do))
        xt, dt, gt = (torch.tensor(a, device=dev) for a in (x, d, g))
        xn, fn_, s, trials = fv.line_search(xt, dt, f, gt)
        assert trials > 0
        x = xn.cpu().numpy()
    f, g, _, _ = fv.eval_sum_of_squares_with_derivatives(x)
    assert abs(f - 4.0) <= 0.1 and np.abs(g).max() <= 0.1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grid_residuals_match_oracle(ctx, res_ops, precision):
    """40 x 40 grid (3,042 triangles, 8 residuals each, plus 2 pins): r and
    J against the oracle (J is FP64 in both modes: 1e-12), f, g = 2 J^T r and
    the Gauss-Newton Hessian 2 J^T J against the oracle's (scipy product)
    over the scalar function's Hessian pattern; device inputs give the same
    bits as host inputs."""
    import torch
    P = configs.grid_residuals(40, seed=5)
    fv = gpu_vector_function(P, res_ops, ctx, precision)
    fo = oracle.VectorFunction.from_problem(P)
    tol = 1e-10 if precision == "fp64" else 2e-5
    ro, Jo = fo.eval_with_jacobian(P.x0)
    r, J = fv.eval_with_jacobian(P.x0)
    assert np.array_equal(J.indptr, Jo.indptr) and np.array_equal(J.indices, Jo.indices)
    assert nrel(r, ro) <= 1e-13 and nrel(J.data, Jo.data) <= 1e-12
    fref, gref = float(ro @ ro), 2.0 * (Jo.T @ ro)
    Href = (2.0 * (Jo.T @ Jo)).tocsr()
    Href.sort_indices()
    f, g, H = fv.eval_gauss_newton(P.x0)
    assert abs(f - fref) <= 1e-12 * abs(fref) and abs(f - fo.eval_sum_of_squares(P.x0)) <= 1e-12 * abs(fref)
    assert nrel(g, gref) <= tol
    assert np.array_equal(H.indptr, Href.indptr) and np.array_equal(H.indices, Href.indices)
    assert nrel(H.data, Href.data) <= tol
    xt = torch.tensor(P.x0, device="cuda:0")
    rd, Jd = fv.eval_with_jacobian(xt)
    assert np.array_equal(rd.cpu().numpy(), r) and np.array_equal(Jd.cpu().numpy(), J.data)
    fd, gd, Hd = fv.eval_gauss_newton(xt)
    assert fd == f and np.array_equal(gd.cpu().numpy(), g) and np.array_equal(Hd.cpu().numpy(), H.data)


def tet_residual_problem(m=4, seed=3):
    """twist_cube's energy as residuals: sqrt(w) per tet and per pin."""
    P = configs.twist_cube(m=m, seed=seed)
    tet, pin = P.terms
    pt = tet.params.copy()
    pt[:, 9] = np.sqrt(pt[:, 9])
    pp = pin.params.copy()
    pp[:, 3] = np.sqrt(pp[:, 3])
    return P, [configs.Term("tet", tet.conn, pt), configs.Term(configs.R_PIN, pin.conn, pp)]


def test_tet_residuals_gauss_newton(ctx, res_ops):
    """18 residuals per tet (K = 12, the hyper-dual passes with two second
    directions): the sum of squares is the catalog's symmetric Dirichlet +
    penalty energy (1e-12) and g = 2 J^T r its gradient (1e-10); J matches
    central differences of r on random directions; J^T J is positive
    semi-definite; the direction equals a direct solve of the Gauss-Newton
    system; 8 Gauss-Newton steps decrease f with Armijo steps."""
    import torch
    P, terms = tet_residual_problem()
    fv = tad.VectorFunction(3, P.n_handles, ctx)
    fv.add_elements_device(res_ops["tet"], terms[0].conn, terms[0].params)
    fv.add_elements_device(res_ops[(configs.R_PIN, 3)], terms[1].conn, terms[1].params)
    fs = tad.ScalarFunction.from_problem(P, ctx)
    x = P.x0.copy()
    f_s, g_s = fs.eval_with_gradient(x)
    f_v, g_v = fv.eval_with_gradient(x)
    assert abs(f_v - f_s) <= 1e-12 * abs(f_s) and nrel(g_v, g_s) <= 1e-10
    r, J = fv.eval_with_jacobian(x)
    rng = np.random.default_rng(2)
    for _ in range(4):
        v = rng.standard_normal(x.size)
        h = 1e-6
        fd = (fv.eval(x + h * v) - fv.eval(x - h * v)) / (2 * h)
        assert nrel(J @ v, fd) <= 1e-7
    _, _, H = fv.eval_gauss_newton(x)
    Hd = H.toarray()
    assert np.abs(Hd - Hd.T).max() <= 1e-12 * np.abs(Hd).max()
    assert np.linalg.eigvalsh(Hd).min() >= -1e-10 * np.abs(Hd).max()
    assert nrel(H.data, (2.0 * (J.T @ J)).tocsr().sorted_indices().data) <= 1e-12
    # the direction against a direct solve of (w I + J^T J) d = -J^T r
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    d = fv.gauss_newton_direction(x, 1e-12, tol=1e-12)
    d_ref = spla.spsolve((J.T @ J + 1e-12 * sp.identity(x.size)).tocsc(), -(J.T @ r))
    assert nrel(d, d_ref) <= 1e-8
    # Gauss-Newton steps (tad_newton_step on a vector function): Armijo steps,
    # each starting at the energy the previous one accepted
    xt = torch.tensor(x, device="cuda:0")
    f_prev = f_v
    opts = tad.newton_default_opts(solve_tol=1e-10)
    for _ in range(8):
        st = fv.newton_step(xt, opts)
        assert st["solve_converged"] and not st["ls_failed"]
        assert st["f"] == f_prev and st["f_new"] < st["f"]
        f_prev = st["f_new"]


def test_vector_errors_and_inverted(ctx, res_ops):
    """An inverted triangle gives +inf residuals with finite (zero) Jacobian
    rows and f = +inf; its gradient 2 J^T r = inf * 0 is not finite and
    fails like any non-finite derivative; scalar and residual elements do
    not mix; the residual entry points reject scalar functions."""
    Pv = configs.gauss_newton_test_problem()
    fv = gpu_vector_function(Pv, res_ops, ctx)
    x = Pv.x0.copy()
    x[2:4], x[4:6] = Pv.x0[4:6], Pv.x0[2:4]
    r, J = fv.eval_with_jacobian(x)
    assert np.all(np.isinf(r[:8])) and np.all(np.isfinite(J.data))
    assert np.all(J.data[: 8 * 6] == 0.0)
    assert fv.eval_sum_of_squares(x) == float("inf")
    with pytest.raises(tad.TadError) as ei:
        fv.eval_sum_of_squares_with_derivatives(x)
    assert ei.value.status == 3
    assert np.isfinite(fv.eval_sum_of_squares(Pv.x0))
    with pytest.raises(tad.TadError):
        fv.add_elements(3, tad.SYMDIR_TRI, Pv.terms[0].conn, Pv.terms[0].params)
    fs = tad.ScalarFunction(2, 6, ctx)
    assert tad.lib().tad_fn_add_residuals_device(fs._h, res_ops[configs.R_SYMDIR_TRI], 0, None, None) == 1
    r = np.zeros(1)
    assert tad.lib().tad_fn_eval_residuals(fs._h, Pv.x0.ctypes.data, r.ctypes.data, 0) == 1
