"""GPU adjoint through the Newton optimum x*(theta), theta = rest positions.

- rest-shape mixed products (d2E/dtheta dx)^T lambda against the oracle's
  Dual<24> truncated-block restatement: <= 1e-11 normwise;
- set_params against a function built from the same parameter rows: bitwise;
- dL/dtheta of L = ||x*(theta) - x_target||^2 (autograd.Function) against the
  oracle-computed gradient (oracle Hessian at x*, sparse LU solve, oracle
  mixed product): <= 1e-8 normwise (PCG residual 1e-13 times cond(H));
- the same gradient against central finite differences (eps = 1e-6) of L
  through full Newton re-solves, on random unit directions: relative error
  <= FD_TOL of |dL/dtheta . v| + 1e-9 |dL/dtheta|_2 (FD cancellation floor).
"""
import math
import time

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
import tinyad_amd as tad
from tinyad_amd import configs
from tinyad_amd.adjoint import RestShapeProblem, optimum, tet_rest_params

pytestmark = pytest.mark.gpu

FD_TOL = 1e-5


def nrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def cuda(a):
    import torch
    return torch.tensor(np.asarray(a, dtype=np.float64), device="cuda:0")


def perturbed(m=4, seed=31):
    P = configs.twist_cube(m=m)
    h = P.info["h"]
    theta = P.info["rest"] + 0.05 * h * configs.normal(seed, P.n_vars, stream=6)
    x = P.info["rest"] + 0.1 * h * configs.normal(seed + 1, P.n_vars, stream=7)
    lam = configs.normal(seed + 2, P.n_vars, stream=8)
    return P, theta, x, lam


def test_rest_vjp_matches_oracle(ctx):
    P, theta, x, lam = perturbed()
    T = P.terms[0].conn
    ref = oracle.tet_rest_mixed_vjp(T, x, theta, lam, 0.7)
    fn = tad.ScalarFunction.from_problem(P, ctx)
    host = fn.rest_vjp(0, x, theta, lam, 0.7)
    dev = fn.rest_vjp(0, cuda(x), cuda(theta), cuda(lam), 0.7).cpu().numpy()
    err = nrel(host, ref)
    print(f"rest_vjp vs oracle: {err:.3e}")
    assert err <= 1e-11
    assert np.array_equal(host, dev)
    # the mixed-precision function runs the same FP64 kernel on a side buffer
    fm = tad.ScalarFunction.from_problem(P, ctx, precision="mixed")
    assert np.array_equal(fm.rest_vjp(0, x, theta, lam, 0.7), host)
    with pytest.raises(tad.TadError):
        fn.rest_vjp(1, x, theta, lam, 0.7)  # penalty term: no rest shape


def test_rest_vjp_inverted_elements_add_nothing(ctx):
    P, theta, x, lam = perturbed()
    T = P.terms[0].conn
    X = x.reshape(-1, 3).copy()
    v0 = T[0, 0]
    X[v0] = 2 * X[T[0, 1:]].mean(axis=0) - X[v0]
    x = X.ravel()
    fn = tad.ScalarFunction.from_problem(P, ctx)
    out = fn.rest_vjp(0, x, theta, lam, 1.0)
    assert math.isinf(fn.eval(x))
    assert nrel(out, oracle.tet_rest_mixed_vjp(T, x, theta, lam, 1.0)) <= 1e-11


def test_set_params_equals_fresh_function(ctx):
    P, theta, x, _ = perturbed()
    T = P.terms[0].conn
    params = tet_rest_params(theta, T, 2.5)
    fa = tad.ScalarFunction.from_problem(P, ctx)
    fa.eval_with_hessian_proj(x)
    fa.set_params(0, params)
    fb = tad.ScalarFunction(3, P.n_handles, ctx)
    fb.add_elements(4, tad.SYMDIR_TET, T, params)
    fb.add_elements(1, tad.PENALTY, P.terms[1].conn, P.terms[1].params)
    a = fa.eval_with_hessian_proj(x)
    b = fb.eval_with_hessian_proj(x)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2].data, b[2].data)
    fa.set_params(0, cuda(params))
    assert fa.eval(x) == b[0]


def _adjoint_setup(ctx, m, seed=5):
    import torch
    P = configs.twist_cube(m=m)
    rp = RestShapeProblem(P, ctx=ctx)
    x = cuda(P.x0)
    t0 = time.time()
    hist = rp.minimize(x)
    fin = hist[-1]
    print(f"m={m}: {len(hist)} Newton iterations ({sum(h['phase'] == 'exact' for h in hist)} unprojected), "
          f"f={fin['f_new']:.15e}, |g|={fin['grad_norm']:.3e}, {time.time() - t0:.2f} s")
    h = P.info["h"]
    xt = x + cuda(0.02 * h * configs.normal(seed, P.n_vars, stream=10))
    theta0 = cuda(rp.theta0)
    return P, rp, x, xt, theta0, torch


def test_adjoint_matches_oracle_gradient(ctx):
    P, rp, xs, xt, theta0, torch = _adjoint_setup(ctx, m=6)
    theta = theta0.clone().requires_grad_(True)
    x = optimum(theta, rp, xs)
    L = ((x - xt) ** 2).sum()
    L.backward()
    g_gpu = theta.grad.cpu().numpy()
    print(f"adjoint solve: {rp.last_solve.iters} PCG iterations, rel. residual {rp.last_solve.rel_residual:.2e}")
    # oracle: unprojected Hessian at the same x*, direct solve, Dual<24> mixed block
    T = P.terms[0].conn
    xn = x.detach().cpu().numpy()
    fo = oracle.Function(3, P.n_handles)
    fo.add_elements(configs.E_SYMDIR_TET, T, tet_rest_params(rp.theta0, T, rp.w_scale))
    fo.add_elements(configs.E_PENALTY, P.terms[1].conn, P.terms[1].params)
    _, g_o, H_o = fo.eval_with_derivatives(xn)
    _, g_start = fo.eval_with_gradient(P.x0)
    assert np.linalg.norm(g_o) <= 1e-9 * np.linalg.norm(g_start)  # x* is stationary for the oracle too
    lam = spla.spsolve(H_o.tocsc(), 2.0 * (xn - xt.cpu().numpy()))
    ref = -oracle.tet_rest_mixed_vjp(T, xn, rp.theta0, lam, rp.w_scale)
    err = nrel(g_gpu, ref)
    print(f"dL/dtheta vs oracle: {err:.3e}")
    assert err <= 1e-8


def _fd_check(ctx, m, n_dirs, eps=1e-6):
    P, rp, xs, xt, theta0, torch = _adjoint_setup(ctx, m=m)
    theta = theta0.clone().requires_grad_(True)
    x = optimum(theta, rp, xs)
    ((x - xt) ** 2).sum().backward()
    g = theta.grad.cpu().numpy()

    def loss(th):
        xx = optimum(th, rp, xs)
        h = rp.last_history
        info.append((len(h), h[-1]["grad_norm"], h[-1]["step"]))
        return float(((xx - xt) ** 2).sum())

    info = []

    t0 = time.time()
    worst = 0.0
    for k in range(n_dirs):
        v = configs.normal(1000 + k, P.n_vars, stream=11)
        v /= np.linalg.norm(v)
        vt = cuda(v)
        fd = (loss(theta0 + eps * vt) - loss(theta0 - eps * vt)) / (2 * eps)
        ad = float(g @ v)
        err = abs(fd - ad) / (abs(ad) + 1e-9 * np.linalg.norm(g))
        worst = max(worst, err)
        assert err <= FD_TOL, (k, fd, ad, err, info[-2:])
    print(f"m={m}: {n_dirs} directions, worst relative FD error {worst:.3e}, {time.time() - t0:.1f} s")


def test_adjoint_matches_central_fd_small(ctx):
    _fd_check(ctx, m=6, n_dirs=8)


def test_adjoint_matches_central_fd_cube30(ctx):
    """30^3 x 6 = 162,000 tets, theta = 89,373 rest coordinates, 64 directions."""
    _fd_check(ctx, m=30, n_dirs=64)
