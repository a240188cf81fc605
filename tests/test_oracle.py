"""Pins the CPU oracle (oracle/tad_oracle.cpp) before it is trusted as the
parity reference: reference known answers, closed forms, invariants, finite
differences, torch autograd, LAPACK eigendecomposition and scipy assembly."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from tinyad_amd import configs


def load_bunny():
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bunny_closed.json")))
    return np.array(d["V"], dtype=np.float64), np.array(d["F"], dtype=np.int32)


def reference_laplace(n, F):
    """tests/DynamicElementsTest.cc:31-52: L(v1,v1) += 1, L(v1,v2) -= 1 per directed face edge."""
    L = sp.lil_matrix((n, n))
    for f in F:
        for i in range(3):
            v1, v2 = f[i], f[(i + 1) % 3]
            L[v1, v1] += 1.0
            L[v1, v2] -= 1.0
    return L.tocsr()


def quad_params(A, b, c):
    return np.concatenate([np.asarray(A, float).ravel(), np.asarray(b, float).ravel(), [float(c)]])


# ---------------------------------------------------------------------------
# Dual operators (ScalarTestBinaryOperators.cc / ScalarTestUnaryOperators.cc)
# ---------------------------------------------------------------------------

def _grad_hess_complex_step(f, x, y):
    """Gradient by complex step (exact to rounding), Hessian by central
    differences of the complex-step gradient."""
    hc = 1e-30

    def grad(xv, yv):
        return np.array([f(xv + 1j * hc, yv).imag / hc, f(xv, yv + 1j * hc).imag / hc])

    h = 1e-5
    g = grad(x, y)
    H = np.stack([(grad(x + h, y) - grad(x - h, y)) / (2 * h), (grad(x, y + h) - grad(x, y - h)) / (2 * h)])
    return f(x, y).real, g, H


def test_dual_div2d_known_answer():
    f, g, H = oracle.dual_known_answer(0, -1.0, -0.5)  # ScalarTestBinaryOperators.cc:623-642
    assert f == -2.0
    assert np.array_equal(g, [4.0, -4.0])
    assert np.array_equal(H, [[-4.0, 8.0], [8.0, -16.0]])


@pytest.mark.parametrize("x,y", [(5.0, 0.0), (1.0, 1.0), (0.0, 5.0), (-1.0, 1.0), (-5.0, 0.0), (-1.0, -1.0),
                                 (0.0, -5.0), (1.0, -1.0)])
def test_dual_div2d_2_against_complex_step(x, y):
    # the eight points of ScalarTestBinaryOperators.cc:691-698
    def fun(a, b):
        num = 0.5 * a * a - b * b + 2.0 * a - b
        den = -(a - 2.0) ** 2 - (b - 3.0) ** 2 + 1.0
        return num / den
    f, g, H = oracle.dual_known_answer(1, x, y)
    fr, gr, Hr = _grad_hess_complex_step(fun, x, y)
    assert abs(f - fr) <= 1e-14 * max(1, abs(fr))
    assert np.allclose(g, gr, rtol=1e-12, atol=1e-13)
    assert np.allclose(H, Hr, rtol=1e-6, atol=1e-7)
    assert abs(H[0, 1] - H[1, 0]) <= 1e-15 * max(1.0, abs(H[0, 1]))  # TINYAD_ASSERT_SYMMETRIC


def test_dual_plus_minus_mult_div_known_answer():
    f, g, H = oracle.dual_known_answer(2, 1.0, 1.5)  # ScalarTestBinaryOperators.cc:711-727
    assert abs(f - 3.0) < 1e-12 and abs(g[0] - 4.5) < 1e-12 and abs(H[0, 0] - 3.0) < 1e-12
    assert abs(H[0, 1] - H[1, 0]) <= 1e-15 * max(1.0, abs(H[0, 1]))  # TINYAD_ASSERT_SYMMETRIC


def test_dual_linear_mix():
    # -(x y) + x - y + 6 x - 1: g = (-y + 7, -x - 1), H = [[0, -1], [-1, 0]]
    f, g, H = oracle.dual_known_answer(3, 0.7, -1.3)
    assert abs(f - (-(0.7 * -1.3) + 0.7 + 1.3 + 4.2 - 1.0)) < 1e-14
    assert np.allclose(g, [1.3 + 7.0, -0.7 - 1.0], atol=1e-14)
    assert np.array_equal(H, [[0.0, -1.0], [-1.0, 0.0]])


# ---------------------------------------------------------------------------
# ScalarFunction known answers (tests/ScalarFunctionTest.cc)
# ---------------------------------------------------------------------------

def test_basic_known_answer():
    fo = oracle.Function(2, 1)
    fo.add_elements(4, [[0]], quad_params(2 * np.eye(2), [0, 0], 0))
    x = np.array([1.0, 2.0])
    assert fo.eval(x) == 5.0
    f, g = fo.eval_with_gradient(x)
    assert f == 5.0 and np.array_equal(g, [2.0, 4.0])
    f, g, H = fo.eval_with_derivatives(x)
    assert H[0, 0] == 2.0 and H[1, 1] == 2.0


def test_1d_known_answer():
    fo = oracle.Function(1, 1)
    fo.add_elements(4, [[0]], quad_params([[4.0]], [1.0], 1.0))
    f, g, H = fo.eval_with_hessian_proj(np.array([1.0]))
    assert (f, g[0], H[0, 0]) == (4.0, 5.0, 4.0)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_2d_known_answer(sign):
    fo = oracle.Function(2, 1)
    A = sign * np.array([[4.0, 2.0], [2.0, 2.0]])
    fo.add_elements(4, [[0]], quad_params(A, sign * np.array([1.0, 0.0]), sign))
    x = np.array([1.0, 2.0])
    f, g, H = fo.eval_with_derivatives(x)
    f2, g2, Hp = fo.eval_with_hessian_proj(x)
    assert f == sign * 12 and f2 == sign * 12
    assert np.array_equal(g, sign * np.array([9.0, 6.0]))
    assert np.array_equal(H.toarray(), A)
    if sign > 0:
        assert np.array_equal(Hp.toarray(), A)
    else:
        assert np.linalg.eigvalsh(Hp.toarray()).min() > 0


# ---------------------------------------------------------------------------
# Newton test (tests/NewtonTest.cc:12-111) with a direct solve as SimplicialLDLT
# ---------------------------------------------------------------------------

def newton_oracle(iters=10):
    P = configs.newton_test_problem()
    fo = oracle.Function.from_problem(P)
    x = P.x0.copy()
    for _ in range(iters):
        f, g, H = fo.eval_with_hessian_proj(x)
        d = spla.spsolve(H.tocsc(), -g)
        x, _, _, _ = fo.line_search(x, d, f, g)
    return fo, x


def test_newton_known_answer():
    fo, x = newton_oracle()
    f, g, H = fo.eval_with_hessian_proj(x)
    assert H.nnz == 4 * 6 + 8 * (6 + 4 - 1)
    assert fo.eval(x) == f
    assert abs(f - 4.0) <= 1e-15
    assert np.abs(g).max() <= 1e-15


def test_newton_deterministic():
    xs = [newton_oracle()[1] for _ in range(3)]
    assert all(np.array_equal(x, xs[0]) for x in xs)


# ---------------------------------------------------------------------------
# Energies: closed forms, invariances, finite differences, torch autograd
# ---------------------------------------------------------------------------

def tet_problem(m=2, **kw):
    P = configs.twist_cube(m=m, **kw)
    P.terms = P.terms[:1]
    return P


def test_tet_energy_at_rest_and_uniform_scale():
    P = tet_problem()
    fo = oracle.Function.from_problem(P)
    rest = P.info["rest"]
    f, g = fo.eval_with_gradient(rest)
    assert abs(f - 6.0) <= 1e-12          # ||I||^2 + ||I||^2 = 6 per unit weight
    assert np.abs(g).max() <= 1e-12
    s = 1.7
    assert abs(fo.eval(s * rest) - 3.0 * (s * s + 1.0 / (s * s))) <= 1e-12
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    xr = (rest.reshape(-1, 3) @ R.T).ravel()
    assert abs(fo.eval(xr) - 6.0) <= 1e-12


def test_tri_energy_isometry():
    V_rest, _, F, _, _ = configs.planar_test_mesh()
    fo = oracle.Function(2, 6)
    fo.add_elements(1, F, configs.planar_rest_params(V_rest, F, 0.25))
    assert abs(fo.eval(V_rest.ravel()) - 4.0) <= 1e-14
    s = 0.5
    assert abs(fo.eval(s * V_rest.ravel()) - 2.0 * (s * s + 1 / (s * s))) <= 1e-13


def perturbed(P, sigma, seed):
    rest = P.info["rest"]
    return rest + sigma * P.info["h"] * configs.normal(seed, rest.size, stream=5)


def test_tet_element_hessian_invariants_and_fd():
    P = tet_problem()
    fo = oracle.Function.from_problem(P)
    x = perturbed(P, 0.2, 3)
    T = P.terms[0]
    for e in [0, 5, 17, 33]:
        val, g, H, idx = fo.element(0, e, x, 12)
        assert np.array_equal(idx, (3 * T.conn[e][:, None] + np.arange(3)).ravel())
        assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
        for c in range(3):  # translations are in the kernel
            t = np.zeros(12)
            t[c::3] = 1.0
            assert np.abs(H @ t).max() <= 1e-10 * np.abs(H).max()
            assert abs(g @ t) <= 1e-10 * np.abs(g).max()
        # central differences of value and gradient
        h = 1e-6
        for k in range(12):
            xp = x.copy(); xp[idx[k]] += h
            xm = x.copy(); xm[idx[k]] -= h
            vp, gp, _, _ = fo.element(0, e, xp, 12)
            vm, gm, _, _ = fo.element(0, e, xm, 12)
            assert abs((vp - vm) / (2 * h) - g[k]) <= 1e-6 * max(1.0, np.abs(g).max())
            assert np.abs((gp - gm) / (2 * h) - H[:, k]).max() <= 1e-5 * max(1.0, np.abs(H).max())


def test_tet_element_matches_torch_hessian():
    torch = pytest.importorskip("torch")
    P = tet_problem()
    fo = oracle.Function.from_problem(P)
    x = perturbed(P, 0.25, 8)
    T = P.terms[0]
    for e in [1, 9, 40]:
        p = torch.tensor(T.params[e], dtype=torch.float64)

        def energy(xe):
            X = xe.reshape(4, 3)
            M = torch.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], dim=1)
            J = M @ p[:9].reshape(3, 3)
            return p[9] * ((J * J).sum() + (torch.linalg.inv(J) ** 2).sum())

        xe = torch.tensor(x.reshape(-1, 3)[T.conn[e]].ravel(), dtype=torch.float64)
        Ht = torch.func.hessian(energy)(xe).numpy()
        gt = torch.func.grad(energy)(xe).numpy()
        val, g, H, _ = fo.element(0, e, x, 12)
        assert abs(val - float(energy(xe))) <= 1e-13 * abs(val)
        assert np.abs(g - gt).max() <= 1e-11 * np.abs(gt).max()
        assert np.abs(H - Ht).max() <= 1e-11 * np.abs(Ht).max()


def test_tri_element_matches_torch_hessian():
    torch = pytest.importorskip("torch")
    P = configs.grid_param(8, seed=2)
    fo = oracle.Function.from_problem(P)
    T = P.terms[0]
    for e in [0, 13, 50]:
        p = torch.tensor(T.params[e], dtype=torch.float64)

        def energy(xe):
            X = xe.reshape(3, 2)
            M = torch.stack([X[1] - X[0], X[2] - X[0]], dim=1)
            J = M @ p[:4].reshape(2, 2)
            return p[4] * ((J * J).sum() + (torch.linalg.inv(J) ** 2).sum())

        xe = torch.tensor(P.x0.reshape(-1, 2)[T.conn[e]].ravel(), dtype=torch.float64)
        Ht = torch.func.hessian(energy)(xe).numpy()
        val, g, H, _ = fo.element(0, e, P.x0, 6)
        assert np.abs(H - Ht).max() <= 1e-11 * np.abs(Ht).max()


def test_inverted_element_semantics():
    # +inf energy, zero gradient, projected Hessian eps I (NewtonTest.cc:40-41 + HessianProjection.hh:84-98)
    P = tet_problem(m=1)
    fo = oracle.Function.from_problem(P)
    X = P.info["rest"].reshape(-1, 3).copy()
    T = P.terms[0].conn
    X[[T[0, 1], T[0, 2]]] = X[[T[0, 2], T[0, 1]]]
    val, g, H, _ = fo.element(0, 0, X.ravel(), 12, project=True, eps=1e-9)
    assert val == np.inf and not g.any()
    assert np.array_equal(H, 1e-9 * np.eye(12))
    val, g, H, _ = fo.element(0, 0, X.ravel(), 12, project=True, eps=-1.0)
    assert not H.any()


# ---------------------------------------------------------------------------
# Projection against LAPACK
# ---------------------------------------------------------------------------

def spectral_projection(A, eps):
    lam, V = np.linalg.eigh(A)
    lam = np.abs(lam) if eps < 0 else np.maximum(lam, eps)
    return (V * lam) @ V.T


@pytest.mark.parametrize("case", ["indefinite", "clustered_zero", "pd", "pdd"])
@pytest.mark.parametrize("eps", [1e-9, -1.0, 0.3])
def test_projection_matches_eigh(case, eps):
    rng = np.random.default_rng(hash(case) % 2**32)
    k = 6
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    if case == "indefinite":
        lam = np.array([-2.0, -0.5, 0.1, 0.7, 1.5, 3.0])
    elif case == "clustered_zero":
        lam = np.array([0.0, 0.0, 0.0, 1.0, 2.0, -1.0])
    elif case == "pd":
        lam = np.array([0.5, 0.8, 1.0, 1.5, 2.0, 4.0])
    else:
        lam = None
    A = (Q * lam) @ Q.T if lam is not None else np.eye(k) * 5 + 0.1 * (lambda B: B + B.T)(rng.standard_normal((k, k)))
    A = 0.5 * (A + A.T)
    fo = oracle.Function(2, 3)
    fo.add_elements(4, [[0, 1, 2]], quad_params(A, np.zeros(k), 0.0))
    _, _, Hp, _ = fo.element(0, 0, np.zeros(6), k, project=True, eps=eps)
    ref = spectral_projection(A, eps)
    assert np.abs(Hp - ref).max() <= 1e-12 * np.abs(ref).max()
    if np.linalg.eigvalsh(A).min() >= (eps if eps >= 0 else 0.0):
        assert np.array_equal(Hp, A)  # unchanged bits when nothing is clamped


@pytest.mark.parametrize("eps", [1e-9, -1.0])
def test_tet_projection_matches_eigh(eps):
    """12 x 12 tet Hessians (three null directions, mixed signs) at strongly
    perturbed positions: the projected element Hessian is the LAPACK spectral
    projection of the unprojected one."""
    P = tet_problem()
    fo = oracle.Function.from_problem(P)
    x = perturbed(P, 0.35, 9)
    clamped = 0
    for e in range(0, 48, 3):
        val, _, H, _ = fo.element(0, e, x, 12)
        if not np.isfinite(val):
            continue
        _, _, Hp, _ = fo.element(0, e, x, 12, project=True, eps=eps)
        ref = spectral_projection(0.5 * (H + H.T), eps)
        assert np.abs(Hp - ref).max() <= 1e-12 * np.abs(ref).max(), e
        clamped += np.linalg.eigvalsh(H).min() < -1e-9 * np.abs(H).max()
    assert clamped >= 3  # indefinite elements were among the cases


# ---------------------------------------------------------------------------
# Assembly
# ---------------------------------------------------------------------------

def test_assembly_matches_scipy_coo():
    P = configs.twist_cube(m=2, seed=4)
    fo = oracle.Function.from_problem(P)
    x = P.x0
    f, g, H = fo.eval_with_hessian_proj(x)
    rows, cols, vals = [], [], []
    gref = np.zeros_like(x)
    fsum = 0.0
    for ti, t in enumerate(P.terms):
        K = 3 * t.valence
        for e in range(t.conn.shape[0]):
            v, ge, He, idx = fo.element(ti, e, x, K, project=True, eps=1e-9)
            fsum += v
            np.add.at(gref, idx, ge)
            rows.append(np.repeat(idx, K)); cols.append(np.tile(idx, K)); vals.append(He.ravel())
    Href = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=H.shape).tocsr()
    Href.sum_duplicates()
    assert abs(f - fsum) <= 1e-13 * abs(f)
    assert np.abs(g - gref).max() <= 1e-13 * np.abs(gref).max()
    assert H.nnz == Href.nnz
    assert np.abs((H - Href).toarray()).max() <= 1e-13 * np.abs(Href.data).max()


def test_laplacian_known_answer():
    # Hessian of 1/4 sum (edge differences)^2 on the closed bunny = graph Laplacian
    # (DynamicElementsTest.cc:60-91): same nnz, Frobenius difference <= 1e-12
    V, F = load_bunny()
    n = V.shape[0]
    fo = oracle.Function(1, n)
    fo.add_elements(4, F, quad_params(0.5 * (3 * np.eye(3) - np.ones((3, 3))), np.zeros(3), 0.0))
    _, _, L = fo.eval_with_derivatives(np.zeros(n))
    Lr = reference_laplace(n, F)
    assert L.shape == Lr.shape and L.nnz == Lr.nnz
    assert sp.linalg.norm(L - Lr) <= 1e-12


# ---------------------------------------------------------------------------
# Solver utilities
# ---------------------------------------------------------------------------

def test_pcg_matches_spsolve():
    # pinned cube: the projected Hessian is well conditioned (a free
    # parametrization has eps-sized translation eigenvalues, where any two
    # solvers legitimately differ by rounding / eps)
    P = configs.twist_cube(m=4, seed=6)
    fo = oracle.Function.from_problem(P)
    f, g, H = fo.eval_with_hessian_proj(P.x0)
    ref = spla.spsolve(H.tocsc(), g)
    for block in (1, 3):
        x, it, rel = oracle.pcg(H, g, block=block, tol=1e-13)
        assert rel <= 1e-13 and it > 0
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()
    x, it, rel = oracle.pcg(H, g, w=0.5, block=3, tol=1e-13)
    ref = spla.spsolve((H + 0.5 * sp.identity(H.shape[0])).tocsc(), g)
    assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()


def test_line_search_armijo_and_failure():
    P = configs.grid_param(10, seed=1)
    fo = oracle.Function.from_problem(P)
    f, g, H = fo.eval_with_hessian_proj(P.x0)
    d = spla.spsolve(H.tocsc(), -g)
    xn, fn_, s, n = fo.line_search(P.x0, d, f, g)
    assert n > 0 and fn_ <= f + 1e-4 * s * (d @ g)
    assert fo.eval(xn) == fn_
    # an ascent direction never satisfies Armijo: x0 comes back after 64 trials
    xn, fn_, s, n = fo.line_search(P.x0, g, f, g)
    assert n == -64 and np.array_equal(xn, P.x0) and s == 0.0


# ---------------------------------------------------------------------------
# Rest-shape mixed products (adjoint reference)
# ---------------------------------------------------------------------------

def _perturbed_rest_problem(m=3, seed=21):
    P = configs.twist_cube(m=m)
    h = P.info["h"]
    theta = P.info["rest"] + 0.05 * h * configs.normal(seed, P.n_vars, stream=6)
    lam = configs.normal(seed + 1, P.n_vars, stream=7)
    return P, theta, lam


def _torch_tet_energy(x, theta, T, w_scale):
    import torch
    X = x.reshape(-1, 3)
    R = theta.reshape(-1, 3)
    M = torch.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], dim=2)
    Mr = torch.stack([R[T[:, k]] - R[T[:, 0]] for k in (1, 2, 3)], dim=2)
    J = M @ torch.linalg.inv(Mr)
    w = w_scale * torch.linalg.det(Mr) / 6.0
    return (w * ((J * J).sum((1, 2)) + (torch.linalg.inv(J) ** 2).sum((1, 2)))).sum()


def test_rest_mixed_vjp_matches_torch_double_backward():
    import torch
    P, theta, lam = _perturbed_rest_problem()
    T = P.terms[0].conn
    ref = oracle.tet_rest_mixed_vjp(T, P.x0, theta, lam, 0.7)
    xt = torch.tensor(P.x0, dtype=torch.float64, requires_grad=True)
    tt = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    E = _torch_tet_energy(xt, tt, torch.tensor(T, dtype=torch.long), 0.7)
    (gx,) = torch.autograd.grad(E, xt, create_graph=True)
    (v,) = torch.autograd.grad(gx @ torch.tensor(lam), tt)
    assert np.abs(ref - v.numpy()).max() <= 1e-12 * np.abs(ref).max()


def test_rest_mixed_vjp_matches_fd_of_gradient():
    """d/dtheta (grad_x E . lambda) by central differences of the oracle gradient."""
    from tinyad_amd.adjoint import tet_rest_params
    P, theta, lam = _perturbed_rest_problem()
    T = P.terms[0].conn
    ws = 0.7
    ref = oracle.tet_rest_mixed_vjp(T, P.x0, theta, lam, ws)

    def phi(th):
        fo = oracle.Function(3, P.n_handles)
        fo.add_elements(configs.E_SYMDIR_TET, T, tet_rest_params(th, T, ws))
        return fo.eval_with_gradient(P.x0)[1] @ lam

    eps = 1e-6
    for k in range(4):
        v = configs.normal(100 + k, P.n_vars, stream=8)
        v /= np.linalg.norm(v)
        fd = (phi(theta + eps * v) - phi(theta - eps * v)) / (2 * eps)
        assert abs(fd - ref @ v) <= 1e-7 * np.abs(ref).sum() * np.abs(v).max()


def test_rest_mixed_vjp_inverted_and_translation():
    P, theta, lam = _perturbed_rest_problem()
    T = P.terms[0].conn
    x = P.x0.copy().reshape(-1, 3)
    # flip one vertex through the opposite face plane: the elements it inverts add nothing
    v0 = T[0, 0]
    x[v0] = 2 * x[T[0, 1:]].mean(axis=0) - x[v0]
    x = x.ravel()
    X = x.reshape(-1, 3)
    det = np.linalg.det(np.stack([X[T[:, k]] - X[T[:, 0]] for k in (1, 2, 3)], axis=2))
    ok = det > 0
    assert not ok.all()
    full = oracle.tet_rest_mixed_vjp(T, x, theta, lam, 1.0)
    part = oracle.tet_rest_mixed_vjp(T[ok], x, theta, lam, 1.0)
    assert np.array_equal(full, part)
    # the energy depends on theta through edge vectors only: a rest translation is free
    assert np.abs(full.reshape(-1, 3).sum(axis=0)).max() <= 1e-12 * np.abs(full).sum()


# ---------------------------------------------------------------------------
# Vector functions and the Gauss-Newton direction (VectorFunction.hh,
# VectorObjectiveTerm.hh, GaussNewtonDirection.hh)
# ---------------------------------------------------------------------------

def armijo(f_of, x, d, f, g, s_max=1.0, shrink=0.8, max_iters=64, c=1e-4):
    """LineSearch.hh:26-65 on an energy callable (x0 on failure)."""
    dg = float(d @ g)
    s = s_max
    for i in range(max_iters):
        xn = x + s * d
        fn = f_of(xn)
        if fn <= f + c * s * dg:
            return xn
        s = 1.0 if (s > 1.0 and s * shrink < 1.0) else s * shrink
    return x


def test_vector_sum_of_squares_matches_scalar_along_gauss_newton():
    """tests/GaussNewtonTest.cc:113-162: along 20 Gauss-Newton iterations
    (w = 1e-12, Armijo on eval_sum_of_squares) the residual form's f and
    g = 2 J^T r equal the scalar form's (symmetric Dirichlet / F.rows() plus
    unit pins) to 1e-12; the iteration reaches f = 4 +- 0.1 with |g| <= 0.1."""
    Pv = configs.gauss_newton_test_problem()
    fv = oracle.VectorFunction.from_problem(Pv)
    fs = oracle.Function.from_problem(configs.newton_test_problem())
    assert fv.n_outputs == 4 * 8 + 2 * 2
    x = Pv.x0.copy()
    for _ in range(20):
        f_ref, g_ref, _ = fs.eval_with_hessian_proj(x)
        assert abs(fs.eval(x) - f_ref) <= 1e-12
        assert abs(fv.eval_sum_of_squares(x) - f_ref) <= 1e-12
        f, g, r, J = fv.eval_sum_of_squares_with_derivatives(x)
        assert abs(f - f_ref) <= 1e-12 and np.abs(g - g_ref).max() <= 1e-12
        assert np.abs(r - fv.eval(x)).max() <= 1e-15 * np.abs(r).max()
        d = oracle.gauss_newton_direction(r, J, 1e-12)
        x = armijo(fv.eval_sum_of_squares, x, d, f, g)
    f, g, _, _ = fv.eval_sum_of_squares_with_derivatives(x)
    assert abs(f - 4.0) <= 0.1 and np.abs(g).max() <= 0.1


def torch_grid_residuals(P, xt):
    """Independent restatement of R_SYMDIR_TRI and R_PIN in torch."""
    import torch
    tri, pin = P.terms
    F = torch.as_tensor(tri.conn.astype(np.int64))
    p = torch.as_tensor(tri.params)
    X = xt.reshape(-1, 2)
    a, b, c = X[F[:, 0]], X[F[:, 1]], X[F[:, 2]]
    M = torch.stack([b - a, c - a], dim=2)
    Ri = p[:, :4].reshape(-1, 2, 2)
    J = M @ Ri
    Ji = torch.linalg.inv(Ri) @ torch.linalg.inv(M)
    rt = p[:, 4:5] * torch.cat([J.reshape(-1, 4), Ji.reshape(-1, 4)], dim=1)
    rp = torch.as_tensor(pin.params[:, 2:3]) * (torch.as_tensor(pin.params[:, :2]) - X[torch.as_tensor(
        pin.conn[:, 0].astype(np.int64))])
    return torch.cat([rt.reshape(-1), rp.reshape(-1)])


def test_vector_jacobian_matches_torch():
    """r and J of a 7 x 7 grid against torch.func.jacfwd of an independent
    torch restatement (1e-12); J's pattern is each residual's element
    variables, columns sorted; J^T J has the scalar function's pattern."""
    import torch
    P = configs.grid_residuals(7, seed=4)
    fv = oracle.VectorFunction.from_problem(P)
    r, J = fv.eval_with_jacobian(P.x0)
    xt = torch.tensor(P.x0)
    rt = torch_grid_residuals(P, xt)
    Jt = torch.func.jacfwd(lambda z: torch_grid_residuals(P, z))(xt).numpy()
    assert np.abs(r - rt.numpy()).max() <= 1e-12 * np.abs(rt.numpy()).max()
    assert np.abs(J.toarray() - Jt).max() <= 1e-12 * np.abs(Jt).max()
    assert J.shape == (fv.n_outputs, P.n_vars) and J.has_sorted_indices
    assert np.all(np.diff(J.indptr)[: 8 * P.terms[0].conn.shape[0]] == 6)
    S = oracle.Function(2, P.n_handles)
    S.add_elements(configs.E_SYMDIR_TRI, P.terms[0].conn, P.terms[0].params)
    S.add_elements(configs.E_PENALTY, P.terms[1].conn, P.terms[1].params)
    H = S.eval_with_derivatives(P.x0)[2]
    JtJ = (J.T @ J).tocsr()
    JtJ.sort_indices()
    assert np.array_equal(JtJ.indptr, H.indptr) and np.array_equal(JtJ.indices, H.indices)


def test_vector_inverted_and_pins_only():
    """An inverted triangle makes all 8 residuals +inf (the reference returns
    Vector<T, 8>::Constant(INFINITY)); for pin residuals alone J = -s I on the
    pinned variables and one Gauss-Newton step (w = 0) lands on the targets."""
    P = configs.gauss_newton_test_problem()
    fv = oracle.VectorFunction.from_problem(P)
    x = P.x0.copy()
    x[2:4], x[4:6] = P.x0[4:6], P.x0[2:4]  # swap vertices 1 and 2: face 0 inverted
    r = fv.eval(x)
    assert np.all(np.isinf(r[:8])) and np.isinf(fv.eval_sum_of_squares(x))
    pins = oracle.VectorFunction(2, 3)
    tgt = np.array([[1.0, -2.0], [0.5, 0.25], [3.0, 4.0]])
    pins.add_elements(oracle.R_PIN, np.arange(3, dtype=np.int32)[:, None],
                      np.concatenate([tgt, np.full((3, 1), 2.0)], axis=1))
    x0 = np.zeros(6)
    r, J = pins.eval_with_jacobian(x0)
    assert np.array_equal(J.toarray(), -2.0 * np.eye(6))
    d = oracle.gauss_newton_direction(r, J, 0.0)
    assert np.abs(x0 + d - tgt.ravel()).max() <= 1e-15


def test_reduced_basis_direction_pins():
    """newton_direction_reduced_basis (NewtonDirection.hh:76-105) restated:
    the reference's SwitchSolverTest answer (H = 2I, g = 2, B = I gives
    (-1, -1), SwitchSolverTest.cc:20-36); B = I equals the full Newton
    direction with w; a selection basis equals the solve on the kept rows and
    columns with zeros elsewhere; a rescaled basis B D gives the same d (the
    subspace is the same); and d is the minimizer of the quadratic model over
    the subspace (its reduced gradient B^T (H d + g) vanishes)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    H = 2.0 * sp.identity(2, format="csr")
    g = np.array([2.0, 2.0])
    assert np.allclose(oracle.newton_direction_reduced_basis(g, H, sp.identity(2)), [-1.0, -1.0], atol=1e-14)
    rng = np.random.default_rng(4)
    n = 40
    A = sp.random(n, n, density=0.1, random_state=5)
    H = (A @ A.T + 0.5 * sp.identity(n)).tocsr()
    g = rng.standard_normal(n)
    w = 0.3
    full = spla.spsolve((H + w * sp.identity(n)).tocsc(), -g)
    assert np.allclose(oracle.newton_direction_reduced_basis(g, H, sp.identity(n), w), full, rtol=1e-12, atol=1e-12)
    keep = np.sort(rng.choice(n, 25, replace=False))
    S = sp.csr_matrix((np.ones(keep.size), (keep, np.arange(keep.size))), shape=(n, keep.size))
    d = oracle.newton_direction_reduced_basis(g, H, S, w)
    Hk = H[keep][:, keep] + w * sp.identity(keep.size)
    ref = np.zeros(n)
    ref[keep] = spla.spsolve(Hk.tocsc(), -g[keep])
    assert np.allclose(d, ref, rtol=1e-12, atol=1e-12)
    B = sp.random(n, 12, density=0.15, random_state=6) + sp.eye(n, 12)
    d1 = oracle.newton_direction_reduced_basis(g, H, B, 0.0)
    D = sp.diags(rng.uniform(0.5, 2.0, 12))
    d2 = oracle.newton_direction_reduced_basis(g, H, B @ D, 0.0)
    assert np.allclose(d1, d2, rtol=1e-10, atol=1e-12)
    assert np.abs(B.T @ (H @ d1 + g)).max() <= 1e-10 * np.abs(B.T @ g).max()


# ---------------------------------------------------------------------------
# Repeated handles within an element (Element.hh:208-243)
# ---------------------------------------------------------------------------

def reference_projection(H, eps):
    """project_positive_definite (HessianProjection.hh:48-101) with numpy:
    the positive-diagonal-dominance early out, then V clamp(lambda) V^T
    unless nothing was clamped."""
    off = np.abs(H).sum(axis=1) - np.abs(np.diag(H))
    if not np.any(np.diag(H) < off + eps):
        return H
    lam, V = np.linalg.eigh(H)
    lc = np.maximum(lam, eps) if eps >= 0 else np.abs(lam)
    return H if np.array_equal(lc, lam) else (V * lc) @ V.T


def merge_matrix(conn, d):
    """S with y = S z: element coordinates y (one block per entry) from the
    element's local variables z (one block per distinct handle, in first-access
    order); returns S and the distinct handles in that order."""
    order = list(dict.fromkeys(int(h) for h in conn))
    S = np.zeros((len(conn) * d, len(order) * d))
    for a, h in enumerate(conn):
        j = order.index(int(h))
        S[a * d:(a + 1) * d, j * d:(j + 1) * d] = np.eye(d)
    return S, order


@pytest.mark.parametrize("eps", [1e-9, -1.0, 0.3])
def test_repeated_handles_merge_before_projection(eps):
    """An element that lists a handle twice gives that handle one set of local
    variables (first-access order), so for a quadratic element E(y) = y^T A y / 2
    + b^T y + c with y = S z its gradient is S^T (A y + b), its Hessian
    M = S^T A S, and the projection acts on M padded with zero rows and columns
    for the unused local variables (ScalarObjectiveTerm.hh:242-276); only the
    used ones are scattered.  Projecting A first and merging after gives a
    different matrix, which this case tells apart."""
    rng = np.random.default_rng(17)
    d, k = 2, 6
    Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
    A = (Q * np.array([-2.0, -0.7, 0.2, 0.9, 1.6, 3.1])) @ Q.T
    A = 0.5 * (A + A.T)
    b = rng.standard_normal(k)
    conn = [2, 0, 2]
    x = rng.standard_normal(4 * d)
    S, order = merge_matrix(conn, d)
    cols = np.concatenate([np.arange(h * d, (h + 1) * d) for h in order])
    y = np.concatenate([x[h * d:(h + 1) * d] for h in conn])
    M = S.T @ A @ S
    pad = np.zeros((k, k))
    pad[:M.shape[0], :M.shape[0]] = M
    P_ref = reference_projection(pad, eps)[:M.shape[0], :M.shape[0]]
    assert np.abs(S.T @ reference_projection(A, eps) @ S - P_ref).max() > 1e-3
    fo = oracle.Function(d, 4)
    fo.add_elements(4, [conn], quad_params(A, b, 0.25))
    f, g, H = fo.eval_with_derivatives(x)
    assert abs(f - (0.5 * y @ A @ y + b @ y + 0.25)) <= 1e-13
    g_ref = np.zeros(4 * d)
    g_ref[cols] = S.T @ (A @ y + b)
    assert np.abs(g - g_ref).max() <= 1e-13
    Hd = H.toarray()
    assert np.abs(Hd[np.ix_(cols, cols)] - M).max() <= 1e-13
    assert H.nnz == M.size  # only the used local variables are scattered
    _, gp, Hp = fo.eval_with_hessian_proj(x, eps)
    assert np.array_equal(gp, g)
    assert np.abs(Hp.toarray()[np.ix_(cols, cols)] - P_ref).max() <= 1e-12 * np.abs(P_ref).max()
    _, _, He, idx = fo.element(0, 0, x, k, project=True, eps=eps)
    assert list(idx) == list(cols) + [-1] * (k - cols.size)
    # the same function written as one valence-2 element over the distinct
    # handles: identical unprojected derivatives, equal projection (the
    # padded rows only add eigenvalues that are dropped)
    f2 = oracle.Function(d, 4)
    f2.add_elements(4, [order], quad_params(M, S.T @ b, 0.25))
    fb, gb, Hb = f2.eval_with_derivatives(x)
    assert abs(fb - f) <= 1e-13 and np.abs(gb - g).max() <= 1e-13
    assert np.abs(Hb.toarray() - Hd).max() <= 1e-13
    _, _, Hbp = f2.eval_with_hessian_proj(x, eps)
    assert np.abs(Hbp.toarray() - Hp.toarray()).max() <= 1e-12 * np.abs(P_ref).max()
