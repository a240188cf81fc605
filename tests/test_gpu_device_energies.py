"""Energies compiled as generic device functors (tad_fn_add_elements_device,
include/tinyad_amd/DeviceElements.hh), the form the reference gives every
energy: a generic lambda over the element's active scalars
(ScalarFunction.hh:72-78, tests/NewtonTest.cc:28-44).

The functors live in tests/cpp/device_energies.hip, built into
tests/cpp/bin/libtad_test_device_energies.so.  Pins:
  * symmetric Dirichlet (triangles, tets) against the oracle and against the
    library's built-in catalog at the FP64 bar (1e-10 normwise);
  * a dense quadratic against the catalog's quadratic element;
  * an energy using sqrt / exp / log / sin / cos / pow against central
    differences of its own passive energy, and its projected Hessian against
    numpy's eigen-projection of its Hessian (HessianProjection.hh:48-101);
  * the same helpers on a K = 8 element (hyper-dual passes) in FP64 and with
    FP32 derivatives;
  * powers at base 0 with double and int exponents (Scalar.hh:266-306) on
    both derivative paths, against their closed forms;
  * mixed precision within the tolerance of the catalog's mixed mode, also
    for meshes translated far from the origin;
  * a Newton trajectory equal to the catalog's to the converged-energy bar;
  * elements that list a handle more than once (merged into one set of
    variables before the projection, Element.hh:208-243) against the oracle,
    closed forms and an explicit merge of distinct-handle derivatives.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
import tinyad_amd as tad
from tinyad_amd import configs, _lib

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
PLUGIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "libtad_test_device_energies.so")


def nrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300))


@pytest.fixture(scope="module")
def plugin():
    if not os.path.exists(PLUGIN):
        raise RuntimeError(f"{PLUGIN} is missing: run python -m tinyad_amd.build first")
    L = ctypes.CDLL(PLUGIN)
    ops = {}
    for name in ["symdir_tri", "symdir_tet", "symdir_tri_inv", "symdir_tet_inv", "false_invariant", "quadratic4",
                 "smooth4", "smooth8"]:
        f = getattr(L, f"tad_test_{name}_ops")
        f.restype = ctypes.c_void_p
        ops[name] = f()
    ops["_lib"] = L
    return ops


def device_function(P, plugin, ctx, precision="fp64", suffix=""):
    """P with its symmetric-Dirichlet terms as device functors, other terms from
    the catalog; suffix "_inv": the functors that declare translation invariance."""
    fn = tad.ScalarFunction(P.d, P.n_handles, ctx, precision)
    for t in P.terms:
        if t.energy == tad.SYMDIR_TRI:
            fn.add_elements_device(plugin["symdir_tri" + suffix], t.conn, t.params)
        elif t.energy == tad.SYMDIR_TET:
            fn.add_elements_device(plugin["symdir_tet" + suffix], t.conn, t.params)
        else:
            fn.add_elements(t.valence, t.energy, t.conn, t.params)
    return fn


def perturbed_cube(m, sigma, seed):
    P = configs.twist_cube(m=m, seed=seed)
    rest = P.info["rest"]
    P.x0 = rest + sigma * P.info["h"] * configs.normal(seed + 1, rest.size, stream=9)
    return P


CASES = {
    "grid_param_12": lambda: configs.grid_param(12, seed=5),
    "newton_test": configs.newton_test_problem,
    "twist_cube_3": lambda: configs.twist_cube(m=3, seed=5),
    "cube_perturbed": lambda: perturbed_cube(3, 0.15, 13),
}


@pytest.mark.parametrize("suffix", ["", "_inv"])
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("eps", [1e-9, -1.0])
def test_device_symdir_matches_oracle_and_catalog(ctx, plugin, name, eps, suffix):
    """Projected through the whole K x K eigensolve, or (suffix _inv, the
    functors declaring translation invariance) through the deflated one."""
    P = CASES[name]()
    fo = oracle.Function.from_problem(P)
    fc = tad.ScalarFunction.from_problem(P, ctx)
    fd = device_function(P, plugin, ctx, suffix=suffix)
    f1, g1, H1 = fo.eval_with_hessian_proj(P.x0, eps)
    f2, g2, H2 = fd.eval_with_hessian_proj(P.x0, eps)
    f3, g3, H3 = fc.eval_with_hessian_proj(P.x0, eps)
    assert np.array_equal(H2.indptr, H3.indptr) and np.array_equal(H2.indices, H3.indices)
    assert abs(f2 - f1) <= 1e-12 * abs(f1)
    assert nrel(g2, g1) <= FP64_TOL and nrel(H2.toarray(), H1.toarray()) <= FP64_TOL
    assert nrel(g2, g3) <= FP64_TOL and nrel(H2.data, H3.data) <= FP64_TOL
    assert abs(fd.eval(P.x0) - f2) <= 1e-14 * abs(f2)
    fa, ga = fd.eval_with_gradient(P.x0)
    fb, gb, Hb = fd.eval_with_derivatives(P.x0)
    fo2, go2, Ho2 = fo.eval_with_derivatives(P.x0)
    assert fa == f2 and fb == f2
    assert nrel(ga, g1) <= FP64_TOL and nrel(gb, go2) <= FP64_TOL
    assert nrel(Hb.toarray(), Ho2.toarray()) <= FP64_TOL


@pytest.mark.parametrize("suffix", ["", "_inv"])
def test_device_inverted_element(ctx, plugin, suffix):
    """+inf energy, zero gradient and eps I for an inverted tet (NewtonTest.cc:40-41)."""
    P = configs.twist_cube(m=2, seed=3)
    X = P.x0.reshape(-1, 3).copy()
    a, b = P.terms[0].conn[0, 1], P.terms[0].conn[0, 2]
    X[[a, b]] = X[[b, a]]
    x = X.ravel()
    fo = oracle.Function.from_problem(P)
    fd = device_function(P, plugin, ctx, suffix=suffix)
    f1, g1, H1 = fo.eval_with_hessian_proj(x)
    f2, g2, H2 = fd.eval_with_hessian_proj(x)
    assert f1 == np.inf and f2 == np.inf and fd.eval(x) == np.inf
    assert nrel(g2, g1) <= FP64_TOL and nrel(H2.toarray(), H1.toarray()) <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_declared_invariance_is_checked(ctx, plugin, precision):
    """A functor that declares translation invariance but is not (|x_0|^2 +
    |x_1 - x_0|^2) is reported as non-finite by the projected evaluation, not
    projected on the wrong subspace; unprojected evaluations are unaffected."""
    fn = tad.ScalarFunction(2, 2, ctx, precision)
    fn.add_elements_device(plugin["false_invariant"], [[0, 1]], np.zeros(0))
    x = np.array([0.3, -0.2, 0.7, 0.4])
    f, g, H = fn.eval_with_derivatives(x)
    assert abs(f - (x[0] ** 2 + x[1] ** 2 + (x[2] - x[0]) ** 2 + (x[3] - x[1]) ** 2)) <= 1e-15
    with pytest.raises(tad.TadError, match="NONFINITE"):
        fn.eval_with_hessian_proj(x, 0.5)  # H = [4I -2I; -2I 2I] is not diagonally dominant by 0.5


def quad_problem(seed, n_handles=7, n_elem=12):
    rng = np.random.default_rng(seed)
    conn = np.array([rng.choice(n_handles, 2, replace=False) for _ in range(n_elem)], dtype=np.int32)
    A = rng.standard_normal((n_elem, 4, 4))  # not symmetric, indefinite
    b = rng.standard_normal((n_elem, 4))
    c = rng.standard_normal((n_elem, 1))
    params = np.concatenate([A.reshape(n_elem, 16), b, c], axis=1)
    x = rng.standard_normal(2 * n_handles)
    return conn, params, x


@pytest.mark.parametrize("eps", [1e-9, -1.0, 0.3])
def test_device_quadratic_matches_catalog(ctx, plugin, eps):
    conn, params, x = quad_problem(4)
    fc = tad.ScalarFunction(2, 7, ctx)
    fc.add_elements(2, tad.QUADRATIC, conn, params)
    fd = tad.ScalarFunction(2, 7, ctx)
    fd.add_elements_device(plugin["quadratic4"], conn, params)
    r1 = fc.eval_with_hessian_proj(x, eps)
    r2 = fd.eval_with_hessian_proj(x, eps)
    assert abs(r2[0] - r1[0]) <= 1e-13 * max(1.0, abs(r1[0]))
    assert nrel(r2[1], r1[1]) <= 1e-13
    assert nrel(r2[2].data, r1[2].data) <= 1e-12


def smooth_energy(x, p):
    """The passive energy of Smooth4 (tests/cpp/device_energies.hip) in numpy."""
    return (p[0] * np.exp(0.3 * x[0] * x[3]) + np.log(1 + x[1] ** 2) + np.sqrt(1 + x[0] ** 2 + x[2] ** 2)
            + np.sin(x[1]) * np.cos(x[2]) + p[1] * (1 + x[3] ** 2) ** 1.5 + 1 / (2 + x[0] ** 2))


def reference_projection(H, eps):
    """project_positive_definite (HessianProjection.hh:48-101) with numpy:
    unchanged when diagonally dominant with H_ii >= sum_j |H_ij| + eps (the
    early out, which for eps < 0 also accepts some indefinite matrices) or
    when no eigenvalue is clamped; else V clamp(lambda) V^T."""
    off = np.abs(H).sum(axis=1) - np.abs(np.diag(H))
    if not np.any(np.diag(H) < off + eps):
        return H
    lam, V = np.linalg.eigh(H)
    lc = np.maximum(lam, eps) if eps >= 0 else np.abs(lam)
    return H if np.array_equal(lc, lam) else (V * lc) @ V.T


def test_device_transcendental_derivatives(ctx, plugin):
    """One Smooth4 element: value against numpy, gradient and Hessian against
    central differences (steps 1e-5: error ~1e-10 relative), projection
    against numpy's eigendecomposition."""
    rng = np.random.default_rng(9)
    p = np.array([0.7, 0.4])
    fd = tad.ScalarFunction(2, 2, ctx)
    fd.add_elements_device(plugin["smooth4"], [[0, 1]], p)
    for trial in range(4):
        x = rng.uniform(-1.2, 1.2, 4)
        f, g, H = fd.eval_with_derivatives(x)
        assert abs(f - smooth_energy(x, p)) <= 1e-14 * abs(f)
        h = 1e-5
        gfd = np.array([(smooth_energy(x + h * e, p) - smooth_energy(x - h * e, p)) / (2 * h) for e in np.eye(4)])
        assert nrel(g, gfd) <= 1e-9
        Hd = H.toarray()
        Hfd = np.array([(fd.eval_with_gradient(x + h * e)[1] - fd.eval_with_gradient(x - h * e)[1]) / (2 * h)
                        for e in np.eye(4)])
        assert np.array_equal(Hd, Hd.T)
        assert nrel(Hd, Hfd) <= 1e-8
        for eps in [1e-9, -1.0, -0.01, 0.5]:
            _, _, Hp = fd.eval_with_hessian_proj(x, eps)
            assert nrel(Hp.toarray(), reference_projection(Hd, eps)) <= 1e-12, (trial, eps, np.linalg.eigvalsh(Hd))


def smooth8_energy(x, p):
    """The passive energy of Smooth8 (tests/cpp/device_energies.hip) in numpy."""
    return (p[0] * np.exp(0.3 * x[0] * x[7]) + np.log(1 + x[1] ** 2) + np.sqrt(1 + x[2] ** 2 + x[4] ** 2)
            + np.sin(x[3]) * np.cos(x[5]) + p[1] * (1 + x[6] ** 2) ** 1.5 + 1 / (2 + x[0] ** 2)
            + x[5] / (1.5 + x[7] ** 2) + (x[2] - x[6]) ** 3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_transcendental_derivatives_k8(ctx, plugin, precision):
    """One Smooth8 element (K = 8: hyper-dual passes, FP64 or FP32
    derivative parts): value against numpy (FP64 in both precisions),
    gradient against central differences of the numpy energy, Hessian
    against central differences of the FP64 gradient, projection against
    numpy's eigen-projection of the element's own Hessian.  Bars: FP64 1e-9
    (g) / 1e-8 (H) / 1e-12 (projection); FP32 derivatives 2e-5 normwise."""
    rng = np.random.default_rng(19)
    p = np.array([0.7, 0.4])
    fd = tad.ScalarFunction(2, 4, ctx, precision)
    fd.add_elements_device(plugin["smooth8"], [[0, 1, 2, 3]], p)
    f64 = tad.ScalarFunction(2, 4, ctx)
    f64.add_elements_device(plugin["smooth8"], [[0, 1, 2, 3]], p)
    tol_g, tol_h = (1e-9, 1e-8) if precision == "fp64" else (2e-5, 2e-5)
    for trial in range(4):
        x = rng.uniform(-1.2, 1.2, 8)
        f, g, H = fd.eval_with_derivatives(x)
        assert abs(f - smooth8_energy(x, p)) <= 1e-14 * abs(f)
        h = 1e-5
        gfd = np.array([(smooth8_energy(x + h * e, p) - smooth8_energy(x - h * e, p)) / (2 * h) for e in np.eye(8)])
        assert nrel(g, gfd) <= tol_g
        Hd = H.toarray()
        Hfd = np.array([(f64.eval_with_gradient(x + h * e)[1] - f64.eval_with_gradient(x - h * e)[1]) / (2 * h)
                        for e in np.eye(8)])
        assert np.array_equal(Hd, Hd.T)
        assert nrel(Hd, Hfd) <= tol_h
        for eps in [1e-9, -1.0, 0.5]:
            _, _, Hp = fd.eval_with_hessian_proj(x, eps)
            tol_p = 1e-12 if precision == "fp64" else 1e-5
            assert nrel(Hp.toarray(), reference_projection(Hd, eps)) <= tol_p, (trial, eps)


@pytest.mark.parametrize("handles", [3, 4])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_pow_base_zero(ctx, plugin, handles, precision):
    """pow at base 0 (Scalar.hh:266-306): x^2 and x^3 with double exponents
    and x^1, x^2, x^3, x^0 with int exponents stay finite and exact at x = 0
    (f = 1, g = e_2, H = diag(2, 0, 0, 2, 0, 0, 2, 2)), on the one-pass Dual
    path (3 handles, K = 6) and the hyper-dual path (4 handles, K = 8); at a
    random point they match the closed forms."""
    L = plugin["_lib"]
    L.tad_test_powers_ops.restype = ctypes.c_void_p
    L.tad_test_powers_ops.argtypes = [ctypes.c_int]
    K = 2 * handles
    fn = tad.ScalarFunction(2, handles, ctx, precision)
    fn.add_elements_device(L.tad_test_powers_ops(handles - 3), [list(range(handles))])
    f, g, H = fn.eval_with_derivatives(np.zeros(K))
    assert f == 1.0
    assert np.array_equal(g, np.eye(K)[2])
    assert np.array_equal(H.toarray(), np.diag([2.0, 0, 0, 2.0, 0, 0, 2.0, 2.0][:K]))
    fp, gp, Hp = fn.eval_with_hessian_proj(np.zeros(K))
    assert fp == 1.0 and np.array_equal(gp, g)
    assert np.allclose(Hp.toarray(), np.diag([2.0, 1e-9, 1e-9, 2.0, 1e-9, 1e-9, 2.0, 2.0][:K]), rtol=1e-6, atol=0)
    x = np.random.default_rng(handles).uniform(-1.5, 1.5, K)
    f, g, H = fn.eval_with_derivatives(x)
    fr = x[0] ** 2 + x[1] ** 3 + x[2] + x[3] ** 2 + x[4] ** 3 + 1.0 + (x[6:] ** 2).sum()
    gr = np.array([2 * x[0], 3 * x[1] ** 2, 1.0, 2 * x[3], 3 * x[4] ** 2, 0.0] + list(2 * x[6:]))
    Hr = np.diag([2.0, 6 * x[1], 0.0, 2.0, 6 * x[4], 0.0] + [2.0] * (K - 6))
    tol = 1e-14 if precision == "fp64" else 1e-6
    assert abs(f - fr) <= 1e-14 * abs(fr)
    assert nrel(g, gr) <= tol and nrel(H.toarray(), Hr) <= tol


def translated(P, offset):
    """P moved by `offset` in every coordinate: positions and penalty targets
    (element rest parameters are translation invariant)."""
    import copy
    Q = copy.deepcopy(P)
    Q.x0 = P.x0 + offset
    for t in Q.terms:
        if t.energy == tad.PENALTY:
            t.params = t.params.copy()
            t.params[:, :P.d] += offset
    return Q


@pytest.mark.parametrize("suffix", ["", "_inv"])
@pytest.mark.parametrize("kind", ["tets", "triangles"])
def test_device_mixed_precision_far_from_origin(ctx, plugin, kind, suffix):
    """Device functors with FP32 derivatives on a mesh translated by 1e3
    (|x| / h up to 4e3 for tets, 1.1e4 for triangles) against the FP64
    catalog: the functor's values stay FP64, so the mixed bar 2e-5 holds as
    at the origin.  The catalog's own mixed mode is checked the same way."""
    P = perturbed_cube(4, 0.1, 17) if kind == "tets" else configs.grid_param(12, seed=5)
    Q = translated(P, 1000.0)
    fc = tad.ScalarFunction.from_problem(Q, ctx)
    fa, ga, Ha = fc.eval_with_hessian_proj(Q.x0)
    for fm in (device_function(Q, plugin, ctx, precision="mixed", suffix=suffix),
               tad.ScalarFunction.from_problem(Q, ctx, precision="mixed")):
        fb, gb, Hb = fm.eval_with_hessian_proj(Q.x0)
        assert abs(fb - fa) <= 1e-12 * abs(fa)
        print(f"{kind} at 1e3: g {nrel(gb, ga):.2e} H {nrel(Hb.data, Ha.data):.2e}")
        assert nrel(gb, ga) <= 2e-5 and nrel(Hb.data, Ha.data) <= 2e-5


def test_device_mixed_precision(ctx, plugin):
    P = configs.grid_param(16, seed=2)
    f64 = device_function(P, plugin, ctx)
    fmx = device_function(P, plugin, ctx, precision="mixed")
    fa, ga, Ha = f64.eval_with_hessian_proj(P.x0)
    fb, gb, Hb = fmx.eval_with_hessian_proj(P.x0)
    assert fa == fb  # the value is the passive FP64 energy in both modes
    assert nrel(gb, ga) <= 2e-5 and nrel(Hb.data, Ha.data) <= 2e-5


def test_device_newton_matches_catalog(ctx, plugin):
    """20 projected-Newton iterations (the 2D parametrization workload, smaller
    grid) with the device functor and with the catalog reach the same energy."""
    import torch
    P = configs.grid_param(30, seed=1)
    fc = tad.ScalarFunction.from_problem(P, ctx)
    fd = device_function(P, plugin, ctx)
    xc = torch.tensor(P.x0, device="cuda:0")
    xd = xc.clone()
    for _ in range(20):
        sc = fc.newton_step(xc)
        sd = fd.newton_step(xd)
    fc_, fd_ = fc.eval(xc), fd.eval(xd)
    print(f"catalog f {fc_:.15g} device functor f {fd_:.15g}")
    assert abs(fd_ - fc_) <= 1e-8 * abs(fc_)
    assert sd["solve_converged"] and sc["solve_converged"]


def test_device_ops_validation(ctx, plugin):
    fn = tad.ScalarFunction(3, 5, ctx)
    with pytest.raises(tad.TadError):  # 2D energy on a 3D function
        fn.add_elements_device(plugin["symdir_tri"], [[0, 1, 2]], np.zeros(5))
    fn2 = tad.ScalarFunction(2, 5, ctx)
    with pytest.raises(tad.TadError):  # parameter row of the wrong length
        fn2.add_elements_device(plugin["symdir_tri"], [[0, 1, 2]], np.zeros(4))
    with pytest.raises(tad.TadError, match="twice"):  # degenerate catalog simplex
        fn2.add_elements(3, tad.SYMDIR_TRI, [[0, 1, 1]], np.zeros(5))
    bad = _lib.DeviceOps.from_buffer_copy(_lib.DeviceOps.from_address(plugin["symdir_tri"]))
    bad.version = 999
    with pytest.raises(tad.TadError, match="TAD_DEVICE_OPS_VERSION"):
        fn2.add_elements_device(ctypes.addressof(bad), [[0, 1, 2]], np.zeros(5))


# ---------------------------------------------------------------------------
# add_elements_dynamic (ScalarFunctionImpl.hh:102-214, DynamicElementsTest.cc)
# ---------------------------------------------------------------------------

def dyn_ops(plugin, name, count):
    f = getattr(plugin["_lib"], f"tad_test_{name}_ops")
    f.restype = ctypes.c_void_p
    f.argtypes = [ctypes.c_int]
    return [f(i) for i in range(count)]


def test_dynamic_elements_grouping(ctx, plugin):
    """DynamicElementsTest.cc:9-33: element e of 4 accesses handles 0..e-1 of a
    2D function; valences <3, 1> give two terms of two elements each.  With
    x = 1: f = sum_e ||sum_{v<e} x_v||^2 = 0 + 2 + 8 + 18, and the derivatives
    in closed form."""
    fn = tad.ScalarFunction(2, 4, ctx)
    fn.add_elements_dynamic_device(dyn_ops(plugin, "sum_squares", 2), [list(range(e)) for e in range(4)])
    assert fn.terms() == [(tad.DEVICE, 3, 2), (tad.DEVICE, 1, 2)]
    x = np.ones(8)
    f, g, H = fn.eval_with_derivatives(x)
    assert f == 28.0
    assert np.array_equal(g, np.repeat([12.0, 10.0, 6.0, 0.0], 2))
    Href = np.zeros((8, 8))
    for v in range(4):
        for w in range(4):
            Href[2 * v:2 * v + 2, 2 * w:2 * w + 2] = 2.0 * sum(1 for e in range(4) if e > max(v, w)) * np.eye(2)
    assert np.array_equal(H.toarray(), Href)
    assert H.indptr[-1] - H.indptr[-3] == 0  # handle 3 is used by no element: empty rows
    # projected: each element's K x K Hessian (K = group valence * 2) has zero
    # rows and columns for the slots it does not access, as in the reference,
    # and is projected as a whole; only the accessed blocks are kept
    fp, gp, Hp = fn.eval_with_hessian_proj(x)
    assert fp == f and np.array_equal(gp, g)
    Hpref = np.zeros((8, 8))
    for e in range(4):
        Vs = 3 if e > 1 else 1
        He = np.zeros((2 * Vs, 2 * Vs))
        He[:2 * e, :2 * e] = 2.0 * np.kron(np.ones((e, e)), np.eye(2))
        Pe = reference_projection(He, 1e-9)
        Hpref[:2 * e, :2 * e] += Pe[:2 * e, :2 * e]
    assert nrel(Hp.toarray(), Hpref) <= 1e-12


def bunny_one_ring():
    import json
    d = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bunny_closed.json")))
    V = np.array(d["V"], dtype=np.float64)
    F = np.array(d["F"], dtype=np.int32)
    nbrs = [[] for _ in range(V.shape[0])]
    for f in F:
        for i in range(3):
            nbrs[f[i]].append(int(f[(i + 1) % 3]))
    return V, F, nbrs


def reference_laplace(n, F):
    """DynamicElementsTest.cc:37-56: L(v1, v1) += 1, L(v1, v2) -= 1 per directed face edge."""
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for f in F:
        for i in range(3):
            v1, v2 = int(f[i]), int(f[(i + 1) % 3])
            rows += [v1, v1]
            cols += [v1, v2]
            vals += [1.0, -1.0]
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def test_dynamic_one_ring_laplacian(ctx, plugin):
    """DynamicElementsTest.cc:58-135: the Hessian of the one-ring Dirichlet
    energy with elements of valence 6, 7, 8 and 10 (the bunny's vertex
    valences plus one, padded up) equals the cotangent-free Laplacian of its
    faces; the pattern holds extra explicit zeros (neighbour-neighbour pairs),
    as the reference notes."""
    V, F, nbrs = bunny_one_ring()
    n = V.shape[0]
    counts = np.array([1 + len(nb) for nb in nbrs], dtype=np.float64)
    assert counts.max() <= 10
    fn = tad.ScalarFunction(1, n, ctx)
    fn.add_elements_dynamic_device(dyn_ops(plugin, "one_ring", 4), [[v] + nbrs[v] for v in range(n)],
                                   counts[:, None])
    got = {va: ne for _, va, ne in fn.terms()}
    hist = {}
    for c in counts.astype(int):
        g = min(va for va in (6, 7, 8, 10) if va >= c)
        hist[g] = hist.get(g, 0) + 1
    assert got == hist
    L = fn.eval_hessian_of_quadratic()
    Lref = reference_laplace(n, F)
    assert L.shape == Lref.shape and L.nnz >= Lref.nnz
    assert abs(L - Lref).max() <= 1e-12


def test_dynamic_newton_step(ctx, plugin):
    """A Newton step on the one-ring energy plus a pin (catalog penalty)
    reaches the minimum (every value at the pin's target): exercises the
    padded energy kernel of the line search and the solve."""
    import torch
    V, F, nbrs = bunny_one_ring()
    n = V.shape[0]
    counts = np.array([1 + len(nb) for nb in nbrs], dtype=np.float64)
    fn = tad.ScalarFunction(1, n, ctx)
    fn.add_elements_dynamic_device(dyn_ops(plugin, "one_ring", 4), [[v] + nbrs[v] for v in range(n)],
                                   counts[:, None])
    fn.add_elements(1, tad.PENALTY, [[0]], [0.3, 1.0])
    x = torch.tensor(np.random.default_rng(5).standard_normal(n), device="cuda:0")
    # unprojected Hessian (projection_eps = NaN): the exact Newton step of a quadratic
    st = fn.newton_step(x, tad.newton_default_opts(projection_eps=float("nan"), solve_tol=1e-13))
    assert st["step"] == 1.0 and st["ls_failed"] == 0 and st["f"] > 1.0
    print(f"one-ring Newton: f {st['f']:.6g} -> {st['f_new']:.3g}, max |x - 0.3| "
          f"{float((x - 0.3).abs().max()):.2e}")
    assert torch.allclose(x, torch.full_like(x, 0.3), atol=1e-6, rtol=0)
    assert fn.eval(x) <= 1e-10


def test_dynamic_validation(ctx, plugin):
    ops = dyn_ops(plugin, "sum_squares", 2)
    fn = tad.ScalarFunction(2, 6, ctx)
    with pytest.raises(tad.TadError, match="exceeds"):
        fn.add_elements_dynamic_device(ops, [[0, 1], [0, 1, 2, 3]])
    with pytest.raises(tad.TadError, match="unique"):
        fn.add_elements_dynamic_device([ops[0], ops[0]], [[0, 1]])
    with pytest.raises(tad.TadError, match="out of range"):
        fn.add_elements_dynamic_device(ops, [[0, -1]])
    assert fn.terms() == []  # failed calls add nothing


def test_device_empty_terms(ctx, plugin):
    """Device terms without elements (and dynamic calls without elements) add
    nothing to f, g or the pattern; they sit beside catalog terms."""
    fn = tad.ScalarFunction(2, 3, ctx)
    fn.add_elements_device(plugin["symdir_tri"], np.zeros((0, 3), dtype=np.int32), np.zeros((0, 5)))
    fn.add_elements_dynamic_device(dyn_ops(plugin, "sum_squares", 2), [])
    fn.add_elements(1, tad.PENALTY, [[1]], [0.5, -0.25, 2.0])
    assert fn.terms() == [(tad.DEVICE, 3, 0), (tad.PENALTY, 1, 1)]
    x = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    f, g, H = fn.eval_with_hessian_proj(x)
    assert abs(f - 2.0 * ((0.3 - 0.5) ** 2 + (0.4 + 0.25) ** 2)) <= 1e-15
    assert H.nnz == 4 and np.count_nonzero(g) == 2
    fd = tad.ScalarFunction(2, 3, ctx)
    fd.add_elements_dynamic_device(dyn_ops(plugin, "sum_squares", 2), [[], [2]])  # valence 0 and 1
    assert fd.terms() == [(tad.DEVICE, 1, 2)]
    f, g, H = fd.eval_with_derivatives(x)
    assert f == 0.5 ** 2 + 0.6 ** 2 and np.array_equal(g, [0, 0, 0, 0, 1.0, 1.2])
    assert np.array_equal(H.toarray()[4:, 4:], 2 * np.eye(2)) and H.nnz == 4


@pytest.mark.parametrize("suffix", ["", "_inv"])
def test_device_tet_newton_matches_catalog(ctx, plugin, suffix):
    """Projected Newton on a twisted cube with the tet energy as a device
    functor (K = 12: hyper-dual passes) and from the catalog: same energies
    along the way and at the end (converged-energy bar 1e-8)."""
    import torch
    P = configs.twist_cube(m=8, seed=2)
    fc = tad.ScalarFunction.from_problem(P, ctx)
    fd = device_function(P, plugin, ctx, suffix=suffix)
    xc = torch.tensor(P.x0, device="cuda:0")
    xd = xc.clone()
    for it in range(12):
        sc = fc.newton_step(xc)
        sd = fd.newton_step(xd)
        assert abs(sd["f"] - sc["f"]) <= 1e-10 * abs(sc["f"]), it
    fc_, fd_ = fc.eval(xc), fd.eval(xd)
    print(f"tets: catalog f {fc_:.15g} device functor f {fd_:.15g}")
    assert abs(fd_ - fc_) <= 1e-8 * abs(fc_)


@pytest.mark.parametrize("suffix", ["", "_inv"])
def test_device_tet_mixed_precision(ctx, plugin, suffix):
    """Tets as device functors with FP32 contributions (hyper-dual passes in
    FP32) against the FP64 catalog: f identical (passive FP64 value), g and
    H_proj within the mixed-mode bar 2e-5, the Hessian unprojected too."""
    P = perturbed_cube(4, 0.1, 17)
    fc = tad.ScalarFunction.from_problem(P, ctx)
    fm = device_function(P, plugin, ctx, precision="mixed", suffix=suffix)
    fa, ga, Ha = fc.eval_with_hessian_proj(P.x0)
    fb, gb, Hb = fm.eval_with_hessian_proj(P.x0)
    assert fa == fb
    assert nrel(gb, ga) <= 2e-5 and nrel(Hb.data, Ha.data) <= 2e-5
    fa, ga, Ha = fc.eval_with_derivatives(P.x0)
    fb, gb, Hb = fm.eval_with_derivatives(P.x0)
    assert nrel(gb, ga) <= 2e-5 and nrel(Hb.data, Ha.data) <= 2e-5


# ---------------------------------------------------------------------------
# Repeated handles within an element (Element.hh:208-243)
# ---------------------------------------------------------------------------

def repeated_quad_problem(seed, N, n_handles, n_elem):
    """Quadratic elements (non-symmetric, indefinite A) of valence N, about half
    of them listing some handle more than once (adjacent and not, all equal)."""
    rng = np.random.default_rng(seed)
    conn = []
    for e in range(n_elem):
        c = rng.choice(n_handles, N, replace=False)
        if e % 2 == 0:
            a, b = sorted(rng.choice(N, 2, replace=False))
            c[b] = c[a]
        if e % 7 == 3:
            c[:] = c[0]
        conn.append(c)
    conn = np.array(conn, dtype=np.int32)
    K = 2 * N
    A = rng.standard_normal((n_elem, K, K))
    b = rng.standard_normal((n_elem, K))
    c = rng.standard_normal((n_elem, 1))
    params = np.concatenate([A.reshape(n_elem, K * K), b, c], axis=1)
    return conn, params, rng.standard_normal(2 * n_handles)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("eps", [1e-9, -1.0, 0.3])
def test_repeated_handles_match_oracle(ctx, plugin, eps, precision):
    """Catalog quadratics of valence 3 and device quadratics of valence 2 with
    repeated handles: the entries of a repeated handle share its variables, the
    merged element Hessian is projected, and only the distinct handles' blocks
    are assembled, against the oracle (which restates Element::variables).
    Same sparsity pattern; FP64 1e-10, FP32 contributions 2e-5 normwise."""
    conn3, params3, x = repeated_quad_problem(21, 3, 8, 24)
    conn2, params2, _ = repeated_quad_problem(22, 2, 8, 16)
    assert any(len(set(r)) < 3 for r in conn3) and any(len(set(r)) == 1 for r in conn2)
    fn = tad.ScalarFunction(2, 8, ctx, precision)
    fn.add_elements(3, tad.QUADRATIC, conn3, params3)
    fn.add_elements_device(plugin["quadratic4"], conn2, params2)
    fo = oracle.Function(2, 8)
    fo.add_elements(tad.QUADRATIC, conn3, params3)
    fo.add_elements(tad.QUADRATIC, conn2, params2)
    tol = FP64_TOL if precision == "fp64" else 2e-5
    f1, g1, H1 = fo.eval_with_hessian_proj(x, eps)
    f2, g2, H2 = fn.eval_with_hessian_proj(x, eps)
    assert np.array_equal(H2.indptr, H1.indptr) and np.array_equal(H2.indices, H1.indices)
    assert abs(f2 - f1) <= 1e-12 * abs(f1)
    assert nrel(g2, g1) <= tol and nrel(H2.data, H1.data) <= tol
    fa, ga = fn.eval_with_gradient(x)
    fb, gb, Hb = fn.eval_with_derivatives(x)
    fo2, go2, Ho2 = fo.eval_with_derivatives(x)
    assert fa == f2 and fb == f2 and abs(fn.eval(x) - f1) <= 1e-12 * abs(f1)
    assert nrel(ga, g1) <= tol and nrel(gb, go2) <= tol
    assert np.array_equal(Hb.indices, Ho2.indices) and nrel(Hb.data, Ho2.data) <= tol


def test_repeated_handles_hyper_dual_element(ctx, plugin):
    """A K = 8 device element (hyper-dual passes) listing handle 0 twice,
    against the same energy on four distinct handles with handle 3 placed at
    handle 0 and then merged (M = S^T H S, gradient S^T g); the projection of
    M padded with two zero rows and columns (the reference's unused local
    variables) against numpy.  FP64 bars 1e-12."""
    rng = np.random.default_rng(23)
    p = np.array([0.7, 0.4])
    x = rng.uniform(-1.0, 1.0, 6)
    fr = tad.ScalarFunction(2, 3, ctx)
    fr.add_elements_device(plugin["smooth8"], [[0, 1, 0, 2]], p)
    fs = tad.ScalarFunction(2, 4, ctx)
    fs.add_elements_device(plugin["smooth8"], [[0, 1, 3, 2]], p)
    xs = np.concatenate([x, x[:2]])
    _, gs, Hs = fs.eval_with_derivatives(xs)
    S = np.zeros((8, 6))
    S[:6, :6] = np.eye(6)
    S[6:8, 0:2] = np.eye(2)
    M = S.T @ Hs.toarray() @ S
    f, g, H = fr.eval_with_derivatives(x)
    assert abs(f - smooth8_energy(np.concatenate([x[:4], x[:2], x[4:]]), p)) <= 1e-14 * abs(f)
    assert nrel(g, S.T @ gs) <= 1e-12 and nrel(H.toarray(), M) <= 1e-12
    for eps in [1e-9, -1.0, 0.5]:
        _, gp, Hp = fr.eval_with_hessian_proj(x, eps)
        pad = np.zeros((8, 8))
        pad[:6, :6] = H.toarray()
        assert np.array_equal(gp, g)
        assert nrel(Hp.toarray(), reference_projection(pad, eps)[:6, :6]) <= 1e-12, eps


def test_repeated_handles_dynamic_and_newton(ctx, plugin):
    """Dynamic-valence elements with repeated handles (sum of squares of the
    listed 2D values, DynamicElementsTest.cc:9-33): [2, 2] is ||2 x_2||^2 and
    [1, 2, 1] is ||2 x_1 + x_2||^2, in closed form.  Then a Newton step on
    repeated-handle quadratics that are convex after merging lands on the
    minimizer of the merged quadratic."""
    fn = tad.ScalarFunction(2, 3, ctx)
    fn.add_elements_dynamic_device(dyn_ops(plugin, "sum_squares", 2), [[2, 2], [1, 2, 1]])
    x = np.array([0.3, -0.2, 0.5, 0.25, -0.4, 0.7])
    X = x.reshape(3, 2)
    f, g, H = fn.eval_with_derivatives(x)
    s = 2 * X[1] + X[2]
    assert abs(f - (4 * X[2] @ X[2] + s @ s)) <= 1e-15
    gref = np.zeros((3, 2))
    gref[2] = 8 * X[2] + 2 * s
    gref[1] = 4 * s
    assert np.abs(g - gref.ravel()).max() <= 1e-15
    Href = np.zeros((6, 6))
    I2 = np.eye(2)
    Href[4:6, 4:6] = 8 * I2 + 2 * I2
    Href[2:4, 2:4] = 8 * I2
    Href[2:4, 4:6] = Href[4:6, 2:4] = 4 * I2
    assert np.array_equal(H.toarray(), Href)
    # convex merged quadratics: A = L L^T + I on distinct-handle blocks of the
    # merged variables, written on elements that list a handle twice
    rng = np.random.default_rng(29)
    n = 6
    conn, params = [], []
    for e in range(10):
        h = rng.choice(n, 2, replace=False)
        c = np.array([h[0], h[1], h[0]], dtype=np.int32)
        L = rng.standard_normal((4, 4))
        Mz = L @ L.T + np.eye(4)  # merged Hessian over (h0, h1)
        S = np.zeros((6, 4))
        S[0:2, 0:2] = S[4:6, 0:2] = np.eye(2)
        S[2:4, 2:4] = np.eye(2)
        A = np.zeros((6, 6))
        A[:4, :4] = Mz  # y^T A y = z^T M z when entry 2 repeats entry 0
        conn.append(c)
        params.append(np.concatenate([A.ravel(), rng.standard_normal(6), [0.0]]))
    conn, params = np.array(conn), np.array(params)
    fq = tad.ScalarFunction(2, n, ctx)
    fq.add_elements(3, tad.QUADRATIC, conn, params)
    fo = oracle.Function(2, n)
    fo.add_elements(tad.QUADRATIC, conn, params)
    x0 = np.zeros(2 * n)
    _, g0, H0 = fo.eval_with_derivatives(x0)
    used = np.unique(conn)
    cols = np.concatenate([np.arange(2 * v, 2 * v + 2) for v in used])
    xs = np.zeros(2 * n)
    xs[cols] = np.linalg.solve(H0.toarray()[np.ix_(cols, cols)], -g0[cols])
    import torch
    xt = torch.zeros(2 * n, dtype=torch.float64, device="cuda:0")
    st = fq.newton_step(xt, tad.newton_default_opts(projection_eps=float("nan"), solve_tol=1e-13))
    assert st["step"] == 1.0 and st["ls_failed"] == 0
    assert np.abs(xt.cpu().numpy() - xs).max() <= 1e-8 * max(1.0, np.abs(xs).max())
