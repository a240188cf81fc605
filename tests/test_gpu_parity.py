"""GPU parity: the HIP library (through the C ABI) against the CPU oracle and
against the reference's own known answers.

Tolerances (FP64 mode): normwise relative error <= 1e-10 for g and H_proj
(max |a - b| / max |b|), <= 1e-12 for f; identical sparsity patterns.  Mixed
mode (FP32 element derivatives): f identical to the FP64 path (values are
always FP64), g and H within 2e-5 normwise.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401

import oracle
import tinyad_amd as tad
from tinyad_amd import configs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def nrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(initial=0.0), 1e-300)
    return float(np.abs(a - b).max(initial=0.0) / den)


def assert_same_pattern(Hg, Ho):
    Hg = Hg.tocsr()
    Ho = Ho.tocsr()
    Hg.sort_indices()
    Ho.sort_indices()
    assert Hg.shape == Ho.shape
    assert Hg.nnz == Ho.nnz
    assert np.array_equal(Hg.indptr, Ho.indptr)
    assert np.array_equal(Hg.indices, Ho.indices)


def perturbed_cube(m=3, sigma=0.15, seed=11):
    P = configs.twist_cube(m=m, seed=seed)
    h = P.info["h"]
    rest = P.info["rest"]
    x = rest + sigma * h * configs.normal(seed + 1, rest.size, stream=9)
    P.x0 = x
    return P


# ---------------------------------------------------------------------------
# Reference known answers (tests/ScalarFunctionTest.cc)
# ---------------------------------------------------------------------------

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
    A = np.asarray(A, dtype=np.float64)
    return np.concatenate([A.ravel(), np.asarray(b, dtype=np.float64).ravel(), [c]])


def test_scalar_function_basic(ctx):
    # x0^2 + x1^2 at (1, 2): f 5, g (2, 4), H diag(2, 2)   (ScalarFunctionTest.cc:10-42)
    func = tad.scalar_function(2, 1, ctx)
    func.add_elements(1, tad.QUADRATIC, [[0]], quad_params(2 * np.eye(2), [0, 0], 0))
    x = np.array([1.0, 2.0])
    assert func.eval(x) == 5.0
    f, g = func.eval_with_gradient(x)
    assert f == 5.0 and g[0] == 2.0 and g[1] == 4.0
    f, g, H = func.eval_with_derivatives(x)
    assert f == 5.0 and H[0, 0] == 2.0 and H[1, 1] == 2.0


def test_scalar_function_1d(ctx):
    # 2 x^2 + x + 1 at 1: f 4, g 5, H_proj 4   (ScalarFunctionTest.cc:44-66)
    func = tad.scalar_function(1, 1, ctx)
    func.add_elements(1, tad.QUADRATIC, [[0]], quad_params([[4.0]], [1.0], 1.0))
    f, g, H = func.eval_with_hessian_proj(np.array([1.0]))
    assert abs(f - 4.0) <= 1e-16 and abs(g[0] - 5.0) <= 1e-16 and abs(H[0, 0] - 4.0) <= 1e-16


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_scalar_function_2d(ctx, sign):
    # +-(2 x0^2 + 2 x0 x1 + x1^2 + x0 + 1) at (1, 2)   (ScalarFunctionTest.cc:72-151)
    func = tad.scalar_function(2, 1, ctx)
    A = sign * np.array([[4.0, 2.0], [2.0, 2.0]])
    func.add_elements(1, tad.QUADRATIC, [[0]], quad_params(A, sign * np.array([1.0, 0.0]), sign * 1.0))
    x = np.array([1.0, 2.0])
    f, g, H = func.eval_with_derivatives(x)
    f2, g2, Hp = func.eval_with_hessian_proj(x)
    assert abs(f - sign * 12) <= 1e-16 and abs(f2 - sign * 12) <= 1e-16
    assert np.allclose(g, sign * np.array([9.0, 6.0]), atol=1e-16, rtol=0)
    assert np.allclose(g2, sign * np.array([9.0, 6.0]), atol=1e-16, rtol=0)
    assert np.allclose(H.toarray(), A, atol=1e-16, rtol=0)
    if sign > 0:
        assert np.array_equal(Hp.toarray(), A)  # PD: projection leaves it unchanged
    else:
        ev = np.linalg.eigvalsh(Hp.toarray())
        assert ev.min() > 0.0


def test_element_variables_unused_handle(ctx):
    # 2 handles, the element touches only handle 0 (ScalarFunctionTest.cc:153-178)
    func = tad.scalar_function(2, 2, ctx)
    func.add_elements(1, tad.QUADRATIC, [[0]], quad_params(np.zeros((2, 2)), [0, 0], 0.0))
    x = np.zeros(4)
    f, g, H = func.eval_with_hessian_proj(x)
    assert f == 0.0 and func.eval(x) == 0.0
    assert H.shape == (4, 4) and H.nnz == 4


def test_sum_of_variables(ctx):
    # v.sum() at (1, 1) = 2   (ScalarFunctionTest.cc:181-215, Move1)
    func = tad.scalar_function(2, 1, ctx)
    func.add_elements(1, tad.QUADRATIC, [[0]], quad_params(np.zeros((2, 2)), [1, 1], 0.0))
    assert func.eval(np.ones(2)) == 2.0


@pytest.mark.parametrize("which,x,y", [(0, -1.0, -0.5), (1, 5.0, 0.0), (1, 1.0, 1.0), (1, -1.0, -1.0),
                                       (2, 1.0, 1.5), (3, 0.7, -1.3)])
def test_dual_ops_against_oracle(ctx, which, x, y):
    fo, go, Ho = oracle.dual_known_answer(which, x, y)
    fg, gg, Hg = tad.selftest_dual(which, x, y, ctx)
    assert abs(fg - fo) <= 1e-13 * max(1, abs(fo))
    assert np.allclose(gg, go, rtol=1e-12, atol=1e-12)
    assert np.allclose(Hg, Ho, rtol=1e-12, atol=1e-12)
    if which == 0:  # ScalarTestBinaryOperators.cc:623-642
        assert fg == -2.0 and np.allclose(gg, [4, -4]) and np.allclose(Hg, [[-4, 8], [8, -16]])
    if which == 2:  # ScalarTestBinaryOperators.cc:711-727
        assert abs(fg - 3.0) < 1e-12 and abs(gg[0] - 4.5) < 1e-12 and abs(Hg[0, 0] - 3.0) < 1e-12


# ---------------------------------------------------------------------------
# Parity against the oracle
# ---------------------------------------------------------------------------

def problems():
    return {
        "grid_param_12": configs.grid_param(12, seed=5),
        "twist_cube_3": configs.twist_cube(m=3, seed=5),
        "cube_perturbed": perturbed_cube(3, 0.15, 13),
        "newton_test": configs.newton_test_problem(),
    }


@pytest.mark.parametrize("name", ["grid_param_12", "twist_cube_3", "cube_perturbed", "newton_test"])
@pytest.mark.parametrize("eps", [1e-9, -1.0, 0.05])
def test_eval_with_hessian_proj_matches_oracle(ctx, name, eps):
    P = problems()[name]
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f1, g1, H1 = fo.eval_with_hessian_proj(P.x0, eps)
    f2, g2, H2 = fg.eval_with_hessian_proj(P.x0, eps)
    assert_same_pattern(H2, H1)
    assert np.isfinite(f1)
    assert abs(f2 - f1) <= 1e-12 * abs(f1)
    assert nrel(g2, g1) <= FP64_TOL
    assert nrel(H2.toarray(), H1.toarray()) <= FP64_TOL
    assert abs(fg.eval(P.x0) - f1) <= 1e-12 * abs(f1)


@pytest.mark.parametrize("name", ["grid_param_12", "cube_perturbed", "newton_test"])
def test_eval_with_derivatives_and_gradient_match_oracle(ctx, name):
    P = problems()[name]
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f1, g1, H1 = fo.eval_with_derivatives(P.x0)
    f2, g2, H2 = fg.eval_with_derivatives(P.x0)
    assert_same_pattern(H2, H1)
    assert abs(f2 - f1) <= 1e-12 * abs(f1)
    assert nrel(g2, g1) <= FP64_TOL
    assert nrel(H2.toarray(), H1.toarray()) <= FP64_TOL
    f3, g3 = fg.eval_with_gradient(P.x0)
    assert f3 == f2 and nrel(g3, g1) <= FP64_TOL


def test_inverted_elements(ctx):
    # an inverted element contributes +inf, zero gradient and eps I (NewtonTest.cc:40-41)
    P = configs.twist_cube(m=2, seed=3)
    X = P.x0.reshape(-1, 3).copy()
    T = P.terms[0].conn
    a, b = T[0, 1], T[0, 2]
    X[[a, b]] = X[[b, a]]
    x = X.ravel()
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f1, g1, H1 = fo.eval_with_hessian_proj(x)
    f2, g2, H2 = fg.eval_with_hessian_proj(x)
    assert f1 == np.inf and f2 == np.inf and fg.eval(x) == np.inf
    assert nrel(g2, g1) <= FP64_TOL
    assert nrel(H2.toarray(), H1.toarray()) <= FP64_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["tets", "triangles"])
def test_grouped_and_per_element_sums_agree(ctx, precision, kind):
    """Hessian blocks summed per wave of 64 consecutive elements first (the
    elements in mesh order: few distinct blocks per group) and one slot per
    element block (the same elements shuffled: too many distinct blocks per
    group, so the term keeps per-element slots) give the same f, g and H up to
    summation order, and both match the oracle (ScalarFunctionImpl.hh:378-416)."""
    import dataclasses
    if kind == "tets":
        P = configs.twist_cube(m=8, seed=5)
    else:
        P, _ = configs.concat(configs.disk_batch(count=8, seed=5, n_min=20, n_max=40))
    rng = np.random.default_rng(3)
    perm = rng.permutation(P.terms[0].conn.shape[0])
    t0 = P.terms[0]
    Q = dataclasses.replace(P, terms=[dataclasses.replace(t0, conn=np.ascontiguousarray(t0.conn[perm]),
                                                          params=np.ascontiguousarray(t0.params[perm]))]
                            + list(P.terms[1:]))
    fa = tad.ScalarFunction.from_problem(P, ctx, precision=precision)
    fb = tad.ScalarFunction.from_problem(Q, ctx, precision=precision)
    ra = fa.eval_with_hessian_proj(P.x0)
    rb = fb.eval_with_hessian_proj(P.x0)
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert abs(ra[0] - rb[0]) <= 1e-12 * abs(ra[0])
    assert nrel(ra[1], rb[1]) <= tol and nrel(ra[2].data, rb[2].data) <= tol
    fo = oracle.Function.from_problem(P)
    fo_, go_, Ho_ = fo.eval_with_hessian_proj(P.x0)
    bar = FP64_TOL if precision == "fp64" else 2e-5
    for r in (ra, rb):
        dH = abs(r[2] - Ho_).max() / abs(Ho_).max()
        assert nrel(r[1], go_) <= bar and dH <= bar, (dH, nrel(r[1], go_))
    for shape in range(4):  # every assembly shape sums the group partials in the same order
        fa.assembly_shape = shape
        f2, g2, H2 = fa.eval_with_hessian_proj(P.x0)
        assert f2 == ra[0] and np.array_equal(g2, ra[1]) and np.array_equal(H2.data, ra[2].data)


def test_deterministic_bits(ctx):
    P = perturbed_cube(4, 0.2, 21)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    r1 = fg.eval_with_hessian_proj(P.x0)
    r2 = fg.eval_with_hessian_proj(P.x0)
    assert r1[0] == r2[0]
    assert np.array_equal(r1[1], r2[1])
    assert np.array_equal(r1[2].data, r2[2].data)


def test_device_arrays(ctx):
    import torch
    P = configs.twist_cube(m=3, seed=2)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    fh, gh, Hh = fg.eval_with_hessian_proj(P.x0)
    xd = torch.from_numpy(P.x0).cuda()
    fd, gd, Hd = fg.eval_with_hessian_proj(xd)
    assert fd == fh
    assert np.array_equal(gd.cpu().numpy(), gh)
    assert np.array_equal(Hd.cpu().numpy(), Hh.data)


@pytest.mark.parametrize("name", ["grid_param_12", "cube_perturbed"])
def test_mixed_precision_errors(ctx, name):
    P = problems()[name]
    fo = oracle.Function.from_problem(P)
    fm = tad.ScalarFunction.from_problem(P, ctx, precision="mixed")
    f1, g1, H1 = fo.eval_with_hessian_proj(P.x0)
    f2, g2, H2 = fm.eval_with_hessian_proj(P.x0)
    assert abs(f2 - f1) <= 1e-12 * abs(f1)
    assert nrel(g2, g1) <= 2e-5
    assert nrel(H2.toarray(), H1.toarray()) <= 2e-5


def test_laplacian_known_answer(ctx):
    # Hessian of 1/4 sum (edge differences)^2 per triangle on the closed bunny =
    # graph Laplacian (DynamicElementsTest.cc:60-91, reference_laplace :31-52)
    V, F = load_bunny()
    n = V.shape[0]
    A = 0.5 * (3 * np.eye(3) - np.ones((3, 3)))
    func = tad.scalar_function(1, n, ctx)
    func.add_elements(3, tad.QUADRATIC, F, quad_params(A, np.zeros(3), 0.0))
    L = func.eval_hessian_of_quadratic()
    Lr = reference_laplace(n, F)
    assert L.shape == Lr.shape and L.nnz == Lr.nnz
    assert sp.linalg.norm(L - Lr) <= 1e-12


# ---------------------------------------------------------------------------
# Errors (ExceptionTest.cc semantics as status codes)
# ---------------------------------------------------------------------------

def test_nonfinite_derivatives_raise_and_recover(ctx):
    P = configs.twist_cube(m=1, seed=1)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    X = P.info["rest"].reshape(-1, 3).copy()
    T = P.terms[0].conn
    # squash one tet to det(M) ~ 1e-100 > 0: the Hessian of ||J^-1||^2 overflows
    v = T[0]
    X[v[3]] = X[v[0]] + 1e-100 * (X[v[3]] - X[v[0]]) + 1e-100 * (X[v[1]] - X[v[0]])
    x = X.ravel()
    for _ in range(2):
        with pytest.raises(tad.TadError) as ei:
            fg.eval_with_hessian_proj(x)
        assert ei.value.status == 3
    f, g, H = fg.eval_with_hessian_proj(P.info["rest"])
    assert np.isfinite(f)


def test_size_mismatch_raises(ctx):
    fg = tad.ScalarFunction.from_problem(configs.newton_test_problem(), ctx)
    with pytest.raises(tad.TadError) as ei:
        fg.eval(np.zeros(5))
    assert ei.value.status == 1


def test_dense_quadratic_k8_and_limit(ctx):
    """K = N d = 8 dense elements project in registers (indefinite A, all eps
    modes, against the oracle); K > 8 is rejected as unsupported."""
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((8, 8)))
    A = (Q * np.array([-3.0, -1e-3, 0.0, 1e-12, 0.5, 1.0, 2.0, 10.0])) @ Q.T
    A = 0.5 * (A + A.T)
    b = rng.standard_normal(8)
    x = rng.standard_normal(8)
    prm = quad_params(A, b, 0.25)
    fg = tad.ScalarFunction(2, 4, ctx)
    fg.add_elements(4, tad.QUADRATIC, [[0, 1, 2, 3]], prm)
    fo = oracle.Function(2, 4)
    fo.add_elements(4, [[0, 1, 2, 3]], prm)
    for eps in (1e-9, -1.0, 0.05):
        f, g, H = fg.eval_with_hessian_proj(x, eps)
        fr, gr, Hr = fo.eval_with_hessian_proj(x, eps)
        assert abs(f - fr) <= 1e-12 * abs(fr) and nrel(g, gr) <= 1e-12
        assert nrel(H.toarray(), Hr.toarray()) <= 1e-12
    f3 = tad.ScalarFunction(3, 3, ctx)
    with pytest.raises(tad.TadError) as ei:
        f3.add_elements(3, tad.QUADRATIC, [[0, 1, 2]], quad_params(np.eye(9), np.zeros(9), 0.0))
    assert ei.value.status == 6


# ---------------------------------------------------------------------------
# Solver and Newton
# ---------------------------------------------------------------------------

def test_pcg_matches_direct_solve(ctx):
    import scipy.sparse.linalg as spla
    P = configs.twist_cube(m=4, seed=3)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    d, st = fg.solve(g, tol=1e-13)
    ref = spla.spsolve(H.tocsc(), g)
    assert st.converged
    assert nrel(d, ref) <= 1e-9
    d2, it, rel = oracle.pcg(H, g, block=3, tol=1e-13)
    assert nrel(d, d2) <= 1e-9
    d3, st3 = fg.solve(g, tol=1e-13, w_identity=0.5)
    ref3 = spla.spsolve((H + 0.5 * sp.identity(H.shape[0])).tocsc(), g)
    assert nrel(d3, ref3) <= 1e-9


def test_solve_reads_lower_triangle(ctx):
    """The solve uses H's scalar lower triangle, as the reference's
    SimplicialLDLT (LinearSolver.hh:14, Eigen's default Lower): scaling the
    strictly upper entries of a symmetric H leaves the solution unchanged;
    doubling the lower triangle and the diagonal only solves with 2 H
    (block-Jacobi alone: the coarse space's Galerkin sums read both
    triangles)."""
    import scipy.sparse.linalg as spla
    P = configs.twist_cube(m=4, seed=3)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    H = H.tocsr()
    H.sort_indices()
    rows = np.repeat(np.arange(H.shape[0]), np.diff(H.indptr))
    upper = H.indices > rows
    d_ref, st = fg.solve(g, H=H.data.copy(), tol=1e-13, coarse_aggregate=0)
    assert st.converged and nrel(d_ref, spla.spsolve(H.tocsc(), g)) <= 1e-9
    Hu = H.data.copy()
    Hu[upper] *= 1.7
    d_u, st_u = fg.solve(g, H=Hu, tol=1e-13, coarse_aggregate=0)
    assert st_u.converged and nrel(d_u, d_ref) <= 1e-9
    Hl = H.data.copy()
    Hl[~upper] *= 2.0
    d_l, st_l = fg.solve(g, H=Hl, tol=1e-13, coarse_aggregate=0)
    assert st_l.converged and nrel(d_l, 0.5 * d_ref) <= 1e-9


def test_two_level_preconditioner(ctx):
    """Block-Jacobi plus the aggregate coarse correction: the direct solution,
    at least 2x fewer iterations than block-Jacobi alone, bitwise
    reproducible; regularized systems; a translations-only coarse space when
    the function has not evaluated a Hessian yet (no positions for the
    rotations)."""
    import scipy.sparse.linalg as spla
    P = configs.twist_cube(m=14, seed=3)  # 15^3 = 3,375 handles
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    ref = spla.spsolve(H.tocsc(), g)
    d0, s0 = fg.solve(g, tol=1e-12, coarse_aggregate=0)
    d1, s1 = fg.solve(g, tol=1e-12, coarse_aggregate=64)
    d2, s2 = fg.solve(g, tol=1e-12, coarse_aggregate=64)
    print(f"block-Jacobi {s0.iters} iterations, two-level (64 handles per aggregate) {s1.iters}")
    assert s0.converged and s1.converged
    assert s0.coarse_unknowns == 0 and s1.coarse_unknowns == 6 * (s1.coarse_unknowns // 6) > 6
    assert nrel(d0, ref) <= 1e-8 and nrel(d1, ref) <= 1e-8
    assert np.array_equal(d1, d2) and s1.iters == s2.iters
    assert 2 * s1.iters <= s0.iters
    d3, s3 = fg.solve(g, tol=1e-12, coarse_aggregate=64, w_identity=0.5)
    ref3 = spla.spsolve((H + 0.5 * sp.identity(H.shape[0])).tocsc(), g)
    assert s3.converged and nrel(d3, ref3) <= 1e-8
    fb = tad.ScalarFunction.from_problem(P, ctx)
    d4, s4 = fb.solve(g, H=H, tol=1e-12, coarse_aggregate=64)
    print(f"translations-only coarse space: {s4.iters} iterations")
    assert s4.converged and nrel(d4, ref) <= 1e-8 and s4.iters < s0.iters
    assert s4.coarse_unknowns * 2 == s1.coarse_unknowns  # 3 instead of 6 unknowns per aggregate


def pinned_cube_mask(P):
    """Free-variable mask of a cube with its bottom and top layers held fixed."""
    rest = P.info["rest"].reshape(-1, 3)
    h = P.info["h"]
    fixed = (rest[:, 2] < 0.5 * h) | (rest[:, 2] > 1.0 - 0.5 * h)
    return np.repeat(~fixed, 3).astype(np.uint8)


def test_reduced_basis_solve(ctx):
    """newton_direction_reduced_basis with a selection basis B (NewtonDirection.hh:76-105):
    sol = B (B^T H B + w I)^-1 B^T rhs against a direct solve of the free block,
    with and without the coarse correction, regularized, host and device."""
    import scipy.sparse.linalg as spla
    import torch
    P = configs.twist_cube(m=12, seed=6)
    free = pinned_cube_mask(P)
    fi = np.nonzero(free)[0]
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    Hff = H.tocsr()[fi][:, fi]
    for w in [0.0, 0.3]:
        ref = np.zeros_like(g)
        ref[fi] = spla.spsolve((Hff + w * sp.identity(fi.size)).tocsc(), -g[fi])
        for coarse in [0, 64]:
            d, st = fg.solve(-g, free=free, tol=1e-12, w_identity=w, coarse_aggregate=coarse)
            print(f"w={w} coarse={coarse}: {st.iters} iterations, coarse size {st.coarse_unknowns}")
            assert st.converged and nrel(d, ref) <= 1e-8
            assert np.all(d[free == 0] == 0.0)
            assert (st.coarse_unknowns > 0) == (coarse > 0)
    dd, st = fg.solve(torch.tensor(-g, device="cuda:0"), free=torch.tensor(free, device="cuda:0"), tol=1e-12)
    assert st.converged and nrel(dd.cpu().numpy(), fg.solve(-g, free=free, tol=1e-12)[0]) <= 1e-12
    with pytest.raises(tad.TadError):
        fg.solve(-g, free=free[:-1])


def test_general_basis_solve(ctx):
    """newton_direction_reduced_basis with a general sparse basis B
    (NewtonDirection.hh:76-105): sol = B (B^T H B + w I)^-1 B^T rhs against the
    oracle's direct solve, for a random sparse B (one to three entries per
    column, plus a unit entry so it has full column rank), a selection basis
    scaled column by column (the same subspace as the selection solve), B = I
    (the full solve), host and device, and the errors."""
    import torch
    P = configs.twist_cube(m=6, seed=6)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    n = g.size
    rng = np.random.default_rng(12)
    m = n // 3
    rows = np.concatenate([rng.choice(n, m, replace=False), rng.integers(0, n, 2 * m)])
    cols = np.concatenate([np.arange(m), rng.integers(0, m, 2 * m)])
    vals = np.concatenate([np.ones(m), rng.uniform(-1.0, 1.0, 2 * m)])
    B = sp.csr_matrix((vals, (rows, cols)), shape=(n, m))
    for w in [0.0, 0.25]:
        ref = oracle.newton_direction_reduced_basis(g, H, B, w)
        d, st = fg.solve_basis(-g, B, tol=1e-13, w_identity=w, max_iters=20000)
        print(f"general basis n={n} m={m} w={w}: {st.iters} iterations, rel residual {st.rel_residual:.2e}")
        assert st.converged and st.coarse_unknowns == 0
        assert nrel(d, ref) <= 1e-8
        d2, _ = fg.solve_basis(-g, B, H=H, tol=1e-13, w_identity=w, max_iters=20000)
        assert np.array_equal(d, d2)  # deterministic, and H=None is the last evaluated Hessian
    free = pinned_cube_mask(P)
    fi = np.nonzero(free)[0]
    S = sp.csr_matrix((rng.uniform(0.5, 2.0, fi.size), (fi, np.arange(fi.size))), shape=(n, fi.size))
    ds, st = fg.solve_basis(-g, S, tol=1e-13, max_iters=20000)
    dsel, _ = fg.solve(-g, free=free, tol=1e-13)
    assert st.converged and nrel(ds, dsel) <= 1e-8 and np.all(ds[free == 0] == 0.0)
    dI, st = fg.solve_basis(-g, sp.identity(n, format="csr"), tol=1e-13, max_iters=20000)
    assert st.converged and nrel(dI, fg.solve(-g, tol=1e-13)[0]) <= 1e-8
    dd, st = fg.solve_basis(torch.tensor(-g, device="cuda:0"), B, tol=1e-13, max_iters=20000)
    assert st.converged and np.array_equal(dd.cpu().numpy(), fg.solve_basis(-g, B, tol=1e-13, max_iters=20000)[0])
    with pytest.raises(tad.TadError):
        fg.solve_basis(-g, B[:-1])
    Bz = sp.hstack([B, sp.csr_matrix((n, 1))]).tocsr()  # an empty column: B^T H B + w I is singular
    with pytest.raises(tad.TadError):
        fg.solve_basis(-g, Bz)
    _, st = fg.solve_basis(-g, B, tol=1e-13, max_iters=3)
    assert not st.converged and st.iters == 3


def test_reduced_basis_newton_keeps_pins(ctx):
    """Projected Newton with hard pins (reduced-basis directions and the
    library's line search): pinned coordinates never move, the energy
    decreases monotonically, and the free gradient drops to the solver level."""
    import torch
    P = configs.twist_cube(m=6, seed=8)
    free = pinned_cube_mask(P)
    fg = tad.ScalarFunction(3, P.n_handles, ctx)
    T = P.terms[0]
    fg.add_elements(4, tad.SYMDIR_TET, T.conn, T.params)
    x = torch.tensor(P.x0, device="cuda:0")
    x0 = x.clone()
    fd = torch.tensor(free, device="cuda:0")
    fprev = float("inf")
    for it in range(12):
        f, g, H = fg.eval_with_hessian_proj(x)
        d, st = fg.solve(-g, H=H, free=fd, tol=1e-12)
        assert st.converged
        x, f_new, step, trials = fg.line_search(x, d, f, g)
        assert f_new <= f <= fprev
        fprev = f_new
    f, g, H = fg.eval_with_hessian_proj(x)
    gf = g.cpu().numpy()[free == 1]
    print(f"reduced Newton: f {f:.12g}, |g_free| {np.abs(gf).max():.2e}")
    assert torch.equal(x[fd == 0], x0[fd == 0])
    assert np.abs(gf).max() <= 1e-8 * max(1.0, abs(f))


def test_coarse_reuse(ctx):
    """coarse_reuse keeps the previous compatible coarse inverse: the solve
    still converges to tolerance, incompatible solves (other w, reduced)
    rebuild, and a refresh schedule replayed from the same point gives the
    same bits."""
    import torch
    P = configs.twist_cube(m=12, seed=4)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    x0 = torch.tensor(P.x0, device="cuda:0")

    def run(k, steps=5):
        x = x0.clone()
        o = tad.newton_default_opts(solve_coarse_aggregate=128)
        sts = []
        for i in range(steps):
            o.solve.coarse_reuse = 0 if i % k == 0 else 1
            sts.append(fg.newton_step(x, o))
        return x, sts

    xa, sa = run(3)
    assert [s["coarse_reused"] for s in sa] == [0, 1, 1, 0, 1]
    assert all(s["solve_converged"] and s["coarse_unknowns"] > 0 for s in sa)
    xb, sb = run(3)
    assert torch.equal(xa, xb) and [s["cg_iters"] for s in sa] == [s["cg_iters"] for s in sb]
    # the same schedule kept by the library (tad_newton_opts.coarse_refresh) on
    # a fresh function gives the same bits
    fg2 = tad.ScalarFunction.from_problem(P, ctx)
    x2 = x0.clone()
    o2 = tad.newton_default_opts(solve_coarse_aggregate=128, coarse_refresh=3)
    s2 = [fg2.newton_step(x2, o2) for _ in range(5)]
    assert [s["coarse_reused"] for s in s2] == [0, 1, 1, 0, 1]
    assert torch.equal(x2, xa) and [s["cg_iters"] for s in s2] == [s["cg_iters"] for s in sa]
    x1, s1 = run(1)
    assert not any(s["coarse_reused"] for s in s1)
    f1, fa = s1[-1]["f_new"], sa[-1]["f_new"]
    print(f"refresh 1: f {f1:.15g}, cg {[s['cg_iters'] for s in s1]}; refresh 3: f {fa:.15g}, "
          f"cg {[s['cg_iters'] for s in sa]}")
    assert abs(f1 - fa) <= 1e-8 * abs(f1)

    f, g, H = fg.eval_with_hessian_proj(P.x0)
    ref = spla_solve(H, g)
    _, st = fg.solve(g, tol=1e-12, coarse_aggregate=128)
    assert st.coarse_reused == 0
    d, st = fg.solve(g, tol=1e-12, coarse_aggregate=128, coarse_reuse=1)
    assert st.coarse_reused == 1 and st.converged and nrel(d, ref) <= 1e-8
    _, st = fg.solve(g, tol=1e-12, coarse_aggregate=128, coarse_reuse=1, w_identity=0.5)
    assert st.coarse_reused == 0
    free = np.ones(g.size, dtype=np.uint8)
    free[:3] = 0
    _, st = fg.solve(g, tol=1e-12, coarse_aggregate=128, coarse_reuse=1, free=free)
    assert st.coarse_reused == 0 and st.coarse_unknowns > 0
    _, st = fg.solve(g, tol=1e-12, coarse_aggregate=128, coarse_reuse=1)
    assert st.coarse_reused == 0  # the reduced solve's factor is not kept
    _, st = fg.solve(g, tol=1e-12, coarse_reuse=1, coarse_aggregate=32)
    assert st.coarse_reused == 0


def spla_solve(H, b):
    import scipy.sparse.linalg as spla
    return spla.spsolve(H.tocsc(), b)


def test_two_level_coarse_size_cap(ctx):
    """A requested aggregate size that would give more than 8,192 coarse
    unknowns is raised until the dense coarse inverse fits that bound."""
    P = configs.twist_cube(m=30, seed=4)  # 29,791 handles
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    d, st = fg.solve(g, tol=1e-10, coarse_aggregate=4)
    res = np.linalg.norm(H @ d - g) / np.linalg.norm(g)
    print(f"coarse size {st.coarse_unknowns}, {st.iters} iterations, residual {res:.2e}")
    assert st.converged and res <= 2e-10
    assert 6 * 1000 <= st.coarse_unknowns <= 8192 and st.coarse_unknowns % 6 == 0


def test_solve_graph_replays_match_direct_launches(ctx):
    """CG iterations replayed from the recorded HIP graph give the same bits
    as direct launches: block-Jacobi and two-level, with w, with a reduced
    basis, on a tet cube and a triangle grid, solved repeatedly with eager work
    and synchronizations in between; operands that change (another H,
    another coarse space) are recorded again, and Newton trajectories agree
    to the bit."""
    import torch
    for P in [configs.twist_cube(m=6, seed=3), configs.grid_param(20, seed=2)]:
        fa = tad.ScalarFunction.from_problem(P, ctx)
        fb = tad.ScalarFunction.from_problem(P, ctx)
        fa.set_solve_graphs(0)
        fb.set_solve_graphs(1)
        f, g, H = fa.eval_with_hessian_proj(P.x0)
        fb.eval_with_hessian_proj(P.x0)  # the positions of the coarse space's rotations
        free = np.ones(P.d * P.n_handles, dtype=np.uint8)
        free[: 2 * P.d] = 0
        cases = [dict(coarse_aggregate=0), dict(coarse_aggregate=16), dict(coarse_aggregate=16, w_identity=0.3),
                 dict(coarse_aggregate=-1, free=free)]
        for rep in range(3):
            for kw in cases:
                da, sa = fa.solve(g, H=H, tol=1e-11, **kw)
                db, sb = fb.solve(g, H=H, tol=1e-11, **kw)
                assert sa.converged and sb.converged
                assert sa.iters == sb.iters and np.array_equal(da, db), (P.name, rep, kw)
                # eager work on the device between replays
                fa.eval(P.x0)
                torch.cuda.synchronize()
        fa.close()
        fb.close()
    P = configs.twist_cube(m=5, seed=4)
    fa = tad.ScalarFunction.from_problem(P, ctx)
    fb = tad.ScalarFunction.from_problem(P, ctx)
    fa.set_solve_graphs(0)
    fb.set_solve_graphs(1)
    xa = torch.tensor(P.x0, device="cuda:0")
    xb = xa.clone()
    for it in range(6):
        sa = fa.newton_step(xa)
        sb = fb.newton_step(xb)
        assert sa["cg_iters"] == sb["cg_iters"] and sa["f_new"] == sb["f_new"], it
        assert torch.equal(xa, xb), it


def test_two_level_preconditioner_2d_and_batches(ctx):
    """Triangle meshes (2 translations + 1 rotation per aggregate) and a batch
    of independent disks (aggregates never span two meshes).  Without pinned
    vertices the Hessian is singular up to the eps of the projection along
    the rigid motions, so the check is the residual of the computed solution."""
    for P in [configs.grid_param(60, seed=2), configs.concat(configs.disk_batch(count=5, seed=7, n_min=20,
                                                                                     n_max=40))[0]]:
        fg = tad.ScalarFunction.from_problem(P, ctx)
        f, g, H = fg.eval_with_hessian_proj(P.x0)
        d0, s0 = fg.solve(g, tol=1e-10, coarse_aggregate=0)
        d1, s1 = fg.solve(g, tol=1e-10, coarse_aggregate=64)
        res = np.linalg.norm(H @ d1 - g) / np.linalg.norm(g)
        print(f"n={P.n_vars}: block-Jacobi {s0.iters}, two-level {s1.iters}, residual {res:.2e}")
        assert s1.converged and res <= 1e-9 and s1.iters < s0.iters


def run_newton_test(ctx, iters=10):
    P = configs.newton_test_problem()
    fg = tad.ScalarFunction.from_problem(P, ctx)
    x = P.x0.copy()
    opts = tad.newton_default_opts(solve_tol=1e-14)
    for _ in range(iters):
        fg.newton_step(x, opts)
    return fg, x


def test_newton_test_converges(ctx):
    # NewtonTest.cc:12-95: nnz = 4V + 8(V + F - 1) = 96; f -> 4, |g|_inf -> 0
    fg, x = run_newton_test(ctx)
    assert fg.nnz == 4 * 6 + 8 * (6 + 4 - 1)
    f, g, H = fg.eval_with_hessian_proj(x)
    assert abs(fg.eval(x) - f) == 0.0
    assert abs(f - 4.0) <= 1e-12
    assert np.abs(g).max() <= 1e-10


def test_newton_deterministic(ctx):
    # NewtonTest.cc:97-111: repeated optimizations give identical x
    xs = [run_newton_test(ctx)[1] for _ in range(3)]
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])


def test_newton_matches_oracle_trajectory_grid(ctx):
    # grid parametrization: 8 Newton iterations on GPU vs oracle + direct solves
    import scipy.sparse.linalg as spla
    P = configs.grid_param(20, seed=4)
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    xo = P.x0.copy()
    xg = P.x0.copy()
    opts = tad.newton_default_opts(solve_tol=1e-13)
    for _ in range(8):
        f, g, H = fo.eval_with_hessian_proj(xo)
        d = spla.spsolve(H.tocsc(), -g)
        xo, _, _, _ = fo.line_search(xo, d, f, g)
        fg.newton_step(xg, opts)
    fo_end = fo.eval(xo)
    fg_end = fg.eval(xg)
    assert abs(fg_end - fo_end) <= 1e-8 * abs(fo_end)


def test_grid100_twenty_newton_iterations(ctx):
    """100 x 100 grid parametrization (19,602 triangles): 20 projected-Newton
    iterations on the GPU (PCG 1e-13) and on the oracle (direct solves as the
    SimplicialLDLT stand-in) end at the same energy within 1e-8 relative."""
    import scipy.sparse.linalg as spla
    P = configs.grid_param(100, seed=1)
    assert P.terms[0].conn.shape[0] == 19602
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    xo = P.x0.copy()
    xg = P.x0.copy()
    opts = tad.newton_default_opts(solve_tol=1e-13, solve_max_iters=100000)
    for _ in range(20):
        f, g, H = fo.eval_with_hessian_proj(xo)
        d = spla.spsolve(H.tocsc(), -g)
        xo, _, _, _ = fo.line_search(xo, d, f, g)
        fg.newton_step(xg, opts)
    fo_end = fo.eval(xo)
    fg_end = fg.eval(xg)
    print(f"grid 100: oracle f = {fo_end:.15e}, GPU f = {fg_end:.15e}")
    assert abs(fg_end - fo_end) <= 1e-8 * abs(fo_end)


def test_mixed_newton_energy_matches_fp64(ctx):
    """Mixed mode (FP32 element derivatives, FP64 reductions and solve): after
    10 projected-Newton iterations on a 20^3 twisted cube (48,000 tets) the
    energy equals the FP64 path's within 1e-8 relative; the energy itself is
    FP64 in both modes, so f(x0) agrees exactly."""
    P = configs.twist_cube(m=20, seed=7)
    f64 = tad.ScalarFunction.from_problem(P, ctx, precision="fp64")
    f32 = tad.ScalarFunction.from_problem(P, ctx, precision="mixed")
    assert f64.eval(P.x0) == f32.eval(P.x0)
    opts = tad.newton_default_opts(solve_tol=1e-10)
    xa = P.x0.copy()
    xb = P.x0.copy()
    for _ in range(10):
        sa = f64.newton_step(xa, opts)
        sb = f32.newton_step(xb, opts)
    fa, fb = f64.eval(xa), f64.eval(xb)
    print(f"twist 20: fp64 f = {fa:.15e}, mixed f = {fb:.15e}, rel {abs(fa - fb) / abs(fa):.2e}")
    assert abs(fa - fb) <= 1e-8 * abs(fa)


# ---------------------------------------------------------------------------
# Batches of independent meshes, autotuned launch shapes
# ---------------------------------------------------------------------------

def test_segments_match_per_mesh_oracle(ctx):
    disks = configs.disk_batch(count=5, seed=9, n_min=8, n_max=14)
    P, off = configs.concat(disks)
    fg = tad.ScalarFunction.from_problem(P, ctx, precision="fp64")
    fg.set_segments(off)
    f, g, H = fg.eval_with_hessian_proj(P.x0)
    fs = fg.segment_energies()
    for k, D in enumerate(disks):
        fo = oracle.Function.from_problem(D)
        f1, g1, H1 = fo.eval_with_hessian_proj(D.x0)
        assert abs(fs[k] - f1) <= 1e-12 * abs(f1)
        sl = slice(2 * off[k], 2 * off[k + 1])
        assert nrel(g[sl], g1) <= FP64_TOL
        assert nrel(H[sl, sl].toarray(), H1.toarray()) <= FP64_TOL
    assert abs(fs.sum() - f) <= 1e-12 * abs(f)
    fe = fg.eval(P.x0)
    assert np.array_equal(fg.segment_energies(), fs) and fe == f


@pytest.mark.parametrize("kind", ["cube", "disks"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_bits(ctx, precision, kind):
    """Every element launch shape and every Hessian assembly shape sums in the
    same order: autotuning never changes a bit."""
    if kind == "cube":
        P = perturbed_cube(6, 0.15, 31)
    else:
        P, _ = configs.concat(configs.disk_batch(count=6, seed=5, n_min=10, n_max=30))
    fg = tad.ScalarFunction.from_problem(P, ctx, precision=precision)
    f1, g1, H1 = fg.eval_with_hessian_proj(P.x0)
    rep = fg.autotune(P.x0)
    assert "best" in rep
    f2, g2, H2 = fg.eval_with_hessian_proj(P.x0)
    assert f1 == f2 and np.array_equal(g1, g2) and np.array_equal(H1.data, H2.data)
    tuned = fg.launch_shape(0), fg.assembly_shape
    for shape in range(4):
        fg.assembly_shape = shape
        for ls in [(0, 0, 0), (64, 8, 1), (128, 3, 0), (256, 1024, 1)]:
            fg.set_launch_shape(0, *ls)
            assert fg.launch_shape(0) == ls
            f3, g3, H3 = fg.eval_with_hessian_proj(P.x0)
            assert f1 == f3 and np.array_equal(g1, g3) and np.array_equal(H1.data, H3.data), (shape, ls)
    fg.set_launch_shape(0, *tuned[0])
    fg.assembly_shape = tuned[1]
    for bad in [(32, 8, 0), (64, 0, 0), (64, 12, 1), (64, 8, 2)]:
        with pytest.raises(tad.TadError):
            fg.set_launch_shape(0, *bad)
    with pytest.raises(tad.TadError):
        fg.assembly_shape = 4
    with pytest.raises(tad.TadError):
        fg.set_launch_shape(len(P.terms), 0)
    if P.terms[-1].energy != tad.SYMDIR_TET and P.terms[-1].energy != tad.SYMDIR_TRI:
        with pytest.raises(tad.TadError):
            fg.set_launch_shape(len(P.terms) - 1, 64, 8, 0)
    assert (fg.launch_shape(0), fg.assembly_shape) == tuned

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pipelined_evaluation_keeps_bits(ctx, precision):
    """Pipelined second-order evaluation (element chunks on the context's
    stream, the rows each chunk completes assembled on a second stream):
    identical f, g and H for 1 (off), 2, 3, 4, 8 and 64 chunks, every
    assembly shape, plain and persistent element launches, projected and
    unprojected Hessians, host and device outputs; the plan is kept until the
    elements or the request change."""
    import torch
    P = configs.twist_cube(m=10, seed=3)   # 6,000 tets + pins
    fg = tad.ScalarFunction.from_problem(P, ctx, precision=precision)
    fg.pipeline = 1
    assert fg.pipeline == (1, 1)
    ref = fg.eval_with_hessian_proj(P.x0)
    ref_d = fg.eval_with_derivatives(P.x0)
    xt = torch.tensor(P.x0, device="cuda:0")
    for chunks in (2, 3, 4, 8, 64):
        fg.pipeline = chunks
        assert fg.pipeline == (chunks, min(chunks, 6000 // 256))
        for shape in range(4):
            fg.assembly_shape = shape
            for ls in [(0, 0, 0), (64, 64, 0), (128, 16, 1)]:
                fg.set_launch_shape(0, *ls)
                f, g, H = fg.eval_with_hessian_proj(P.x0)
                assert f == ref[0] and np.array_equal(g, ref[1]) and np.array_equal(H.data, ref[2].data), \
                    (chunks, shape, ls)
        f, g, H = fg.eval_with_derivatives(P.x0)
        assert f == ref_d[0] and np.array_equal(g, ref_d[1]) and np.array_equal(H.data, ref_d[2].data)
        fd, gd, Hd = fg.eval_with_hessian_proj(xt)
        assert fd == ref[0] and np.array_equal(gd.cpu().numpy(), ref[1]) and np.array_equal(Hd.cpu().numpy(), ref[2].data)
    fg.pipeline = 0
    assert fg.pipeline == (0, 1)  # automatic: off
    with pytest.raises(tad.TadError):
        fg.pipeline = 65


# ---------------------------------------------------------------------------
# Line search and Newton decrement (LineSearch.hh:14-65, NewtonDecrement.hh:20-26)
# ---------------------------------------------------------------------------

def quadratic_chain(n=60, seed=3):
    """A strictly convex quadratic over n 2D handles: SPD 4x4 blocks on
    consecutive handle pairs plus unit penalties (catalog QUADRATIC and
    PENALTY elements), so f(x + s d) is an exact parabola in s."""
    rng = np.random.default_rng(seed)
    conn = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)
    rows = []
    for _ in range(n - 1):
        Q, _ = np.linalg.qr(rng.standard_normal((4, 4)))
        A = (Q * rng.uniform(0.5, 3.0, 4)) @ Q.T
        rows.append(quad_params(0.5 * (A + A.T), rng.standard_normal(4), 0.0))
    pins = np.arange(n, dtype=np.int32)[:, None]
    pen = np.concatenate([rng.standard_normal((n, 2)), np.full((n, 1), 0.5)], axis=1)
    x0 = 3.0 * rng.standard_normal(2 * n)
    return conn, np.array(rows), pins, pen, x0


def quadratic_functions(ctx, n=60, seed=3):
    conn, rows, pins, pen, x0 = quadratic_chain(n, seed)
    fg = tad.ScalarFunction(2, n, ctx)
    fg.add_elements(2, tad.QUADRATIC, conn, rows)
    fg.add_elements(1, tad.PENALTY, pins, pen)
    fo = oracle.Function(2, n)
    fo.add_elements(tad.QUADRATIC, conn, rows)
    fo.add_elements(tad.PENALTY, pins, pen)
    return fg, fo, x0


def both_line_searches(fg, fo, x0, d, s_max=1.0):
    """The same (x0, d, f, g) through the GPU search and the oracle's."""
    import torch
    f, g, _ = fo.eval_with_hessian_proj(x0)
    xo, fo_new, so, no = fo.line_search(x0, d, f, g, s_max=s_max)
    dev = torch.device("cuda", 0)
    xt, dt, gt = (torch.tensor(a, device=dev) for a in (x0, d, g))
    xg, fg_new, sg, ng = fg.line_search(xt, dt, f, gt, opts=tad.newton_default_opts(ls_s_max=s_max))
    return (xo, fo_new, so, no), (xg.cpu().numpy(), fg_new, sg, ng), f, g


@pytest.mark.parametrize("alpha,s_max,step,trials", [
    (1.0, 1.0, 1.0, 1),          # the full Newton step of a quadratic
    (1.98, 1.0, 1.0, 1),         # s alpha = 1.98 <= 2 (1 - 1e-4): accepted at once
    (1.98, 2.0, 1.0, 5),         # s_max > 1: 2, 1.6, 1.28, 1.024 rejected, then exactly 1 (LineSearch.hh:40, :55-58)
    (50.0, 1.0, 15, 16),         # 15 shrinks (4 candidate batches of 4)
])
def test_line_search_matches_oracle_on_quadratic(ctx, alpha, s_max, step, trials):
    """Armijo backtracking on an exact quadratic, direction alpha * d_Newton:
    f(x + s d) - f = -s alpha q + (s alpha)^2 q / 2 with q = d_N^T H d_N, so
    the Armijo test holds iff s alpha <= 2 (1 - 1e-4).  The GPU search and the
    oracle accept the same step after the same number of trials; the step is
    the closed-form one; f_new to 1e-12, x_new to 1e-14.  An integer `step`
    means that many shrinks of s = 1 by 0.8, in the search's own arithmetic."""
    import scipy.sparse.linalg as spla
    if isinstance(step, int):
        k, step = step, 1.0
        for _ in range(k):
            step *= 0.8
    fg, fo, x0 = quadratic_functions(ctx)
    f, g, H = fo.eval_with_hessian_proj(x0)
    dN = -spla.spsolve(H.tocsc(), g)
    (xo, fo_new, so, no), (xg, fg_new, sg, ng), f, g = both_line_searches(fg, fo, x0, alpha * dN, s_max)
    print(f"alpha {alpha} s_max {s_max}: oracle step {so} trials {no}, GPU step {sg} trials {ng}")
    assert no == trials and ng == trials
    assert so == step and sg == step
    assert abs(fg_new - fo_new) <= 1e-12 * abs(fo_new)
    assert fg_new <= f + 1e-4 * sg * float(alpha * dN @ g)
    assert nrel(xg, xo) <= 1e-14


def test_line_search_failure_and_errors(ctx):
    """An ascent direction: no step passes Armijo within 64 trials; the search
    returns x0 bit for bit, step 0 and the energy f (LineSearch.hh:62-64),
    as the oracle does.  s_max <= 0 is rejected (LineSearch.hh:36-37); a NaN
    energy is an error (LineSearch.hh:52)."""
    import scipy.sparse.linalg as spla
    import torch
    fg, fo, x0 = quadratic_functions(ctx)
    f, g, H = fo.eval_with_hessian_proj(x0)
    dN = -spla.spsolve(H.tocsc(), g)
    (xo, fo_new, so, no), (xg, fg_new, sg, ng), f, g = both_line_searches(fg, fo, x0, -dN)
    assert no == -64 and ng == -64
    assert so == 0.0 and sg == 0.0 and fg_new == f and fo_new == f
    assert np.array_equal(xg, x0) and np.array_equal(xo, x0)
    dev = torch.device("cuda", 0)
    xt, dt, gt = (torch.tensor(a, device=dev) for a in (x0, dN, g))
    with pytest.raises(tad.TadError) as ei:
        fg.line_search(xt, dt, f, gt, opts=tad.newton_default_opts(ls_s_max=0.0))
    assert ei.value.status == 1
    with pytest.raises(oracle.OracleError):
        fo.line_search(x0, dN, f, g, s_max=0.0)
    dnan = dt.clone()
    dnan[3] = float("nan")
    with pytest.raises(tad.TadError) as ei:
        fg.line_search(xt, dnan, f, gt)
    assert ei.value.status == 5
    with pytest.raises(oracle.OracleError):
        fo.line_search(x0, dnan.cpu().numpy(), f, g)
    with pytest.raises(tad.TadError):  # a direction of the wrong length
        fg.line_search(xt, dt[:-1].contiguous(), f, gt)


def test_line_search_matches_oracle_on_tets(ctx):
    """The projected Newton direction of a twisted cube (symmetric Dirichlet:
    not quadratic, steps < 1 at x0): same accepted step and trial count as the
    oracle, f_new to 1e-12, x_new to 1e-14."""
    import scipy.sparse.linalg as spla
    P = configs.twist_cube(m=6, seed=9)
    fo = oracle.Function.from_problem(P)
    fg = tad.ScalarFunction.from_problem(P, ctx)
    f, g, H = fo.eval_with_hessian_proj(P.x0)
    d = -spla.spsolve(H.tocsc(), g)
    for scale in (1.0, 3.0):
        (xo, fo_new, so, no), (xg, fg_new, sg, ng), f, g = both_line_searches(fg, fo, P.x0, scale * d)
        print(f"tets scale {scale}: oracle step {so} trials {no}, GPU step {sg} trials {ng}")
        assert no == ng and so == sg and no > 0
        assert abs(fg_new - fo_new) <= 1e-12 * abs(fo_new)
        assert nrel(xg, xo) <= 1e-14


def test_newton_decrement(ctx):
    """newton_decrement(d, g) = -0.5 d^T g (NewtonDecrement.hh:20-26): against
    an exactly rounded sum (math.fsum) to 1e-14; for an exact quadratic and
    the exact Newton direction it equals f(x) - f(x + d) (the drop of the
    quadratic model, the reference's definition); and tad_newton_step's
    reported decrement is the one of its own direction, bit for bit."""
    import math
    import scipy.sparse.linalg as spla
    import torch
    fg, fo, x0 = quadratic_functions(ctx)
    rng = np.random.default_rng(1)
    g = rng.standard_normal(fg.n_vars)
    d = -g * rng.uniform(0.1, 2.0, fg.n_vars)
    ref = -0.5 * math.fsum(d * g)
    assert abs(fg.newton_decrement(d, g) - ref) <= 1e-14 * abs(ref)
    dev = torch.device("cuda", 0)
    assert fg.newton_decrement(torch.tensor(d, device=dev), torch.tensor(g, device=dev)) == fg.newton_decrement(d, g)
    f, g, H = fo.eval_with_hessian_proj(x0)
    dN = -spla.spsolve(H.tocsc(), g)
    lam = fg.newton_decrement(dN, g)
    assert abs(lam - (f - fo.eval(x0 + dN))) <= 1e-10 * abs(lam)
    # the Newton step's own decrement
    P = configs.twist_cube(m=5, seed=4)
    ft = tad.ScalarFunction.from_problem(P, ctx)
    xt = torch.tensor(P.x0, device=dev)
    opts = tad.newton_default_opts(solve_tol=1e-12)
    fv, gv, Hv = ft.eval_with_hessian_proj(xt)
    sol, st = ft.solve(gv, H=Hv, tol=1e-12)
    dec = ft.newton_decrement(-sol, gv)
    stn = ft.newton_step(xt.clone(), opts)
    print(f"decrement {dec:.17g}, newton_step {stn['newton_decrement']:.17g}")
    assert stn["newton_decrement"] == dec and dec > 0.0
