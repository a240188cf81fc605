"""The benchmark workloads at their full sizes, on the GPU, against the CPU
oracle (the restated reference path, ScalarFunctionImpl.hh:378-416):

* twisted Kuhn cube, 60^3 cells x 6 = 1,296,000 tets, FP64:
  eval_with_hessian_proj at x0 and after 30 projected-Newton iterations
  against the oracle (f 1e-12, g and H_proj 1e-10 normwise, identical CSR
  patterns, CSR value arrays compared directly; at the converged point the
  gradient's error is measured against the gradient scale at x0); the 30
  iterations themselves (NewtonTest.cc:69-75 loop body) checked step by step: Armijo
  with the reported step and decrement (LineSearch.hh:14-24), monotone
  energy, each step starting at the energy the previous one accepted, a
  converged solve, the gradient driven down; two runs bit-identical;
* the same cube in mixed precision (FP32 element derivatives and
  projection, FP64 reductions and solve): g and H_proj within 2e-5 of the
  oracle at x0 and at its own Newton end point, f identical to the FP64
  path, and the energy after 30 iterations within 1e-8 of the FP64 run's;
* 256 radius-clipped height-field disks (n drawn from 20..160, seed 3), one
  function with segments and autotuned launch shapes: per-disk f, g and
  H_proj against a separate oracle function per disk (1e-10 FP64, 2e-5
  mixed), and every element launch shape and assembly shape giving the
  same bits.

The 60^3 oracle evaluation takes about 10 s on 16 host threads and 6 GB of
host memory.
"""
import numpy as np
import pytest

import oracle
import tinyad_amd as tad
from tinyad_amd import configs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 2e-5
NEWTON_ITERS = 30


def nrel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300))


def log(msg):
    print(msg, flush=True)


def compare_with_oracle(fg_res, fo_res, tol, f_tol=1e-12, g_scale=None):
    """f relative, g and the CSR value arrays normwise; patterns identical.
    g_scale: the magnitude g's error is measured against (default max|g|).
    At a converged point every entry of g is a sum of O(1) element
    contributions that cancel to ~1e-17, so its error is judged against the
    gradient scale of the same mesh before convergence instead."""
    f1, g1, H1 = fo_res
    f2, g2, H2 = fg_res
    assert H1.has_sorted_indices and H2.has_sorted_indices
    assert np.array_equal(H2.indptr, H1.indptr) and np.array_equal(H2.indices, H1.indices)
    gs = np.abs(g1).max() if g_scale is None else g_scale
    ef, eg, eh = abs(f2 - f1) / abs(f1), float(np.abs(g2 - g1).max() / gs), nrel(H2.data, H1.data)
    log(f"  f rel {ef:.2e}  g {eg:.2e} (of {gs:.3e})  H_proj {eh:.2e}  (nnz {H1.nnz})")
    assert ef <= f_tol and eg <= tol and eh <= tol
    return ef, eg, eh


def newton_run(fn, x, iters, opts):
    """iters Newton steps on the device tensor x with per-step property checks."""
    sts = []
    for i in range(iters):
        st = fn.newton_step(x, opts)
        dg = -2.0 * st["newton_decrement"]  # d^T g of the step's direction
        assert st["solve_converged"] and not st["ls_failed"], (i, st)
        assert dg < 0.0, (i, st)
        assert st["f_new"] <= st["f"] + opts.ls_armijo * st["step"] * dg, (i, st)  # LineSearch.hh:23
        if sts:
            assert st["f"] == sts[-1]["f_new"], (i, st["f"], sts[-1]["f_new"])
        sts.append(st)
    return sts


@pytest.fixture(scope="module")
def cube60():
    P = configs.twist_cube(m=60, seed=7)
    assert P.terms[0].conn.shape[0] == 1296000 and P.n_handles == 226981
    return P


@pytest.fixture(scope="module")
def cube60_fp64_run(ctx, cube60):
    """The FP64 function and its 30-iteration Newton run from x0 (bench options)."""
    import torch
    fn = tad.ScalarFunction.from_problem(cube60, ctx)
    x = torch.tensor(cube60.x0, device="cuda:0")
    opts = tad.newton_default_opts(solve_tol=1e-10, coarse_refresh=5)
    sts = newton_run(fn, x, NEWTON_ITERS, opts)
    return fn, x, sts, opts


def test_cube60_fp64_eval_matches_oracle(ctx, cube60, cube60_fp64_run):
    P = cube60
    fn = cube60_fp64_run[0]
    log("oracle eval_with_hessian_proj at x0 (1,296,000 tets) ...")
    fo = oracle.Function.from_problem(P)
    ro = fo.eval_with_hessian_proj(P.x0)
    assert fn.nnz == 29845269 == ro[2].nnz
    compare_with_oracle(fn.eval_with_hessian_proj(P.x0), ro, FP64_TOL)
    g0_scale = float(np.abs(ro[1]).max())
    assert fn.eval(P.x0) == ro[0] or abs(fn.eval(P.x0) - ro[0]) <= 1e-12 * abs(ro[0])
    # the unprojected Hessian and the gradient-only path at x0
    compare_with_oracle(fn.eval_with_derivatives(P.x0), fo.eval_with_derivatives(P.x0), FP64_TOL)
    fgr, ggr = fn.eval_with_gradient(P.x0)
    assert abs(fgr - ro[0]) <= 1e-12 * abs(ro[0]) and nrel(ggr, ro[1]) <= FP64_TOL
    # after the 30 Newton iterations: a twisted, converged state
    x30 = cube60_fp64_run[1].cpu().numpy()
    log("oracle eval_with_hessian_proj after 30 Newton iterations ...")
    compare_with_oracle(fn.eval_with_hessian_proj(x30), fo.eval_with_hessian_proj(x30), FP64_TOL, g_scale=g0_scale)


def test_cube60_fp64_newton(ctx, cube60, cube60_fp64_run):
    """30 projected-Newton iterations (the workload's run): every step
    satisfies Armijo, the energy never rises, each step starts at the energy
    the previous one accepted; the gradient falls by 1e-6 or more; a second
    run from x0 gives the same bits and CG counts."""
    import torch
    fn, x, sts, opts = cube60_fp64_run
    fs = [sts[0]["f"]] + [s["f_new"] for s in sts]
    log("f: " + " ".join(f"{v:.10g}" for v in fs[:8]) + " ... " + f"{fs[-1]:.15g}")
    log("cg: " + str([s["cg_iters"] for s in sts]))
    log(f"|g|: {sts[0]['grad_norm']:.3e} -> {sts[-1]['grad_norm']:.3e}, steps {[round(s['step'], 4) for s in sts[:6]]}")
    assert all(b <= a for a, b in zip(fs, fs[1:]))
    assert fs[-1] < 0.05 * fs[0]
    assert sts[-1]["grad_norm"] <= 1e-6 * sts[0]["grad_norm"]
    assert sum(1 for s in sts if s["coarse_reused"] == 0) == NEWTON_ITERS // opts.coarse_refresh
    x2 = torch.tensor(cube60.x0, device="cuda:0")
    sts2 = newton_run(fn, x2, NEWTON_ITERS, opts)
    assert torch.equal(x2, x)
    assert [s["cg_iters"] for s in sts2] == [s["cg_iters"] for s in sts]
    assert [s["f_new"] for s in sts2] == [s["f_new"] for s in sts]


def test_cube60_mixed(ctx, cube60, cube60_fp64_run):
    """Mixed precision on the 60^3 cube: errors against the oracle at x0 and
    at its own Newton end point within 2e-5 (f identical to the FP64 path: the
    energy is FP64 in both modes); the energy after 30 iterations within 1e-8
    of the FP64 run's."""
    import torch
    P = cube60
    f64 = cube60_fp64_run[0]
    fm = tad.ScalarFunction.from_problem(P, ctx, precision="mixed")
    fo = oracle.Function.from_problem(P)
    log("oracle eval_with_hessian_proj at x0 ...")
    ro = fo.eval_with_hessian_proj(P.x0)
    rm = fm.eval_with_hessian_proj(P.x0)
    assert rm[0] == f64.eval_with_hessian_proj(P.x0)[0]
    compare_with_oracle(rm, ro, MIXED_TOL)
    g0_scale = float(np.abs(ro[1]).max())
    x = torch.tensor(P.x0, device="cuda:0")
    sts = newton_run(fm, x, NEWTON_ITERS, cube60_fp64_run[3])
    fa = f64.eval(cube60_fp64_run[1])
    fb = f64.eval(x)
    log(f"after {NEWTON_ITERS} iterations: fp64 f {fa:.15e}, mixed f {fb:.15e}, rel {abs(fa - fb) / abs(fa):.2e}")
    assert abs(fa - fb) <= 1e-8 * abs(fa)
    xm = x.cpu().numpy()
    log("oracle eval_with_hessian_proj at the mixed run's end point ...")
    compare_with_oracle(fm.eval_with_hessian_proj(xm), fo.eval_with_hessian_proj(xm), MIXED_TOL, g_scale=g0_scale)


@pytest.fixture(scope="module")
def disks256():
    disks = configs.disk_batch(count=256, seed=3)
    P, off = configs.concat(disks)
    ns = [D.info["n"] for D in disks]
    assert min(ns) >= 20 and max(ns) <= 160
    log(f"256 disks: {P.terms[0].conn.shape[0]} triangles, {P.n_handles} vertices, n {min(ns)}..{max(ns)}")
    return disks, P, off


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_disks256_match_per_disk_oracles(ctx, disks256, precision):
    """One function over all 256 disks (segments, autotuned shapes) against
    one oracle function per disk: per-disk energies, gradient slices and the
    CSR rows of each disk (their columns shifted by the disk's offset)."""
    disks, P, off = disks256
    tol = FP64_TOL if precision == "fp64" else MIXED_TOL
    fn = tad.ScalarFunction.from_problem(P, ctx, precision=precision)
    fn.set_segments(off)
    r0 = fn.eval_with_hessian_proj(P.x0)
    rep = fn.autotune(P.x0)
    log(rep.splitlines()[0])
    f, g, H = fn.eval_with_hessian_proj(P.x0)
    assert f == r0[0] and np.array_equal(g, r0[1]) and np.array_equal(H.data, r0[2].data)
    fs = fn.segment_energies()
    worst = [0.0, 0.0, 0.0]
    for k, D in enumerate(disks):
        fo = oracle.Function.from_problem(D)
        f1, g1, H1 = fo.eval_with_hessian_proj(D.x0)
        r0_, r1_ = 2 * off[k], 2 * off[k + 1]
        a, b = H.indptr[r0_], H.indptr[r1_]
        assert np.array_equal(H.indptr[r0_:r1_ + 1] - a, H1.indptr)
        assert np.array_equal(H.indices[a:b] - r0_, H1.indices)
        ef = abs(fs[k] - f1) / abs(f1)
        eg = nrel(g[r0_:r1_], g1)
        eh = nrel(H.data[a:b], H1.data)
        worst = [max(worst[0], ef), max(worst[1], eg), max(worst[2], eh)]
        assert ef <= 1e-12 and eg <= tol and eh <= tol, (k, ef, eg, eh)
    log(f"{precision}: worst per-disk f {worst[0]:.2e} g {worst[1]:.2e} H_proj {worst[2]:.2e}")
    assert abs(fs.sum() - f) <= 1e-12 * abs(f)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_disks256_shapes_keep_bits(ctx, disks256, precision):
    """Every assembly shape and a spread of element launch shapes (plain
    grid, persistent grids of 64/128/256 lanes, strided and per-XCD orders)
    give bit-identical f, g and H_proj on the 256-disk batch."""
    _, P, off = disks256
    fn = tad.ScalarFunction.from_problem(P, ctx, precision=precision)
    fn.set_segments(off)
    f1, g1, H1 = fn.eval_with_hessian_proj(P.x0)
    for shape in range(4):
        fn.assembly_shape = shape
        for ls in [(0, 0, 0), (64, 2048, 0), (128, 1024, 1), (256, 512, 1), (64, 8, 0)]:
            fn.set_launch_shape(0, *ls)
            f2, g2, H2 = fn.eval_with_hessian_proj(P.x0)
            assert f2 == f1 and np.array_equal(g2, g1) and np.array_equal(H2.data, H1.data), (shape, ls)
