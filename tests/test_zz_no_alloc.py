"""tad_fn_eval_noalloc: the energy with no library allocation (partial sums in
a static device array), against tad_fn_eval (eval / operator(),
ScalarFunctionImpl.hh:256-282) on the same inputs.  The two sum in different
orders, so they agree to rounding (1e-13 relative); two calls in a row, the
second on a smaller function, must not see each other's partials."""
import numpy as np
import pytest

import tinyad_amd as tad
from tinyad_amd import configs

pytestmark = pytest.mark.gpu

REL = 1e-13


def noalloc(fn, x):
    import torch
    f = torch.full((1,), float("nan"), dtype=torch.float64, device=x.device)
    fn.eval_noalloc(x, f)
    fn.ctx.synchronize()
    return float(f.cpu()[0])


def close(a, b):
    return a == b or abs(a - b) <= REL * abs(b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_eval_twice_then_smaller(ctx, precision):
    import torch
    dev = torch.device("cuda", 0)
    big = configs.twist_cube(m=12, seed=5)       # 10,368 tets + pins: 41 + 2 partial sums
    small = configs.twist_cube(m=3, seed=6)      # 162 tets + pins
    fb = tad.ScalarFunction.from_problem(big, ctx, precision=precision)
    fs = tad.ScalarFunction.from_problem(small, ctx, precision=precision)
    xb = torch.tensor(big.x0, device=dev)
    xs = torch.tensor(small.x0, device=dev)
    for fn, x in [(fb, xb), (fs, xs), (fb, xb)]:
        want = fn.eval(x)
        got = noalloc(fn, x)
        print(f"{precision} n_elem {fn.n_elements}: eval {want!r} noalloc {got!r}")
        assert close(got, want), (got, want)
    # a second point of the large function, then the small one again
    x2 = xb + 0.01 * torch.sin(torch.arange(xb.numel(), device=dev, dtype=torch.float64))
    assert close(noalloc(fb, x2), fb.eval(x2))
    assert close(noalloc(fs, xs), fs.eval(xs))


def test_noalloc_triangles_quadratics_and_inf(ctx):
    """Grid triangles with penalties, dense quadratic terms, an empty
    function, and an inverted element (+inf energy, as eval gives)."""
    import torch
    dev = torch.device("cuda", 0)
    P = configs.grid_param(24, seed=2)
    fn = tad.ScalarFunction.from_problem(P, ctx)
    x = torch.tensor(P.x0, device=dev)
    assert close(noalloc(fn, x), fn.eval(x))
    xi = P.x0.copy()
    a, b, c = P.terms[0].conn[7]
    xi[2 * b:2 * b + 2], xi[2 * c:2 * c + 2] = P.x0[2 * c:2 * c + 2], P.x0[2 * b:2 * b + 2]
    xit = torch.tensor(xi, device=dev)
    assert fn.eval(xit) == float("inf") and noalloc(fn, xit) == float("inf")
    # dense quadratics over pairs of 2D handles plus unit penalties
    rng = np.random.default_rng(4)
    n = 700
    conn = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)
    A = rng.standard_normal((n - 1, 4, 4))
    rows = np.concatenate([(A + A.transpose(0, 2, 1)).reshape(n - 1, 16), rng.standard_normal((n - 1, 5))], axis=1)
    fq = tad.ScalarFunction(2, n, ctx)
    fq.add_elements(2, tad.QUADRATIC, conn, rows)
    fq.add_elements(1, tad.PENALTY, np.arange(n, dtype=np.int32)[:, None],
                    np.concatenate([rng.standard_normal((n, 2)), np.full((n, 1), 0.5)], axis=1))
    xq = torch.tensor(rng.standard_normal(2 * n), device=dev)
    assert close(noalloc(fq, xq), fq.eval(xq))
    fe = tad.ScalarFunction(2, 4, ctx)
    assert noalloc(fe, torch.zeros(8, dtype=torch.float64, device=dev)) == 0.0


def test_noalloc_rejects_sizes_beyond_the_scratch(ctx):
    """16,777,217 elements need 65,537 partial sums: rejected before any
    launch with a message naming the limit; 16,777,216 still run."""
    import torch
    dev = torch.device("cuda", 0)
    n_h = 64
    for n_e, ok in [(16777217, False), (16777216, True)]:
        fn = tad.ScalarFunction(1, n_h, ctx)
        conn = (np.arange(n_e, dtype=np.int64) % n_h).astype(np.int32)[:, None]
        params = np.empty((n_e, 2))
        params[:, 0] = 1.0
        params[:, 1] = 0.25
        fn.add_elements(1, tad.PENALTY, conn, params)
        del conn, params
        x = torch.zeros(n_h, dtype=torch.float64, device=dev)
        if ok:
            got = noalloc(fn, x)
            assert got == 0.25 * n_e
        else:
            with pytest.raises(tad.TadError) as ei:
                noalloc(fn, x)
            assert ei.value.status == 6 and "65536" in str(ei.value)
        fn.close()
