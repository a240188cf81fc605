"""Throughput of energies written as device functors against the library's
catalog kernels, on the same problems (not a test: run it directly).

Uses the functors of tests/cpp/device_energies.hip (symmetric Dirichlet
triangles and tets, built into tests/cpp/bin/libtad_test_device_energies.so
by `python -m tinyad_amd.build`).  Prints one JSON line per case and
precision: time per eval_with_hessian_proj, element-kernel time, elements/s,
and whether both paths return the same energy.

    python tests/cpp/device_energies_bench.py [--cube 60] [--grid 100] [--reps 3]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cube", type=int, default=60, help="tet cube cells per side")
    ap.add_argument("--grid", type=int, default=100, help="triangle grid vertices per side")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()

    import torch
    import tinyad_amd as tad
    from tinyad_amd import configs

    plugin = os.path.join(ROOT, "tests", "cpp", "bin", "libtad_test_device_energies.so")
    if not os.path.exists(plugin):
        sys.exit(f"{plugin} is missing: run python -m tinyad_amd.build first")
    L = ctypes.CDLL(plugin)
    ops = {}
    for name, energy in [("symdir_tri", tad.SYMDIR_TRI), ("symdir_tet", tad.SYMDIR_TET)]:
        f = getattr(L, f"tad_test_{name}_ops")
        f.restype = ctypes.c_void_p
        ops[energy] = f()
    ctx = tad.Context(0)

    def device_function(P, precision):
        fn = tad.ScalarFunction(P.d, P.n_handles, ctx, precision)
        for t in P.terms:
            if t.energy in ops:
                fn.add_elements_device(ops[t.energy], t.conn, t.params)
            else:
                fn.add_elements(t.valence, t.energy, t.conn, t.params)
        return fn

    def measure(fn, x, reps):
        f = fn.eval_with_hessian_proj(x)[0]
        ctx.synchronize()
        ctx.reset_stats()
        ctx.enable_timing(True)
        for _ in range(reps):
            fn.eval_with_hessian_proj(x)
        ctx.synchronize()
        ctx.enable_timing(False)
        n_el, el_ms = ctx.kernel_stats("element")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn.eval_with_hessian_proj(x)
        ctx.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, el_ms / max(n_el, 1), f

    cases = [(f"grid_{args.grid}", configs.grid_param(args.grid, seed=3)),
             (f"twist_cube_{args.cube}", configs.twist_cube(m=args.cube, seed=7))]
    for name, P in cases:
        x = torch.tensor(P.x0, device="cuda:0")
        n = P.terms[0].conn.shape[0]
        for prec in ["fp64", "mixed"]:
            row = {"case": name, "precision": prec, "elements": n}
            for kind in ["catalog", "device"]:
                fn = tad.ScalarFunction.from_problem(P, ctx, prec) if kind == "catalog" else device_function(P, prec)
                ms, el_ms, f = measure(fn, x, args.reps)
                row[kind] = {"ms": round(ms, 4), "element_ms": round(el_ms, 4), "M_elements_s": round(n / ms / 1e3, 2),
                             "f": f}
                fn.close()
            row["same_f"] = row["catalog"]["f"] == row["device"]["f"]
            print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
