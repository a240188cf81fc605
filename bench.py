#!/usr/bin/env python3
"""Benchmark of the sparse second-order path on one MI355X.

Headline workload (default): the twisted Kuhn cube, 60^3 cells x 6 =
1,296,000 tets, 226,981 vertices (n = 680,943 variables, nnz = 29,845,269),
symmetric Dirichlet per tet + quadratic pins on the bottom / top faces, FP64.

Timed phases (inputs resident in HBM, device sync on both sides):
  A. --steps evaluations of eval_with_hessian_proj at x0
     -> value = M elements/s, ms_per_step
  B. --steps projected-Newton iterations (eval_with_hessian_proj, two-level
     PCG solve, batched Armijo line search) continuing from x0 after --warmup
     iterations -> newton.iters_per_s
Per-kernel times come from an identical second pass that records HIP events
on the library's stream around every launch.

The default run also reports, on the same box and in the same JSON line:
  * mixed_precision: the same cube with FP32 element derivatives and
    projection (FP64 reductions and solve): its eval and Newton timings, the
    per-quantity errors against the FP64 path and the energy after
    --newton-iters iterations against the FP64 run's;
  * disk_batch: 256 radius-clipped height-field disks (n from 20..160,
    seed 3) evaluated by one launch per kernel with autotuned shapes, FP32
    elements / FP64 reductions;
  * cpu_baseline: the restated reference path (the oracle, OpenMP) timed on
    this host: eval_with_hessian_proj on a bounded sample of the tets, and a
    Newton iteration (eval + block-Jacobi PCG to the same tolerance + line
    search energies);
  * roofline (the dominant kernel of the timed work, HBM), roofline_element
    (the element kernel, VALU: FP64 operations per launch from the committed
    rocprofv3 counter profile / the live launch time) and eval_hbm (the whole
    eval against the algorithmic bytes).

--workload disk_batch runs the disk batch as the headline, --workload
device_functors compares energies written as device functors with the
catalog's kernels (one JSON line per case).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M elements/s for eval_with_hessian_proj; Newton iters/s at 1.3M tets"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_table(name):
    """Rows of the markdown table in profiles/<name> as lists of cells (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    rows = []
    for line in open(path):
        if line.startswith("|") and not line.startswith("|---"):
            rows.append([c.strip() for c in line.strip().strip("|").split("|")])
    return rows


def pmc_traffic(tag, kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic_<tag>.md
    (FETCH_SIZE and WRITE_SIZE, scaled by the plain-copy calibration)."""
    rows = profile_table(f"traffic_{tag}.md")
    for cells in rows or []:
        if len(cells) >= 6 and cells[0] == f"`{kernel}`":
            return (float(cells[4]) + float(cells[5])) * 1e6
    return None


def pmc_valu(tag, kernel):
    """(FP operations per launch, VALU issue utilization) of `kernel` from
    profiles/valu_<tag>.md, or (None, None)."""
    rows = profile_table(f"valu_{tag}.md")
    for cells in rows or []:
        if len(cells) >= 4 and cells[0] == f"`{kernel}`":
            return float(cells[1]), float(cells[3])
    return None, None


class Timer:
    """Wall time of a region bracketed by device synchronization."""

    def __init__(self, ctx):
        self.ctx = ctx

    def __enter__(self):
        import torch
        torch.cuda.synchronize()
        self.ctx.synchronize()
        self.t0 = time.perf_counter()
        return self

    def __exit__(self, *a):
        import torch
        self.ctx.synchronize()
        torch.cuda.synchronize()
        self.s = time.perf_counter() - self.t0


def kernel_table(ctx, names):
    return {k: ctx.kernel_stats(k) for k in names}


EVAL_KERNELS = ["element", "assemble_gradient", "assemble_hessian", "energy_sum"]
NEWTON_KERNELS = EVAL_KERNELS + ["spmv", "pcg_update", "pcg_coarse", "pcg_setup", "coarse_setup", "coarse_galerkin",
                                 "coarse_inverse", "pcg_init", "energy"]


def eval_phase(ctx, fn, x0, steps, warmup):
    """Timed eval_with_hessian_proj calls (library buffers as outputs), then
    an identical pass with per-kernel events.  Returns (seconds, stats)."""
    from tinyad_amd import _lib
    L = _lib.lib()
    f = ctypes.c_double()

    def ev():
        _lib.check(L.tad_fn_eval_with_hessian_proj(fn._h, x0.data_ptr(), 1e-9, ctypes.byref(f), None, None,
                                                   _lib.MEM_DEVICE))

    for _ in range(warmup):
        ev()
    with Timer(ctx) as t:
        for _ in range(steps):
            ev()
    ctx.reset_stats()
    ctx.enable_timing(True)
    for _ in range(steps):
        ev()
    ctx.synchronize()
    ctx.enable_timing(False)
    return t.s, kernel_table(ctx, EVAL_KERNELS)


def newton_steps(fn, x, opts, count, refresh):
    """count Newton iterations on x; the coarse space is rebuilt on every
    refresh-th one (counted from the first of this call)."""
    out = []
    for i in range(count):
        opts.solve.coarse_reuse = 0 if i % refresh == 0 else 1
        out.append(fn.newton_step(x, opts))
    return out


def newton_phase(ctx, fn, x0, steps, warmup, opts, refresh):
    """warmup untimed iterations from x0, then `steps` timed ones, then the
    same `steps` again from the same point with per-kernel events (checked to
    reproduce the timed pass bit for bit).  Returns (seconds, stats, x_start, sts)."""
    import torch
    x = x0.clone()
    for i, st in enumerate(newton_steps(fn, x, opts, warmup, refresh)):
        log(f"[bench] warmup newton {i}: f {st['f']:.10g} -> {st['f_new']:.10g} cg {st['cg_iters']} "
            f"ls {st['ls_trials']} step {st['step']:.3g}")
    ctx.set_stream(None)
    x_start = x.clone()
    with Timer(ctx) as t:
        sts = newton_steps(fn, x, opts, steps, refresh)
    xi = x_start.clone()
    ctx.reset_stats()
    ctx.enable_timing(True)
    sts_i = newton_steps(fn, xi, opts, steps, refresh)
    ctx.synchronize()
    ctx.enable_timing(False)
    if [s["cg_iters"] for s in sts_i] != [s["cg_iters"] for s in sts] or not torch.equal(xi, x):
        raise RuntimeError("instrumented Newton pass diverged from the timed pass")
    for st in sts:
        log(f"[bench] newton: f {st['f']:.10g} -> {st['f_new']:.10g} cg {st['cg_iters']} ls {st['ls_trials']} "
            f"step {st['step']:.3g}")
    return t.s, kernel_table(ctx, NEWTON_KERNELS), x_start, sts


def step_bytes(P, fn, sts, steps, s_r=8):
    """Algorithmic bytes of the Newton phase: steps * B_eval_hess +
    coarse_rebuilds * B_coarse_setup + n_cg * (B_cg + B_coarse) + n_ls * B_energy
    (B_cg with scalar CSR indices, 12 nnz + 4 (n + 1) + 104 n)."""
    n, nnz = fn.n_vars, fn.nnz
    n_e = sum(t.conn.shape[0] for t in P.terms if t.valence == 4)
    B_eval = n_e * (4 * 4 + 10 * s_r) + 16 * n + 8 + 8 * nnz
    B_energy = n_e * (4 * 4 + 10 * s_r) + 16 * n
    B_cg = 12 * nnz + 4 * (n + 1) + 104 * n
    nc = sts[-1]["coarse_unknowns"]
    B_coarse = (40 * n + 8 * P.n_handles + 8 * nc * nc + 16 * nc) if nc else 0
    B_coarse_setup = (8 * nnz + 24 * (nnz // 9) + 32 * nc * nc) if nc else 0
    n_cg = sum(s["cg_iters"] for s in sts)
    n_ls = sum(s["ls_trials"] for s in sts)
    n_rebuilt = sum(1 for s in sts if not s["coarse_reused"])
    return (steps * B_eval + n_rebuilt * B_coarse_setup + n_cg * (B_cg + B_coarse) + n_ls * B_energy), B_eval


def merged(*tables):
    out = {}
    for t in tables:
        for k, v in t.items():
            a = out.get(k, (0, 0.0))
            out[k] = (a[0] + v[0], a[1] + v[1])
    return out


def run_device_functors(args):
    """Energies written as device functors (tests/cpp/device_energies.hip)
    against the catalog's kernels on the same problems: time per
    eval_with_hessian_proj and element-kernel time, FP64 and mixed."""
    import torch
    import tinyad_amd as tad
    from tinyad_amd import configs

    plugin = os.path.join(ROOT, "tests", "cpp", "bin", "libtad_test_device_energies.so")
    if not os.path.exists(plugin):
        sys.exit(f"{plugin} is missing: run python -m tinyad_amd.build first")
    L = ctypes.CDLL(plugin)
    ops = {}
    # the functors declaring translation invariance (deflated eigensolve, as a
    # geometry energy would declare it)
    for name, energy in [("symdir_tri_inv", tad.SYMDIR_TRI), ("symdir_tet_inv", tad.SYMDIR_TET)]:
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

    rows = []
    for name, P in [(f"grid_param_{args.grid}", configs.grid_param(args.grid, seed=3)),
                    (f"twist_cube_{args.m}", configs.twist_cube(m=args.m, seed=7))]:
        x = torch.tensor(P.x0, device="cuda:0")
        n = P.terms[0].conn.shape[0]
        for prec in ["fp64", "mixed"]:
            row = {"case": name, "precision": prec, "elements": n}
            for kind in ["catalog", "device"]:
                fn = tad.ScalarFunction.from_problem(P, ctx, prec) if kind == "catalog" else device_function(P, prec)
                fv = fn.eval_with_hessian_proj(x)[0]
                sec, st = eval_phase(ctx, fn, x, args.steps, 1)
                ms = sec / args.steps * 1e3
                row[kind] = {"ms": round(ms, 4), "element_ms": round(st["element"][1] / max(st["element"][0], 1), 4),
                             "M_elements_s": round(n / ms / 1e3, 2), "f": fv}
                fn.close()
            row["same_f"] = row["catalog"]["f"] == row["device"]["f"]
            log(json.dumps(row))
            rows.append(row)
    tets = [r for r in rows if r["case"].startswith("twist") and r["precision"] == "fp64"][0]
    out = {"metric": "M elements/s for eval_with_hessian_proj, tets as a device functor (FP64)",
           "value": tets["device"]["M_elements_s"], "unit": "M elements/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_step": tets["device"]["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"device functors vs catalog: grid_param {args.grid}x{args.grid}, "
                                  f"twist_cube m={args.m}"},
           "cases": rows}
    print(json.dumps(out), flush=True)
    ctx.close()


def cpu_baseline(P, n_tets, x0_host, newton_ctx, args):
    """The oracle (restated reference path) on this host.  eval: a bounded
    sample of the tets, reference default thread count (and all threads).
    Newton iteration: the oracle eval time scaled from the sample to the
    whole mesh, plus the oracle's block-Jacobi PCG solved to the same
    tolerance on the full system (the GPU's projected Hessian at the start of
    the timed Newton phase: same values to 1e-10), plus the line-search
    energies at full size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    t_wall = time.time()
    S = min(args.cpu_sample, n_tets)
    T = P.terms[0]
    fo = oracle.Function(3, P.n_handles)
    fo.add_elements(T.energy, T.conn[:S], T.params[:S])
    sec = fo.time_eval_with_hessian_proj(P.x0, 1e-9, reps=2)
    default_threads = fo.threads
    n_all = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    fa = oracle.Function(3, P.n_handles, n_threads=n_all)
    fa.add_elements(T.energy, T.conn[:S], T.params[:S])
    sec_all = fa.time_eval_with_hessian_proj(P.x0, 1e-9, reps=1)
    del fo, fa
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"value": S / sec / 1e6, "unit": "M elements/s", "cores": default_threads, "kind": "port",
           "sample": f"first {S} of the {n_tets} tets, eval_with_hessian_proj (dense Double<12>, per-element "
                     f"Jacobi eigen-projection, serial triplets + linear-time CSR), median of 2 after warm-up, "
                     f"{default_threads} threads = the reference's default (available threads - 1)",
           "all_threads": {"threads": n_all, "value": S / sec_all / 1e6},
           "host_cpu": cpu_model}
    log(f"[bench] cpu eval: {out['value']:.4f} M elements/s on {default_threads} threads, "
        f"{S / sec_all / 1e6:.4f} on {n_all}")
    if newton_ctx is not None:
        H, g, tol, n_ls, gpu_cg = newton_ctx
        fo_full = oracle.Function.from_problem(P)
        t0 = time.perf_counter()
        x_sol, it, rel = oracle.pcg(H, g, block=3, tol=tol, max_iters=20000)
        t_solve = time.perf_counter() - t0
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            fo_full.eval(x0_host)
        t_energy = (time.perf_counter() - t0) / reps
        t_eval = sec * n_tets / S
        t_iter = t_eval + t_solve + n_ls * t_energy
        out["newton"] = {"iters_per_s": 1.0 / t_iter, "ms_per_iter": t_iter * 1e3, "cores": default_threads,
                         "eval_s": t_eval, "solve_s": t_solve, "cg_iters": it, "cg_rel_residual": rel,
                         "energy_s": t_energy, "ls_trials_per_iter": n_ls, "gpu_cg_iters": gpu_cg,
                         "sample": f"one Newton iteration at the full {n_tets}-tet size: eval_with_hessian_proj "
                                   f"time of the {S}-tet sample scaled by {n_tets / S:.3f}, oracle block-Jacobi "
                                   f"PCG to rel. residual {tol:g} on the full projected Hessian at the start of the "
                                   f"timed Newton phase ({it} iterations), {n_ls:g} full-mesh energy evaluations "
                                   f"(the timed phase's mean line-search trials), {default_threads} threads"}
        log(f"[bench] cpu newton: {t_iter * 1e3:.0f} ms per iteration (eval {t_eval:.2f} s, pcg {t_solve:.2f} s "
            f"in {it} iterations, energy {t_energy * 1e3:.1f} ms x {n_ls:g})")
    out["wall_s"] = round(time.time() - t_wall, 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=["twist_cube", "disk_batch", "device_functors"], default="twist_cube")
    ap.add_argument("--m", type=int, default=60, help="cube cells per side")
    ap.add_argument("--grid", type=int, default=100, help="grid vertices per side (device_functors workload)")
    ap.add_argument("--disks", type=int, default=256, help="disks in the disk_batch workload")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick the element-kernel launch shape first (untimed; default on)")
    ap.add_argument("--pcg-tol", type=float, default=1e-10)
    ap.add_argument("--coarse-aggregate", type=int, default=None,
                    help="handles per aggregate of the two-level preconditioner (0: block-Jacobi only; "
                         "default: the library's automatic size)")
    ap.add_argument("--coarse-refresh", type=int, default=5,
                    help="rebuild the coarse space of the direction solve every this many Newton iterations "
                         "(reusing the previous inverse coarse matrix in between; 1: every iteration)")
    ap.add_argument("--solve-graphs", type=int, default=0, choices=[0, 1],
                    help="CG iterations replayed from a recorded HIP graph (1) or launched directly (0)")
    ap.add_argument("--cpu-sample", type=int, default=400000, help="tets in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--extras", action=argparse.BooleanOptionalAction, default=True,
                    help="with the FP64 cube: also the mixed-precision cube and the 256-disk batch (default on)")
    ap.add_argument("--newton-iters", type=int, default=30,
                    help="Newton iterations of the converged-energy comparison (mixed precision)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MB (a plain copy kernel of known "
                         "size, used to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    coarse_kw = {} if args.coarse_aggregate is None else {"solve_coarse_aggregate": args.coarse_aggregate}
    if args.coarse_refresh < 1:
        ap.error("--coarse-refresh must be at least 1")

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is a design note, not implemented)"}), flush=True)
        sys.exit(2)
    if args.workload == "device_functors":
        run_device_functors(args)
        return

    import numpy as np
    import torch
    import tinyad_amd as tad
    from tinyad_amd import configs

    if args.copy_calibration_mb:
        nel = args.copy_calibration_mb * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration: 3 copies of {args.copy_calibration_mb} MiB")
        del src, dst

    def disk_problem():
        disks = configs.disk_batch(count=args.disks, seed=3)
        return configs.concat(disks, name=f"disk_batch_{args.disks}")

    t_setup = time.time()
    seg_off = None
    if args.workload == "disk_batch":
        P, seg_off = disk_problem()
        args.no_newton = True
    else:
        P = configs.twist_cube(m=args.m, seed=7)
    n_tets = P.terms[0].conn.shape[0]
    log(f"[bench] {P.name}: {n_tets} elements, {P.n_handles} vertices, precision {args.precision}")
    ctx = tad.Context(0)
    fn = tad.ScalarFunction.from_problem(P, ctx, precision=args.precision)
    fn.set_solve_graphs(args.solve_graphs)
    if seg_off is not None:
        fn.set_segments(seg_off)
    nnz = fn.nnz
    n = fn.n_vars
    dev = torch.device("cuda", 0)
    x0 = torch.from_numpy(P.x0).to(dev)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    log(f"[bench] setup {time.time() - t_setup:.1f}s, nnz {nnz}")
    if args.autotune:
        log("[bench] autotune:\n" + fn.autotune(x0))

    # ---------------- phase A: eval_with_hessian_proj ----------------
    tA, statsA = eval_phase(ctx, fn, x0, args.steps, args.warmup)
    ms_eval = tA / args.steps * 1e3
    melem = n_tets * args.steps / tA / 1e6
    log(f"[bench] phase A: {ms_eval:.3f} ms per eval_with_hessian_proj, {melem:.2f} M elements/s")

    # ---------------- phase B: Newton iterations ----------------
    newton = None
    statsB = {}
    newton_cpu_ctx = None
    if not args.no_newton:
        opts = tad.newton_default_opts(solve_tol=args.pcg_tol, **coarse_kw)
        tB, statsB, x_start, sts = newton_phase(ctx, fn, x0, args.steps, args.warmup, opts, args.coarse_refresh)
        s_r = 8
        B_step_total, _ = step_bytes(P, fn, sts, args.steps, s_r)
        n_cg = sum(s["cg_iters"] for s in sts)
        n_ls = sum(s["ls_trials"] for s in sts)
        newton = {
            "iters_per_s": args.steps / tB,
            "ms_per_iter": tB / args.steps * 1e3,
            "cg_iters_total": n_cg,
            "ls_trials_total": n_ls,
            "f_final": sts[-1]["f_new"],
            "step_roofline_frac": B_step_total / tB / (HBM_PEAK_GBS * 1e9),
            "step_bytes_formula": ("steps*B_eval_hess + coarse_rebuilds*B_coarse_setup + n_cg*(B_cg + B_coarse) "
                                   "+ n_ls*B_energy (scalar-CSR B_cg)"),
            "coarse_unknowns": sts[-1]["coarse_unknowns"],
            "coarse_refresh": args.coarse_refresh,
            "coarse_rebuilds": sum(1 for s in sts if not s["coarse_reused"]),
            "cg_iters_per_step": [s["cg_iters"] for s in sts],
        }
        log(f"[bench] phase B: {newton['ms_per_iter']:.2f} ms per Newton iteration, cg {n_cg}, ls {n_ls}")
        if args.cpu_sample > 0 and args.workload == "twist_cube":
            # the system of the first timed Newton step, for the CPU solve
            fs, gs, Hs = fn.eval_with_hessian_proj(x_start.cpu().numpy())
            newton_cpu_ctx = (Hs, gs, args.pcg_tol, n_ls / args.steps, sts[0]["cg_iters"])

    # ---------------- rooflines ----------------
    stats = merged(statsA, statsB)
    dom = max(stats, key=lambda k: stats[k][1])
    cnt, tot_ms = stats[dom]
    avg_s = tot_ms / max(cnt, 1) / 1e3
    nb = nnz // 9
    prec_tag = args.precision
    tag = f"{args.workload}_{args.m if args.workload == 'twist_cube' else args.disks}_{prec_tag}"
    if dom == "spmv":
        # s = (w I + H) z from the solver's packed lower triangle: its blocks
        # once (every vertex has its diagonal block here, so n_b + n_v blocks
        # make both triangles), per CSR block its column and its stored-block
        # index, block-row offsets, z read once, s written
        n_sym = (nb + P.n_handles) // 2
        traffic_alg = 72 * n_sym + 8 * nb + 4 * (P.n_handles + 1) + 2 * 8 * n
        roof = {"kernel": "k_spmv_sym<3,8,true>", "bound": "hbm", "achieved": traffic_alg / avg_s / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_launch": traffic_alg,
                "pmc_name": "tad::k_spmv_sym<3, 8, true>"}
    elif dom == "assemble_hessian":
        tsz = 8 if args.precision == "fp64" else 4
        d = P.d
        nblk = sum(t.conn.shape[0] * t.valence * (t.valence + 1) // 2 for t in P.terms)
        per_launch = nblk * d * d * tsz + 8 * nnz
        kbase = ["k_assemble_hessian", "k_assemble_hessian_tpb", "k_assemble_hessian_upper",
                 "k_assemble_hessian_upper_tpb"][fn.assembly_shape]
        tname = "double" if tsz == 8 else "float"
        roof = {"kernel": f"{kbase}<{tname},{d}>", "bound": "hbm", "achieved": per_launch / avg_s / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_launch": per_launch,
                "pmc_name": f"tad::{kbase}<{tname}, {d}>"}
    else:
        Nv = P.terms[0].valence
        r = P.terms[0].params.shape[1]
        per_launch = n_tets * (4 * Nv + r * 8) + 8 * n
        roof = {"kernel": dom, "bound": "hbm", "achieved": per_launch / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "bytes_per_launch": per_launch,
                "note": "element inputs only; the kernel is VALU-bound (roofline_element)"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["launches"] = cnt
    roof["traffic"] = pmc_traffic(tag, roof.pop("pmc_name", None))
    # the element kernel: executed FP operations per launch from the counter
    # profile of this workload, over its live average launch time
    ecnt, ems = statsA["element"]
    e_avg_s = ems / max(ecnt, 1) / 1e3
    Nv = P.terms[0].valence
    ekernel = f"k_element<SymdirElem<{P.d}>,{'double' if args.precision == 'fp64' else 'float'}>"
    flops, valu_util = pmc_valu(tag, ekernel)
    peak_tf = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
    roof_el = {"kernel": ekernel, "bound": "valu", "unit": "TFLOP/s", "peak": peak_tf,
               "avg_launch_ms": e_avg_s * 1e3, "launches": ecnt,
               "achieved": (flops / e_avg_s / 1e12) if flops else None,
               "flops_per_launch": flops, "valu_issue_utilization": valu_util,
               "source": f"profiles/valu_{tag}.md (SQ_INSTS_VALU_{{ADD,MUL,FMA,TRANS}}_F{{64,32}} x 64 lanes, "
                         "FMA counted twice)"}
    roof_el["frac"] = roof_el["achieved"] / peak_tf if roof_el["achieved"] else None
    B_eval = n_tets * (4 * Nv + P.terms[0].params.shape[1] * 8) + 16 * n + 8 + 8 * nnz
    eval_hbm = {"bytes_per_eval": B_eval, "achieved": B_eval / (ms_eval / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": B_eval / (ms_eval / 1e3) / 1e9 / HBM_PEAK_GBS,
                "formula": "n_e (4N + r s_r) + 16 n + 8 + 8 nnz"}
    kernel_ms = {k: {"launches": v[0], "total_ms": round(v[1], 4)} for k, v in stats.items() if v[0]}

    # ---------------- mixed precision (same cube, FP32 elements) ----------------
    prec = None
    extras = args.extras and args.workload == "twist_cube" and args.precision == "fp64"
    if extras or (args.precision == "mixed" and args.workload == "twist_cube"):
        f64, fmx = (fn, tad.ScalarFunction.from_problem(P, ctx, precision="mixed")) if args.precision == "fp64" \
            else (tad.ScalarFunction.from_problem(P, ctx, precision="fp64"), fn)
        f64.set_solve_graphs(args.solve_graphs)
        fmx.set_solve_graphs(args.solve_graphs)
        if args.precision == "fp64":
            if args.autotune:
                fmx.autotune(x0)
            tAm, statsAm = eval_phase(ctx, fmx, x0, args.steps, args.warmup)
        else:
            tAm, statsAm = tA, statsA
        fm_, gm, Hm = fmx.eval_with_hessian_proj(x0)
        fd_, gd, Hd = f64.eval_with_hessian_proj(x0)

        def errs(a, b):
            dlt = (a - b).abs()
            return {"max_rel": float(dlt.max() / b.abs().max()),
                    "rms_rel": float(torch.sqrt((dlt * dlt).mean()) / torch.sqrt((b * b).mean()))}

        prec = {"eval_ms": tAm / args.steps * 1e3, "M_elements_s": n_tets * args.steps / tAm / 1e6,
                "element_kernel_ms": statsAm["element"][1] / max(statsAm["element"][0], 1),
                "reference": "FP64 GPU path at x0 (equal to the oracle within 1e-10 normwise in the full-size "
                             "parity tests)",
                "f_rel": abs(fm_ - fd_) / abs(fd_), "g": errs(gm, gd), "H_proj": errs(Hm, Hd)}
        if newton is not None:
            opts_m = tad.newton_default_opts(solve_tol=args.pcg_tol, **coarse_kw)
            if args.precision == "fp64":
                tBm, _, _, stsm = newton_phase(ctx, fmx, x0, args.steps, args.warmup, opts_m, args.coarse_refresh)
                prec["newton_ms_per_iter"] = tBm / args.steps * 1e3
                prec["newton_iters_per_s"] = args.steps / tBm
                prec["newton_speedup_vs_fp64"] = newton["ms_per_iter"] / prec["newton_ms_per_iter"]
                Bm, _ = step_bytes(P, fmx, stsm, args.steps, 8)
                prec["step_roofline_frac"] = Bm / tBm / (HBM_PEAK_GBS * 1e9)
            xa, xb = x0.clone(), x0.clone()
            o64 = tad.newton_default_opts(solve_tol=args.pcg_tol, **coarse_kw)
            newton_steps(f64, xa, o64, args.newton_iters, args.coarse_refresh)
            newton_steps(fmx, xb, opts_m, args.newton_iters, args.coarse_refresh)
            fa, fb = f64.eval(xa), f64.eval(xb)
            prec["converged_newton_iters"] = args.newton_iters
            prec["converged_f_fp64"] = fa
            prec["converged_f_mixed"] = fb
            prec["converged_f_rel"] = abs(fa - fb) / abs(fa)
        if args.precision == "fp64":
            prec["eval_speedup_vs_fp64"] = ms_eval / prec["eval_ms"]
        log(f"[bench] mixed: {json.dumps(prec)}")
        if args.precision == "fp64":
            fmx.close()
        else:
            f64.close()

    # ---------------- 256-disk batch, FP32 elements ----------------
    disk = None
    if extras:
        Pd, offd = disk_problem()
        nd = Pd.terms[0].conn.shape[0]
        fdk = tad.ScalarFunction.from_problem(Pd, ctx, precision="mixed")
        fdk.set_segments(offd)
        xd = torch.from_numpy(Pd.x0).to(dev)
        if args.autotune:
            fdk.autotune(xd)
        tD, statsD = eval_phase(ctx, fdk, xd, args.steps, args.warmup)
        disk = {"triangles": nd, "disks": args.disks, "vertices": Pd.n_handles, "precision": "mixed",
                "eval_ms": tD / args.steps * 1e3, "M_elements_s": nd * args.steps / tD / 1e6,
                "launch_shape": list(fdk.launch_shape(0)), "assembly_shape": fdk.assembly_shape,
                "kernels_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in statsD.items() if v[0]}}
        log(f"[bench] disk batch: {disk['eval_ms']:.3f} ms, {disk['M_elements_s']:.1f} M elements/s")
        fdk.close()

    # ---------------- CPU baseline (oracle on a bounded sample) ----------------
    cpu = None
    if args.cpu_sample > 0 and args.workload == "twist_cube":
        cpu = cpu_baseline(P, n_tets, P.x0, newton_cpu_ctx, args)
        if newton is not None and "newton" in cpu:
            cpu["newton"]["gpu_speedup"] = newton["iters_per_s"] / cpu["newton"]["iters_per_s"]

    out = {
        "metric": METRIC,
        "value": melem,
        "unit": "M elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_eval,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-element/f64-reduce",
        "data": "synthetic",
        "config": {"workload": (f"twist_cube m={args.m}: {n_tets} tets, {P.n_handles} vertices, n={n}, nnz={nnz}, "
                                "symmetric Dirichlet + pins, eval_with_hessian_proj (eps 1e-9)")
                   if args.workload == "twist_cube" else
                   (f"disk_batch: {args.disks} radius-clipped height-field disks, {n_tets} triangles, "
                    f"{P.n_handles} vertices, nnz={nnz}, one launch per term, per-disk energies"),
                   "precision": args.precision, "pcg_tol": args.pcg_tol,
                   "coarse_aggregate": ("auto" if args.coarse_aggregate is None
                                        and tad.solve_default_opts().coarse_aggregate < 0
                                        else (args.coarse_aggregate if args.coarse_aggregate is not None
                                              else tad.solve_default_opts().coarse_aggregate)),
                   "coarse_refresh": args.coarse_refresh,
                   "solve_graphs": args.solve_graphs,
                   "autotuned": bool(args.autotune),
                   "launch_shape": list(fn.launch_shape(0)), "assembly_shape": fn.assembly_shape},
        "roofline": roof,
        "roofline_element": roof_el,
        "eval_hbm": eval_hbm,
        "cpu_baseline": cpu,
        "newton": newton,
        "mixed_precision": prec,
        "disk_batch": disk,
        "kernels": kernel_ms,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
