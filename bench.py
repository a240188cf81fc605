#!/usr/bin/env python3
"""Benchmark of the sparse second-order path on one MI355X.

Workload (default): the twisted Kuhn cube, 60^3 cells x 6 = 1,296,000 tets,
226,981 vertices (n = 680,943 variables, nnz = 29,845,269), symmetric
Dirichlet per tet + quadratic pins on the bottom / top faces, FP64.

Timed phases (inputs resident in HBM, barrier + device sync on both sides):
  A. --steps evaluations of eval_with_hessian_proj at x0
     -> value = M elements/s, ms_per_step
  B. --steps projected-Newton iterations (eval_with_hessian_proj, block-Jacobi
     PCG solve, batched Armijo line search) continuing from x0 after --warmup
     iterations -> newton.iters_per_s
Per-kernel times come from HIP events recorded by the library on its own
stream around every launch inside the timed phases.

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


def pmc_traffic(args, pmc_name):
    """HBM bytes per launch of `pmc_name` from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, scaled by the plain
    512 MiB copy), or None when no summary exists for it."""
    if not pmc_name:
        return None
    tag = f"{args.workload}_{args.m if args.workload == 'twist_cube' else args.disks}_{args.precision}"
    path = os.path.join(ROOT, "profiles", f"traffic_{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == f"`{pmc_name}`":
            return (float(cells[4]) + float(cells[5])) * 1e6
    return None
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--workload", choices=["twist_cube", "disk_batch"], default="twist_cube")
    ap.add_argument("--m", type=int, default=60, help="cube cells per side")
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
    ap.add_argument("--cpu-sample", type=int, default=400000, help="tets in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--newton-iters", type=int, default=30,
                    help="Newton iterations of the converged-energy comparison (mixed precision)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MB (a plain copy kernel of known "
                         "size, used to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    coarse_kw = {} if args.coarse_aggregate is None else {"solve_coarse_aggregate": args.coarse_aggregate}
    if args.coarse_refresh < 1:
        ap.error("--coarse-refresh must be at least 1")

    def newton_steps(f, x, opts, count):
        """count Newton iterations on x; the coarse space is rebuilt on every
        coarse_refresh-th one (counted from the first of this call)."""
        out = []
        for i in range(count):
            opts.solve.coarse_reuse = 0 if i % args.coarse_refresh == 0 else 1
            out.append(f.newton_step(x, opts))
        return out

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is a design note, not implemented)"}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    import tinyad_amd as tad
    from tinyad_amd import configs, _lib

    if args.copy_calibration_mb:
        nel = args.copy_calibration_mb * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration: 3 copies of {args.copy_calibration_mb} MiB")
        del src, dst

    t_setup = time.time()
    seg_off = None
    if args.workload == "disk_batch":
        disks = configs.disk_batch(count=args.disks, seed=3)
        P, seg_off = configs.concat(disks, name=f"disk_batch_{args.disks}")
        args.no_newton = True
    else:
        P = configs.twist_cube(m=args.m, seed=7)
    n_tets = P.terms[0].conn.shape[0]
    log(f"[bench] {P.name}: {n_tets} elements, {P.n_handles} vertices, precision {args.precision}")
    ctx = tad.Context(0)
    fn = tad.ScalarFunction.from_problem(P, ctx, precision=args.precision)
    if seg_off is not None:
        fn.set_segments(seg_off)
    nnz = fn.nnz
    n = fn.n_vars
    dev = torch.device("cuda", 0)
    x0 = torch.from_numpy(P.x0).to(dev)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    L = _lib.lib()
    f = ctypes.c_double()

    def eval_hp(xt):
        _lib.check(L.tad_fn_eval_with_hessian_proj(fn._h, xt.data_ptr(), 1e-9, ctypes.byref(f), None, None,
                                                   _lib.MEM_DEVICE))
        return f.value

    log(f"[bench] setup {time.time() - t_setup:.1f}s, nnz {nnz}")
    tune_report = None
    if args.autotune:
        tune_report = fn.autotune(x0)
        log("[bench] autotune:\n" + tune_report)
    for _ in range(args.warmup):
        eval_hp(x0)
    ctx.synchronize()

    # ---------------- phase A: eval_with_hessian_proj ----------------
    # timed region without per-kernel events; an identical second pass records
    # HIP events around every launch (on the library's stream) for the kernel
    # table and the roofline (events add gaps between kernels, not to them)
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eval_hp(x0)
    ctx.synchronize()
    torch.cuda.synchronize()
    tA = time.perf_counter() - t0
    ctx.reset_stats()
    ctx.enable_timing(True)
    for _ in range(args.steps):
        eval_hp(x0)
    ctx.synchronize()
    ctx.enable_timing(False)
    statsA = {k: ctx.kernel_stats(k) for k in ["element", "assemble_gradient", "assemble_hessian", "energy_sum"]}
    ms_eval = tA / args.steps * 1e3
    melem = n_tets * args.steps / tA / 1e6
    log(f"[bench] phase A: {ms_eval:.3f} ms per eval_with_hessian_proj, {melem:.2f} M elements/s")

    # ---------------- phase B: Newton iterations ----------------
    newton = None
    statsB = {}
    if not args.no_newton:
        opts = tad.newton_default_opts(solve_tol=args.pcg_tol, **coarse_kw)
        x = x0.clone()
        for i, st in enumerate(newton_steps(fn, x, opts, args.warmup)):
            log(f"[bench] warmup newton {i}: f {st['f']:.10g} -> {st['f_new']:.10g} cg {st['cg_iters']} "
                f"ls {st['ls_trials']} step {st['step']:.3g}")
        ctx.set_stream(None)
        x_start = x.clone()
        torch.cuda.synchronize()
        ctx.synchronize()
        t0 = time.perf_counter()
        sts = newton_steps(fn, x, opts, args.steps)
        ctx.synchronize()
        torch.cuda.synchronize()
        tB = time.perf_counter() - t0
        # the same iterations again from the same point with per-kernel events
        # (deterministic: identical CG counts and bits)
        xi = x_start.clone()
        ctx.reset_stats()
        ctx.enable_timing(True)
        sts_i = newton_steps(fn, xi, opts, args.steps)
        ctx.synchronize()
        ctx.enable_timing(False)
        if [s["cg_iters"] for s in sts_i] != [s["cg_iters"] for s in sts] or not torch.equal(xi, x):
            raise RuntimeError("instrumented Newton pass diverged from the timed pass")
        for st in sts:
            log(f"[bench] newton: f {st['f']:.10g} -> {st['f_new']:.10g} cg {st['cg_iters']} ls {st['ls_trials']} "
                f"step {st['step']:.3g}")
        statsB = {k: ctx.kernel_stats(k) for k in ["element", "assemble_gradient", "assemble_hessian", "energy_sum",
                                                   "spmv", "pcg_update", "pcg_coarse", "pcg_setup",
                                                   "coarse_setup", "coarse_galerkin", "coarse_potrf",
                                                   "coarse_potri", "pcg_init", "energy"]}
        n_cg = sum(s["cg_iters"] for s in sts)
        n_ls = sum(s["ls_trials"] for s in sts)
        s_r = 8
        B_eval = n_tets * (4 * 4 + 10 * s_r) + 16 * n + 8 + 8 * nnz
        B_energy = n_tets * (4 * 4 + 10 * s_r) + 16 * n
        B_cg = 12 * nnz + 4 * (n + 1) + 104 * n
        # two-level preconditioner: per CG iteration the restriction (r, lever
        # arms, member list), the dense coarse product and the prolongation (z
        # read and written, lever arms, aggregate ids); per solve the Galerkin
        # sums (H, block indices) and the dense coarse matrix written, factored
        # and inverted in place (counted as four passes over it)
        nc = sts[-1]["coarse_unknowns"]
        nb_blocks = nnz // 9
        B_coarse = (40 * n + 8 * P.n_handles + 8 * nc * nc + 16 * nc) if nc else 0
        B_coarse_setup = (8 * nnz + 24 * nb_blocks + 32 * nc * nc) if nc else 0
        n_rebuilt = sum(1 for s in sts if not s["coarse_reused"])
        B_step_total = (args.steps * B_eval + n_rebuilt * B_coarse_setup + n_cg * (B_cg + B_coarse)
                        + n_ls * B_energy)
        newton = {
            "iters_per_s": args.steps / tB,
            "ms_per_iter": tB / args.steps * 1e3,
            "cg_iters_total": n_cg,
            "ls_trials_total": n_ls,
            "f_final": sts[-1]["f_new"],
            "step_roofline_frac": B_step_total / tB / (HBM_PEAK_GBS * 1e9),
            "step_bytes_formula": ("steps*B_eval_hess + coarse_rebuilds*B_coarse_setup + n_cg*(B_cg + B_coarse) "
                                   "+ n_ls*B_energy (scalar-CSR B_cg)"),
            "coarse_unknowns": nc,
            "coarse_refresh": args.coarse_refresh,
            "coarse_rebuilds": n_rebuilt,
            "cg_iters_per_step": [s["cg_iters"] for s in sts],
        }
        log(f"[bench] phase B: {newton['ms_per_iter']:.2f} ms per Newton iteration, cg {n_cg}, ls {n_ls}")

    # ---------------- roofline of the dominant kernel ----------------
    stats = dict(statsA)
    for k, v in statsB.items():
        a = stats.get(k, (0, 0.0))
        stats[k] = (a[0] + v[0], a[1] + v[1])
    dom = max(stats, key=lambda k: stats[k][1])
    cnt, tot_ms = stats[dom]
    avg_s = tot_ms / max(cnt, 1) / 1e3
    nb = (nnz // 9)
    if dom == "spmv":
        # s = (w I + H) z: CSR values, block columns, block-row offsets, the
        # first scalar-row offset of each block row, z read once, s written
        traffic_alg = 8 * nnz + 4 * nb + 4 * (P.n_handles + 1) + 4 * P.n_handles + 2 * 8 * n
        roof = {"kernel": "k_spmv<3,16>", "bound": "hbm", "achieved": traffic_alg / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "bytes_per_launch": traffic_alg, "pmc_name": "tad::k_spmv<3, 16>"}
    elif dom == "assemble_hessian":
        # contributions read once (n_e N(N+1)/2 upper blocks of d x d in the
        # element precision) and the CSR values written
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
    elif dom == "element":
        Nv = P.terms[0].valence
        r = P.terms[0].params.shape[1]
        per_launch = n_tets * (4 * Nv + r * 8) + 8 * n
        kname = "k_element<SymdirElem<3>>" if Nv == 4 else "k_element<SymdirElem<2>>"
        roof = {"kernel": kname, "bound": "hbm", "achieved": per_launch / avg_s / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_launch": per_launch,
                "note": "VALU-bound kernel; HBM figure from algorithmic input bytes"}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    if roof.get("achieved") is not None:
        roof["frac"] = roof["achieved"] / roof["peak"]
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["launches"] = cnt
    roof["traffic"] = pmc_traffic(args, roof.pop("pmc_name", None))
    kernel_ms = {k: {"launches": v[0], "total_ms": round(v[1], 4)} for k, v in stats.items() if v[0]}

    # ---------------- CPU baseline (oracle on a bounded sample) ----------------
    cpu = None
    if args.cpu_sample > 0 and args.workload == "twist_cube":
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        S = min(args.cpu_sample, n_tets)
        T = P.terms[0]
        fo = oracle.Function(3, P.n_handles)
        fo.add_elements(T.energy, T.conn[:S], T.params[:S])
        t0 = time.time()
        sec = fo.time_eval_with_hessian_proj(P.x0, 1e-9, reps=2)
        default_threads = fo.threads
        # the same sample on every hardware thread the process may use
        n_all = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        fa = oracle.Function(3, P.n_handles, n_threads=n_all)
        fa.add_elements(T.energy, T.conn[:S], T.params[:S])
        sec_all = fa.time_eval_with_hessian_proj(P.x0, 1e-9, reps=2)
        cpu_model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": S / sec / 1e6, "unit": "M elements/s", "cores": default_threads, "kind": "port",
               "sample": f"first {S} of the {n_tets} tets, eval_with_hessian_proj (dense Double<12>, "
                         f"per-element Jacobi eigen-projection, serial triplets + linear-time CSR), median of 2 after warm-up, "
                         f"{default_threads} threads = the reference's default (available threads - 1)",
               "all_threads": {"threads": n_all, "value": S / sec_all / 1e6},
               "host_cpu": cpu_model,
               "wall_s": round(time.time() - t0, 2)}
        log(f"[bench] cpu baseline: {cpu['value']:.4f} M elements/s on {cpu['cores']} threads, "
            f"{S / sec_all / 1e6:.4f} on {n_all}")

    # ---------------- mixed precision: errors against the FP64 path ----------------
    prec = None
    if args.precision == "mixed":
        f64 = tad.ScalarFunction.from_problem(P, ctx, precision="fp64")
        fm, gm, Hm = fn.eval_with_hessian_proj(x0)
        fd, gd, Hd = f64.eval_with_hessian_proj(x0)

        def errs(a, b):
            dlt = (a - b).abs()
            return {"max_rel": float(dlt.max() / b.abs().max()),
                    "rms_rel": float(torch.sqrt((dlt * dlt).mean()) / torch.sqrt((b * b).mean()))}

        prec = {"reference": "FP64 GPU path at x0 (itself equal to the oracle within 1e-14 normwise in the parity "
                             "tests)",
                "f_rel": abs(fm - fd) / abs(fd), "g": errs(gm, gd), "H_proj": errs(Hm, Hd)}
        if newton is not None:
            opts64 = tad.newton_default_opts(solve_tol=args.pcg_tol, **coarse_kw)
            x64 = x0.clone()
            newton_steps(f64, x64, opts64, args.warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st64 = newton_steps(f64, x64, opts64, args.steps)
            ctx.synchronize()
            t64 = time.perf_counter() - t0
            prec["newton_ms_per_iter_fp64"] = t64 / args.steps * 1e3
            prec["newton_speedup_vs_fp64"] = (t64 / args.steps * 1e3) / newton["ms_per_iter"]
            # converged energies: the Newton run of the workload (30 iterations
            # from x0) in both precisions, untimed
            xa, xb = x0.clone(), x0.clone()
            newton_steps(f64, xa, opts64, args.newton_iters)
            newton_steps(fn, xb, opts, args.newton_iters)
            fa, fb = f64.eval(xa), f64.eval(xb)
            prec["converged_newton_iters"] = args.newton_iters
            prec["converged_f_fp64"] = fa
            prec["converged_f_mixed"] = fb
            prec["converged_f_rel"] = abs(fa - fb) / abs(fa)
        log(f"[bench] mixed vs fp64: {json.dumps(prec)}")
        f64.close()

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
                   "autotuned": bool(args.autotune),
                   "launch_shape": list(fn.launch_shape(0)), "assembly_shape": fn.assembly_shape},
        "roofline": roof,
        "cpu_baseline": cpu,
        "newton": newton,
        "mixed_precision": prec,
        "kernels": kernel_ms,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
