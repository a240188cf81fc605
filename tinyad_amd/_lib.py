"""ctypes binding of libtinyad_amd.so (the C ABI in include/tinyad_amd.h).

The library is loaded from the package directory only; a missing or
unloadable library raises immediately (there is no fallback path).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtinyad_amd.so")

TAD_OK = 0
STATUS_NAMES = {0: "TAD_OK", 1: "TAD_ERR_INVALID", 2: "TAD_ERR_HIP", 3: "TAD_ERR_NONFINITE",
                4: "TAD_ERR_SOLVE", 5: "TAD_ERR_NAN_ENERGY", 6: "TAD_ERR_UNSUPPORTED"}
MEM_HOST, MEM_DEVICE = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

# Every symbol declared in include/tinyad_amd.h.
EXPORTED = [
    "tad_last_error", "tad_version",
    "tad_ctx_create", "tad_ctx_destroy", "tad_ctx_set_stream", "tad_ctx_get_stream", "tad_ctx_synchronize",
    "tad_ctx_enable_timing", "tad_ctx_kernel_stats", "tad_ctx_reset_stats",
    "tad_fn_create", "tad_fn_destroy", "tad_fn_add_elements", "tad_fn_add_elements_device",
    "tad_fn_add_elements_dynamic_device", "tad_fn_n_terms", "tad_fn_term_info", "tad_fn_n_vars",
    "tad_fn_n_elements",
    "tad_fn_pattern", "tad_fn_eval", "tad_fn_eval_noalloc", "tad_fn_eval_with_gradient", "tad_fn_eval_with_derivatives",
    "tad_fn_eval_with_hessian_proj", "tad_fn_device_buffers",
    "tad_solve_default_opts", "tad_fn_solve", "tad_fn_solve_reduced", "tad_fn_solve_basis", "tad_newton_default_opts", "tad_newton_step",
    "tad_line_search", "tad_newton_decrement",
    "tad_fn_set_segments", "tad_fn_segment_energies", "tad_fn_autotune",
    "tad_fn_get_launch_shape", "tad_fn_set_launch_shape", "tad_fn_get_assembly_shape", "tad_fn_set_assembly_shape",
    "tad_fn_get_pipeline", "tad_fn_set_pipeline", "tad_fn_set_solve_graphs",
    "tad_fn_set_params", "tad_fn_rest_vjp",
    "tad_fn_create_vector", "tad_fn_add_residuals_device", "tad_fn_n_outputs", "tad_fn_jacobian_pattern",
    "tad_fn_eval_residuals", "tad_fn_eval_with_jacobian", "tad_fn_eval_sum_of_squares_with_derivatives",
    "tad_selftest_dual",
]


class TadError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SolveOpts(ctypes.Structure):
    _fields_ = [("w_identity", ctypes.c_double), ("tol", ctypes.c_double), ("max_iters", ctypes.c_int),
                ("check_every", ctypes.c_int), ("coarse_aggregate", ctypes.c_int), ("coarse_reuse", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("converged", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("coarse_unknowns", ctypes.c_int), ("coarse_reused", ctypes.c_int)]


class DeviceOps(ctypes.Structure):
    """tad_device_ops (read-only view; the function pointers are opaque)."""
    _fields_ = [("version", ctypes.c_int), ("valence", ctypes.c_int), ("dim", ctypes.c_int),
                ("n_params", ctypes.c_int), ("element", ctypes.c_void_p), ("energy", ctypes.c_void_p)]


class DeviceResidualOps(ctypes.Structure):
    """tad_device_residual_ops (read-only view; the function pointers are opaque)."""
    _fields_ = [("version", ctypes.c_int), ("valence", ctypes.c_int), ("dim", ctypes.c_int),
                ("n_params", ctypes.c_int), ("n_outputs", ctypes.c_int), ("element", ctypes.c_void_p),
                ("energy", ctypes.c_void_p), ("residuals", ctypes.c_void_p)]


class NewtonOpts(ctypes.Structure):
    _fields_ = [("projection_eps", ctypes.c_double), ("solve", SolveOpts), ("ls_s_max", ctypes.c_double),
                ("ls_shrink", ctypes.c_double), ("ls_max_iters", ctypes.c_int), ("ls_armijo", ctypes.c_double),
                ("ls_batch", ctypes.c_int), ("coarse_refresh", ctypes.c_int)]


class NewtonStats(ctypes.Structure):
    _fields_ = [("f", ctypes.c_double), ("f_new", ctypes.c_double), ("step", ctypes.c_double),
                ("newton_decrement", ctypes.c_double), ("grad_norm", ctypes.c_double), ("cg_iters", ctypes.c_int),
                ("ls_trials", ctypes.c_int), ("ls_failed", ctypes.c_int), ("solve_converged", ctypes.c_int),
                ("coarse_unknowns", ctypes.c_int), ("coarse_reused", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    # PyTorch ships its own HIP runtime with the same SONAME as ROCm's
    # (libamdhip64.so.7).  Loading torch first makes this library bind to
    # that already-loaded runtime, so one process never holds two HIP
    # runtimes (which breaks device discovery for whichever loads second).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tinyad_amd.build` "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, dbl, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
    pvp = ctypes.POINTER(ctypes.c_void_p)
    pi64 = ctypes.POINTER(ctypes.c_int64)
    pdbl = ctypes.POINTER(ctypes.c_double)
    L.tad_last_error.restype = ctypes.c_char_p
    L.tad_version.restype = ctypes.c_char_p
    sig = {
        "tad_ctx_create": [i32, pvp],
        "tad_ctx_destroy": [vp],
        "tad_ctx_set_stream": [vp, vp],
        "tad_ctx_get_stream": [vp, pvp],
        "tad_ctx_synchronize": [vp],
        "tad_ctx_enable_timing": [vp, i32],
        "tad_ctx_kernel_stats": [vp, ctypes.c_char_p, pi64, pdbl],
        "tad_ctx_reset_stats": [vp],
        "tad_fn_create": [vp, i32, i64, i32, pvp],
        "tad_fn_destroy": [vp],
        "tad_fn_add_elements": [vp, i32, i32, i64, vp, vp, i32],
        "tad_fn_add_elements_device": [vp, vp, i64, vp, vp],
        "tad_fn_add_elements_dynamic_device": [vp, vp, i32, i64, vp, vp, vp],
        "tad_fn_n_terms": [vp, vp],
        "tad_fn_term_info": [vp, i32, vp, vp, vp],
        "tad_fn_n_vars": [vp, pi64],
        "tad_fn_n_elements": [vp, pi64],
        "tad_fn_pattern": [vp, pi64, pvp, pvp, vp, vp],
        "tad_fn_eval": [vp, vp, i32, pdbl],
        "tad_fn_eval_noalloc": [vp, vp, vp],
        "tad_fn_create_vector": [vp, i32, i64, i32, vp],
        "tad_fn_add_residuals_device": [vp, vp, i64, vp, vp],
        "tad_fn_n_outputs": [vp, vp],
        "tad_fn_jacobian_pattern": [vp, vp, vp, vp, vp, vp],
        "tad_fn_eval_residuals": [vp, vp, vp, i32],
        "tad_fn_eval_with_jacobian": [vp, vp, vp, vp, i32],
        "tad_fn_eval_sum_of_squares_with_derivatives": [vp, vp, pdbl, vp, vp, vp, i32],
        "tad_fn_eval_with_gradient": [vp, vp, pdbl, vp, i32],
        "tad_fn_eval_with_derivatives": [vp, vp, pdbl, vp, vp, i32],
        "tad_fn_eval_with_hessian_proj": [vp, vp, dbl, pdbl, vp, vp, i32],
        "tad_fn_device_buffers": [vp, pvp, pvp],
        "tad_solve_default_opts": [ctypes.POINTER(SolveOpts)],
        "tad_fn_solve": [vp, vp, vp, vp, i32, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
        "tad_fn_solve_reduced": [vp, vp, vp, vp, i32, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
        "tad_fn_solve_basis": [vp, vp, i64, vp, vp, vp, vp, vp, i32, ctypes.POINTER(SolveOpts),
                               ctypes.POINTER(SolveStats)],
        "tad_newton_default_opts": [ctypes.POINTER(NewtonOpts)],
        "tad_newton_step": [vp, vp, i32, ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonStats)],
        "tad_line_search": [vp, vp, vp, dbl, vp, vp, i32, ctypes.POINTER(NewtonOpts), pdbl, pdbl,
                            ctypes.POINTER(ctypes.c_int)],
        "tad_newton_decrement": [vp, vp, vp, i32, pdbl],
        "tad_fn_set_segments": [vp, i64, vp],
        "tad_fn_segment_energies": [vp, vp, i32],
        "tad_fn_autotune": [vp, vp, i32, dbl, ctypes.c_char_p, ctypes.c_size_t],
        "tad_fn_get_launch_shape": [vp, i32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int)],
        "tad_fn_set_launch_shape": [vp, i32, i32, i32, i32],
        "tad_fn_get_assembly_shape": [vp, ctypes.POINTER(ctypes.c_int)],
        "tad_fn_set_assembly_shape": [vp, i32],
        "tad_fn_get_pipeline": [vp, vp, vp],
        "tad_fn_set_pipeline": [vp, i32],
        "tad_fn_set_solve_graphs": [vp, i32],
        "tad_fn_set_params": [vp, i32, vp, i32],
        "tad_fn_rest_vjp": [vp, i32, vp, vp, vp, dbl, vp, i32],
        "tad_selftest_dual": [vp, i32, dbl, dbl, pdbl],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = L
    return L


def check(status: int):
    if status != TAD_OK:
        raise TadError(status, lib().tad_last_error().decode())
