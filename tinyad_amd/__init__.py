"""tinyad_amd -- MI355X (gfx950) sparse second-order path with TinyAD's
ScalarFunction interface.

Python mirror of the reference's C++ API (include/TinyAD/ScalarFunction.hh):

    import tinyad_amd as tad
    func = tad.scalar_function(2, n_handles)                 # scalar_function<2>(range(n))
    func.add_elements(3, tad.SYMDIR_TRI, F, params)          # add_elements<3>(range(nF), lambda)
    x = func.x_from_data(V)                                  # x_from_data
    f, g, H = func.eval_with_hessian_proj(x)                 # eval_with_hessian_proj

Element functions come from the energy catalog (tad.SYMDIR_TRI, SYMDIR_TET,
PENALTY, QUADRATIC) with per-element parameter rows instead of a C++ lambda.
Arrays may be numpy (host) or torch CUDA tensors (device-resident: nothing
crosses PCIe).  All compute runs in libtinyad_amd.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import TadError, lib, check, NewtonOpts, NewtonStats, SolveOpts, SolveStats

SYMDIR_TRI = 1
SYMDIR_TET = 2
PENALTY = 3
QUADRATIC = 4
DEVICE = 100  # terms added from device energies

DEFAULT_HESSIAN_PROJECTION_EPS = 1e-9  # HessianProjection.hh:16

__all__ = ["Context", "ScalarFunction", "scalar_function", "TadError", "SYMDIR_TRI", "SYMDIR_TET", "PENALTY",
           "QUADRATIC", "newton_default_opts", "solve_default_opts", "version"]


def version() -> str:
    return lib().tad_version().decode()


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _order_after_torch(ctx, device):
    """Queue library work behind torch's producer kernels: share torch's
    current stream, or (torch's legacy default stream) wait for it."""
    import torch
    st = torch.cuda.current_stream(device)
    if st.cuda_stream:
        ctx.set_stream(st.cuda_stream)
    else:
        st.synchronize()
        ctx.set_stream(None)


class Context:
    """A device plus one HIP stream (tad_ctx)."""

    _default = None

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().tad_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @classmethod
    def default(cls):
        if cls._default is None:
            import atexit
            cls._default = Context(0)
            atexit.register(cls._default.close)
        return cls._default

    def close(self):
        if getattr(self, "_h", None):
            lib().tad_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(lib().tad_ctx_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def set_stream(self, stream_ptr: int | None):
        check(lib().tad_ctx_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))

    def synchronize(self):
        check(lib().tad_ctx_synchronize(self._h))

    def enable_timing(self, on: bool = True):
        check(lib().tad_ctx_enable_timing(self._h, 1 if on else 0))

    def reset_stats(self):
        check(lib().tad_ctx_reset_stats(self._h))

    def kernel_stats(self, name: str):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(lib().tad_ctx_kernel_stats(self._h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value


def solve_default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    check(lib().tad_solve_default_opts(ctypes.byref(o)))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def newton_default_opts(**kw) -> NewtonOpts:
    o = NewtonOpts()
    check(lib().tad_newton_default_opts(ctypes.byref(o)))
    for k, v in kw.items():
        if k.startswith("solve_"):
            setattr(o.solve, k[6:], v)
        else:
            setattr(o, k, v)
    return o


class ScalarFunction:
    """ScalarFunction<d, double, int> backed by the HIP library."""

    def __init__(self, d: int, n_handles: int, ctx: Context | None = None, precision: str = "fp64"):
        self.ctx = ctx or Context.default()
        self.d = int(d)
        self.n_handles = int(n_handles)
        self.n_vars = self.d * self.n_handles
        self.precision = precision
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        h = ctypes.c_void_p()
        check(lib().tad_fn_create(self.ctx._h, self.d, self.n_handles, prec, ctypes.byref(h)))
        self._h = h
        self.n_elements = 0
        self._pattern = None

    def close(self):
        if getattr(self, "_h", None):
            lib().tad_fn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- construction ------------------------------------------------------
    def add_elements(self, valence: int, energy: int, conn, params):
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, valence)
        params = np.ascontiguousarray(params, dtype=np.float64)
        if params.ndim == 1:
            params = np.ascontiguousarray(np.broadcast_to(params, (conn.shape[0], params.size)))
        check(lib().tad_fn_add_elements(self._h, energy, valence, conn.shape[0], conn.ctypes.data,
                                        params.ctypes.data, params.shape[1]))
        self.n_elements += conn.shape[0]
        self._pattern = None

    def add_elements_device(self, ops, conn, params=None):
        """add_elements<N>(range, lambda) for an energy compiled on the device.

        ops: address of a tad_device_ops (an int or ctypes pointer), e.g. the
        return value of a function in the caller's hipcc-built library that
        returns &ops for tinyad_amd::device_element<F, N, d, NP>()
        (include/tinyad_amd/DeviceElements.hh).  conn: n_elem x N handles;
        params: n_elem x NP rows (one row is broadcast)."""
        addr = ctypes.cast(ops, ctypes.c_void_p).value if not isinstance(ops, int) else ops
        if not addr:
            raise TadError(1, "null device ops")
        o = _lib.DeviceOps.from_address(addr)
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, o.valence)
        n = conn.shape[0]
        if o.n_params > 0:
            params = np.ascontiguousarray(params, dtype=np.float64)
            if params.ndim == 1:
                params = np.ascontiguousarray(np.broadcast_to(params, (n, params.size)))
            if params.shape != (n, o.n_params):
                raise TadError(1, f"device energy expects {o.n_params} parameters per element")
            pp = params.ctypes.data
        else:
            pp = None
        check(lib().tad_fn_add_elements_device(self._h, addr, n, conn.ctypes.data, pp))
        self.n_elements += n
        self._pattern = None

    def add_elements_dynamic_device(self, ops_list, elements, params=None):
        """add_elements_dynamic<V1, V2, ...>(range, lambda) (ScalarFunctionImpl.hh:102-214)
        for a device energy compiled at several valences.

        ops_list: tad_device_ops addresses, one per valence; elements: a
        sequence of handle lists (any lengths up to the largest valence);
        params: n_elem x NP rows (one row is broadcast).  Each element joins the
        smallest valence that holds it; one term per non-empty group."""
        addrs = [ctypes.cast(o, ctypes.c_void_p).value if not isinstance(o, int) else o for o in ops_list]
        if not addrs or not all(addrs):
            raise TadError(1, "null device ops")
        np_ = _lib.DeviceOps.from_address(addrs[0]).n_params
        lens = np.array([len(e) for e in elements], dtype=np.int64)
        ptr = np.zeros(lens.size + 1, dtype=np.int64)
        np.cumsum(lens, out=ptr[1:])
        flat = np.ascontiguousarray(np.concatenate([np.asarray(e, dtype=np.int32).ravel() for e in elements])
                                    if lens.sum() else np.zeros(1, dtype=np.int32), dtype=np.int32)
        n = lens.size
        if np_ > 0:
            params = np.ascontiguousarray(params, dtype=np.float64)
            if params.ndim == 1:
                params = np.ascontiguousarray(np.broadcast_to(params, (n, params.size)))
            if params.shape != (n, np_):
                raise TadError(1, f"device energy expects {np_} parameters per element")
            pp = params.ctypes.data
        else:
            pp = None
        arr = (ctypes.c_void_p * len(addrs))(*addrs)
        check(lib().tad_fn_add_elements_dynamic_device(self._h, arr, len(addrs), n, ptr.ctypes.data,
                                                       flat.ctypes.data, pp))
        self.n_elements += n
        self._pattern = None

    def terms(self):
        """(energy, valence, n_elements) per objective term, in insertion order
        (the reference's objective_terms, ScalarFunction.hh:229)."""
        n = ctypes.c_int()
        check(lib().tad_fn_n_terms(self._h, ctypes.byref(n)))
        out = []
        for t in range(n.value):
            en, va, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
            check(lib().tad_fn_term_info(self._h, t, ctypes.byref(en), ctypes.byref(va), ctypes.byref(ne)))
            out.append((en.value, va.value, ne.value))
        return out

    @classmethod
    def from_problem(cls, problem, ctx=None, precision="fp64"):
        f = cls(problem.d, problem.n_handles, ctx, precision)
        for t in problem.terms:
            f.add_elements(t.valence, t.energy, t.conn, t.params)
        return f

    def x_from_data(self, data) -> np.ndarray:
        """x_from_data (ScalarFunctionImpl.hh:216-235): rows of `data` are the handle values."""
        x = np.ascontiguousarray(data, dtype=np.float64).reshape(self.n_handles, self.d).ravel().copy()
        if not np.all(np.isfinite(x)):
            raise TadError(1, "x_from_data: non-finite entries")
        return x

    def x_to_data(self, x) -> np.ndarray:
        """x_to_data (ScalarFunctionImpl.hh:237-254)."""
        x = np.asarray(x, dtype=np.float64)
        if x.size != self.n_vars:
            raise TadError(1, "x_to_data: size mismatch")
        return x.reshape(self.n_handles, self.d).copy()

    # -- pattern -----------------------------------------------------------
    def pattern(self):
        """(row_ptr, col_idx) host copies of the Hessian CSR pattern."""
        if self._pattern is None:
            nnz = ctypes.c_int64()
            check(lib().tad_fn_pattern(self._h, ctypes.byref(nnz), None, None, None, None))
            rp = np.zeros(self.n_vars + 1, dtype=np.int32)
            ci = np.zeros(nnz.value, dtype=np.int32)
            check(lib().tad_fn_pattern(self._h, None, None, None, rp.ctypes.data, ci.ctypes.data))
            self._pattern = (rp, ci)
        return self._pattern

    @property
    def nnz(self) -> int:
        return int(self.pattern()[1].size)

    def pattern_device(self):
        """(nnz, row_ptr_dev, col_idx_dev) raw device pointers."""
        nnz = ctypes.c_int64()
        rp = ctypes.c_void_p()
        ci = ctypes.c_void_p()
        check(lib().tad_fn_pattern(self._h, ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci), None, None))
        return nnz.value, rp.value, ci.value

    # -- evaluation --------------------------------------------------------
    def _x(self, x):
        if _is_torch(x):
            import torch
            if x.device.type != "cuda" or x.dtype != torch.float64 or not x.is_contiguous():
                raise TadError(1, "device x must be a contiguous float64 CUDA tensor")
            if x.numel() != self.n_vars:
                raise TadError(1, f"x has {x.numel()} entries, expected {self.n_vars}")
            _order_after_torch(self.ctx, x.device)
            return x.data_ptr(), _lib.MEM_DEVICE
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.n_vars:
            raise TadError(1, f"x has {x.size} entries, expected {self.n_vars}")
        self._xkeep = x
        return x.ctypes.data, _lib.MEM_HOST

    def eval(self, x) -> float:
        ptr, mem = self._x(x)
        f = ctypes.c_double()
        check(lib().tad_fn_eval(self._h, ptr, mem, ctypes.byref(f)))
        return f.value

    __call__ = eval

    def eval_noalloc(self, x, f_out) -> None:
        """tad_fn_eval_noalloc: the energy at device x into the device scalar
        f_out (a float64 CUDA tensor with at least one entry), with no
        allocation by the library.  Asynchronous on the context's stream."""
        import torch
        ptr, mem = self._x(x)
        if mem != _lib.MEM_DEVICE:
            raise TadError(1, "eval_noalloc needs x on the device")
        if (not isinstance(f_out, torch.Tensor) or f_out.device.type != "cuda" or f_out.dtype != torch.float64
                or f_out.numel() < 1):
            raise TadError(1, "f_out must be a float64 CUDA tensor with at least one entry")
        check(lib().tad_fn_eval_noalloc(self._h, ptr, f_out.data_ptr()))

    def _outputs(self, mem, with_h):
        if mem == _lib.MEM_DEVICE:
            import torch
            dev = torch.device("cuda", self.ctx.device)
            g = torch.empty(self.n_vars, dtype=torch.float64, device=dev)
            H = torch.empty(self.nnz, dtype=torch.float64, device=dev) if with_h else None
            return g, H, g.data_ptr(), (H.data_ptr() if with_h else None)
        g = np.zeros(self.n_vars)
        H = np.zeros(self.nnz) if with_h else None
        return g, H, g.ctypes.data, (H.ctypes.data if with_h else None)

    def _wrap_h(self, H, mem):
        if mem == _lib.MEM_DEVICE:
            return H
        import scipy.sparse as sp
        rp, ci = self.pattern()
        return sp.csr_matrix((H, ci, rp), shape=(self.n_vars, self.n_vars))

    def eval_with_gradient(self, x):
        ptr, mem = self._x(x)
        g, _, gp, _ = self._outputs(mem, False)
        f = ctypes.c_double()
        check(lib().tad_fn_eval_with_gradient(self._h, ptr, ctypes.byref(f), gp, mem))
        return f.value, g

    def eval_with_derivatives(self, x):
        ptr, mem = self._x(x)
        g, H, gp, hp = self._outputs(mem, True)
        f = ctypes.c_double()
        check(lib().tad_fn_eval_with_derivatives(self._h, ptr, ctypes.byref(f), gp, hp, mem))
        return f.value, g, self._wrap_h(H, mem)

    def eval_hessian(self, x):
        return self.eval_with_derivatives(x)[2]

    def eval_hessian_of_quadratic(self):
        """Hessian at x = 0 (ScalarFunctionImpl.hh:369-376)."""
        return self.eval_hessian(np.zeros(self.n_vars))

    def eval_with_hessian_proj(self, x, eps: float = DEFAULT_HESSIAN_PROJECTION_EPS):
        ptr, mem = self._x(x)
        g, H, gp, hp = self._outputs(mem, True)
        f = ctypes.c_double()
        check(lib().tad_fn_eval_with_hessian_proj(self._h, ptr, float(eps), ctypes.byref(f), gp, hp, mem))
        return f.value, g, self._wrap_h(H, mem)

    # -- parameters and rest-shape products ---------------------------------
    def set_params(self, term: int, params):
        """Replaces the parameter rows of term `term` (the pattern is kept)."""
        if _is_torch(params):
            import torch
            params = params.detach().to(torch.float64).contiguous()
            _order_after_torch(self.ctx, params.device)
            check(lib().tad_fn_set_params(self._h, int(term), params.data_ptr(), _lib.MEM_DEVICE))
            return
        params = np.ascontiguousarray(params, dtype=np.float64)
        check(lib().tad_fn_set_params(self._h, int(term), params.ctypes.data, _lib.MEM_HOST))

    def rest_vjp(self, term: int, x, rest, lam, w_scale: float):
        """sum_e (d2E_e / dtheta dx)^T lam for a symmetric-Dirichlet tet term whose
        parameters come from rest positions theta (w_e = w_scale det(Mr_e) / 6)."""
        if _is_torch(x):
            import torch
            xs = [self._x(x)[0]]
            for a in (rest, lam):
                if not (_is_torch(a) and a.dtype == torch.float64 and a.is_contiguous() and a.numel() == self.n_vars):
                    raise TadError(1, "rest and lam must be contiguous float64 CUDA tensors like x")
            out = torch.empty_like(x)
            check(lib().tad_fn_rest_vjp(self._h, int(term), xs[0], rest.data_ptr(), lam.data_ptr(), float(w_scale),
                                        out.data_ptr(), _lib.MEM_DEVICE))
            return out
        ptr, _ = self._x(x)
        rest = np.ascontiguousarray(rest, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        if rest.size != self.n_vars or lam.size != self.n_vars:
            raise TadError(1, "rest and lam must have n_vars entries")
        out = np.zeros(self.n_vars)
        check(lib().tad_fn_rest_vjp(self._h, int(term), ptr, rest.ctypes.data, lam.ctypes.data, float(w_scale),
                                    out.ctypes.data, _lib.MEM_HOST))
        return out

    # -- batches of independent problems ----------------------------------
    def set_segments(self, handle_offsets):
        """Declares handle ranges [off[s], off[s+1]) as independent problems."""
        off = np.ascontiguousarray(handle_offsets, dtype=np.int64)
        check(lib().tad_fn_set_segments(self._h, off.size - 1, off.ctypes.data))
        self.n_segments = off.size - 1

    def segment_energies(self) -> np.ndarray:
        """Per-segment energies of the last evaluation."""
        f = np.zeros(self.n_segments)
        check(lib().tad_fn_segment_energies(self._h, f.ctypes.data, _lib.MEM_HOST))
        return f

    def autotune(self, x, eps: float = DEFAULT_HESSIAN_PROJECTION_EPS) -> str:
        """Picks the fastest element-kernel launch shape on x; returns the timing table."""
        ptr, mem = self._x(x)
        buf = ctypes.create_string_buffer(16384)
        check(lib().tad_fn_autotune(self._h, ptr, mem, float(eps), buf, len(buf)))
        return buf.value.decode()

    def launch_shape(self, term: int):
        """(block, grid, map) of term's projected-Hessian element launch (block 0: plain grid)."""
        b, g, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().tad_fn_get_launch_shape(self._h, int(term), ctypes.byref(b), ctypes.byref(g), ctypes.byref(m)))
        return b.value, g.value, m.value

    def set_launch_shape(self, term: int, block: int, grid: int = 0, map: int = 0):
        check(lib().tad_fn_set_launch_shape(self._h, int(term), int(block), int(grid), int(map)))

    @property
    def assembly_shape(self) -> int:
        """Hessian assembly shape 0..3 (bit-identical results; the autotuner picks one)."""
        v = ctypes.c_int()
        check(lib().tad_fn_get_assembly_shape(self._h, ctypes.byref(v)))
        return v.value

    @assembly_shape.setter
    def assembly_shape(self, shape: int):
        check(lib().tad_fn_set_assembly_shape(self._h, int(shape)))

    @property
    def pipeline(self) -> tuple:
        """(requested, used) element chunks of the pipelined second-order
        evaluation (requested 0: automatic, 1: off); identical results for all."""
        r, u = ctypes.c_int(), ctypes.c_int()
        check(lib().tad_fn_get_pipeline(self._h, ctypes.byref(r), ctypes.byref(u)))
        return r.value, u.value

    @pipeline.setter
    def pipeline(self, chunks: int):
        check(lib().tad_fn_set_pipeline(self._h, int(chunks)))

    def set_solve_graphs(self, mode: int):
        """CG iterations replayed from a recorded HIP graph: 1 on, 0 off (the
        default); identical results."""
        check(lib().tad_fn_set_solve_graphs(self._h, int(mode)))

    # -- solver ------------------------------------------------------------
    def _dev_vec(self, a, n: int, name: str):
        """a checked as a contiguous float64 CUDA tensor of n entries on the context's device."""
        import torch
        if not (_is_torch(a) and a.device.type == "cuda" and a.dtype == torch.float64 and a.is_contiguous()):
            raise TadError(1, f"{name} must be a contiguous float64 CUDA tensor")
        if a.device.index != self.ctx.device:
            raise TadError(1, f"{name} is on cuda:{a.device.index}, the function on cuda:{self.ctx.device}")
        if a.numel() != n:
            raise TadError(1, f"{name} has {a.numel()} entries, expected {n}")
        return a

    def _host_vec(self, a, n: int, name: str):
        """a as a contiguous float64 numpy array of n entries (a sparse matrix gives its values)."""
        if hasattr(a, "tocsr") and not isinstance(a, np.ndarray):
            a = a.data
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != n:
            raise TadError(1, f"{name} has {a.size} entries, expected {n}")
        return a.ravel()

    def solve(self, rhs, H=None, free=None, **opts):
        """Solves (w I + H) sol = rhs with the library's PCG; H=None uses the last evaluated Hessian.

        rhs: n_vars values (numpy, or a float64 CUDA tensor: then H, if
        given, is the CUDA tensor of the nnz CSR values); H on the host: the
        nnz values in the pattern order or a scipy matrix over the pattern.
        free: optional per-variable mask (nonzero = free) of a reduced-basis
        solve (newton_direction_reduced_basis with a selection basis): CG on
        the free variables, sol = 0 on the fixed ones."""
        o = solve_default_opts(**opts)
        st = SolveStats()
        if _is_torch(rhs):
            import torch
            self._dev_vec(rhs, self.n_vars, "rhs")
            hp = self._dev_vec(H, self.nnz, "H").data_ptr() if H is not None else None
            _order_after_torch(self.ctx, rhs.device)
            sol = torch.empty_like(rhs)
            if free is None:
                check(lib().tad_fn_solve(self._h, hp, rhs.data_ptr(), sol.data_ptr(), _lib.MEM_DEVICE,
                                         ctypes.byref(o), ctypes.byref(st)))
            else:
                fm = torch.as_tensor(free, device=rhs.device).to(torch.uint8).contiguous()
                if fm.numel() != rhs.numel():
                    raise TadError(1, "free mask must have one entry per variable")
                check(lib().tad_fn_solve_reduced(self._h, hp, rhs.data_ptr(), sol.data_ptr(), _lib.MEM_DEVICE,
                                                 fm.data_ptr(), ctypes.byref(o), ctypes.byref(st)))
            return sol, st
        rhs = self._host_vec(rhs, self.n_vars, "rhs")
        sol = np.zeros_like(rhs)
        hv = self._host_vec(H, self.nnz, "H") if H is not None else None
        hptr = hv.ctypes.data if hv is not None else None
        if free is None:
            check(lib().tad_fn_solve(self._h, hptr, rhs.ctypes.data, sol.ctypes.data, _lib.MEM_HOST,
                                     ctypes.byref(o), ctypes.byref(st)))
        else:
            fm = np.ascontiguousarray(np.asarray(free) != 0, dtype=np.uint8)
            if fm.size != rhs.size:
                raise TadError(1, "free mask must have one entry per variable")
            check(lib().tad_fn_solve_reduced(self._h, hptr, rhs.ctypes.data, sol.ctypes.data, _lib.MEM_HOST,
                                             fm.ctypes.data, ctypes.byref(o), ctypes.byref(st)))
        return sol, st

    def solve_basis(self, rhs, B, H=None, **opts):
        """sol = B y with (B^T H B + w I_m) y = B^T rhs for a general sparse basis B (scipy
        matrix, n_vars x m): newton_direction_reduced_basis (NewtonDirection.hh:76-105) with
        rhs = -g.  rhs numpy (H then numpy or None) or a float64 CUDA tensor (H then a CUDA
        tensor of the nnz values or None); H=None uses the last evaluated Hessian."""
        import scipy.sparse as sp
        Bc = sp.csr_matrix(B)
        Bc.sort_indices()
        if Bc.shape[0] != self.n_vars or Bc.shape[1] < 1:
            raise TadError(1, f"B must be {self.n_vars} x m with m >= 1, got {Bc.shape}")
        m = Bc.shape[1]
        rp = np.ascontiguousarray(Bc.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(Bc.indices, dtype=np.int32)
        bv = np.ascontiguousarray(Bc.data, dtype=np.float64)
        o = solve_default_opts(**opts)
        st = SolveStats()
        if _is_torch(rhs):
            import torch
            self._dev_vec(rhs, self.n_vars, "rhs")
            hp = self._dev_vec(H, self.nnz, "H").data_ptr() if H is not None else None
            dev = rhs.device
            rpt = torch.from_numpy(rp).to(dev)
            cit = torch.from_numpy(ci).to(dev) if ci.size else torch.zeros(1, dtype=torch.int32, device=dev)
            bvt = torch.from_numpy(bv).to(dev) if bv.size else torch.zeros(1, dtype=torch.float64, device=dev)
            _order_after_torch(self.ctx, dev)
            sol = torch.empty_like(rhs)
            check(lib().tad_fn_solve_basis(self._h, hp, m, rpt.data_ptr(), cit.data_ptr(), bvt.data_ptr(),
                                           rhs.data_ptr(), sol.data_ptr(), _lib.MEM_DEVICE, ctypes.byref(o),
                                           ctypes.byref(st)))
            self.ctx.synchronize()
            return sol, st
        rhs = self._host_vec(rhs, self.n_vars, "rhs")
        sol = np.zeros_like(rhs)
        hv = self._host_vec(H, self.nnz, "H") if H is not None else None
        hptr = hv.ctypes.data if hv is not None else None
        ci_ = ci if ci.size else np.zeros(1, dtype=np.int32)
        bv_ = bv if bv.size else np.zeros(1, dtype=np.float64)
        check(lib().tad_fn_solve_basis(self._h, hptr, m, rp.ctypes.data, ci_.ctypes.data, bv_.ctypes.data,
                                       rhs.ctypes.data, sol.ctypes.data, _lib.MEM_HOST, ctypes.byref(o),
                                       ctypes.byref(st)))
        return sol, st

    def newton_step(self, x, opts: NewtonOpts | None = None) -> dict:
        """One projected-Newton iteration in place on x (NewtonTest.cc:69-75 loop body)."""
        o = opts or newton_default_opts()
        st = NewtonStats()
        if _is_torch(x):
            self._x(x)
            ptr, mem = x.data_ptr(), _lib.MEM_DEVICE
        else:
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
                raise TadError(1, "host x must be a contiguous float64 numpy array (updated in place)")
            ptr, mem = x.ctypes.data, _lib.MEM_HOST
        check(lib().tad_newton_step(self._h, ptr, mem, ctypes.byref(o), ctypes.byref(st)))
        return st.as_dict()

    def line_search(self, x0, d, f, g, out=None, opts: NewtonOpts | None = None):
        """line_search (LineSearch.hh:26-65) on device tensors; returns (x_new, f_new, step, trials).

        trials is negative when no step was accepted: then x_new equals x0
        and step is 0 (the reference returns x0, LineSearch.hh:62-64)."""
        import torch
        o = opts or newton_default_opts()
        self._x(x0)
        for a, name in ((d, "d"), (g, "g")):
            self._dev_vec(a, self.n_vars, name)
        out = torch.empty_like(x0) if out is None else self._dev_vec(out, self.n_vars, "out")
        fo, so, tr = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().tad_line_search(self._h, x0.data_ptr(), d.data_ptr(), float(f), g.data_ptr(), out.data_ptr(),
                                    _lib.MEM_DEVICE, ctypes.byref(o), ctypes.byref(fo), ctypes.byref(so),
                                    ctypes.byref(tr)))
        return out, fo.value, so.value, tr.value

    def newton_decrement(self, d, g) -> float:
        """newton_decrement(d, g) = -0.5 d^T g (NewtonDecrement.hh:20-26), with the
        library's fixed-order FP64 dot product (numpy arrays or CUDA tensors)."""
        out = ctypes.c_double()
        if _is_torch(d) or _is_torch(g):
            self._dev_vec(d, self.n_vars, "d")
            self._dev_vec(g, self.n_vars, "g")
            _order_after_torch(self.ctx, d.device)
            check(lib().tad_newton_decrement(self._h, d.data_ptr(), g.data_ptr(), _lib.MEM_DEVICE,
                                             ctypes.byref(out)))
            return out.value
        d = self._host_vec(d, self.n_vars, "d")
        g = self._host_vec(g, self.n_vars, "g")
        check(lib().tad_newton_decrement(self._h, d.ctypes.data, g.ctypes.data, _lib.MEM_HOST, ctypes.byref(out)))
        return out.value


class VectorFunction(ScalarFunction):
    """VectorFunction<d, double, int> (VectorFunction.hh:19-192) backed by the
    HIP library.  Residual elements are device functors
    (tinyad_amd::device_residual<F, N, d, M, NP>() in
    include/tinyad_amd/DeviceElements.hh).  The inherited ScalarFunction
    methods act on the sum of squares f = r . r: eval_with_gradient gives
    g = 2 J^T r, the second-order evaluations the Gauss-Newton Hessian
    2 J^T J, line_search searches on f and newton_step is a Gauss-Newton step
    (solve w = 2 w_identity for the reference's w_identity)."""

    def __init__(self, d: int, n_handles: int, ctx: Context | None = None, precision: str = "fp64"):
        self.ctx = ctx or Context.default()
        self.d = int(d)
        self.n_handles = int(n_handles)
        self.n_vars = self.d * self.n_handles
        self.precision = precision
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        h = ctypes.c_void_p()
        check(lib().tad_fn_create_vector(self.ctx._h, self.d, self.n_handles, prec, ctypes.byref(h)))
        self._h = h
        self.n_elements = 0
        self._pattern = None
        self._jpattern = None

    def add_elements(self, *a, **kw):
        raise TadError(1, "vector functions take residual elements (add_elements_device with residual ops)")

    def add_elements_device(self, ops, conn, params=None):
        """add_elements<N, M>(range, lambda) (VectorFunctionImpl.hh:58-90) for
        residuals compiled on the device: ops is the address of a
        tad_device_residual_ops (tinyad_amd::device_residual<F, N, d, M, NP>()),
        conn n_elem x N handles, params n_elem x NP rows (one row is broadcast)."""
        addr = ctypes.cast(ops, ctypes.c_void_p).value if not isinstance(ops, int) else ops
        if not addr:
            raise TadError(1, "null residual ops")
        o = _lib.DeviceResidualOps.from_address(addr)
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, o.valence)
        n = conn.shape[0]
        pp = None
        if o.n_params > 0:
            params = np.ascontiguousarray(params, dtype=np.float64)
            if params.ndim == 1:
                params = np.ascontiguousarray(np.broadcast_to(params, (n, params.size)))
            if params.shape != (n, o.n_params):
                raise TadError(1, f"residual expects {o.n_params} parameters per element")
            pp = params.ctypes.data
        check(lib().tad_fn_add_residuals_device(self._h, addr, n, conn.ctypes.data, pp))
        self.n_elements += n
        self._pattern = None
        self._jpattern = None

    def add_elements_dynamic_device(self, *a, **kw):
        raise TadError(1, "vector functions take fixed-valence residual elements")

    @property
    def n_outputs(self) -> int:
        m = ctypes.c_int64()
        check(lib().tad_fn_n_outputs(self._h, ctypes.byref(m)))
        return m.value

    def jacobian_pattern(self):
        """(row_ptr, col_idx) host copies of the Jacobian CSR pattern (n_outputs x n_vars)."""
        if self._jpattern is None:
            nnz = ctypes.c_int64()
            check(lib().tad_fn_jacobian_pattern(self._h, ctypes.byref(nnz), None, None, None, None))
            rp = np.zeros(self.n_outputs + 1, dtype=np.int32)
            ci = np.zeros(nnz.value, dtype=np.int32)
            check(lib().tad_fn_jacobian_pattern(self._h, None, None, None, rp.ctypes.data, ci.ctypes.data))
            self._jpattern = (rp, ci)
        return self._jpattern

    def _r_J(self, mem, with_j):
        m = self.n_outputs
        nnz_j = int(self.jacobian_pattern()[1].size) if with_j else 0
        if mem == _lib.MEM_DEVICE:
            import torch
            dev = torch.device("cuda", self.ctx.device)
            r = torch.empty(m, dtype=torch.float64, device=dev)
            J = torch.empty(nnz_j, dtype=torch.float64, device=dev) if with_j else None
            return r, J, r.data_ptr(), (J.data_ptr() if with_j else None)
        r = np.zeros(m)
        J = np.zeros(nnz_j) if with_j else None
        return r, J, r.ctypes.data, (J.ctypes.data if with_j else None)

    def _wrap_j(self, J, mem):
        if mem == _lib.MEM_DEVICE:
            return J
        import scipy.sparse as sp
        rp, ci = self.jacobian_pattern()
        return sp.csr_matrix((J, ci, rp), shape=(self.n_outputs, self.n_vars))

    def eval(self, x):
        """r(x) (VectorFunctionImpl.hh:130-147)."""
        ptr, mem = self._x(x)
        r, _, rp, _ = self._r_J(mem, False)
        check(lib().tad_fn_eval_residuals(self._h, ptr, rp, mem))
        return r

    __call__ = eval

    def eval_with_jacobian(self, x):
        """(r, J) (VectorFunctionImpl.hh:157-176); J is a scipy CSR matrix for
        host x, the CUDA tensor of its values (jacobian_pattern order) for device x."""
        ptr, mem = self._x(x)
        r, J, rp, jp = self._r_J(mem, True)
        check(lib().tad_fn_eval_with_jacobian(self._h, ptr, rp, jp, mem))
        return r, self._wrap_j(J, mem)

    def eval_sum_of_squares(self, x) -> float:
        """sum_e r_e . r_e (VectorFunctionImpl.hh:228-239)."""
        return ScalarFunction.eval(self, x)

    def eval_sum_of_squares_with_derivatives(self, x):
        """(f, g, r, J) with f = r . r and g = 2 J^T r (VectorFunctionImpl.hh:258-275)."""
        ptr, mem = self._x(x)
        g, _, gp, _ = self._outputs(mem, False)
        r, J, rp, jp = self._r_J(mem, True)
        f = ctypes.c_double()
        check(lib().tad_fn_eval_sum_of_squares_with_derivatives(self._h, ptr, ctypes.byref(f), gp, rp, jp, mem))
        return f.value, g, r, self._wrap_j(J, mem)

    def eval_gauss_newton(self, x):
        """(f, g, H) with g = 2 J^T r and H = 2 J^T J over the Hessian pattern."""
        return self.eval_with_derivatives(x)

    def gauss_newton_direction(self, x, w_identity: float = 0.0, **solve_opts):
        """d with (w_identity I + J^T J) d = -J^T r at x (GaussNewtonDirection.hh:24-47):
        the library's PCG on 2 (w_identity I + J^T J) d = -2 J^T r."""
        _, g, H = self.eval_with_derivatives(x)
        rhs = -g
        sol, st = self.solve(rhs, H if _is_torch(rhs) else H.data, w_identity=2.0 * float(w_identity), **solve_opts)
        if not st.converged:
            raise TadError(4, "Gauss-Newton solve did not converge")
        return sol


def vector_function(d: int, n_handles: int, ctx: Context | None = None, precision: str = "fp64") -> VectorFunction:
    """vector_function<d>(range(n_handles)) (VectorFunctionImpl.hh:300-325)."""
    return VectorFunction(d, n_handles, ctx, precision)


def scalar_function(d: int, n_handles: int, ctx: Context | None = None, precision: str = "fp64") -> ScalarFunction:
    """scalar_function<d>(range(n_handles)) (ScalarFunctionImpl.hh:418-442)."""
    return ScalarFunction(d, n_handles, ctx, precision)


def selftest_dual(which: int, x: float, y: float, ctx: Context | None = None):
    ctx = ctx or Context.default()
    out = (ctypes.c_double * 7)()
    check(lib().tad_selftest_dual(ctx._h, which, x, y, out))
    v = np.array(out[:])
    return v[0], v[1:3], v[3:].reshape(2, 2)
