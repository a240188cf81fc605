"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the CPU oracle (tad_oracle.cpp).

Used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline leg as the
checker; the shipped library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_last_error.restype = ctypes.c_char_p
        L.or_fn_create.restype = ctypes.c_void_p
        L.or_fn_create.argtypes = [ctypes.c_int, ctypes.c_long]
        L.or_fn_destroy.argtypes = [ctypes.c_void_p]
        L.or_fn_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_fn_threads.argtypes = [ctypes.c_void_p]
        L.or_fn_add_elements.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                         _ip, _dp, ctypes.c_int]
        L.or_eval.restype = ctypes.c_double
        L.or_eval.argtypes = [ctypes.c_void_p, _dp]
        L.or_eval_with_gradient.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.or_eval_second_order.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                           ctypes.POINTER(ctypes.c_long)]
        L.or_get_csr.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        L.or_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long, _dp, ctypes.c_int,
                                 ctypes.c_double, _dp, _dp, _dp, _ip]
        L.or_time_eval_with_hessian_proj.restype = ctypes.c_double
        L.or_time_eval_with_hessian_proj.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int]
        L.or_pcg.argtypes = [ctypes.c_long, ctypes.c_int, _ip, _ip, _dp, ctypes.c_double, _dp, _dp,
                             ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]
        L.or_line_search.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.or_dual_known_answer.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.or_tet_rest_mixed_vjp.argtypes = [ctypes.c_long, _ip, _dp, _dp, _dp, ctypes.c_double, ctypes.c_long, _dp]
        L.or_vfn_create.restype = ctypes.c_void_p
        L.or_vfn_create.argtypes = [ctypes.c_int, ctypes.c_long]
        L.or_vfn_destroy.argtypes = [ctypes.c_void_p]
        L.or_vfn_add_elements.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_long, _ip, _dp,
                                          ctypes.c_int]
        L.or_vfn_n_outputs.restype = ctypes.c_long
        L.or_vfn_n_outputs.argtypes = [ctypes.c_void_p]
        L.or_vfn_eval.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_vfn_sum_of_squares.restype = ctypes.c_double
        L.or_vfn_sum_of_squares.argtypes = [ctypes.c_void_p, _dp]
        L.or_vfn_eval_with_jacobian.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.POINTER(ctypes.c_long)]
        L.or_vfn_get_jacobian.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class OracleError(RuntimeError):
    pass


class Function:
    """CPU restatement of ScalarFunction (FP64)."""

    def __init__(self, d: int, n_handles: int, n_threads: int = 0):
        L = lib()
        self._h = L.or_fn_create(d, n_handles)
        if not self._h:
            raise OracleError(L.or_last_error().decode())
        self.d = d
        self.n_vars = d * n_handles
        if n_threads:
            L.or_fn_set_threads(self._h, n_threads)

    @classmethod
    def from_problem(cls, problem, n_threads: int = 0):
        f = cls(problem.d, problem.n_handles, n_threads)
        for t in problem.terms:
            f.add_elements(t.energy, t.conn, t.params)
        return f

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_fn_destroy(self._h)
            self._h = None

    @property
    def threads(self) -> int:
        return lib().or_fn_threads(self._h)

    def add_elements(self, energy: int, conn, params):
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        params = np.ascontiguousarray(params, dtype=np.float64)
        if conn.ndim == 1:
            conn = conn[:, None]
        if params.ndim == 1:
            params = params[None, :].repeat(conn.shape[0], axis=0)
        rc = lib().or_fn_add_elements(self._h, energy, conn.shape[1], conn.shape[0], _i(conn), _d(params),
                                      params.shape[1])
        if rc:
            raise OracleError(lib().or_last_error().decode())

    def eval(self, x) -> float:
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().or_eval(self._h, _d(x))

    def eval_with_gradient(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.zeros(1)
        g = np.zeros(self.n_vars)
        if lib().or_eval_with_gradient(self._h, _d(x), _d(f), _d(g)):
            raise OracleError(lib().or_last_error().decode())
        return float(f[0]), g

    def _second(self, x, project, eps):
        import scipy.sparse as sp
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.zeros(1)
        g = np.zeros(self.n_vars)
        nnz = ctypes.c_long(0)
        if lib().or_eval_second_order(self._h, _d(x), int(project), float(eps), _d(f), _d(g), ctypes.byref(nnz)):
            raise OracleError(lib().or_last_error().decode())
        rp = np.zeros(self.n_vars + 1, dtype=np.int32)
        ci = np.zeros(nnz.value, dtype=np.int32)
        v = np.zeros(nnz.value)
        lib().or_get_csr(self._h, _i(rp), _i(ci), _d(v))
        return float(f[0]), g, sp.csr_matrix((v, ci, rp), shape=(self.n_vars, self.n_vars))

    def eval_with_derivatives(self, x):
        return self._second(x, False, float("nan"))

    def eval_with_hessian_proj(self, x, eps: float = 1e-9):
        return self._second(x, True, eps)

    def element(self, term: int, e: int, x, K: int, project: bool = False, eps: float = 1e-9):
        x = np.ascontiguousarray(x, dtype=np.float64)
        val = np.zeros(1)
        g = np.zeros(K)
        H = np.zeros(K * K)
        idx = np.zeros(K, dtype=np.int32)
        if lib().or_element(self._h, term, e, _d(x), int(project), float(eps), _d(val), _d(g), _d(H), _i(idx)):
            raise OracleError(lib().or_last_error().decode())
        return float(val[0]), g, H.reshape(K, K), idx

    def time_eval_with_hessian_proj(self, x, eps: float = 1e-9, reps: int = 3) -> float:
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().or_time_eval_with_hessian_proj(self._h, _d(x), eps, reps)

    def line_search(self, x0, d, f, g, s_max=1.0, shrink=0.8, max_iters=64, armijo_const=1e-4):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        xo = np.zeros_like(x0)
        fo = np.zeros(1)
        so = np.zeros(1)
        n = lib().or_line_search(self._h, _d(x0), _d(d), float(f), _d(g), s_max, shrink, max_iters,
                                 armijo_const, _d(xo), _d(fo), _d(so))
        if n == 0:
            raise OracleError(lib().or_last_error().decode())
        return xo, float(fo[0]), float(so[0]), n


# residual ids of the vector-function restatement (tad_oracle.cpp)
R_SYMDIR_TRI = 101
R_PIN = 102


class VectorFunction:
    """CPU restatement of VectorFunction (FP64, VectorFunction.hh:19-192)
    over the residuals of tad_oracle.cpp (R_SYMDIR_TRI, R_PIN)."""

    def __init__(self, d: int, n_handles: int):
        L = lib()
        self._h = L.or_vfn_create(d, n_handles)
        if not self._h:
            raise OracleError(L.or_last_error().decode())
        self.d = d
        self.n_vars = d * n_handles

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_vfn_destroy(self._h)
            self._h = None

    @classmethod
    def from_problem(cls, problem):
        f = cls(problem.d, problem.n_handles)
        for t in problem.terms:
            f.add_elements(t.energy, t.conn, t.params)
        return f

    @property
    def n_outputs(self) -> int:
        return lib().or_vfn_n_outputs(self._h)

    def add_elements(self, rid: int, conn, params):
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        params = np.ascontiguousarray(params, dtype=np.float64)
        if conn.ndim == 1:
            conn = conn[:, None]
        rc = lib().or_vfn_add_elements(self._h, rid, conn.shape[1], conn.shape[0], _i(conn), _d(params),
                                       params.shape[1])
        if rc:
            raise OracleError(lib().or_last_error().decode())

    def eval(self, x):
        """r(x) (VectorFunctionImpl.hh:130-147)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.n_outputs)
        lib().or_vfn_eval(self._h, _d(x), _d(r))
        return r

    def eval_with_jacobian(self, x):
        """(r, J) with J a scipy CSR matrix n_outputs x n_vars (VectorFunctionImpl.hh:157-176)."""
        import scipy.sparse as sp
        x = np.ascontiguousarray(x, dtype=np.float64)
        m = self.n_outputs
        r = np.zeros(m)
        nnz = ctypes.c_long(0)
        if lib().or_vfn_eval_with_jacobian(self._h, _d(x), _d(r), ctypes.byref(nnz)):
            raise OracleError(lib().or_last_error().decode())
        rp = np.zeros(m + 1, dtype=np.int32)
        ci = np.zeros(nnz.value, dtype=np.int32)
        v = np.zeros(nnz.value)
        lib().or_vfn_get_jacobian(self._h, _i(rp), _i(ci), _d(v))
        return r, sp.csr_matrix((v, ci, rp), shape=(m, self.n_vars))

    def eval_sum_of_squares(self, x) -> float:
        """sum_e r_e . r_e in element order (VectorFunctionImpl.hh:228-239)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().or_vfn_sum_of_squares(self._h, _d(x))

    def eval_sum_of_squares_with_derivatives(self, x):
        """(f, g, r, J): f = r . r, g = 2 J^T r (VectorFunctionImpl.hh:260-275)."""
        r, J = self.eval_with_jacobian(x)
        return float(r @ r), 2.0 * (J.T @ r), r, J


def gauss_newton_direction(r, J, w_identity: float = 0.0):
    """d with (w I + J^T J) d = -J^T r, by a direct sparse solve
    (GaussNewtonDirection.hh:24-47 factorizes with SimplicialLDLT)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    A = (J.T @ J + w_identity * sp.identity(J.shape[1])).tocsc()
    return spla.spsolve(A, -(J.T @ r))


def newton_direction_reduced_basis(g, H, B, w_identity: float = 0.0):
    """d = B y with (B^T H B + w I_m) y = -B^T g, by a direct sparse solve
    (NewtonDirection.hh:76-105 factorizes H_reduced with SimplicialLDLT)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    B = sp.csr_matrix(B)
    Hr = (B.T @ H @ B + w_identity * sp.identity(B.shape[1])).tocsc()
    y = spla.spsolve(Hr, -(B.T @ g))
    return B @ np.atleast_1d(y)


def pcg(H, b, w: float = 0.0, block: int = 1, tol: float = 1e-10, max_iters: int = 10000, n_threads: int = 0):
    """Block-Jacobi PCG on (w I + H) x = b with x0 = 0 (the GPU solve's
    recurrences without its coarse correction).  n_threads: OpenMP threads
    (0: the reference's default, available - 1); the result does not depend
    on it."""
    H = H.tocsr()
    rp = np.ascontiguousarray(H.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(H.indices, dtype=np.int32)
    v = np.ascontiguousarray(H.data, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    rel = np.zeros(1)
    it = lib().or_pcg(b.size, block, _i(rp), _i(ci), _d(v), w, _d(b), _d(x), tol, max_iters, _d(rel), n_threads)
    return x, it, float(rel[0])


def dual_known_answer(which: int, x: float, y: float):
    out = np.zeros(7)
    lib().or_dual_known_answer(which, x, y, _d(out))
    return out[0], out[1:3].copy(), out[3:].reshape(2, 2).copy()


def tet_rest_mixed_vjp(conn, x, rest, lam, w_scale):
    """sum_e (d2E_e / dx dtheta)^T lambda_e for symmetric Dirichlet tets with
    w_e = w_scale det(Mr_e(theta)) / 6 (TinyAD-style truncated Hessian block)."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    rest = np.ascontiguousarray(rest, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.zeros_like(x)
    lib().or_tet_rest_mixed_vjp(conn.shape[0], _i(conn), _d(x), _d(rest), _d(lam), float(w_scale), x.size, _d(out))
    return out
