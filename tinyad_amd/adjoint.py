"""Implicit differentiation through the Newton optimum x*(theta).

theta are per-vertex rest positions of a symmetric-Dirichlet tet term: every
element's parameters are Mr(theta)^-1 and w_e = w_scale det(Mr_e(theta)) / 6.
At the optimum grad_x E(x*, theta) = 0, so for a loss L(x*)

    dL/dtheta = -(d2E / dtheta dx)^T lambda,   H(x*) lambda = dL/dx*,

with H the unprojected Hessian (eval_with_derivatives,
ScalarFunctionImpl.hh:316-376).  One FP64 PCG solve gives lambda and the
library's rest-shape product kernel gives the mixed term; `optimum()` wraps
both in a torch.autograd.Function.  The reference has no adjoint: a TinyAD
user would form the truncated mixed block Scalar<24, T, true, 0, 12, 12, 12>
(Scalar.hh:24-47, :188-196) per element and a sparse LDLT solve.
"""
from __future__ import annotations

import math

import numpy as np

from . import ScalarFunction, SYMDIR_TET, TadError, newton_default_opts

__all__ = ["RestShapeProblem", "optimum", "tet_rest_params"]


def _det3(M):
    """det of (..., 3, 3) by cofactor expansion along the first row."""
    return (M[..., 0, 0] * (M[..., 1, 1] * M[..., 2, 2] - M[..., 1, 2] * M[..., 2, 1])
            - M[..., 0, 1] * (M[..., 1, 0] * M[..., 2, 2] - M[..., 1, 2] * M[..., 2, 0])
            + M[..., 0, 2] * (M[..., 1, 0] * M[..., 2, 1] - M[..., 1, 1] * M[..., 2, 0]))


def _rest_edges(theta: np.ndarray, conn: np.ndarray) -> np.ndarray:
    V = theta.reshape(-1, 3)
    a = V[conn[:, 0]]
    return np.stack([V[conn[:, 1]] - a, V[conn[:, 2]] - a, V[conn[:, 3]] - a], axis=2)


def tet_rest_params(theta: np.ndarray, conn: np.ndarray, w_scale: float) -> np.ndarray:
    """Parameter rows [Mr^-1 (row-major), w_scale det(Mr) / 6] of a SYMDIR_TET term."""
    Mr = _rest_edges(np.asarray(theta, dtype=np.float64), conn)
    det = _det3(Mr)
    if not np.all(det != 0.0):
        raise TadError(1, "degenerate rest tetrahedron")
    R = np.linalg.inv(Mr)
    return np.concatenate([R.reshape(-1, 9), (w_scale * det / 6.0)[:, None]], axis=1)


class RestShapeProblem:
    """A ScalarFunction whose SYMDIR_TET term `term` is parameterized by rest positions.

    problem: a configs.Problem with problem.info["rest"] (n_vars rest positions).
    w_scale: weight factor; the default 1 / total rest volume reproduces the
    volume-fraction weights of configs.tet_rest_params at theta0 and is held
    fixed (not differentiated) as theta moves.
    """

    def __init__(self, problem, term: int = 0, ctx=None, w_scale: float | None = None):
        t = problem.terms[term]
        if t.energy != SYMDIR_TET or problem.d != 3:
            raise TadError(6, "rest-shape parameters exist for symmetric-Dirichlet tets in 3D only")
        self.term = term
        self.conn = np.ascontiguousarray(t.conn, dtype=np.int32)
        self.theta0 = np.asarray(problem.info["rest"], dtype=np.float64).copy()
        if w_scale is None:
            w_scale = 1.0 / float((_det3(_rest_edges(self.theta0, self.conn)) / 6.0).sum())
        self.w_scale = float(w_scale)
        self.fn = ScalarFunction.from_problem(problem, ctx, precision="fp64")
        self.n_vars = self.fn.n_vars
        self.set_rest(self.theta0)

    def close(self):
        self.fn.close()

    # -- forward ---------------------------------------------------------------
    def set_rest(self, theta):
        """Re-parameterizes the term from rest positions theta (numpy or tensor)."""
        th = theta.detach().cpu().numpy() if hasattr(theta, "detach") else np.asarray(theta, dtype=np.float64)
        if th.size != self.n_vars:
            raise TadError(1, f"theta has {th.size} entries, expected {self.n_vars}")
        self.fn.set_params(self.term, tet_rest_params(th.reshape(-1), self.conn, self.w_scale))

    def minimize(self, x, max_iters: int = 200, switch_rel: float = 1e-12, proj_tol: float = 1e-10,
                 exact_tol: float = 1e-12):
        """Newton to the optimum in place on the device tensor x.

        Projected Newton steps with line search (newton_step, eps 1e-9) until
        the Newton decrement falls below switch_rel |f|; then full Newton steps
        with the unprojected Hessian (eval_with_derivatives + solve) until the
        gradient norm stops shrinking by at least 4x, i.e. reaches its rounding
        floor.  In that last phase the decrement is far below one ulp of f, so
        an energy-based line search could not tell steps apart; near a strict
        minimum full steps converge quadratically.  Returns per-iteration stats.
        """
        import torch
        proj = newton_default_opts(solve_tol=proj_tol)
        hist = []
        for _ in range(max_iters):
            st = self.fn.newton_step(x, proj)
            st["phase"] = "proj"
            hist.append(st)
            if st["ls_failed"] or st["newton_decrement"] <= switch_rel * max(abs(st["f"]), 1e-300):
                break
        prev_g = math.inf
        x_prev = None
        for _ in range(max_iters):
            f, g, H = self.fn.eval_with_derivatives(x)
            gn = float(torch.linalg.vector_norm(g))
            if gn > 0.25 * prev_g:
                if gn > prev_g:
                    x.copy_(x_prev)  # the last step did not help: keep the better point
                hist.append({"phase": "exact", "f": f, "f_new": f, "grad_norm": gn, "step": 0.0,
                             "newton_decrement": 0.0, "cg_iters": 0, "ls_failed": 0})
                break
            d, sst = self.fn.solve(g, H=H, tol=exact_tol)
            x_prev = x.clone()
            x.sub_(d)
            hist.append({"phase": "exact", "f": f, "f_new": f, "grad_norm": gn, "step": 1.0,
                         "newton_decrement": 0.5 * float(torch.dot(d, g)), "cg_iters": sst.iters,
                         "ls_failed": 0})
            prev_g = gn
        return hist

    # -- adjoint ---------------------------------------------------------------
    def hessian_solve(self, x, rhs, tol: float = 1e-13, max_iters: int = 50000):
        """lambda with H(x) lambda = rhs, H the unprojected Hessian at x (device tensors)."""
        _, _, H = self.fn.eval_with_derivatives(x)
        lam, st = self.fn.solve(rhs, H=H, tol=tol, max_iters=max_iters)
        if not st.converged:
            raise TadError(4, f"adjoint solve did not converge (rel. residual {st.rel_residual:.3e})")
        return lam, st

    def mixed_vjp(self, x, theta, lam):
        """(d2E / dtheta dx)^T lam at (x, theta)."""
        return self.fn.rest_vjp(self.term, x, theta, lam, self.w_scale)

    def theta_gradient(self, x, theta, grad_x, tol: float = 1e-13):
        """dL/dtheta for dL/dx* = grad_x at the optimum x = x*(theta)."""
        lam, st = self.hessian_solve(x, grad_x, tol=tol)
        self.last_solve = st
        return -self.mixed_vjp(x, theta, lam)


def _autograd_function():
    import torch

    class _Optimum(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, rp, x_init):
            rp.set_rest(theta)
            x = x_init.detach().clone().contiguous()
            rp.last_history = rp.minimize(x)
            ctx.rp = rp
            ctx.save_for_backward(theta.detach(), x)
            return x

        @staticmethod
        def backward(ctx, grad_x):
            theta, x = ctx.saved_tensors
            rp = ctx.rp
            rp.set_rest(theta)  # the function may have been re-parameterized since forward
            th = theta.to(torch.float64).contiguous()
            g = rp.theta_gradient(x, th, grad_x.to(torch.float64).contiguous())
            return g.to(theta.dtype), None, None

    return _Optimum


_OPTIMUM = None


def optimum(theta, rp: RestShapeProblem, x_init):
    """x*(theta) (differentiable in theta through the implicit adjoint).

    theta, x_init: float64 CUDA tensors of n_vars entries; x_init is the Newton
    start (e.g. the previous optimum).
    """
    global _OPTIMUM
    if _OPTIMUM is None:
        _OPTIMUM = _autograd_function()
    return _OPTIMUM.apply(theta, rp, x_init)
