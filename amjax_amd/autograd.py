"""Differentiable solves (torch.autograd.Function) on the gfx950 solve phase.

Three forms:

* ``MultilevelSolver.solve(b, A=A, ...)`` with ``A.values_tensor`` (or a dense
  ``A``) requiring grad, or ``b`` requiring grad: the reference's custom VJP
  (amjax/multilevel.py:428-444).  Backward solves ``lam = solve(v, A=A)`` with
  the same (untransposed) cycle, tol and maxiter; ``grad_b = lam`` and
  ``grad_A = -outer(lam, x)``, restricted to A's stored pattern (the dense
  outer product of the reference is N^2 and is never formed; a dense ``A`` is
  treated as a full pattern).  The hierarchy stays fixed.
* ``cycles(ml, b, m)``: x after m cycles from a zero guess, differentiable in
  b.  Backward runs the transposed cycle: Aᵀ products, R and P swapped, pre-
  and post-smoothing exchanged (``MultilevelSolver.transpose``).  The map is
  linear in b, so no forward iterate is stored or recomputed.
* ``amjax_amd.cg(A, b, M=...)`` with A's values requiring grad: the implicit
  adjoint of the converged preconditioned CG solve, ``lam = cg(A, v, M)``
  (A symmetric, as jax's cg assumes through custom_linear_solve), and
  ``grad_values = -lam[row] x[col]``.
"""

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .sparse import SparseMatrix, amg_dtype, _prep


def pattern_grad(A, lam, x, scale=-1.0):
    """g[q] = scale * sum_c lam[row_q, c] x[col_q, c] over A's stored entries (CSR order)."""
    rows, cols, _ = A.pattern()
    lam2 = _prep(lam, A.dtype, A.device).reshape(lam.shape[0], -1)
    x2 = _prep(x, A.dtype, A.device).reshape(x.shape[0], -1)
    g = torch.empty(A.nnz, dtype=A.dtype, device=A.device)
    _lib.check(_lib.load().amg_grad_values(rows.data_ptr(), cols.data_ptr(), A.nnz, amg_dtype(A.dtype),
                                           lam2.contiguous().data_ptr(), x2.contiguous().data_ptr(), lam2.shape[1],
                                           float(scale), g.data_ptr(), _lib.stream_ptr(A.device)),
               "amg_grad_values")
    return g


def as_operator(ml, A):
    """(SparseMatrix, values tensor or None, dense shape or None) for an ``A=`` argument."""
    if isinstance(A, SparseMatrix):
        return A, A.values_tensor, None
    if torch.is_tensor(A) and A.dim() == 2:
        n = A.shape[0]
        full = sp.csr_matrix(np.ones((n, A.shape[1])))
        pat = SparseMatrix.from_scipy(full, ml.dtype, ml.device)
        vals = A.reshape(-1)
        return pat.with_values(vals), vals, tuple(A.shape)
    M = SparseMatrix.from_scipy(A if sp.issparse(A) else sp.csr_matrix(np.asarray(A)), ml.dtype, ml.device)
    return M, None, None


class _SolveVJP(torch.autograd.Function):
    """amjax/multilevel.py:428-444 (_solve_vjp / _solve_fwd / _solve_bwd)."""

    @staticmethod
    def forward(ctx, b, values, ml, Aop, tol, maxiter, cycle, cpl):
        x = ml._solve_nograd(b, Aop, None, tol, maxiter, cycle, cpl)
        ctx.ml, ctx.Aop, ctx.cfg = ml, Aop, (tol, maxiter, cycle, cpl)
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, v):
        (x,) = ctx.saved_tensors
        tol, maxiter, cycle, cpl = ctx.cfg
        lam = ctx.ml._solve_nograd(v.contiguous(), ctx.Aop, None, tol, maxiter, cycle, cpl)
        g_b = lam if ctx.needs_input_grad[0] else None
        g_v = pattern_grad(ctx.Aop, lam, x, -1.0) if ctx.needs_input_grad[1] else None
        return g_b, g_v, None, None, None, None, None, None


class _MatvecFn(torch.autograd.Function):
    """y = A x for A's stored values ``values``: gradients g_i x_j on the pattern
    for the values and A^T g for x."""

    @staticmethod
    def forward(ctx, values, x, Aop):
        ctx.Aop = Aop
        ctx.save_for_backward(x)
        return Aop.matmul(x)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        g_v = pattern_grad(ctx.Aop, g, x, 1.0) if ctx.needs_input_grad[0] else None
        g_x = ctx.Aop.T.matmul(g.contiguous()) if ctx.needs_input_grad[1] else None
        return g_v, g_x, None


def solve_with_grad(ml, b, A, x0, tol, maxiter, cycle, cpl):
    """A start x0 follows amjax/multilevel.py:385-393: b - A x0 is solved and x0
    added back, with gradients through the shift (to b, A's values and x0)."""
    Aop, vals, dense_shape = as_operator(ml, A)
    bt = b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b), device=ml.device)
    if vals is None:
        vals = torch.zeros(0, dtype=ml.dtype, device=ml.device)
    if x0 is not None:
        x0t = x0 if torch.is_tensor(x0) else torch.as_tensor(np.asarray(x0), device=ml.device)
        x0t = x0t.to(device=ml.device, dtype=Aop.dtype).reshape(bt.shape)
        bt = bt - _MatvecFn.apply(vals, x0t, Aop).to(bt.dtype)
    x = _SolveVJP.apply(bt, vals, ml, Aop, tol, maxiter, cycle, cpl)
    if x0 is not None:
        x = x + x0t.to(x.dtype)
    return x


class _CyclesFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, ml, m, cycle, cpl):
        ctx.ml, ctx.m, ctx.cfg = ml, m, (cycle, cpl)
        x = None
        for _ in range(m):
            x = ml._apply_cycle(x, b, cycle, cpl)
        return x if x is not None else torch.zeros_like(b)

    @staticmethod
    def backward(ctx, v):
        # x_m = sum_t M^(m-1-t) B b with M = I - B A, so
        # grad_b = sum_t B^T (M^T)^t v,  M^T mu = mu - A^T (B^T mu)
        ml, m = ctx.ml, ctx.m
        cycle, cpl = ctx.cfg
        mt = ml.transpose()
        At = mt.levels[0].A
        # the F-cycle's transpose visits the coarse levels in reverse order
        tcycle = "T" if cycle == "F" else cycle
        mu = v.contiguous()
        g = torch.zeros_like(mu)
        for t in range(m):
            nu = mt._apply_cycle(None, mu, tcycle, cpl)
            g = g + nu
            if t < m - 1:
                mu = mu - At.matmul(nu)
        return g, None, None, None, None


def cycles(ml, b, m, cycle="V", cycles_per_level=1):
    """x after ``m`` cycles from a zero guess (no convergence test), differentiable in b.

    The backward pass applies the transposed cycle m times (Aᵀ products, R and
    P swapped, pre- and post-smoothing exchanged).
    """
    bt = _prep(b, ml.dtype, ml.device)
    if torch.is_tensor(b) and b.requires_grad:
        bt = b.to(dtype=ml.dtype)
    return _CyclesFn.apply(bt, ml, int(m), cycle, cycles_per_level)


class _PCGFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, values, ml, Aop, tol, atol, maxiter, cycle, cpl, precondition, x0=None):
        if x0 is None:
            x = ml._pcg_with_operator(b, Aop, tol, atol, maxiter, cycle, cpl, precondition)
        else:
            # a start x0 as in the non-differentiable path: the shifted solve with
            # jax's per-column thresholds; the adjoint below needs only the converged x
            from .krylov import _x0_shift, _x0_unshift
            r0, s, zero, x02, vec = _x0_shift(Aop, b, x0, tol, atol, ml.device)
            e = ml._pcg_with_operator(r0, Aop, 0.0, 1.0, maxiter, cycle, cpl, precondition)
            x = _x0_unshift(e, s, zero, x02, vec)
        ctx.ml, ctx.Aop = ml, Aop
        ctx.cfg = (tol, atol, maxiter, cycle, cpl, precondition)
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, v):
        (x,) = ctx.saved_tensors
        tol, atol, maxiter, cycle, cpl, precondition = ctx.cfg
        lam = ctx.ml._pcg_with_operator(v.contiguous(), ctx.Aop, tol, atol, maxiter, cycle, cpl, precondition)
        g_b = lam if ctx.needs_input_grad[0] else None
        g_v = pattern_grad(ctx.Aop, lam, x, -1.0) if ctx.needs_input_grad[1] else None
        return g_b, g_v, None, None, None, None, None, None, None, None, None
