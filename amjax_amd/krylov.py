"""Krylov front ends: ``cg``, ``bicgstab(A, b, x0=None, *, tol, atol, maxiter, M)`` and ``gmres``.

Same call shape as ``jax.scipy.sparse.linalg.cg`` (used by the reference's
benchmark and tests: benchmarks/benchmark.ipynb cell 7,
tests/test_multilevel.py:69,274,278) with ``M = ml.aspreconditioner(cycle)``.
Runs the device PCG of the preconditioner's hierarchy (float64 outer
iteration); returns ``(x, None)`` as jax does.  With ``A`` built by
``SparseMatrix.with_values(values)`` and ``values.requires_grad``, or ``b``
requiring grad, the solve is differentiable through the implicit adjoint of
the converged solve (the adjoint solve is the same PCG, A being symmetric).

``bicgstab`` has the call shape of ``jax.scipy.sparse.linalg.bicgstab`` (the
reference's README roadmap pairs AMJax preconditioning with JAX BiCGSTAB and
GMRES for nonsymmetric operators, README.md:140-149): right-preconditioned, in
float64, on the device, with the preconditioner's cycle as M.  It has no
adjoint.
"""

from collections import OrderedDict

import numpy as np
import scipy.sparse as sp
import torch

from .multilevel import MultilevelSolver, Preconditioner
from .sparse import SparseMatrix

# Device-resident plain (unpreconditioned) solvers of the operators most recently
# passed without M, least recently used first.  Bounded: each entry holds its
# operator and a float64 device copy of it.
_plain = OrderedDict()
_PLAIN_MAX = 4


class _Hier:
    def __init__(self, levels):
        self.levels = levels


class _L:
    def __init__(self, A):
        self.A = A


def _plain_solver(A):
    key = id(A)
    hit = _plain.get(key)
    if hit is not None and hit[0] is A:
        _plain.move_to_end(key)
        return hit[1]
    H = A.host if isinstance(A, SparseMatrix) else sp.csr_matrix(A)
    ml = MultilevelSolver.from_pyamg(_Hier([_L(sp.csr_matrix(H, dtype=np.float64))]), coarse_solver="jacobi",
                                     coarse_solver_kwargs={"iterations": 0}, dtype=np.float64,
                                     device=A.device if isinstance(A, SparseMatrix) else None)
    _plain[key] = (A, ml)
    _plain.move_to_end(key)
    while len(_plain) > _PLAIN_MAX:
        _plain.popitem(last=False)
    return ml


def _x0_shift(Aop, b, x0, tol, atol, dev):
    """A start x0 as a shifted solve.  jax's cg and bicgstab started from x0 stop
    per column once ||b - A x|| <= max(tol ||b||, atol).  Started from 0 on
    r0 = b - A x0 they produce the same residuals, so the solve runs on r0 with
    tol = 0 and one atol: column c of r0 is scaled by s_c = 1 / max(tol ||b_c||,
    atol), which maps each column's own threshold to atol = 1.  A column whose
    threshold is 0 (b_c = 0 and atol = 0) is given the exact solution 0.

    Returns (scaled r0 (n, k), s (k,), zero-threshold mask (k,), x0 (n, k), vec)."""
    bt = b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b))
    bt = bt.to(device=dev, dtype=torch.float64)
    vec = bt.dim() == 1
    b2 = bt.reshape(bt.shape[0], -1).contiguous()
    xt = x0 if torch.is_tensor(x0) else torch.as_tensor(np.asarray(x0))
    x02 = xt.to(device=dev, dtype=torch.float64).reshape(b2.shape).contiguous()
    r0 = b2 - Aop.matmul(x02)
    thr = torch.clamp(tol * torch.linalg.vector_norm(b2, dim=0), min=float(atol))
    zero = thr == 0
    s = torch.where(zero, torch.ones_like(thr), 1.0 / torch.where(zero, torch.ones_like(thr), thr))
    return (r0 * s).contiguous(), s, zero, x02, vec


def _x0_unshift(e, s, zero, x02, vec):
    x = x02 + e.reshape(x02.shape) / s
    x[:, zero] = 0.0
    return x.reshape(-1) if vec else x


def cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Preconditioned conjugate gradients on the GPU; returns ``(x, None)``.

    ``x0`` (optional, shaped like b) starts the iteration as in jax: the same
    residuals and stopping rule, run as a shifted solve on the device."""
    if x0 is not None:
        if M is not None and not isinstance(M, Preconditioner):
            raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
        vals = getattr(A, "values_tensor", None)
        ml = M.ml if M is not None else _plain_solver(A)
        if torch.is_grad_enabled() and ((torch.is_tensor(b) and b.requires_grad)
                                        or (vals is not None and vals.requires_grad)):
            return _cg_grad(A, b, ml, M, tol, atol, maxiter, x0), None
        Aop = ml._op64() if (M is None or A is ml.levels[0].A) else A
        r0, s, zero, x02, vec = _x0_shift(Aop, b, x0, tol, atol, ml.device)
        e, _ = cg(A, r0, tol=0.0, atol=1.0, maxiter=maxiter, M=M)
        return _x0_unshift(e, s, zero, x02, vec), None
    if M is not None and not isinstance(M, Preconditioner):
        raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
    ml = M.ml if M is not None else _plain_solver(A)
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    precondition = M is not None
    vals = getattr(A, "values_tensor", None)
    needs_grad = torch.is_grad_enabled() and (
        (torch.is_tensor(b) and b.requires_grad) or (vals is not None and vals.requires_grad))
    same_op = A is ml.levels[0].A or (M is None)
    if not needs_grad:
        if same_op:
            return ml.pcg(b, tol=tol, atol=atol, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                          precondition=precondition), None
        return ml._pcg_with_operator(b, A, tol, atol, maxiter, cycle, cpl, precondition), None
    return _cg_grad(A, b, ml, M, tol, atol, maxiter, None), None


def _cg_grad(A, b, ml, M, tol, atol, maxiter, x0):
    """cg through the implicit adjoint of the converged solve (x0: a start, which
    gets no gradient)."""
    from .autograd import _PCGFn
    if not isinstance(A, SparseMatrix):
        raise TypeError("gradients need A as an amjax_amd SparseMatrix (see SparseMatrix.with_values)")
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    vals = getattr(A, "values_tensor", None)
    bt = b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b), device=ml.device)
    if vals is None:
        vals = torch.zeros(0, dtype=torch.float64, device=ml.device)
    return _PCGFn.apply(bt, vals, ml, A, tol, atol, maxiter, cycle, cpl, M is not None, x0)


def bicgstab(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Right-preconditioned BiCGSTAB on the GPU; returns ``(x, None)``.  ``A`` is the
    preconditioner hierarchy's finest operator (or any SparseMatrix when M is None).
    ``x0`` starts the iteration as in jax (a shifted solve, as in ``cg``)."""
    if M is not None and not isinstance(M, Preconditioner):
        raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
    ml = M.ml if M is not None else _plain_solver(A)
    if M is not None and A is not ml.levels[0].A:
        raise ValueError("A must be the preconditioner hierarchy's finest operator")
    if x0 is not None:
        r0, s, zero, x02, vec = _x0_shift(ml._op64(), b, x0, tol, atol, ml.device)
        e, _ = bicgstab(A, r0, tol=0.0, atol=1.0, maxiter=maxiter, M=M)
        return _x0_unshift(e, s, zero, x02, vec), None
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    return ml.bicgstab(b, tol=tol, atol=atol, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                       precondition=M is not None), None


def gmres(A, b, x0=None, *, tol=1e-5, atol=0.0, restart=20, maxiter=None, M=None, solve_method="batched"):
    """Restarted GMRES on the GPU (``jax.scipy.sparse.linalg.gmres`` call shape, left
    preconditioning, solve_method 'batched' or 'incremental'); returns ``(x, None)``.
    ``A`` is the preconditioner hierarchy's finest operator (or any SparseMatrix when
    M is None)."""
    if solve_method not in ("batched", "incremental"):
        raise ValueError(f"solve_method {solve_method!r} not supported. Use 'batched' or 'incremental'.")
    if M is not None and not isinstance(M, Preconditioner):
        raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
    ml = M.ml if M is not None else _plain_solver(A)
    if M is not None and A is not ml.levels[0].A:
        raise ValueError("A must be the preconditioner hierarchy's finest operator")
    if x0 is not None:
        # jax stops once ||M (b - A x)|| <= ptol = ||M b|| min(1, max(tol ||b||, atol) / ||b||).
        # The shifted solve on r0 = b - A x0 with tol = 0, atol = 1 stops at
        # ||M r|| <= ||M r0|| / ||r0||; scaling column c of r0 by
        # s_c = ||M r0_c|| / (||r0_c|| ptol_c) makes that threshold ptol_c.
        r0, _, _, x02, vec = _x0_shift(ml._op64(), b, x0, 0.0, 1.0, ml.device)
        bt = (b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b))).to(device=ml.device,
                                                                             dtype=torch.float64)
        b2 = bt.reshape(x02.shape)
        Mf = (lambda v: M(v).to(torch.float64)) if M is not None else (lambda v: v)
        nb = torch.linalg.vector_norm(b2, dim=0)
        nmb = torch.linalg.vector_norm(Mf(b2), dim=0)
        nr0 = torch.linalg.vector_norm(r0, dim=0)
        nmr0 = torch.linalg.vector_norm(Mf(r0), dim=0)
        thr = torch.clamp(tol * nb, min=float(atol))
        ptol = nmb * torch.clamp(thr / torch.where(nb > 0, nb, torch.ones_like(nb)), max=1.0)
        # b_c = 0 (ptol = 0): the iteration tends to the exact solution 0, given directly
        zero = nb == 0
        done = (nmr0 <= ptol) | (nr0 == 0) | zero
        safe = torch.where(done, torch.ones_like(nr0), nr0 * ptol)
        sc = torch.where(done, torch.zeros_like(nr0), nmr0 / safe)
        e, _ = gmres(A, (r0 * sc).contiguous(), tol=0.0, atol=1.0, restart=restart, maxiter=maxiter, M=M,
                     solve_method=solve_method)
        e = e.reshape(x02.shape)
        x = x02 + torch.where(done, torch.zeros_like(sc), 1.0 / torch.where(done, torch.ones_like(sc), sc)) * e
        x[:, zero] = 0.0
        return (x.reshape(-1) if vec else x), None
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    return ml.gmres(b, tol=tol, atol=atol, restart=restart, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                    precondition=M is not None, solve_method=solve_method), None
