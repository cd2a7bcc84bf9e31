"""Krylov front ends: ``cg`` and ``bicgstab(A, b, x0=None, *, tol, atol, maxiter, M)``.

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

import numpy as np
import scipy.sparse as sp
import torch

from .multilevel import MultilevelSolver, Preconditioner
from .sparse import SparseMatrix

_plain = {}


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
        return hit[1]
    H = A.host if isinstance(A, SparseMatrix) else sp.csr_matrix(A)
    ml = MultilevelSolver.from_pyamg(_Hier([_L(sp.csr_matrix(H, dtype=np.float64))]), coarse_solver="jacobi",
                                     coarse_solver_kwargs={"iterations": 0}, dtype=np.float64,
                                     device=A.device if isinstance(A, SparseMatrix) else None)
    _plain[key] = (A, ml)
    return ml


def cg(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Preconditioned conjugate gradients on the GPU; returns ``(x, None)``."""
    if x0 is not None:
        raise NotImplementedError("x0 is not supported; solve for b - A x0 instead")
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
    from .autograd import _PCGFn
    if not isinstance(A, SparseMatrix):
        raise TypeError("gradients need A as an amjax_amd SparseMatrix (see SparseMatrix.with_values)")
    bt = b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b), device=ml.device)
    if vals is None:
        vals = torch.zeros(0, dtype=torch.float64, device=ml.device)
    x = _PCGFn.apply(bt, vals, ml, A, tol, atol, maxiter, cycle, cpl, precondition)
    return x, None


def bicgstab(A, b, x0=None, *, tol=1e-5, atol=0.0, maxiter=None, M=None):
    """Right-preconditioned BiCGSTAB on the GPU; returns ``(x, None)``.  ``A`` is the
    preconditioner hierarchy's finest operator (or any SparseMatrix when M is None)."""
    if x0 is not None:
        raise NotImplementedError("x0 is not supported; solve for b - A x0 instead")
    if M is not None and not isinstance(M, Preconditioner):
        raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
    ml = M.ml if M is not None else _plain_solver(A)
    if M is not None and A is not ml.levels[0].A:
        raise ValueError("A must be the preconditioner hierarchy's finest operator")
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    return ml.bicgstab(b, tol=tol, atol=atol, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                       precondition=M is not None), None


def gmres(A, b, x0=None, *, tol=1e-5, atol=0.0, restart=20, maxiter=None, M=None, solve_method="batched"):
    """Restarted GMRES on the GPU (``jax.scipy.sparse.linalg.gmres`` call shape, its
    default solve_method='batched', left preconditioning); returns ``(x, None)``.
    ``A`` is the preconditioner hierarchy's finest operator (or any SparseMatrix when
    M is None)."""
    if x0 is not None:
        raise NotImplementedError("x0 is not supported; solve for b - A x0 instead")
    if solve_method != "batched":
        raise NotImplementedError("only solve_method='batched' is implemented")
    if M is not None and not isinstance(M, Preconditioner):
        raise TypeError("M must be a MultilevelSolver.aspreconditioner() object")
    ml = M.ml if M is not None else _plain_solver(A)
    if M is not None and A is not ml.levels[0].A:
        raise ValueError("A must be the preconditioner hierarchy's finest operator")
    cycle = M.cycle if M is not None else "V"
    cpl = M.cycles_per_level if M is not None else 1
    return ml.gmres(b, tol=tol, atol=atol, restart=restart, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                    precondition=M is not None), None
