"""amjax_amd: the multigrid solve phase of AMJax, native to MI355X (gfx950).

Public surface mirrors ``amjax/__init__.py:1-10`` (MultilevelSolver,
change_smoothers, jacobi, inverse_diagonal), plus the device sparse matrix,
the in-repo hierarchy builders that stand in for PyAMG, and the Krylov
front end, and the batched solve of many independent problems.
"""

from .multilevel import MultilevelSolver, coarse_grid_solver, multilevel_solver
from .relaxation.smoothing import change_smoothers
from .relaxation.relaxation import jacobi, inverse_diagonal
from .sparse import SparseMatrix
from .krylov import bicgstab, cg, gmres
from .autograd import cycles
from .batch import BatchPCG
from . import gallery, hierarchy

__all__ = [
    "MultilevelSolver",
    "change_smoothers",
    "jacobi",
    "inverse_diagonal",
    "coarse_grid_solver",
    "multilevel_solver",
    "SparseMatrix",
    "cg",
    "bicgstab",
    "gmres",
    "cycles",
    "BatchPCG",
    "gallery",
    "hierarchy",
]
