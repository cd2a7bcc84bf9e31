"""Uneven batches of AMG-preconditioned CG solves.

The reference solves one problem per call: ``jax.vmap`` maps a solve over the
right-hand sides of one hierarchy, and its benchmark loops over problems, one
``jax.scipy.sparse.linalg.cg(A, b, M=ml.aspreconditioner())`` call each
(benchmarks/benchmark.ipynb, JSON lines 148-177 and 255-319).  ``BatchPCG``
takes many independent problems, each with its own hierarchy and its own
number of right-hand sides, and advances them together. Every phase of the
V-cycle PCG iteration is one kernel launch across all problems, and each
problem computes exactly what ``MultilevelSolver.pcg`` computes for it alone.
"""

import ctypes

import torch

from . import _lib
from .sparse import _prep


class BatchPCG:
    """AMG-preconditioned CG over a batch of ``MultilevelSolver`` objects.

    ``solvers[p]`` is used with ``ks[p]`` right-hand sides.  All solvers share
    one dtype (float32 hierarchies give the mixed float32-cycle / float64-CG
    solve); ``cycle`` ('V', 'W' or 'F', with
    ``cycles_per_level`` for F) is the preconditioner's cycle.  The batch is rebuilt
    when a solver's native plan changes (new smoothers, fuse setting).
    ``grid`` > 0 runs every launch as a persistent grid of that many blocks;
    ``split_kp`` runs a phase as one launch per column-lane count;
    ``autotune`` gives the row kernels the autotuner's launch shapes.
    """

    def __init__(self, solvers, ks, grid=0, split_kp=True, autotune=True, cycle="V", cycles_per_level=1):
        if len(solvers) != len(ks) or not solvers:
            raise ValueError("need one k per solver and at least one solver")
        self.solvers = list(solvers)
        self.ks = [int(k) for k in ks]
        if any(k < 1 for k in self.ks):
            raise ValueError("every k must be >= 1")
        dts = {ml.dtype for ml in self.solvers}
        if len(dts) != 1:
            raise ValueError("all solvers of a batch must have the same dtype")
        self.grid = int(grid)
        self.split_kp = bool(split_kp)
        self.autotune = bool(autotune)
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        self.cycle = cycle
        self.cycles_per_level = int(cycles_per_level)
        self._b = None
        self._hs = None

    def _native(self):
        # A solver's handle can be destroyed and a new one created at the same
        # address (new smoothers, fuse setting), so the batch is keyed on each
        # solver's handle generation, which never repeats.
        hs = tuple(ml._native().value for ml in self.solvers)
        key = tuple(ml._generation for ml in self.solvers)
        if self._b is not None and key == self._hs:
            return self._b
        self._destroy()
        lib = _lib.load()
        n = len(hs)
        arr = (ctypes.c_void_p * n)(*hs)
        karr = (ctypes.c_int * n)(*self.ks)
        b = ctypes.c_void_p()
        _lib.check(lib.amg_batch_create(arr, karr, n, ctypes.byref(b)), "amg_batch_create")
        self._b, self._hs = b, key
        if self.grid:
            _lib.check(lib.amg_batch_set_option(b, _lib.AMG_BATCH_OPT_GRID, self.grid), "amg_batch_set_option")
        if not self.split_kp:
            _lib.check(lib.amg_batch_set_option(b, _lib.AMG_BATCH_OPT_SPLIT_KP, 0), "amg_batch_set_option")
        if not self.autotune:
            _lib.check(lib.amg_batch_set_option(b, _lib.AMG_BATCH_OPT_AUTOTUNE, 0), "amg_batch_set_option")
        if self.cycle != "V":
            _lib.check(lib.amg_batch_set_option(b, _lib.AMG_BATCH_OPT_CYCLE, ord(self.cycle)), "amg_batch_set_option")
        if self.cycles_per_level != 1:
            _lib.check(lib.amg_batch_set_option(b, _lib.AMG_BATCH_OPT_CPL, self.cycles_per_level),
                       "amg_batch_set_option")
        return b

    def _destroy(self):
        if self._b is not None:
            _lib.load().amg_batch_destroy(self._b)
        self._b = None
        self._hs = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def info(self):
        """(kernel launches, virtual blocks) of one batched PCG iteration."""
        nl, nb = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().amg_batch_info(self._native(), ctypes.byref(nl), ctypes.byref(nb)), "amg_batch_info")
        return nl.value, nb.value

    def pcg(self, Bs, tol=1e-5, atol=0.0, maxiter=None, flexible=False, return_info=False):
        """Solve A_p X_p = B_p for every problem (``MultilevelSolver.pcg`` semantics per column).

        ``Bs[p]`` has shape (n_p,) or (n_p, ks[p]); returns the list of float64
        solutions, plus per-problem lists of iteration counts and true relative
        residuals with ``return_info``.
        """
        if len(Bs) != len(self.solvers):
            raise ValueError("need one right-hand side block per solver")
        dev = self.solvers[0].device
        bs, xs, vec = [], [], []
        for ml, k, B in zip(self.solvers, self.ks, Bs):
            b = _prep(B, torch.float64, dev)
            vec.append(b.dim() == 1)
            b2 = b.reshape(b.shape[0], -1)
            if b2.shape != (ml.levels[0].A.shape[0], k):
                raise ValueError(f"right-hand side of shape {tuple(b.shape)} does not match n = "
                                 f"{ml.levels[0].A.shape[0]}, k = {k}")
            bs.append(b2)
            xs.append(torch.empty_like(b2))
        if maxiter is None:
            maxiter = -1  # per problem: 10 n_p, as a single MultilevelSolver.pcg
        n = len(bs)
        kt = sum(self.ks)
        its = (ctypes.c_int * kt)()
        rel = (ctypes.c_double * kt)()
        barr = (ctypes.c_void_p * n)(*[b.data_ptr() for b in bs])
        xarr = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        _lib.check(_lib.load().amg_batch_pcg(self._native(), barr, xarr, float(tol), float(atol), int(maxiter),
                                             int(bool(flexible)), its, rel, _lib.stream_ptr(dev)), "amg_batch_pcg")
        out = [x.reshape(-1) if v else x for x, v in zip(xs, vec)]
        if not return_info:
            return out
        it_l, rr_l, o = [], [], 0
        for k in self.ks:
            it_l.append(list(its[o:o + k]))
            rr_l.append(list(rel[o:o + k]))
            o += k
        return out, it_l, rr_l
