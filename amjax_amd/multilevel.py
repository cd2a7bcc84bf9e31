"""MultilevelSolver: the multigrid solve phase on MI355X (mirror of amjax/multilevel.py).

Host code on PyTorch-ROCm tensors.  ``from_pyamg`` packs a PyAMG-shaped
hierarchy (any object with ``levels[i].A/.P/.R`` as scipy sparse matrices) once
into level-packed sliced-ELL buffers on the GPU; ``solve``,
``aspreconditioner`` and ``pcg`` call the gfx950 kernels through the C-ABI
(include/amjax_amd.h).  There is no CPU execution path: without the built
library or a GPU these methods raise.

Precision: the hierarchy dtype (``dtype=``, default the finest A's) is the
arithmetic type of the cycle.  A float32 hierarchy keeps a float64 copy of the
finest operator, so ``solve`` on a float64 right-hand side runs iterative
refinement with float64 residuals and ``pcg`` runs float64 CG with the
float32 cycle as preconditioner ("mixed").
"""

import ctypes
from warnings import warn

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import torch

from . import _lib
from ._pack import Packer
from .relaxation import relaxation, smoothing
from .sparse import SparseMatrix, as_torch_dtype, amg_dtype, default_device, np_dtype, _prep


class _Level:
    """One level of the hierarchy (amjax/multilevel.py:16-58)."""

    def __init__(self):
        self.A = None
        self.Dinv = None
        self.A_inv = None
        self.A_dense = None
        self.lu_factor = None
        self.piv = None
        self.Q = None
        self.R_mat = None
        self.P = None
        self.R = None
        self.presmoother = None
        self.postsmoother = None
        self._presmoother_spec = None
        self._postsmoother_spec = None


# Reliable-update CG: an update when the recurrence residual has fallen by this
# factor since the last one.
RU_DELTA = 1e-3

_NOT_IMPLEMENTED_SOLVERS = {
    "cholesky", "pinv2", "splu", "cg", "gmres", "bicgstab",
    "bicg", "cgs", "qmr", "minres", "fgmres", "gauss_seidel",
    "sor", "block_jacobi", "block_gauss_seidel",
}
_COARSE = ("jacobi", "pinv", "lu", "qr")


def _check_coarse_name(name):
    if name in _NOT_IMPLEMENTED_SOLVERS:
        raise NotImplementedError(f"'{name}' coarse solver is not yet implemented in amjax_amd. "
                                  f"Available: {list(_COARSE)}")
    if name not in _COARSE:
        raise ValueError(f"Unknown coarse solver {name!r}. Available: {list(_COARSE)}")


def _dense_solve(kind, f0, f1, b):
    """x = solve(b) with a dense factor on the GPU (pinv / lu / qr)."""
    dt, dev = f0.dtype, f0.device
    b = _prep(b, dt, dev)
    k = 1 if b.dim() == 1 else b.shape[1]
    x = torch.empty_like(b)
    desc = _lib.CoarseDesc(kind=_lib.COARSE_KIND[kind], n=int(f0.shape[0]), f0=f0.data_ptr(),
                           f1=(f1.data_ptr() if f1 is not None else None), iters=0, omega=1.0)
    _lib.check(_lib.load().amg_dense_solve(ctypes.byref(desc), amg_dtype(dt), b.data_ptr(), x.data_ptr(), k,
                                           _lib.stream_ptr(dev)), "amg_dense_solve")
    return x


def coarse_grid_solver(solver_name, lvl, **kwargs):
    """Coarse-grid solver ``(A, x, b) -> x`` (amjax/multilevel.py:455-498).

    'pinv' applies ``lvl.A_inv``, 'lu' ``lvl.lu_factor``/``lvl.piv``, 'qr'
    ``lvl.Q``/``lvl.R_mat`` (dense GPU kernels); 'jacobi' runs
    ``iterations`` (default 10) Jacobi sweeps with ``omega`` (default 1.0).
    """
    _check_coarse_name(solver_name)
    if solver_name == "jacobi":
        iterations = kwargs.get("iterations", 10)
        omega = kwargs.get("omega", 1.0)
        Dinv = lvl.Dinv

        def solve(A, x, b):
            return relaxation.jacobi(A, x, b, Dinv, iterations=iterations, omega=omega)
    elif solver_name == "pinv":
        A_inv = lvl.A_inv

        def solve(A, x, b):
            del A, x
            return _dense_solve("pinv", _contig(A_inv), None, b)
    elif solver_name == "lu":
        lu, piv = lvl.lu_factor, lvl.piv

        def solve(A, x, b):
            del A, x
            return _dense_solve("lu", _contig(lu), _contig(piv, torch.int32), b)
    else:
        Q, R_mat = lvl.Q, lvl.R_mat

        def solve(A, x, b):
            del A, x
            return _dense_solve("qr", _contig(Q), _contig(R_mat), b)
    solve.__name__ = solver_name
    return solve


def _contig(t, dtype=None):
    if not torch.is_tensor(t):
        t = torch.as_tensor(np.asarray(t), device=default_device())
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _pinv(A_dense, rtol=None, rcond=None, hermitian=False):
    """SVD pseudo-inverse with ``jnp.linalg.pinv``'s default cutoff
    ``rtol = 10 * max(m, n) * eps(dtype)`` relative to the largest singular value."""
    A = np.asarray(A_dense)
    dt = A.dtype
    if rtol is None:
        rtol = rcond if rcond is not None else 10 * max(A.shape) * np.finfo(dt).eps
    U, s, Vt = np.linalg.svd(A.astype(np.float64), full_matrices=False, hermitian=hermitian)
    cutoff = rtol * (s.max() if s.size else 0.0)
    sinv = np.where(s > cutoff, 1.0 / np.where(s > cutoff, s, 1.0), 0.0)
    return ((Vt.T * sinv) @ U.T).astype(dt)


class Preconditioner:
    """``matvec(b) -> x``: one cycle from a zero guess (amjax/multilevel.py:265-288)."""

    def __init__(self, ml, cycle="V", cycles_per_level=1):
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        self.ml = ml
        self.cycle = cycle
        self.cycles_per_level = cycles_per_level

    def __call__(self, b):
        return self.ml._apply_cycle(None, b, self.cycle, self.cycles_per_level)

    matvec = __call__


class MultilevelSolver:
    """Algebraic multigrid hierarchy with a gfx950 solve phase (amjax/multilevel.py:61-394)."""

    Level = _Level
    _handles_created = 0  # process-wide count: handle generations never repeat

    def __init__(self, levels, coarse_solver_fn, coarse_solver_name="pinv", coarse_solver_kwargs=None):
        self.levels = levels
        self.coarse_solver = coarse_solver_fn
        self._coarse_solver_name = coarse_solver_name
        self._coarse_solver_kwargs = coarse_solver_kwargs or {}
        self.symmetric_smoothing = False
        self._outer = None          # float64 level-0 operator of a float32 hierarchy
        self._handle = None
        self._generation = 0        # bumped for every native handle created
        self._key = None
        self._keep = None
        self._derived = {}
        # Execution strategy (no effect on the mathematics): fuse the zero-guess
        # pre-sweep, residual and restriction into one product with the
        # precomputed operator R (I - w A D^-1).  Changes only the rounding order.
        self.fuse = True
        # Launch shapes of the row kernels from the measured autotuner cache
        # (column lanes per row group, wide-row variant) instead of the fixed rule.
        self.autotune = False
        # Row kernels read 16 bytes of a row per lane where k allows it (same
        # row sums bit for bit; dot products grouped differently).
        self.vector = True
        # Largest first level of the one-launch coarse tail of zero-guess
        # V-cycles (None: the library default; 0: one launch per level
        # operation).  Same sums, bit for bit.
        self.tail = None
        for lvl in self.levels[:-1]:
            if getattr(lvl, "R", None) is None:
                lvl.R = lvl.P.T

    # ------------------------------------------------------------ construction
    @classmethod
    def from_pyamg(cls, pyamg_solver, presmoother=("jacobi", {"iterations": 1, "omega": 1.0}),
                   postsmoother=("jacobi", {"iterations": 1, "omega": 1.0}), coarse_solver="pinv",
                   coarse_solver_kwargs=None, dtype=None, device=None):
        """Upload a PyAMG-shaped hierarchy (amjax/multilevel.py:125-186, 557-590).

        Same arguments as the reference, plus ``device`` (default: the current
        ROCm device).  Level matrices are cast to ``dtype`` (default: the finest
        A's); ``R`` defaults to ``P^H``; the coarse factor is computed once from
        the densified coarsest A.  Smoother and coarse-solver names are checked
        before anything is uploaded.
        """
        kw = dict(coarse_solver_kwargs or {})
        _check_coarse_name(coarse_solver)
        n = len(pyamg_solver.levels) - 1
        if n > 0:
            for spec in (smoothing._expand_to_n(smoothing._as_list(presmoother), n)
                         + smoothing._expand_to_n(smoothing._as_list(postsmoother), n)):
                smoothing._setup_call(smoothing._unpack_arg(spec)[0])
        levels, outer = _convert_hierarchy(pyamg_solver, dtype, coarse_solver, kw, device)
        fn = coarse_grid_solver(coarse_solver, levels[-1], **kw)
        ml = cls(levels, fn, coarse_solver_name=coarse_solver, coarse_solver_kwargs=kw)
        ml._outer = outer
        smoothing.change_smoothers(ml, presmoother, postsmoother)
        return ml

    # ------------------------------------------------------------ diagnostics
    def __repr__(self):
        total = sum(l.A.nnz for l in self.levels)
        out = "MultilevelSolver\n"
        out += f"Number of Levels: {len(self.levels)}\n"
        out += f"Operator Complexity: {self.operator_complexity():6.3f}\n"
        out += f"Grid Complexity: {self.grid_complexity():6.3f}\n"
        out += f"Cycle Complexity: {self.cycle_complexity():6.3f}\n"
        out += f"Coarse Solver: {self._coarse_solver_name!r}\n"
        out += "level unknowns nonzeros\n"
        for i, lvl in enumerate(self.levels):
            out += f"{i:>6} {lvl.A.shape[0]:>11} {lvl.A.nnz:>12} [{100 * lvl.A.nnz / total:2.2f}%]\n"
        return out

    def operator_complexity(self):
        """sum_l nnz(A_l) / nnz(A_0) (amjax/multilevel.py:204-214)."""
        nnz = [_nnz(l.A) for l in self.levels]
        return sum(nnz) / float(nnz[0])

    def grid_complexity(self):
        """sum_l n_l / n_0."""
        n = [l.A.shape[0] for l in self.levels]
        return sum(n) / float(n[0])

    def cycle_complexity(self, cycle="V"):
        """Work of one cycle in units of nnz(A_0) (amjax/multilevel.py:216-263)."""
        cycle = str(cycle).upper()
        nnz = [_nnz(l.A) for l in self.levels]
        L = len(nnz)

        def V(l):
            if L == 1:
                return nnz[0]
            if l == L - 2:
                return 2 * nnz[l] + nnz[l + 1]
            return 2 * nnz[l] + V(l + 1)

        def W(l):
            if L == 1:
                return nnz[0]
            if l == L - 2:
                return 2 * nnz[l] + nnz[l + 1]
            return 2 * nnz[l] + 2 * W(l + 1)

        def F(l):
            if L == 1:
                return nnz[0]
            if l == L - 2:
                return 2 * nnz[l] + nnz[l + 1]
            return 2 * nnz[l] + F(l + 1) + V(l + 1)

        fn = {"V": V, "W": W, "F": F}.get(cycle)
        if fn is None:
            raise NotImplementedError(f"Cycle complexity for {cycle!r} not implemented.")
        return float(fn(0)) / float(nnz[0])

    # ------------------------------------------------------------ native plan
    def _invalidate(self):
        if self._handle is not None:
            _lib.load().amg_destroy(self._handle)
        self._handle = None
        self._key = None
        self._keep = None

    def __del__(self):
        try:
            self._invalidate()
        except Exception:
            pass

    @property
    def dtype(self):
        return self.levels[0].A.dtype

    @property
    def device(self):
        return self.levels[0].A.device

    def tail_level(self):
        """First level whose zero-guess V-cycle visits run, with every visit
        below, as one coarse-tail launch (-1: none)."""
        lvl = ctypes.c_int(-1)
        _lib.check(_lib.load().amg_tail_level(self._native(), ctypes.byref(lvl)), "amg_tail_level")
        return lvl.value

    def _native(self):
        """Create (or reuse) the C-ABI handle for the current levels and smoothers."""
        L = len(self.levels)
        dt = self.dtype
        name = self._coarse_solver_name
        specs = []
        keep = []
        key = [id(self._outer), name, tuple(sorted(self._coarse_solver_kwargs.items()))]
        descs = (_lib.LevelDesc * L)()
        for i, lvl in enumerate(self.levels):
            if not isinstance(lvl.A, SparseMatrix) or lvl.A.dtype != dt:
                raise TypeError(f"level {i}: A must be an amjax_amd SparseMatrix of dtype {dt}")
            d = descs[i]
            d.A = lvl.A._desc
            dinv = _contig(lvl.Dinv, dt)
            keep.append(dinv)
            d.dinv = dinv.data_ptr()
            key += [id(lvl.A), dinv.data_ptr()]
            if i < L - 1:
                for nm in ("P", "R"):
                    M = getattr(lvl, nm)
                    if not isinstance(M, SparseMatrix) or M.dtype != dt:
                        raise TypeError(f"level {i}: {nm} must be an amjax_amd SparseMatrix of dtype {dt}")
                    key.append(id(M))
                d.P = lvl.P._desc
                d.R = lvl.R._desc
                for which, sm in (("pre", lvl.presmoother), ("post", lvl.postsmoother)):
                    kind = getattr(sm, "_kind", "none")
                    its = int(getattr(sm, "_iterations", 0))
                    om = float(getattr(sm, "_omega", 1.0))
                    code = {"jacobi": _lib.AMG_SMOOTH_JACOBI, "poly": _lib.AMG_SMOOTH_POLY}.get(
                        kind, _lib.AMG_SMOOTH_NONE)
                    setattr(d, which + "_kind", code)
                    setattr(d, which + "_iters", its)
                    setattr(d, which + "_omega", om)
                    coefs = tuple(getattr(sm, "_coefficients", ()) if kind == "poly" else ())
                    if coefs:
                        arr = (ctypes.c_double * len(coefs))(*coefs)
                        keep.append(arr)
                        setattr(d, which + "_coef", ctypes.cast(arr, ctypes.c_void_p).value)
                        setattr(d, which + "_ncoef", len(coefs))
                    specs.append((kind, its, om, coefs))
        c = self.levels[-1]
        cd = _lib.CoarseDesc(kind=_lib.COARSE_KIND[name], n=int(c.A.shape[0]), f0=None, f1=None, iters=0,
                             omega=1.0)
        fac = self._coarse_factors64() if dt == torch.float32 and name != "jacobi" else None
        if fac is not None:
            # float32 cycle: the coarse solve applies float64 factors (computed from
            # the float64 coarse matrix) and rounds only its result
            cd.factor_f64 = 1
            keep += list(fac)
            cd.f0 = fac[0].data_ptr()
            cd.f1 = fac[1].data_ptr() if len(fac) > 1 else None
        elif name == "pinv":
            f0 = _contig(c.A_inv, dt)
            keep.append(f0)
            cd.f0 = f0.data_ptr()
        elif name == "lu":
            f0, f1 = _contig(c.lu_factor, dt), _contig(c.piv, torch.int32)
            keep += [f0, f1]
            cd.f0, cd.f1 = f0.data_ptr(), f1.data_ptr()
        elif name == "qr":
            f0, f1 = _contig(c.Q, dt), _contig(c.R_mat, dt)
            keep += [f0, f1]
            cd.f0, cd.f1 = f0.data_ptr(), f1.data_ptr()
        else:
            cd.iters = int(self._coarse_solver_kwargs.get("iterations", 10))
            cd.omega = float(self._coarse_solver_kwargs.get("omega", 1.0))
        key += [cd.f0, cd.f1] + specs + [self.fuse, self.autotune, self.vector, self.tail]
        key = tuple(key)
        if self._handle is not None and key == self._key:
            return self._handle
        self._invalidate()
        for i in range(L - 1):
            D = self._down_operator(i) if self.fuse else None
            if D is not None:
                descs[i].D = D._desc
                keep.append(D)
        outer = None
        if self._outer is not None:
            outer = ctypes.byref(self._outer._desc)
        h = ctypes.c_void_p()
        lib = _lib.load()
        _lib.check(lib.amg_create(descs, L, ctypes.byref(cd), amg_dtype(dt), outer, ctypes.byref(h)), "amg_create")
        self._handle, self._key, self._keep = h, key, keep
        MultilevelSolver._handles_created += 1
        self._generation = MultilevelSolver._handles_created
        _lib.check(lib.amg_set_option(h, _lib.AMG_OPT_FUSE, int(bool(self.fuse))), "amg_set_option")
        _lib.check(lib.amg_set_option(h, _lib.AMG_OPT_AUTOTUNE, int(bool(self.autotune))), "amg_set_option")
        _lib.check(lib.amg_set_option(h, _lib.AMG_OPT_VECTOR, int(bool(self.vector))), "amg_set_option")
        if self.tail is not None:
            _lib.check(lib.amg_set_option(h, _lib.AMG_OPT_TAIL, int(self.tail)), "amg_set_option")
        return h

    def _coarse_factors64(self):
        """float64 coarse factors of the coarsest A (pinv: A_inv; lu: LU, piv; qr: Q, R)."""
        c = self.levels[-1]
        key = (id(c.A), self._coarse_solver_name, tuple(sorted(self._coarse_solver_kwargs.items())))
        hit = getattr(self, "_fac64", None)
        if hit is not None and hit[0] == key:
            return hit[1]
        Ac = sp.csr_matrix(getattr(c, "_A64", None) if getattr(c, "_A64", None) is not None else c.A.host,
                           dtype=np.float64).toarray()
        dev = self.device
        name = self._coarse_solver_name
        if name == "pinv":
            fac = (torch.as_tensor(_pinv(Ac, **self._coarse_solver_kwargs), device=dev),)
        elif name == "lu":
            lu, piv = sla.lu_factor(Ac)
            fac = (torch.as_tensor(lu, device=dev), torch.as_tensor(piv.astype(np.int32), device=dev))
        else:
            Q, Rm = np.linalg.qr(Ac)
            fac = (torch.as_tensor(Q, device=dev), torch.as_tensor(Rm, device=dev))
        fac = tuple(f.contiguous() for f in fac)
        self._fac64 = (key, fac)
        return fac

    def _down_operator(self, i):
        """R_i (I - w A_i D_i^-1) for a level whose pre-smoother is one Jacobi sweep
        (the restriction of the residual after a zero-guess sweep), or None."""
        lvl = self.levels[i]
        sm = lvl.presmoother
        if getattr(sm, "_kind", "none") != "jacobi" or getattr(sm, "_iterations", 0) != 1:
            return None
        w = float(sm._omega)
        dinv = _contig(lvl.Dinv, self.dtype)
        key = (id(lvl.A), id(lvl.R), dinv.data_ptr(), w, self.dtype)
        hit = self._derived.get(i)
        if hit is not None and hit[0] == key:
            return hit[1]
        A = sp.csr_matrix(lvl.A.host, dtype=np.float64)
        R = sp.csr_matrix(lvl.R.host, dtype=np.float64)
        d = dinv.detach().cpu().numpy().astype(np.float64)
        D = (R - w * ((R @ A) @ sp.diags(d))).tocsr()
        D.eliminate_zeros()
        D.sort_indices()
        Dm = SparseMatrix.from_scipy(D, self.dtype, self.device)
        self._derived[i] = (key, Dm)
        return Dm

    # ------------------------------------------------------------ cycles
    def aspreconditioner(self, cycle="V", cycles_per_level=1):
        """Preconditioner applying one cycle from a zero guess (amjax/multilevel.py:265-288)."""
        return Preconditioner(self, cycle, cycles_per_level)

    def _apply_cycle(self, x, b, cycle, cycles_per_level, A=None):
        # 'T' (internal): the transposed F-cycle, used by the adjoint of fixed F-cycles
        if cycle not in ("V", "W", "F", "T"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        dt, dev = self.dtype, self.device
        b = _prep(b, dt, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        k = b2.shape[1]
        out = torch.empty_like(b2)
        h = self._native()
        with _Level0(self, h, A):
            if x is None:
                code = _lib.load().amg_cycle(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(),
                                             out.data_ptr(), k, 1, _lib.stream_ptr(dev))
            else:
                out.copy_(_prep(x, dt, dev).reshape(b2.shape))
                code = _lib.load().amg_cycle(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(),
                                             out.data_ptr(), k, 0, _lib.stream_ptr(dev))
        _lib.check(code, "amg_cycle")
        return out.reshape(-1) if vec else out

    def _cycle(self, x, b, cycle="V", cycles_per_level=1, A=None):
        """One cycle ``x <- cycle(x, b)`` (amjax/multilevel.py:290-339)."""
        return self._apply_cycle(x, b, cycle, cycles_per_level, A=A)

    # ------------------------------------------------------------ solves
    def solve(self, b, A=None, x0=None, tol=1e-5, maxiter=100, cycle="V", cycles_per_level=1,
              return_info=False):
        """Solve ``A x = b`` by repeated cycles (amjax/multilevel.py:341-425).

        ``b`` is (n,) or (n, k); each column is an independent solve that stops
        at ``||b - A x|| < tol ||b||`` or after ``maxiter`` cycles (a converged
        column is frozen, as under ``jax.vmap``).  ``A`` (optional,
        SparseMatrix of the same shape) replaces the finest operator of the
        smoothing, residuals and stopping test while the rest of the hierarchy
        stays fixed.  With a float32 hierarchy and a float64 ``b`` the loop is
        iterative refinement with float64 residuals.  ``return_info`` adds
        per-column iteration counts and final relative residuals.
        """
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        if _requires_grad(b, A) or (torch.is_grad_enabled() and torch.is_tensor(x0) and x0.requires_grad):
            if return_info:
                raise ValueError("return_info is not available on the differentiable path")
            from .autograd import solve_with_grad
            return solve_with_grad(self, b, self.levels[0].A if A is None else A, x0, tol, maxiter, cycle,
                                   cycles_per_level)
        return self._solve_nograd(b, A, x0, tol, maxiter, cycle, cycles_per_level, return_info)

    def _solve_nograd(self, b, A, x0, tol, maxiter, cycle, cycles_per_level, return_info=False):
        if A is not None and not isinstance(A, SparseMatrix):
            from .autograd import as_operator
            A = as_operator(self, A)[0]
        dev = self.device
        bt = b if torch.is_tensor(b) else torch.as_tensor(np.asarray(b))
        refine = self.dtype == torch.float32 and bt.dtype == torch.float64 and self._outer is not None
        wdt = torch.float64 if refine else self.dtype
        b = _prep(bt, wdt, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        if x0 is not None:
            x0 = _prep(x0, wdt, dev).reshape(b2.shape)
            op = self._op64() if refine else self._op(A)
            b2 = b2 - op.matmul(x0)
        x, its, rel = self._solve_native(b2, tol, maxiter, cycle, cycles_per_level, refine, A)
        if x0 is not None:
            x = x + x0
        x = x.reshape(-1) if vec else x
        if return_info:
            return x, (its[0] if vec else its), (rel[0] if vec else rel)
        return x

    def _op(self, A=None):
        return self.levels[0].A if A is None else A

    def _op64(self):
        return self._outer if self._outer is not None else self.levels[0].A

    def _solve_native(self, b2, tol, maxiter, cycle, cpl, refine, A=None):
        n, k = b2.shape
        x = torch.empty_like(b2)
        its = (ctypes.c_int * k)()
        rel = (ctypes.c_double * k)()
        h = self._native()
        with _Level0(self, h, A):
            code = _lib.load().amg_solve(h, cycle.encode(), int(cpl), b2.data_ptr(), x.data_ptr(), k, float(tol),
                                         int(maxiter), int(refine), its, rel, _lib.stream_ptr(self.device))
        _lib.check(code, "amg_solve")
        return x, list(its), list(rel)

    def _pcg_with_operator(self, b, Aop, tol, atol, maxiter, cycle, cpl, precondition):
        """PCG on the float64 operator Aop (same pattern and shape as the finest A)."""
        if self.dtype != torch.float64:
            raise NotImplementedError("an operator override in PCG needs a float64 hierarchy")
        h = self._native()
        with _Level0(self, h, Aop):
            return self.pcg(b, tol=tol, atol=atol, maxiter=maxiter, cycle=cycle, cycles_per_level=cpl,
                            precondition=precondition)

    def transpose(self):
        """The solver of the transposed cycle: A_l^T, R'_l = P_l^T, P'_l = R_l^T,
        pre- and post-smoothers exchanged (same resolved omegas), transposed
        coarse factor.  Its zero-guess cycle applies B^T when this one applies B
        (the F-cycle's transpose is its cycle 'T').  When every A_l is symmetric,
        every R_l equals P_l^T and each level's pre- and post-smoother are alike,
        the transposed hierarchy is this one and ``self`` is returned (no second
        upload); otherwise the transposed solver is uploaded once and cached."""
        cached = getattr(self, "_transposed", None)
        if cached is not None and cached[0] == self._key_light():
            return cached[1]

        def spec(sm):
            kind = getattr(sm, "_kind", "none")
            if kind == "poly":
                # p(A)^T = p(A^T): the same coefficients on the transposed operator
                return (sm.__name__, {"iterations": sm._iterations, "coefficients": list(sm._coefficients)})
            if kind != "jacobi":
                return None
            return ("jacobi", {"iterations": sm._iterations, "omega": sm._omega})

        if self._self_transposed(spec):
            self._transposed = (self._key_light(), self)
            return self

        class _H:
            pass

        hier = _H()
        hier.levels = []
        for i, lvl in enumerate(self.levels):
            L = _H()
            A = sp.csr_matrix(lvl.A.host)
            L.A = A if abs(A - A.T).max() == 0 else sp.csr_matrix(A.T)
            if i < len(self.levels) - 1:
                L.P = sp.csr_matrix(lvl.R.host.T)
                L.R = sp.csr_matrix(lvl.P.host.T)
            hier.levels.append(L)

        pre = [spec(l.postsmoother) for l in self.levels[:-1]] or [None]
        post = [spec(l.presmoother) for l in self.levels[:-1]] or [None]
        kw = dict(self._coarse_solver_kwargs)
        mt = MultilevelSolver.from_pyamg(hier, pre, post, coarse_solver=self._coarse_solver_name,
                                         coarse_solver_kwargs=kw, dtype=self.dtype, device=self.device)
        mt.fuse = self.fuse
        mt.autotune = self.autotune
        mt.vector = self.vector
        mt.tail = self.tail
        self._transposed = (self._key_light(), mt)
        return mt

    def _self_transposed(self, spec):
        """True when the transposed hierarchy is this one: every A_l symmetric (the
        Galerkin products R A P are symmetric only to within rounding, so A_l may
        differ from A_l^T by 8 eps max|A_l| entry-wise), every R_l equal to P_l^T and
        each level's pre- and post-smoother alike."""
        def same(X, Y, ulps=0):
            X, Y = sp.csr_matrix(X), sp.csr_matrix(Y)
            if X.shape != Y.shape:
                return False
            D = abs(X - Y)
            if D.nnz == 0:
                return True
            if ulps == 0:
                return False
            return bool(D.max() <= ulps * np.finfo(X.dtype).eps * abs(X).max())

        for i, lvl in enumerate(self.levels):
            A = lvl.A.host
            if not same(A, A.T, ulps=8):
                return False
            if i < len(self.levels) - 1:
                if not same(lvl.R.host, lvl.P.host.T):
                    return False
                if spec(lvl.presmoother) != spec(lvl.postsmoother):
                    return False
        return True

    def _key_light(self):
        return tuple((id(l.A), id(getattr(l, "P", None)), getattr(l.presmoother, "_omega", None),
                      getattr(l.postsmoother, "_omega", None)) for l in self.levels)

    def pcg(self, b, tol=1e-5, atol=0.0, maxiter=None, cycle="V", cycles_per_level=1, flexible=False,
            precondition=True, return_info=False):
        """AMG-preconditioned CG in float64 (``jax.scipy.sparse.linalg.cg`` semantics).

        ``M = aspreconditioner(cycle)``; x0 = 0; each column stops when
        ``r.r <= max(tol^2 b.b, atol^2)`` on the recurrence residual or after
        ``maxiter`` (default 10 n) iterations.  With a float32 hierarchy the
        cycle runs in float32.  ``flexible`` selects the Polak-Ribiere beta
        (for a preconditioner that varies between iterations); the default is
        jax cg's textbook beta r.z / (r.z)_old, with which the float32 cycle
        converges in the same iteration counts here.  Returns x (float64), plus per-column iteration
        counts and true relative residuals with ``return_info``.
        """
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        dev = self.device
        b = _prep(b, torch.float64, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        n, k = b2.shape
        if maxiter is None:
            maxiter = 10 * n
        x = torch.empty_like(b2)
        its = (ctypes.c_int * k)()
        rel = (ctypes.c_double * k)()
        h = self._native()
        # the true residuals cost one more residual product: only when they are returned
        _lib.check(_lib.load().amg_pcg(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(), x.data_ptr(), k,
                                       float(tol), float(atol), int(maxiter), int(bool(flexible)),
                                       int(bool(precondition)), its, rel if return_info else None,
                                       _lib.stream_ptr(dev)), "amg_pcg")
        x = x.reshape(-1) if vec else x
        if return_info:
            return x, (list(its)[0] if vec else list(its)), (list(rel)[0] if vec else list(rel))
        return x

    def bicgstab(self, b, tol=1e-5, atol=0.0, maxiter=None, cycle="V", cycles_per_level=1, precondition=True,
                 return_info=False):
        """Right-preconditioned BiCGSTAB in float64 (``jax.scipy.sparse.linalg.bicgstab``
        semantics) with one cycle as M; per-column stopping on the recurrence residual,
        early exit when s.s < max(tol^2 b.b, atol^2), stop on breakdown.  A float32
        hierarchy gives the float32-cycle / float64-iteration form."""
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        dev = self.device
        b = _prep(b, torch.float64, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        n, k = b2.shape
        if maxiter is None:
            maxiter = 10 * n
        x = torch.empty_like(b2)
        its = (ctypes.c_int * k)()
        rel = (ctypes.c_double * k)()
        h = self._native()
        _lib.check(_lib.load().amg_bicgstab(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(), x.data_ptr(),
                                            k, float(tol), float(atol), int(maxiter), int(bool(precondition)),
                                            its, rel, _lib.stream_ptr(dev)), "amg_bicgstab")
        x = x.reshape(-1) if vec else x
        if return_info:
            return x, (list(its)[0] if vec else list(its)), (list(rel)[0] if vec else list(rel))
        return x

    def gmres(self, b, tol=1e-5, atol=0.0, restart=20, maxiter=None, cycle="V", cycles_per_level=1,
              precondition=True, return_info=False, solve_method="batched"):
        """Restarted GMRES in float64 (``jax.scipy.sparse.linalg.gmres`` semantics,
        solve_method 'batched' (its default) or 'incremental'), left-preconditioned by
        one cycle; per-column stopping on the preconditioned residual.  A float32 hierarchy gives the
        float32-cycle / float64-iteration form.  With ``return_info``: x, restarts per
        column and the true relative residuals ||b - A x|| / ||b||."""
        return _gmres_call(self, b, tol, atol, restart, maxiter, cycle, cycles_per_level, precondition,
                           return_info, solve_method)

    def pcg_ru(self, b, tol=1e-5, atol=0.0, delta=RU_DELTA, maxiter=None, cycle="V", cycles_per_level=1,
               flexible=False, return_info=False):
        """Preconditioned CG with reliable updates: the mixed-precision form without restarts.

        The CG vectors (correction, recurrence residual, p, Ap) are in the
        hierarchy dtype (float32 for a float32 hierarchy), dot products
        accumulate in float64 and one cycle preconditions; x = 0.  When a
        column's recurrence residual has fallen by ``delta`` since its last
        reliable update, reaches the tolerance, or the column reaches
        ``maxiter`` (default 10 n), the residual is replaced by the float64
        residual of x + correction, keeping the search direction: b - A x is
        kept in float64 and each update subtracts A (correction) from it (equal
        to b - A (x + correction) in exact arithmetic), then x += correction in
        float64.  A column stops on that residual:
        ``||b - A x|| <= max(tol ||b||, atol)``.  With ``return_info``: x, CG
        iterations, reliable updates and the relative residuals of the last
        updates.
        """
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        if not 0.0 < delta < 1.0:
            raise ValueError("delta must be in (0, 1)")
        dev = self.device
        b = _prep(b, torch.float64, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        n, k = b2.shape
        if maxiter is None:
            maxiter = 10 * n
        x = torch.empty_like(b2)
        its = (ctypes.c_int * k)()
        ups = (ctypes.c_int * k)()
        rel = (ctypes.c_double * k)()
        h = self._native()
        _lib.check(_lib.load().amg_pcg_ru(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(), x.data_ptr(), k,
                                          float(tol), float(atol), float(delta), int(maxiter), int(bool(flexible)),
                                          its, ups, rel, _lib.stream_ptr(dev)), "amg_pcg_ru")
        x = x.reshape(-1) if vec else x
        if return_info:
            one = (lambda a: list(a)[0]) if vec else list
            return x, one(its), one(ups), one(rel)
        return x

    def pcg_ir(self, b, tol=1e-5, inner_tol=1e-4, maxiter=20, inner_maxiter=None, cycle="V", cycles_per_level=1,
               flexible=False, return_info=False):
        """Float64 iterative refinement around a CG in the hierarchy's dtype.

        Per column: x = 0; while ||b - A x|| >= tol ||b|| (float64 residual with the
        float64 finest operator) and fewer than ``maxiter`` steps, solve A e = r with
        preconditioned CG whose vectors are in the hierarchy dtype (float32 for a
        float32 hierarchy; dot products accumulated in float64) to ``inner_tol``
        relative to its right-hand side, then x += e.  With ``return_info``: x, the
        inner CG iterations summed over the steps, the number of steps and the true
        relative residuals.
        """
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        dev = self.device
        b = _prep(b, torch.float64, dev)
        vec = b.dim() == 1
        b2 = b.reshape(b.shape[0], -1)
        n, k = b2.shape
        if inner_maxiter is None:
            inner_maxiter = 10 * n
        x = torch.empty_like(b2)
        its = (ctypes.c_int * k)()
        steps = (ctypes.c_int * k)()
        rel = (ctypes.c_double * k)()
        h = self._native()
        _lib.check(_lib.load().amg_pcg_ir(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(), x.data_ptr(), k,
                                          float(tol), float(inner_tol), int(maxiter), int(inner_maxiter),
                                          int(bool(flexible)), its, steps, rel, _lib.stream_ptr(dev)), "amg_pcg_ir")
        x = x.reshape(-1) if vec else x
        if return_info:
            one = (lambda a: list(a)[0]) if vec else list
            return x, one(its), one(steps), one(rel)
        return x


def _gmres_call(ml, b, tol, atol, restart, maxiter, cycle, cycles_per_level, precondition, return_info,
                solve_method="batched"):
    if cycle not in ("V", "W", "F"):
        raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
    methods = {"batched": 0, "incremental": 1}
    if solve_method not in methods:
        raise ValueError(f"solve_method {solve_method!r} not supported. Use 'batched' or 'incremental'.")
    dev = ml.device
    b = _prep(b, torch.float64, dev)
    vec = b.dim() == 1
    b2 = b.reshape(b.shape[0], -1)
    n, k = b2.shape
    if maxiter is None:
        maxiter = 10 * n
    x = torch.empty_like(b2)
    its = (ctypes.c_int * k)()
    rel = (ctypes.c_double * k)()
    h = ml._native()
    _lib.check(_lib.load().amg_gmres(h, cycle.encode(), int(cycles_per_level), b2.data_ptr(), x.data_ptr(), k,
                                     float(tol), float(atol), int(restart), int(maxiter), int(bool(precondition)),
                                     methods[solve_method], its, rel, _lib.stream_ptr(dev)), "amg_gmres")
    x = x.reshape(-1) if vec else x
    if return_info:
        return x, (list(its)[0] if vec else list(its)), (list(rel)[0] if vec else list(rel))
    return x


class _Level0:
    """Temporarily replace the handle's level-0 operator (the ``A=`` argument)."""

    def __init__(self, ml, h, A):
        self.ml, self.h, self.A = ml, h, A

    def __enter__(self):
        if self.A is not None:
            A = self.A
            if not isinstance(A, SparseMatrix):
                from .autograd import as_operator
                A = as_operator(self.ml, A)[0]
                self.A = A
            if A.shape != self.ml.levels[0].A.shape or A.dtype != self.ml.dtype:
                raise ValueError("A must match the finest level's shape and dtype")
            _lib.check(_lib.load().amg_set_level0(self.h, ctypes.byref(A._desc)), "amg_set_level0")
        return self

    def __exit__(self, *exc):
        if self.A is not None:
            _lib.check(_lib.load().amg_set_level0(self.h, None), "amg_set_level0")
        return False


def _requires_grad(b, A):
    if not torch.is_grad_enabled():
        return False
    if torch.is_tensor(b) and b.requires_grad:
        return True
    if torch.is_tensor(A) and A.requires_grad:
        return True
    vals = getattr(A, "values_tensor", None)
    return vals is not None and vals.requires_grad


def _nnz(A):
    if hasattr(A, "nnz"):
        return A.nnz
    if hasattr(A, "nse"):
        return A.nse
    raise TypeError(f"Cannot determine nnz for type {type(A)}")


def _host_csr(M, dtype):
    M = sp.csr_matrix(M).astype(dtype)
    M.sum_duplicates()
    M.sort_indices()
    return M


def _convert_hierarchy(pyamg_solver, dtype, coarse_solver, kw, device):
    """Pack and upload all levels (amjax/multilevel.py:557-590 and 173-181)."""
    td = as_torch_dtype(dtype) or as_torch_dtype(sp.csr_matrix(pyamg_solver.levels[0].A).dtype)
    nd = np_dtype(td)
    device = torch.device(device) if device is not None else default_device()
    host = []
    for py in pyamg_solver.levels:
        A = _host_csr(py.A, nd)
        P = _host_csr(py.P, nd) if getattr(py, "P", None) is not None else None
        R = _host_csr(py.R, nd) if getattr(py, "R", None) is not None else None
        host.append([A, P, R])
    for h in host[:-1]:
        if h[2] is None:
            h[2] = _host_csr(h[1].T.conj(), nd)
    pk = Packer(nd)
    recs = []
    for i, (A, P, R) in enumerate(host):
        dinv = relaxation.inverse_diagonal(A)
        rec = {"A": pk.add_sell(A), "dinv": (pk.add_val(dinv), A.shape[0])}
        if i < len(host) - 1:
            rec["P"] = pk.add_sell(P)
            rec["R"] = pk.add_sell(R)
        recs.append(rec)
    Ac = host[-1][0].toarray()
    nc = Ac.shape[0]
    factors = {"A_dense": (pk.add_val(Ac), (nc, nc))}
    if coarse_solver == "pinv":
        factors["A_inv"] = (pk.add_val(_pinv(Ac, **kw)), (nc, nc))
    elif coarse_solver == "lu":
        lu, piv = sla.lu_factor(Ac.astype(np.float64))
        factors["lu_factor"] = (pk.add_val(lu.astype(nd)), (nc, nc))
        piv_off = pk.add_int(piv.astype(np.int32))
    elif coarse_solver == "qr":
        Q, Rm = np.linalg.qr(Ac.astype(np.float64))
        factors["Q"] = (pk.add_val(Q.astype(nd)), (nc, nc))
        factors["R_mat"] = (pk.add_val(Rm.astype(nd)), (nc, nc))
    ibuf = torch.from_numpy(pk.int_buffer()).to(device)
    vbuf = torch.from_numpy(pk.val_buffer()).to(device)
    levels = []
    for i, rec in enumerate(recs):
        lvl = _Level()
        lvl.A = SparseMatrix(host[i][0], ibuf, vbuf, _lib.sell_desc(rec["A"], ibuf, vbuf))
        off, n = rec["dinv"]
        lvl.Dinv = vbuf[off:off + n]
        if "P" in rec:
            lvl.P = SparseMatrix(host[i][1], ibuf, vbuf, _lib.sell_desc(rec["P"], ibuf, vbuf))
            lvl.R = SparseMatrix(host[i][2], ibuf, vbuf, _lib.sell_desc(rec["R"], ibuf, vbuf))
        levels.append(lvl)
    c = levels[-1]
    for name, (off, shape) in factors.items():
        setattr(c, name, vbuf[off:off + shape[0] * shape[1]].view(*shape))
    if coarse_solver == "lu":
        c.piv = ibuf[piv_off:piv_off + nc]
    outer = None
    if td == torch.float32:
        outer = SparseMatrix.from_scipy(sp.csr_matrix(pyamg_solver.levels[0].A).astype(np.float64),
                                        torch.float64, device)
        c._A64 = sp.csr_matrix(pyamg_solver.levels[-1].A).astype(np.float64)
    return levels, outer


class multilevel_solver(MultilevelSolver):  # noqa: N801
    """Deprecated alias for MultilevelSolver (amjax/multilevel.py:730-743)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        warn("multilevel_solver is deprecated. Use MultilevelSolver.", DeprecationWarning, stacklevel=2)
