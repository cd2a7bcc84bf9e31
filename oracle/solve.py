"""Numpy restatement of the reference's multigrid solve phase - TEST INFRASTRUCTURE ONLY.

Every function cites the reference file:line it follows (paths relative to the
reference repository root).  Arithmetic is done in the hierarchy dtype
(float64 or float32), in the same expression order as the reference.

What pins each function (tests/test_oracle.py):
  inverse_diagonal            known answers incl. duplicate entries and zero diagonal
  jacobi                      the reference's known-answer vectors
                              (tests/relaxation/test_relaxation.py:42-80)
  approximate_spectral_radius rho = 1 for a diagonal matrix; within 10 % of the
                              largest eigenvalue of D^-1 A (scipy eigsh)
                              (tests/relaxation/test_relaxation.py:82-101); equal to the
                              closed form ||M^15 v0|| / ||M^14 v0|| (v0 the normalised
                              seed-0 normal vector) on diagonal and non-diagonal M = D^-1 A,
                              where 14 or 16 steps or a Rayleigh quotient differ; float32
                              moves off the float64 value by rounding only
  solve (x0)                  ||b - A x|| < tol ||b - A x0|| with the count of a plain solve
                              of b - A x0; a missing or doubled + x0 fails
  coarse solvers              A x = b on the reference's cases
                              (tests/test_multilevel.py:16-47, 450-478)
  cycle                       two-level V-cycle with exact coarse solve equals the
                              closed-form error propagator S (I - P Ac^-1 R A) S;
                              symmetric V-cycle operator is symmetric; W/F reduce
                              to V on two levels; the solve converges to spsolve
  solve_loop                  stopping rule and iteration count against a direct
                              restatement of the loop on the closed-form operator
  cycle_complexity            the reference's closed forms
                              (tests/test_multilevel.py:110-153)
  pcg                         converges to spsolve; with M = identity and exact
                              arithmetic terminates in <= n steps; iterates match
                              scipy.sparse.linalg.cg on the same problem
  chebyshev_polynomial_       degree 1 gives 2/(a+b); C(0) = 1; C vanishes at the scaled
  coefficients                Chebyshev roots; max |C| on [a, b] equals 1/T_d((b+a)/(b-a))
  bicgstab                    iterates after fixed iteration counts match
                              scipy.sparse.linalg.bicgstab (same algorithm) on a
                              nonsymmetric system, with and without M; converges to spsolve
  polynomial                  one sweep's error propagator equals prod_i (I - A/r_i) over
                              the scaled roots (dense, random SPD); one coefficient gives
                              x + c (b - A x)
  pcg_ir                      in float64, one step with inner_tol = tol is pcg (pinned above);
                              in float32 the true float64 residual meets a tolerance far
                              below float32 precision in 2-4 steps
  pcg_ru                      in float64 its iterates and counts equal pcg's for delta -> 0
                              (one update) and delta = 0.5 (one every other iteration); in
                              float32 the true float64 residual b - A x, computed afresh by
                              the test, meets tol (so the residual the updates keep,
                              s - A xh, tracks it) with about
                              log(tol) / log(delta) updates; maxiter = 3 equals three pcg
                              iterations (the last correction folded)
  gmres                       the Givens least-squares solve equals numpy lstsq; a full-length
                              restart solves in one restart; restarted runs meet the
                              preconditioned tolerance and agree with a direct solve; a zero
                              right-hand side returns 0
"""

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp


# ---------------------------------------------------------------- relaxation

def matrix_diagonal(A):
    """amjax/relaxation/relaxation.py:9-33: diagonal with duplicate entries summed."""
    if sp.issparse(A):
        coo = sp.coo_matrix(A)
        d = np.zeros(A.shape[0], dtype=coo.data.dtype)
        m = coo.row == coo.col
        np.add.at(d, coo.row[m], coo.data[m])
        return d
    return np.diag(np.asarray(A)).copy()


def inverse_diagonal(A):
    """amjax/relaxation/relaxation.py:36-61: 1/d, with 0 where d == 0."""
    d = matrix_diagonal(A)
    one = d.dtype.type(1)
    safe = np.where(d != 0, d, one)
    return np.where(d != 0, one / safe, d.dtype.type(0))


def jacobi(A, x, b, Dinv, iterations=1, omega=1.0):
    """amjax/relaxation/relaxation.py:64-98, same expression order (94-96)."""
    dt = Dinv.dtype
    omega = dt.type(omega)
    one = dt.type(1)
    xk = np.asarray(x, dtype=dt)
    for _ in range(iterations):
        temp = Dinv * (b - A @ xk) + xk
        xk = (one - omega) * xk + omega * temp
    return xk


def approximate_spectral_radius(A, Dinv, n_iter=15, seed=0):
    """amjax/relaxation/relaxation.py:101-132: power iteration on D^-1 A."""
    dt = Dinv.dtype
    rng = np.random.default_rng(seed)
    v = np.asarray(rng.standard_normal(A.shape[0]), dtype=dt)
    v = v / np.linalg.norm(v)
    rho = 1.0
    for _ in range(n_iter):
        w = Dinv * (A @ v)
        rho = float(np.linalg.norm(w))
        if rho == 0.0:
            break
        v = (w / rho).astype(dt)
    return rho


def chebyshev_polynomial_coefficients(a, b, degree):
    """PyAMG's ``pyamg.relaxation.chebyshev.chebyshev_polynomial_coefficients`` (pyamg is
    absent here): the monic polynomial whose roots are the ``degree`` Chebyshev
    points of [a, b], scaled to C(0) = 1, highest power first."""
    if a >= b or a <= 0:
        raise ValueError(f"invalid interval [{a},{b}]")
    std_roots = np.cos(np.pi * (np.arange(degree) + 0.5) / degree)
    scaled_roots = 0.5 * (b - a) * (1 + std_roots) + a
    scaled_poly = np.poly(scaled_roots)
    scaled_poly /= np.polyval(scaled_poly, 0)
    return scaled_poly


def polynomial(A, x, b, coefficients, iterations=1):
    """PyAMG's ``pyamg.relaxation.relaxation.polynomial``, the sweep of its 'chebyshev'
    and 'richardson' smoothers: residual = b - A x (b when x is zero);
    h = c_0 r; h = c_i r + A h for i = 1..; x += h."""
    dt = np.asarray(b).dtype
    cs = [dt.type(c) for c in coefficients]
    xk = np.asarray(x, dtype=dt)
    for _ in range(iterations):
        r = np.asarray(b, dtype=dt) if not np.any(xk) else (b - A @ xk).astype(dt)
        h = cs[0] * r
        for c in cs[1:]:
            h = c * r + A @ h
        xk = xk + h
    return xk


def poly_coefficients(name, rho, kw):
    """PyAMG's setup_chebyshev / setup_richardson (pyamg.relaxation.smoothing):
    chebyshev: -C[:-1] with C = chebyshev_polynomial_coefficients(rho lower_bound,
    rho upper_bound, degree) (defaults 1/30, 1.1, 3); richardson: [omega / rho]
    (default omega = 1).  Explicit ``coefficients`` are taken as given."""
    if kw.get("coefficients") is not None:
        return [float(c) for c in kw["coefficients"]]
    if name == "chebyshev":
        lo, hi, deg = kw.get("lower_bound", 1.0 / 30.0), kw.get("upper_bound", 1.1), kw.get("degree", 3)
        return list(-chebyshev_polynomial_coefficients(rho * lo, rho * hi, deg)[:-1])
    return [kw.get("omega", 1.0) / rho]


# ---------------------------------------------------------------- smoothers

DEFAULT_NITER = 1                      # amjax/relaxation/smoothing.py:30
SYMMETRIC_RELAXATION = ["jacobi", None]  # amjax/relaxation/smoothing.py:33


def _unpack(v):
    """amjax/relaxation/smoothing.py:36-40."""
    if isinstance(v, tuple):
        return v[0], v[1]
    return v, {}


def _as_list(v):
    """amjax/relaxation/smoothing.py:43-49."""
    if isinstance(v, list):
        return v
    if isinstance(v, (str, tuple)) or v is None:
        return [v]
    raise ValueError(f"Unrecognized smoother specification: {v!r}")


def _expand(specs, n):
    """amjax/relaxation/smoothing.py:52-56."""
    if not specs:
        raise ValueError("Empty smoother specification.")
    return (specs + [specs[-1]] * n)[:n]


class _Level:
    pass


def _make_smoother(lvl, name, kw):
    """amjax/relaxation/smoothing.py:85-117 (setup_jacobi / setup_none)."""
    if name is None or name == "none":
        return ("none", 0, 0.0)
    if name in ("chebyshev", "richardson"):
        # rho(A): the power iteration of approximate_spectral_radius with D^-1 = I
        if kw.get("coefficients") is None and not hasattr(lvl, "rho_A"):
            lvl.rho_A = approximate_spectral_radius(lvl.A, np.ones(lvl.A.shape[0], dtype=lvl.Dinv.dtype))
        coefs = poly_coefficients(name, getattr(lvl, "rho_A", None), kw)
        return ("poly", int(kw.get("iterations", DEFAULT_NITER)), coefs)
    if name != "jacobi":
        raise ValueError(f"oracle supports jacobi/chebyshev/richardson/none, got {name!r}")
    iterations = kw.get("iterations", DEFAULT_NITER)
    omega = kw.get("omega", 1.0)
    if kw.get("withrho", False):
        if not hasattr(lvl, "rho"):
            lvl.rho = approximate_spectral_radius(lvl.A, lvl.Dinv)
        omega = omega / lvl.rho
    return ("jacobi", int(iterations), omega)


def pinv_default(A_dense, rtol=None, rcond=None):
    """``jnp.linalg.pinv``: SVD pseudo-inverse, cutoff ``rtol * s_max`` with
    default ``rtol = 10 * max(m, n) * eps(dtype)`` (jax.numpy.linalg.pinv)."""
    A_dense = np.asarray(A_dense)
    dt = A_dense.dtype
    if rtol is None:
        rtol = rcond if rcond is not None else 10 * max(A_dense.shape) * np.finfo(dt).eps
    U, s, Vt = np.linalg.svd(A_dense.astype(np.float64), full_matrices=False)
    cutoff = rtol * (s.max() if s.size else 0.0)
    sinv = np.where(s > cutoff, 1.0 / np.where(s > cutoff, s, 1.0), 0.0)
    return ((Vt.T * sinv) @ U.T).astype(dt)


class OracleMG:
    """Restates ``MultilevelSolver`` (amjax/multilevel.py:61-394) on scipy matrices.

    ``hier`` is any object with ``levels[i].A/.P/.R`` (scipy sparse).
    """

    def __init__(self, hier, presmoother=("jacobi", {"iterations": 1, "omega": 1.0}),
                 postsmoother=("jacobi", {"iterations": 1, "omega": 1.0}),
                 coarse_solver="pinv", coarse_solver_kwargs=None, dtype=None):
        # _convert_hierarchy, amjax/multilevel.py:557-590
        kw = dict(coarse_solver_kwargs or {})
        levels = []
        for py in hier.levels:
            lvl = _Level()
            A = sp.csr_matrix(py.A)
            if dtype is None:
                dtype = A.dtype
            lvl.A = A.astype(dtype)
            lvl.Dinv = inverse_diagonal(lvl.A)
            lvl.P = sp.csr_matrix(py.P).astype(dtype) if getattr(py, "P", None) is not None else None
            lvl.R = sp.csr_matrix(py.R).astype(dtype) if getattr(py, "R", None) is not None else None
            levels.append(lvl)
        for lvl in levels[:-1]:
            if lvl.R is None:
                lvl.R = sp.csr_matrix(lvl.P.T.conj())
        self.dtype = np.dtype(dtype)
        self.levels = levels
        # coarse factor, amjax/multilevel.py:173-181
        c = levels[-1]
        self.coarse_name = coarse_solver
        if coarse_solver in ("pinv", "lu", "qr"):
            c.A_dense = c.A.toarray()
            if coarse_solver == "pinv":
                c.A_inv = pinv_default(c.A_dense, **kw)
            elif coarse_solver == "lu":
                c.lu = sla.lu_factor(c.A_dense)
            else:
                c.Q, c.R_mat = np.linalg.qr(c.A_dense)
        elif coarse_solver == "jacobi":
            c.cj_iterations = kw.get("iterations", 10)   # _coarse_jacobi, :501-507
            c.cj_omega = kw.get("omega", 1.0)
        else:
            raise ValueError(coarse_solver)
        # change_smoothers, amjax/relaxation/smoothing.py:161-228
        n = len(levels) - 1
        self.symmetric_smoothing = False
        if n > 0:
            pre = _expand(_as_list(presmoother), n)
            post = _expand(_as_list(postsmoother), n)
            self.symmetric_smoothing = True
            for i in range(n):
                f1, k1 = _unpack(pre[i])
                f2, k2 = _unpack(post[i])
                levels[i].pre = _make_smoother(levels[i], f1, k1)
                levels[i].post = _make_smoother(levels[i], f2, k2)
                it1 = k1.get("iterations", DEFAULT_NITER)
                it2 = k2.get("iterations", DEFAULT_NITER)
                if it1 != it2 or f1 != f2 or f1 not in SYMMETRIC_RELAXATION:
                    self.symmetric_smoothing = False

    # -- pieces of the cycle
    def _smooth(self, lvl, spec, A, x, b):
        kind, iters, omega = spec
        if kind == "none":
            return x
        if kind == "poly":
            return polynomial(A, x, b, omega, iterations=iters)
        return jacobi(A, x, b, lvl.Dinv, iterations=iters, omega=omega)

    def coarse_solve(self, A, x, b):
        """amjax/multilevel.py:501-536."""
        c = self.levels[-1]
        if self.coarse_name == "pinv":
            return c.A_inv @ b
        if self.coarse_name == "lu":
            return sla.lu_solve(c.lu, b).astype(self.dtype)
        if self.coarse_name == "qr":
            return sla.solve_triangular(c.R_mat, c.Q.T @ b).astype(self.dtype)
        return jacobi(A, x, b, c.Dinv, iterations=c.cj_iterations, omega=c.cj_omega)

    def cycle(self, x, b, cycle="V", cycles_per_level=1, A=None):
        """amjax/multilevel.py:290-339 (_cycle / _recursive)."""
        if cycle not in ("V", "W", "F"):
            raise ValueError(f"Cycle {cycle!r} not supported. Use 'V', 'W', or 'F'.")
        L = len(self.levels)
        if L == 1:
            # the reference reads levels[0].R = None here and fails; a one-level
            # hierarchy is treated as a direct coarse solve
            return self.coarse_solve(self.levels[0].A if A is None else A, x, b)

        def rec(l, x, b, ct):
            lvl = self.levels[l]
            Al = A if (l == 0 and A is not None) else lvl.A
            x = self._smooth(lvl, lvl.pre, Al, x, b)
            cb = lvl.R @ (b - Al @ x)
            cx = np.zeros_like(cb)
            if l == L - 2:
                cx = self.coarse_solve(self.levels[-1].A, cx, cb)
            elif ct == "V":
                cx = rec(l + 1, cx, cb, "V")
            elif ct == "W":
                cx = rec(l + 1, cx, cb, "W")
                cx = rec(l + 1, cx, cb, "W")
            elif ct == "F":
                cx = rec(l + 1, cx, cb, "F")
                for _ in range(cycles_per_level):
                    cx = rec(l + 1, cx, cb, "V")
            x = x + lvl.P @ cx
            x = self._smooth(lvl, lvl.post, Al, x, b)
            return x

        return rec(0, x, b, cycle)

    def aspreconditioner(self, cycle="V", cycles_per_level=1):
        """amjax/multilevel.py:265-288."""
        def matvec(b):
            b = np.ravel(np.asarray(b, dtype=self.dtype))
            return self.cycle(np.zeros_like(b), b, cycle, cycles_per_level)
        return matvec

    def solve_loop(self, b, tol=1e-5, maxiter=100, cycle="V", cycles_per_level=1, A=None):
        """amjax/multilevel.py:407-425.  Returns (x, iterations, [normr per iteration])."""
        A = self.levels[0].A if A is None else A
        normb = np.linalg.norm(b)
        if normb == 0.0:
            normb = normb.dtype.type(1) if hasattr(normb, "dtype") else 1.0
        x = np.zeros_like(b)
        normr = np.linalg.norm(b - A @ x)
        it = 0
        hist = [float(normr)]
        while it < maxiter and normr >= tol * normb:
            x = self.cycle(x, b, cycle, cycles_per_level, A=A)
            normr = np.linalg.norm(b - A @ x)
            it += 1
            hist.append(float(normr))
        return x, it, hist

    def solve(self, b, A=None, x0=None, tol=1e-5, maxiter=100, cycle="V", cycles_per_level=1,
              return_info=False):
        """amjax/multilevel.py:341-394; a 2-D b (n, k) is k independent solves (vmap)."""
        b = np.asarray(b, dtype=self.dtype)
        if b.ndim == 2:
            outs = [self.solve(b[:, j], A, None if x0 is None else np.asarray(x0)[:, j], tol,
                               maxiter, cycle, cycles_per_level, return_info=True)
                    for j in range(b.shape[1])]
            X = np.stack([o[0] for o in outs], axis=1)
            return (X, [o[1] for o in outs]) if return_info else X
        b = np.ravel(b)
        Aop = self.levels[0].A if A is None else A
        if x0 is not None:
            x0 = np.ravel(np.asarray(x0, dtype=self.dtype))
            b = b - Aop @ x0
        x, it, hist = self.solve_loop(b, tol, maxiter, cycle, cycles_per_level, A=A)
        if x0 is not None:
            x = x + x0
        return (x, it) if return_info else x

    def operator_complexity(self):
        """amjax/multilevel.py:204-214."""
        nnz = [l.A.nnz for l in self.levels]
        return sum(nnz) / float(nnz[0])

    def cycle_complexity(self, cycle="V"):
        """amjax/multilevel.py:216-263."""
        return cycle_complexity([l.A.nnz for l in self.levels], cycle)


def cycle_complexity(nnz, cycle="V"):
    """amjax/multilevel.py:216-263 on a list of per-level nonzero counts."""
    cycle = str(cycle).upper()
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


# ---------------------------------------------------------------- Krylov

def bicgstab(A, b, M=None, tol=1e-5, atol=0.0, maxiter=None, return_info=False, x0=None):
    """Right-preconditioned BiCGSTAB as ``jax.scipy.sparse.linalg.bicgstab`` computes it
    (jax is absent; restated from its published algorithm, van der Vorst's
    preconditioned BiCGSTAB with rhat = r0): x0 = 0; while r.r > max(tol^2 b.b,
    atol^2), k < maxiter (default 10 n) and no breakdown:
      rho_ = rhat.r; beta = rho_/rho alpha/omega; p = r + beta (p - omega q);
      phat = M p; q = A phat; alpha = rho_ / rhat.q; s = r - alpha q;
      exit = s.s < atol2; shat = M s; t = A shat; omega = t.s / t.t;
      x += exit ? alpha phat : alpha phat + omega shat; r = exit ? s : s - omega t;
    breakdown when omega or alpha (k = -11) or rho_ (k = -10) is zero.  A start x0
    gives r = b - A x0 (= rhat), with the threshold still from b.  A 2-D b runs
    each column independently (vmap).  Returns x (and the iteration count)."""
    b = np.asarray(b)
    if b.ndim == 2:
        outs = [bicgstab(A, b[:, j], M, tol, atol, maxiter, True, None if x0 is None else np.asarray(x0)[:, j])
                for j in range(b.shape[1])]
        X = np.stack([o[0] for o in outs], axis=1)
        return (X, [o[1] for o in outs]) if return_info else X
    n = b.shape[0]
    if maxiter is None:
        maxiter = 10 * n
    Mf = (lambda v: v) if M is None else M
    atol2 = max(tol ** 2 * float(b @ b), atol ** 2)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=b.dtype)
    r = b.copy() if x0 is None else b - A @ x
    rhat = r.copy()
    alpha = rho = omega = 1.0
    p = r.copy()
    q = r.copy()
    k = 0
    while float(r @ r) > atol2 and k < maxiter and k >= 0:
        rho_ = float(rhat @ r)
        beta = rho_ / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = Mf(p)
        q = A @ phat
        alpha = rho_ / float(rhat @ q)
        s_ = r - alpha * q
        exit_early = float(s_ @ s_) < atol2
        shat = Mf(s_)
        t = A @ shat
        omega = float(t @ s_) / float(t @ t)
        if exit_early:
            x = x + alpha * phat
            r = s_
        else:
            x = x + (alpha * phat + omega * shat)
            r = s_ - omega * t
        k = -11 if (omega == 0.0 or alpha == 0.0) else k + 1
        if rho_ == 0.0:
            k = -10
        rho = rho_
    return (x, k) if return_info else x


def pcg(A, b, M=None, tol=1e-5, atol=0.0, maxiter=None, x0=None, return_info=False):
    """Preconditioned CG as ``jax.scipy.sparse.linalg.cg`` computes it.

    Textbook Hestenes-Stiefel PCG; iteration stops when
    ``r.r <= max(tol^2 b.b, atol^2)`` on the recurrence residual (``gamma``
    is used instead of ``r.r`` when M is the identity), or after maxiter
    iterations (default 10 n).  Call sites: benchmarks/benchmark.ipynb
    cell 7, tests/test_multilevel.py:69,274,278.  A 2-D b runs each column
    independently (vmap).
    """
    b = np.asarray(b)
    if b.ndim == 2:
        outs = [pcg(A, b[:, j], M, tol, atol, maxiter, None if x0 is None else x0[:, j], True)
                for j in range(b.shape[1])]
        X = np.stack([o[0] for o in outs], axis=1)
        return (X, [o[1] for o in outs]) if return_info else X
    n = b.shape[0]
    if maxiter is None:
        maxiter = 10 * n
    ident = M is None
    Mf = (lambda r: r) if ident else M
    bs = float(b @ b)
    atol2 = max(tol ** 2 * bs, atol ** 2)
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, dtype=b.dtype).copy()
    r = b - A @ x
    z = Mf(r)
    p = z
    gamma = r @ z
    k = 0
    while True:
        rs = gamma if ident else r @ r
        if not (rs > atol2 and k < maxiter):
            break
        Ap = A @ p
        alpha = gamma / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = Mf(r)
        gamma_ = r @ z
        beta = gamma_ / gamma
        p = z + beta * p
        gamma = gamma_
        k += 1
    return (x, k) if return_info else x


def _givens_lstsq(H, m, beta, with_residual=False):
    """min ||beta e_1 - H y|| for an (m + 1) x m upper Hessenberg H by Givens
    rotations, truncated at the first zero subdiagonal (a breakdown).  With
    with_residual also the least-squares residual |g_mm| (mm the columns used)."""
    R = np.array(H, dtype=np.float64, copy=True)
    g = np.zeros(m + 1)
    g[0] = beta
    cs, sn = np.zeros(m), np.zeros(m)
    mm = m
    for j in range(m):
        for i in range(j):
            a, b_ = R[i, j], R[i + 1, j]
            R[i, j] = cs[i] * a + sn[i] * b_
            R[i + 1, j] = -sn[i] * a + cs[i] * b_
        a, b_ = R[j, j], R[j + 1, j]
        r = float(np.hypot(a, b_))
        cs[j] = a / r if r > 0 else 1.0
        sn[j] = b_ / r if r > 0 else 0.0
        R[j, j] = r
        R[j + 1, j] = 0.0
        g[j + 1] = -sn[j] * g[j]
        g[j] = cs[j] * g[j]
        if b_ == 0.0:
            mm = j + 1
            break
    y = np.zeros(m)
    for i in range(mm - 1, -1, -1):
        acc = g[i] - R[i, i + 1:mm] @ y[i + 1:mm]
        y[i] = acc / R[i, i] if R[i, i] != 0 else 0.0
    return (y, abs(float(g[mm]))) if with_residual else y


def gmres(A, b, M=None, tol=1e-5, atol=0.0, restart=20, maxiter=None, return_info=False, x0=None,
          solve_method="batched"):
    """Restarted GMRES as ``jax.scipy.sparse.linalg.gmres`` computes it with its default
    solve_method='batched' (jax is absent; restated from its published algorithm):
    left preconditioning, x0 = 0, ptol = ||M b|| min(1, max(tol ||b||, atol) / ||b||);
    while ||M (b - A x)|| > ptol and fewer than maxiter (default 10 n) restarts:
    V_0 = r / ||r||; `restart` Arnoldi steps w = M A V_j with classical Gram-Schmidt
    applied twice, V_{j+1} = w / ||w||; y = argmin ||beta e_1 - H y|| (Givens);
    x += V y.  The README roadmap's nonsymmetric Krylov caller (README.md:140-149).
    A start x0 replaces x = 0 (ptol still from b).  solve_method='incremental'
    (jax's other method) stops a restart's Arnoldi loop after the first step j
    whose least-squares residual |g_{j+1}| is <= ptol and updates x with the
    j + 1 basis vectors built so far.  A 2-D b runs each column independently
    (vmap).  Returns x (and the restarts)."""
    b = np.asarray(b, dtype=np.float64)
    if b.ndim == 2:
        outs = [gmres(A, b[:, j], M, tol, atol, restart, maxiter, True, None if x0 is None else np.asarray(x0)[:, j],
                      solve_method) for j in range(b.shape[1])]
        X = np.stack([o[0] for o in outs], axis=1)
        return (X, [o[1] for o in outs]) if return_info else X
    n = b.shape[0]
    if maxiter is None:
        maxiter = 10 * n
    m = min(restart, n)
    Mf = (lambda v: v) if M is None else (lambda v: np.asarray(M(v), dtype=np.float64))
    bn = float(np.linalg.norm(b))
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    if bn == 0.0 and x0 is None:
        return (x, 0) if return_info else x
    ptol = float(np.linalg.norm(Mf(b))) * min(1.0, max(tol * bn, atol) / bn) if bn > 0.0 else 0.0
    r = Mf(b - A @ x)
    beta = float(np.linalg.norm(r))
    k = 0
    while beta > ptol and k < maxiter:
        V = np.zeros((n, m + 1))
        H = np.zeros((m + 1, m))
        V[:, 0] = r / beta if beta > 0 else 0.0
        for j in range(m):
            w = Mf(A @ V[:, j])
            h1 = V[:, :j + 1].T @ w
            w = w - V[:, :j + 1] @ h1
            h2 = V[:, :j + 1].T @ w
            w = w - V[:, :j + 1] @ h2
            hn = float(np.linalg.norm(w))
            H[:j + 1, j] = h1 + h2
            H[j + 1, j] = hn
            V[:, j + 1] = w / hn if hn > 0 else 0.0
            if solve_method == "incremental" and j + 1 < m:
                _, est = _givens_lstsq(H[:j + 2, :j + 1], j + 1, beta, with_residual=True)
                if est <= ptol:
                    H[j + 2:, :] = 0.0
                    H[:, j + 1:] = 0.0
                    break
        y = _givens_lstsq(H, m, beta)
        x = x + V[:, :m] @ y
        r = Mf(b - A @ x)
        beta = float(np.linalg.norm(r))
        k += 1
    return (x, k) if return_info else x


def pcg_ir(A, b, M, tol=1e-8, inner_tol=1e-4, maxiter=20, inner_maxiter=None, dtype=np.float32, flexible=False,
           return_info=False):
    """Float64 iterative refinement around a preconditioned CG in ``dtype``.

    No reference counterpart: a float32 cycle inside float64 iterative
    refinement, with a CG in the cycle precision between them.  Per
    column: x = 0, r = b; while ||r|| >= tol ||b|| and fewer than maxiter steps:
    e = CG(A_dtype, (dtype) r) with x, r, p, Ap held in ``dtype``, dot products
    in float64, the stopping rule of ``pcg`` at ``inner_tol`` (or once the inner
    residual norm is below tol ||b|| / 2, so a last step does not over-solve)
    and the Polak-Ribiere beta when ``flexible``; x += e; r = b - A x in float64.
    Returns x (and per-column inner iterations summed over the steps, steps).
    With dtype float64, maxiter 1, flexible False and inner_tol >= tol / 2 (the
    floor does not bind) the single step is ``pcg``.
    """
    b = np.asarray(b, dtype=np.float64)
    if b.ndim == 2:
        outs = [pcg_ir(A, b[:, j], M, tol, inner_tol, maxiter, inner_maxiter, dtype, flexible, True)
                for j in range(b.shape[1])]
        X = np.stack([o[0] for o in outs], axis=1)
        return (X, [o[1] for o in outs], [o[2] for o in outs]) if return_info else X
    n = b.shape[0]
    if inner_maxiter is None:
        inner_maxiter = 10 * n
    A = sp.csr_matrix(A)
    At = A.astype(dtype)
    dd = lambda u, v: float(np.dot(u.astype(np.float64), v.astype(np.float64)))
    normb = float(np.linalg.norm(b)) or 1.0
    x = np.zeros(n)
    r = b.copy()
    total = steps = 0
    while float(np.linalg.norm(r)) >= tol * normb and steps < maxiter:
        rt = r.astype(dtype)
        atol2 = max(inner_tol ** 2 * dd(rt, rt), (0.5 * tol * normb) ** 2)
        e = np.zeros(n, dtype=dtype)
        z = np.asarray(M(rt), dtype=dtype)
        p = z.copy()
        gamma = dd(rt, z)
        it = 0
        while dd(rt, rt) > atol2 and it < inner_maxiter:
            Ap = At @ p
            alpha = gamma / dd(p, Ap)
            e = (e.astype(np.float64) + alpha * p.astype(np.float64)).astype(dtype)
            rt = (rt.astype(np.float64) - alpha * Ap.astype(np.float64)).astype(dtype)
            z = np.asarray(M(rt), dtype=dtype)
            g_new = dd(rt, z)
            beta = -alpha * dd(z, Ap) / gamma if flexible else g_new / gamma
            p = (z.astype(np.float64) + beta * p.astype(np.float64)).astype(dtype)
            gamma = g_new
            it += 1
        total += it
        x = x + e.astype(np.float64)
        r = b - A @ x
        steps += 1
    return (x, total, steps) if return_info else x


def pcg_ru(A, b, M, tol=1e-8, atol=0.0, delta=1e-2, maxiter=None, dtype=np.float32, flexible=False,
           return_info=False):
    """Preconditioned CG with reliable updates, its vectors in ``dtype``.

    No reference counterpart: the float32 cycle inside a float64 outer
    iteration, without restarts (Strzodka & Goeddeke 2006; Clark et al. 2010,
    "Solving lattice QCD systems of equations using mixed precision solvers on
    GPUs").  Per column: x = 0 (float64), xh = 0 and rh = (dtype) b; z = M rh,
    p = z, gamma = rh.z, and the reference norm s_ref = rh.rh.  Each iteration:
    Ap = A_dtype p, alpha = gamma / p.Ap, xh += alpha p, rh -= alpha Ap (each
    rounded to dtype from float64), it += 1; an update is due when rh.rh <=
    delta^2 s_ref, rh.rh <= max(tol^2 b.b, atol^2) or it >= maxiter: then
    r = s - A xh in float64, where s = b - A x is kept in float64 (s = b at the
    start, s = r after each update; r equals b - A (x + xh) in exact
    arithmetic), x += xh, xh = 0, rh = (dtype) r, s_ref = r.r, and the column
    stops when r.r <= max(tol^2 b.b, atol^2) or it >= maxiter;
    otherwise z = M rh, beta (textbook, or Polak-Ribiere when ``flexible``),
    p = z + beta p.  Dot products are float64.  Returns x (and per-column
    iterations and updates).  With dtype float64 the updates replace rh by a
    value equal to it in exact arithmetic, so the iterates are those of ``pcg``.
    """
    b = np.asarray(b, dtype=np.float64)
    if b.ndim == 2:
        outs = [pcg_ru(A, b[:, j], M, tol, atol, delta, maxiter, dtype, flexible, True) for j in range(b.shape[1])]
        X = np.stack([o[0] for o in outs], axis=1)
        return (X, [o[1] for o in outs], [o[2] for o in outs]) if return_info else X
    n = b.shape[0]
    if maxiter is None:
        maxiter = 10 * n
    A = sp.csr_matrix(A)
    At = A.astype(dtype)
    f64 = lambda u: u.astype(np.float64)
    dd = lambda u, v: float(np.dot(f64(u), f64(v)))
    atol2 = max(tol ** 2 * float(b @ b), atol ** 2)
    x = np.zeros(n)
    xh = np.zeros(n, dtype=dtype)
    rh = b.astype(dtype)
    s64 = b.copy()
    z = np.asarray(M(rh), dtype=dtype)
    p = z.copy()
    gamma = dd(rh, z)
    s_ref = dd(rh, rh)
    active = s_ref > atol2 and maxiter > 0
    it = ups = 0
    while active:
        Ap = (At @ p).astype(dtype)
        alpha = gamma / dd(p, Ap)
        xh = (f64(xh) + alpha * f64(p)).astype(dtype)
        rh = (f64(rh) - alpha * f64(Ap)).astype(dtype)
        it += 1
        s = dd(rh, rh)
        if s <= delta ** 2 * s_ref or s <= atol2 or it >= maxiter:
            r = s64 - A @ f64(xh)
            s64 = r
            x = x + f64(xh)
            xh = np.zeros(n, dtype=dtype)
            rh = r.astype(dtype)
            s_ref = float(r @ r)
            ups += 1
            active = s_ref > atol2 and it < maxiter
        if not active:
            break
        z = np.asarray(M(rh), dtype=dtype)
        g_new = dd(rh, z)
        beta = -alpha * dd(z, Ap) / gamma if flexible else g_new / gamma
        p = (f64(z) + beta * f64(p)).astype(dtype)
        gamma = g_new
    return (x, it, ups) if return_info else x


def relres(A, b, x):
    """||b - A x|| / ||b|| per column."""
    b = np.asarray(b)
    r = b - A @ x
    if b.ndim == 2:
        return np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)
    return np.linalg.norm(r) / np.linalg.norm(b)
