"""In-repo multigrid hierarchy setup (the step before the solve phase).

The reference never builds a hierarchy itself: ``MultilevelSolver.from_pyamg``
(amjax/multilevel.py:125-186, 557-590) reads ``pyamg_solver.levels[i].A``,
``.P`` and ``.R`` from a hierarchy built by PyAMG's
``smoothed_aggregation_solver`` / ``ruge_stuben_solver``.  PyAMG is not
installed here, so this module provides builders that return an object of the
same shape (``.levels[i].A/.P/.R`` as scipy CSR) for ``from_pyamg`` to accept.

Smoothed aggregation follows the published PyAMG defaults: symmetric strength
with theta=0, standard aggregation, near-null-space candidate ``B = ones``
improved on the finest level by four symmetric Gauss-Seidel sweeps on
``A B = 0`` (coarser levels take the fitted candidates as they are), tentative
prolongator from a per-aggregate normalisation of ``B``, one Jacobi
prolongation-smoothing step with ``omega = (4/3) / rho(D^-1 A)``, ``R = P^T``,
Galerkin ``A_c = R A P``; ``max_levels=10``, ``max_coarse=10``.
Classical (Ruge-Stueben) setup: classical strength (theta=0.25), RS C/F
splitting, classical interpolation, ``R = P^T``, Galerkin product.

The hierarchies are deterministic; they are not bit-identical to PyAMG's.
"""

import ctypes
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_setup_lib = None


def _lib():
    global _setup_lib
    if _setup_lib is None:
        path = os.path.join(_HERE, "libamjax_setup.so")
        if not os.path.exists(path):
            from . import build
            build.build_setup_lib()
        lib = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.amg_setup_standard_aggregation.argtypes = [ctypes.c_int, ip, ip, ip, ip]
        lib.amg_setup_standard_aggregation.restype = ctypes.c_int
        lib.amg_setup_gauss_seidel.argtypes = [ctypes.c_int, ip, ip, dp, dp, dp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.amg_setup_gauss_seidel.restype = None
        lib.amg_setup_rs_cf_splitting.argtypes = [ctypes.c_int, ip, ip, ip, ip, ctypes.c_int, ip]
        lib.amg_setup_rs_cf_splitting.restype = None
        _setup_lib = lib
    return _setup_lib


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Level:
    """One level of a host hierarchy: ``A`` and, except on the coarsest, ``P`` and ``R``."""

    def __init__(self, A):
        self.A = A


class Hierarchy:
    """Minimal PyAMG-shaped hierarchy: ``levels[i].A / .P / .R`` (scipy CSR)."""

    def __init__(self, levels, kind):
        self.levels = levels
        self.kind = kind

    def __repr__(self):
        sizes = ", ".join(str(l.A.shape[0]) for l in self.levels)
        return f"Hierarchy({self.kind}: {sizes})"

    def save(self, path):
        """Write the hierarchy to a versioned ``.npz`` (format ``amjax-hier-1``)."""
        arrs = {"format": np.array("amjax-hier-1"), "kind": np.array(self.kind),
                "nlev": np.array(len(self.levels))}
        for i, lvl in enumerate(self.levels):
            for name in ("A", "P", "R"):
                M = getattr(lvl, name, None)
                if M is None:
                    continue
                M = sp.csr_matrix(M)
                arrs[f"{name}{i}_indptr"] = M.indptr.astype(np.int64)
                arrs[f"{name}{i}_indices"] = M.indices.astype(np.int64)
                arrs[f"{name}{i}_data"] = M.data
                arrs[f"{name}{i}_shape"] = np.array(M.shape, dtype=np.int64)
        np.savez_compressed(path, **arrs)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        if str(z["format"]) != "amjax-hier-1":
            raise ValueError(f"unknown hierarchy format {z['format']!r}")
        levels = []
        for i in range(int(z["nlev"])):
            mats = {}
            for name in ("A", "P", "R"):
                key = f"{name}{i}_indptr"
                if key in z:
                    mats[name] = sp.csr_matrix(
                        (z[f"{name}{i}_data"], z[f"{name}{i}_indices"].astype(np.int32),
                         z[f"{name}{i}_indptr"].astype(np.int32)),
                        shape=tuple(int(s) for s in z[f"{name}{i}_shape"]))
            lvl = Level(mats["A"])
            if "P" in mats:
                lvl.P = mats["P"]
            if "R" in mats:
                lvl.R = mats["R"]
            levels.append(lvl)
        return cls(levels, str(z["kind"]))


def _csr(A):
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def symmetric_strength(A, theta=0.0):
    """Pattern of strong connections: ``|a_ij| >= theta * sqrt(|a_ii a_jj|)`` (diagonal kept)."""
    A = _csr(A)
    if theta == 0.0:
        S = A.copy()
        S.data = np.abs(S.data)
        return S
    coo = A.tocoo()
    d = np.abs(A.diagonal())
    keep = (coo.row == coo.col) | (np.abs(coo.data) >= theta * np.sqrt(d[coo.row] * d[coo.col]))
    S = sp.csr_matrix((np.abs(coo.data[keep]), (coo.row[keep], coo.col[keep])), shape=A.shape)
    return _csr(S)


def classical_strength(A, theta=0.25):
    """Classical strength: ``|a_ij| >= theta * max_{k != i} |a_ik|`` (off-diagonal only)."""
    A = _csr(A)
    coo = A.tocoo()
    off = coo.row != coo.col
    mag = np.abs(coo.data)
    rowmax = np.zeros(A.shape[0])
    np.maximum.at(rowmax, coo.row[off], mag[off])
    keep = off & (mag >= theta * rowmax[coo.row]) & (mag > 0)
    S = sp.csr_matrix((mag[keep], (coo.row[keep], coo.col[keep])), shape=A.shape)
    return _csr(S)


def standard_aggregation(S):
    """Greedy three-pass aggregation; returns (AggOp (n x nagg CSR of ones), roots)."""
    S = _csr(S)
    n = S.shape[0]
    agg = np.empty(n, dtype=np.int32)
    root = np.empty(max(n, 1), dtype=np.int32)
    nagg = _lib().amg_setup_standard_aggregation(n, _iptr(S.indptr), _iptr(S.indices),
                                                 _iptr(agg), _iptr(root))
    rows = np.nonzero(agg >= 0)[0]
    AggOp = sp.csr_matrix((np.ones(rows.size), (rows, agg[rows])), shape=(n, nagg))
    return _csr(AggOp), root[:nagg].copy()


def gauss_seidel(A, x, b, iterations=1, sweep="symmetric"):
    """In-place Gauss-Seidel on ``A x = b`` for x of shape (n,) or (n, m)."""
    A = _csr(A)
    n = A.shape[0]
    x2 = x.reshape(n, -1)
    b2 = np.ascontiguousarray(np.broadcast_to(np.asarray(b, dtype=np.float64).reshape(n, -1), x2.shape))
    xc = np.ascontiguousarray(x2, dtype=np.float64)
    data = np.ascontiguousarray(A.data, dtype=np.float64)
    lib = _lib()
    m = xc.shape[1]
    for _ in range(iterations):
        if sweep in ("forward", "symmetric"):
            lib.amg_setup_gauss_seidel(n, _iptr(A.indptr), _iptr(A.indices), _dptr(data),
                                       _dptr(xc), _dptr(b2), m, 0, n, 1)
        if sweep in ("backward", "symmetric"):
            lib.amg_setup_gauss_seidel(n, _iptr(A.indptr), _iptr(A.indices), _dptr(data),
                                       _dptr(xc), _dptr(b2), m, n - 1, -1, -1)
    x2[...] = xc
    return x


def fit_candidates(AggOp, B):
    """Tentative prolongator T and coarse candidates Bc with ``T Bc = B`` per aggregate.

    For each aggregate the rows of B it owns are orthonormalised (thin QR);
    T holds Q and Bc holds R.  Single candidate: ``T[i,a] = B[i]/||B_a||``.
    """
    AggOp = _csr(AggOp)
    n, nagg = AggOp.shape
    B = np.asarray(B, dtype=np.float64).reshape(n, -1)
    m = B.shape[1]
    agg = np.full(n, -1, dtype=np.int64)
    coo = AggOp.tocoo()
    agg[coo.row] = coo.col
    if m == 1:
        norms2 = np.zeros(nagg)
        owned = agg >= 0
        np.add.at(norms2, agg[owned], B[owned, 0] ** 2)
        norms = np.sqrt(norms2)
        safe = np.where(norms > 0, norms, 1.0)
        rows = np.nonzero(owned)[0]
        vals = np.where(norms[agg[rows]] > 0, B[rows, 0] / safe[agg[rows]], 0.0)
        T = sp.csr_matrix((vals, (rows, agg[rows])), shape=(n, nagg))
        return _csr(T), norms.reshape(nagg, 1)
    rows_all, cols_all, vals_all = [], [], []
    Bc = np.zeros((nagg * m, m))
    order = np.argsort(agg, kind="stable")
    counts = np.bincount(agg[agg >= 0], minlength=nagg)
    start = np.searchsorted(agg[order], 0)
    pos = start
    for a in range(nagg):
        idx = order[pos:pos + counts[a]]
        pos += counts[a]
        Q, Rm = np.linalg.qr(B[idx, :])
        for c in range(m):
            rows_all.append(idx)
            cols_all.append(np.full(idx.size, a * m + c))
            vals_all.append(Q[:, c] if c < Q.shape[1] else np.zeros(idx.size))
        Bc[a * m:a * m + Rm.shape[0], :] = Rm
    T = sp.csr_matrix((np.concatenate(vals_all), (np.concatenate(rows_all), np.concatenate(cols_all))),
                      shape=(n, nagg * m))
    return _csr(T), Bc


def lanczos_spectral_radius(A, Dinv=None, steps=15, seed=0):
    """Largest Ritz value of ``D^-1 A`` from ``steps`` Lanczos steps (symmetric A).

    Works on the similar symmetric operator ``D^-1/2 A D^-1/2`` from a
    ``default_rng(seed)`` start vector.
    """
    A = _csr(A)
    n = A.shape[0]
    if Dinv is None:
        d = A.diagonal()
        Dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    s = np.sqrt(np.abs(Dinv))
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    v /= np.linalg.norm(v)
    alphas, betas = [], []
    v_prev = np.zeros(n)
    beta = 0.0
    for _ in range(min(steps, n)):
        w = s * (A @ (s * v))
        alpha = float(v @ w)
        w = w - alpha * v - beta * v_prev
        alphas.append(alpha)
        beta = float(np.linalg.norm(w))
        if beta < 1e-12 * max(1.0, abs(alpha)):
            break
        betas.append(beta)
        v_prev, v = v, w / beta
    k = len(alphas)
    Tm = np.diag(alphas) + np.diag(betas[:k - 1], 1) + np.diag(betas[:k - 1], -1)
    return float(np.max(np.abs(np.linalg.eigvalsh(Tm))))


def jacobi_prolongation_smoother(A, T, omega=4.0 / 3.0, degree=1):
    """``P = (I - (omega/rho) D^-1 A)^degree T``."""
    A = _csr(A)
    d = A.diagonal()
    Dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0)
    rho = lanczos_spectral_radius(A, Dinv)
    DinvA = sp.diags(Dinv) @ A
    w = omega / rho
    P = T
    for _ in range(degree):
        P = P - w * (DinvA @ P)
    return _csr(P)


def _per_level(spec, level):
    """Entry ``level`` of a per-level list whose last entry repeats (PyAMG's levelize rule)."""
    if not isinstance(spec, (list, tuple)):
        return spec
    if not spec:
        return None
    return spec[min(level, len(spec) - 1)]


def smoothed_aggregation_solver(A, B=None, strength_theta=0.0, smooth_omega=4.0 / 3.0,
                                improve_candidates=(4, None), max_levels=10, max_coarse=10, **unused):
    """Build a smoothed-aggregation hierarchy (see module docstring for the recipe).

    ``improve_candidates`` gives, per level, the number of symmetric Gauss-Seidel
    sweeps on ``A B = 0`` applied to the candidates before they are fitted
    (None or 0: none); the last entry repeats.  The default ``(4, None)`` is
    PyAMG's ``(('block_gauss_seidel', {'sweep': 'symmetric', 'iterations': 4}),
    None)``: the finest level only.  A plain int applies to every level.
    """
    A = _csr(A).astype(np.float64)
    n = A.shape[0]
    B = np.ones((n, 1)) if B is None else np.asarray(B, dtype=np.float64).reshape(n, -1).copy()
    levels = [Level(A)]
    while len(levels) < max_levels and levels[-1].A.shape[0] > max_coarse:
        Al = levels[-1].A
        C = symmetric_strength(Al, strength_theta)
        AggOp, _ = standard_aggregation(C)
        if AggOp.shape[1] == 0 or AggOp.shape[1] >= Al.shape[0]:
            break
        sweeps = _per_level(improve_candidates, len(levels) - 1)
        if sweeps:
            gauss_seidel(Al, B, np.zeros_like(B), iterations=int(sweeps), sweep="symmetric")
        T, Bc = fit_candidates(AggOp, B)
        P = jacobi_prolongation_smoother(Al, T, omega=smooth_omega)
        R = _csr(P.T)
        Ac = _csr(R @ Al @ P)
        levels[-1].P = P
        levels[-1].R = R
        levels.append(Level(Ac))
        B = Bc
    return Hierarchy(levels, "smoothed_aggregation")


def _classical_interpolation(A, S, splitting):
    """Classical (Ruge-Stueben) interpolation.

    For an F point i with strong C neighbours C_i, strong F neighbours F_i and
    weak neighbours W_i:
      ``P_ij = -(a_ij + sum_{k in F_i} a_ik a_kj / sum_{m in C_i} a_km)
               / (a_ii + sum_{n in W_i} a_in)``  for j in C_i,
    where the k-sums use only entries ``a_kj`` of sign opposite to ``a_kk``
    (an F neighbour with no such coupling into C_i is lumped into the
    diagonal).  C points inject.
    """
    A = _csr(A)
    S = _csr(S)
    n = A.shape[0]
    cidx = np.cumsum(splitting) - 1
    Ap, Aj, Ax = A.indptr, A.indices, A.data
    Sp, Sj = S.indptr, S.indices
    rows, cols, vals = [], [], []
    diag = A.diagonal()
    for i in range(n):
        if splitting[i]:
            rows.append(i); cols.append(cidx[i]); vals.append(1.0)
            continue
        strong = set(Sj[Sp[i]:Sp[i + 1]].tolist())
        strong.discard(i)
        Ci = [j for j in strong if splitting[j]]
        if not Ci:
            continue
        Ci_set = set(Ci)
        num = {j: 0.0 for j in Ci}
        denom = diag[i]
        for jj in range(Ap[i], Ap[i + 1]):
            j = Aj[jj]
            a = Ax[jj]
            if j == i:
                continue
            if j in Ci_set:
                num[j] += a
            elif j in strong:
                # strong F neighbour: distribute a_ij over C_i through row j
                akk = diag[j]
                row_c = {}
                tot = 0.0
                for kk in range(Ap[j], Ap[j + 1]):
                    m = Aj[kk]
                    if m in Ci_set and Ax[kk] * akk < 0:
                        row_c[m] = row_c.get(m, 0.0) + Ax[kk]
                        tot += Ax[kk]
                if tot == 0.0:
                    denom += a
                else:
                    for m, v in row_c.items():
                        num[m] += a * v / tot
            else:
                denom += a
        for j in Ci:
            rows.append(i); cols.append(cidx[j]); vals.append(-num[j] / denom)
    nc = int(splitting.sum())
    P = sp.csr_matrix((vals, (rows, cols)), shape=(n, nc))
    return _csr(P)


def rs_cf_splitting(S, second_pass=False):
    S = _csr(S)
    T = _csr(S.T)
    n = S.shape[0]
    split = np.empty(n, dtype=np.int32)
    _lib().amg_setup_rs_cf_splitting(n, _iptr(S.indptr), _iptr(S.indices), _iptr(T.indptr),
                                     _iptr(T.indices), int(bool(second_pass)), _iptr(split))
    return split


def ruge_stuben_solver(A, strength_theta=0.25, second_pass=False, max_levels=10, max_coarse=10,
                       **unused):
    """Build a classical (Ruge-Stueben) hierarchy (see module docstring)."""
    A = _csr(A).astype(np.float64)
    levels = [Level(A)]
    while len(levels) < max_levels and levels[-1].A.shape[0] > max_coarse:
        Al = levels[-1].A
        S = classical_strength(Al, strength_theta)
        split = rs_cf_splitting(S, second_pass)
        nc = int(split.sum())
        if nc == 0 or nc >= Al.shape[0]:
            break
        P = _classical_interpolation(Al, S, split)
        R = _csr(P.T)
        levels[-1].P = P
        levels[-1].R = R
        levels.append(Level(_csr(R @ Al @ P)))
    return Hierarchy(levels, "ruge_stuben")
