"""Relaxation methods (mirror of amjax/relaxation/relaxation.py).

``jacobi`` runs the gfx950 Jacobi-sweep kernel; the diagonal and spectral
radius helpers are setup-time host computations, as in the reference (which
calls ``approximate_spectral_radius`` outside of ``jit``, smoothing.py:89-92).
"""

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from .. import _lib
from ..sparse import SparseMatrix, amg_dtype, _prep


def _host(A):
    if isinstance(A, SparseMatrix):
        return A.host
    if torch.is_tensor(A):
        return A.detach().cpu().numpy()
    return A


def matrix_diagonal(A):
    """Diagonal of A with duplicate entries summed (relaxation.py:9-33)."""
    H = _host(A)
    if sp.issparse(H):
        coo = sp.coo_matrix(H)
        d = np.zeros(H.shape[0], dtype=coo.data.dtype)
        m = coo.row == coo.col
        np.add.at(d, coo.row[m], coo.data[m])
    else:
        d = np.diag(np.asarray(H)).copy()
    if isinstance(A, SparseMatrix) or torch.is_tensor(A):
        dev = A.device
        return torch.as_tensor(d, device=dev)
    return d


def inverse_diagonal(A):
    """Element-wise inverse of the diagonal; zero entries map to zero (relaxation.py:36-61)."""
    d = matrix_diagonal(A)
    if torch.is_tensor(d):
        one = torch.ones((), dtype=d.dtype, device=d.device)
        safe = torch.where(d != 0, d, one)
        return torch.where(d != 0, one / safe, torch.zeros((), dtype=d.dtype, device=d.device))
    one = d.dtype.type(1)
    safe = np.where(d != 0, d, one)
    return np.where(d != 0, one / safe, d.dtype.type(0))


def jacobi(A, x, b, Dinv, iterations=1, omega=1.0):
    """Damped Jacobi: ``iterations`` sweeps of
    ``temp = Dinv*(b - A x) + x;  x = (1-omega) x + omega temp``
    (relaxation.py:64-98) on the GPU.  Returns a new array; inputs are not
    modified.  ``A`` is a :class:`SparseMatrix`; x, b, Dinv are (n,) or (n, k).
    """
    if not isinstance(A, SparseMatrix):
        raise TypeError("jacobi expects an amjax_amd SparseMatrix for A")
    dt, dev = A.dtype, A.device
    x = _prep(x, dt, dev)
    b = _prep(b, dt, dev)
    Dinv = _prep(Dinv, dt, dev)
    k = 1 if x.dim() == 1 else x.shape[1]
    if iterations <= 0:
        return x.clone()
    lib = _lib.load()
    cur = x
    for _ in range(int(iterations)):
        out = torch.empty_like(x)
        _lib.check(lib.amg_sell_jacobi(ctypes.byref(A._desc), Dinv.data_ptr(), amg_dtype(dt),
                                       cur.data_ptr(), b.data_ptr(), out.data_ptr(), k, float(omega),
                                       _lib.stream_ptr(dev)),
                   "amg_sell_jacobi")
        cur = out
    return cur


def polynomial(A, x, b, coefficients, iterations=1):
    """Polynomial smoother sweep (PyAMG relaxation.polynomial): ``iterations`` times
    r = b - A x (b when x is zero), h = c_0 r, h = c_i r + A h, x += h, on the GPU
    (one fused kernel per coefficient).  Returns a new array."""
    if not isinstance(A, SparseMatrix):
        raise TypeError("polynomial expects an amjax_amd SparseMatrix for A")
    dt, dev = A.dtype, A.device
    x = _prep(x, dt, dev)
    b = _prep(b, dt, dev)
    k = 1 if x.dim() == 1 else x.shape[1]
    if iterations <= 0:
        return x.clone()
    coef = (ctypes.c_double * len(coefficients))(*[float(c) for c in coefficients])
    out = torch.empty_like(x)
    work = torch.empty((4,) + tuple(x.shape), dtype=dt, device=dev)
    zero = int(not bool(torch.any(x != 0)))
    _lib.check(_lib.load().amg_sell_poly(ctypes.byref(A._desc), amg_dtype(dt), coef, len(coefficients),
                                         int(iterations), x.data_ptr(), b.data_ptr(), out.data_ptr(),
                                         work.data_ptr(), k, zero, _lib.stream_ptr(dev)), "amg_sell_poly")
    return out


def approximate_spectral_radius(A, Dinv, n_iter=15, seed=0):
    """Power-iteration estimate of rho(D^-1 A) (relaxation.py:101-132).

    Setup-time host computation in the matrix dtype: ``v`` from
    ``np.random.default_rng(seed).standard_normal(n)``, normalised, then
    ``n_iter`` times ``w = Dinv * (A v); rho = ||w||; v = w / rho``.
    """
    H = _host(A)
    D = Dinv.detach().cpu().numpy() if torch.is_tensor(Dinv) else np.asarray(Dinv)
    dt = D.dtype
    rng = np.random.default_rng(seed)
    v = np.asarray(rng.standard_normal(H.shape[0]), dtype=dt)
    v = v / np.linalg.norm(v)
    rho = 1.0
    for _ in range(n_iter):
        w = D * (H @ v)
        rho = float(np.linalg.norm(w))
        if rho == 0.0:
            break
        v = (w / rho).astype(dt)
    return rho


def none(A, x, b):
    """Return x unchanged (no relaxation)."""
    return x
