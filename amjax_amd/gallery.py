"""Model problems for the multigrid solve phase (host side, scipy CSR).

The five/seven-point Poisson operators are the ones ``pyamg.gallery.poisson``
returns (Dirichlet boundary, diagonal ``2*d``, off-diagonals ``-1``, natural
lexicographic ordering, sorted CSR); the reference builds every test and
benchmark matrix with that call (``tests/test_multilevel.py:7,54,76``,
``benchmarks/benchmark.ipynb`` cell 10).  PyAMG is not installed here, so the
operator is assembled directly as a Kronecker sum, which is the same matrix.
"""

import numpy as np
import scipy.sparse as sp


def _lap1d(n, dtype):
    e = np.ones(n, dtype=dtype)
    return sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1], shape=(n, n), format="csr", dtype=dtype)


def poisson(grid, dtype=np.float64, format="csr"):
    """Return the d-dimensional Dirichlet finite-difference Laplacian on ``grid``.

    ``grid`` is a tuple such as ``(n,)``, ``(nx, ny)`` or ``(nx, ny, nz)``.
    The last grid index varies fastest (C order), as in
    ``pyamg.gallery.stencil_grid``.
    """
    grid = tuple(int(g) for g in grid)
    if len(grid) < 1 or min(grid) < 1:
        raise ValueError(f"invalid grid {grid!r}")
    A = None
    for axis, n in enumerate(grid):
        left = int(np.prod(grid[:axis], dtype=np.int64))
        right = int(np.prod(grid[axis + 1:], dtype=np.int64))
        term = sp.kron(sp.kron(sp.identity(left, dtype=dtype, format="csr"), _lap1d(n, dtype)),
                       sp.identity(right, dtype=dtype, format="csr"), format="csr")
        A = term if A is None else (A + term)
    A = sp.csr_matrix(A, dtype=dtype)
    A.eliminate_zeros()
    A.sort_indices()
    return A.asformat(format)


def gaussian_random_field(shape, length, sigma=1.0, seed=1):
    """Stationary Gaussian random field with squared-exponential covariance.

    ``C(r) = sigma^2 exp(-r^2 / (2 length^2))`` (``r`` in cells), sampled by
    spectral synthesis on a periodic grid padded by ``4*length`` cells per axis
    so the periodic wrap does not correlate opposite edges.
    """
    rng = np.random.default_rng(seed)
    pad = [int(s + 4 * length) for s in shape]
    k = np.meshgrid(*[np.fft.fftfreq(p) for p in pad], indexing="ij")
    k2 = sum(ki ** 2 for ki in k)
    d = len(shape)
    # spectral density of the squared-exponential kernel on the unit lattice
    spec = (2 * np.pi * length ** 2) ** (d / 2) * np.exp(-2 * np.pi ** 2 * length ** 2 * k2)
    noise = rng.standard_normal(pad)
    field = np.fft.ifftn(np.fft.fftn(noise) * np.sqrt(spec)).real
    field = field[tuple(slice(0, s) for s in shape)]
    field *= sigma / max(field.std(), 1e-300)
    return field - field.mean()


def diffusion_fv(g):
    """Five-point finite-volume operator ``-div(k grad u)`` on a 2-D cell grid.

    ``g`` (nx, ny) is the log-conductance per cell.  Edge conductances are
    ``exp(g)`` averaged harmonically between the two cells of an interior
    edge; a boundary edge couples to a Dirichlet ghost value with the cell's
    own conductance times two (half-cell distance).  The result is SPD.
    """
    kc = np.exp(np.asarray(g, dtype=np.float64))
    nx, ny = kc.shape
    idx = np.arange(nx * ny).reshape(nx, ny)
    rows, cols, vals = [], [], []
    diag = np.zeros((nx, ny))
    # x-direction interior edges
    kx = 2.0 * kc[:-1, :] * kc[1:, :] / (kc[:-1, :] + kc[1:, :])
    rows += [idx[:-1, :].ravel(), idx[1:, :].ravel()]
    cols += [idx[1:, :].ravel(), idx[:-1, :].ravel()]
    vals += [-kx.ravel(), -kx.ravel()]
    diag[:-1, :] += kx
    diag[1:, :] += kx
    ky = 2.0 * kc[:, :-1] * kc[:, 1:] / (kc[:, :-1] + kc[:, 1:])
    rows += [idx[:, :-1].ravel(), idx[:, 1:].ravel()]
    cols += [idx[:, 1:].ravel(), idx[:, :-1].ravel()]
    vals += [-ky.ravel(), -ky.ravel()]
    diag[:, :-1] += ky
    diag[:, 1:] += ky
    # Dirichlet boundary faces
    diag[0, :] += 2 * kc[0, :]
    diag[-1, :] += 2 * kc[-1, :]
    diag[:, 0] += 2 * kc[:, 0]
    diag[:, -1] += 2 * kc[:, -1]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nx * ny, nx * ny))
    A.sum_duplicates()
    A.sort_indices()
    return A
