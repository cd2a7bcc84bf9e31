"""Failure paths: the reference's silent semantics on bad inputs.

The reference raises nothing when a solve cannot converge: a NaN norm fails
the loop test ``normr >= tol * normb`` and the loop stops with x = 0
(amjax/multilevel.py:409-425); an infinite right-hand side keeps the test true
until its residual turns NaN or ``maxiter``; a zero diagonal entry gives D^-1 = 0 there, so Jacobi leaves
that unknown as it is (amjax/relaxation/relaxation.py:58-61, 94-96); pinv
applies the pseudo-inverse of a singular coarse matrix (multilevel.py:509-516).
Every case here ends normally (maxiter bounds the graph launches) with a
finite iteration count and no error code, and agrees with the oracle.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def test_nan_and_inf_right_hand_sides(cuda):
    """A NaN column stops at once with x = 0; an Inf column runs to maxiter; the
    other columns of the block are untouched by either."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    h = sa_hierarchy((32, 32))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(6).standard_normal((1024, 4))
    clean, its_clean, _ = ml.solve(_t(B, cuda), tol=1e-8, maxiter=50, return_info=True)
    Bbad = B.copy()
    Bbad[17, 1] = np.nan
    Bbad[3, 2] = np.inf
    X, its, _ = ml.solve(_t(Bbad, cuda), tol=1e-8, maxiter=50, return_info=True)
    X = X.cpu().numpy()
    _, its_ref = om.solve(Bbad, tol=1e-8, maxiter=50, return_info=True)
    assert list(its) == list(its_ref)
    assert its[1] == 0 and np.all(X[:, 1] == 0)
    # inf >= tol * inf holds, so the Inf column cycles until its residual is NaN
    assert 1 <= its[2] <= 50
    for c in (0, 3):
        assert its[c] == its_clean[c]
        assert np.array_equal(X[:, c], clean.cpu().numpy()[:, c])
    # PCG (jax cg semantics): a NaN or Inf column does not iterate
    Xp, itp, _ = ml.pcg(_t(Bbad, cuda), tol=1e-8, maxiter=50, return_info=True)
    _, itp_ref = pcg(om.levels[0].A, Bbad, om.aspreconditioner(), tol=1e-8, maxiter=50, return_info=True)
    assert itp[1] == itp_ref[1] == 0 and itp[2] == itp_ref[2] == 0
    assert rel(Xp.cpu().numpy()[:, 0], pcg(om.levels[0].A, B[:, 0], om.aspreconditioner(), tol=1e-8,
                                          maxiter=50)) < 1e-12


def test_zero_diagonal_row(cuda):
    """D^-1 = 0 on a row whose diagonal is zero: fixed cycles equal the oracle's."""
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    from oracle.solve import OracleMG
    A = gallery.poisson((24, 24)).tolil()
    A[100, 100] = 0.0
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    h = hierarchy.smoothed_aggregation_solver(A)
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    assert om.levels[0].Dinv[100] == 0.0
    b = np.random.default_rng(7).standard_normal(A.shape[0])
    x, it, rr = ml.solve(_t(b, cuda), tol=1e-8, maxiter=6, return_info=True)
    x_ref, it_ref = om.solve(b, tol=1e-8, maxiter=6, return_info=True)
    assert it == it_ref and np.isfinite(rr)
    assert rel(x.cpu().numpy(), x_ref) < 1e-12


def _neumann(n):
    """Graph Laplacian of the n x n grid (pure Neumann): singular, null space = constants."""
    L1 = sp.diags([-np.ones(n - 1), -np.ones(n - 1)], [-1, 1], shape=(n, n))
    I = sp.identity(n)
    off = sp.kron(L1, I) + sp.kron(I, L1)
    deg = -np.asarray(off.sum(axis=1)).ravel()
    return sp.csr_matrix(off + sp.diags(deg))


def test_singular_coarse_matrix_under_pinv(cuda):
    """Pure-Neumann Laplacian: the coarsest matrix is singular and pinv applies its
    pseudo-inverse; a mean-zero right-hand side converges as in the oracle."""
    from amjax_amd import MultilevelSolver, hierarchy
    from oracle.solve import OracleMG, pcg
    A = _neumann(44)
    h = hierarchy.smoothed_aggregation_solver(A)
    s = np.linalg.svd(h.levels[-1].A.toarray(), compute_uv=False)
    # numerically singular, well below pinv's cutoff 10 max(m, n) eps s_max
    assert s[-1] / s[0] < 0.5 * 10 * s.size * np.finfo(np.float64).eps
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    b = np.random.default_rng(8).standard_normal(A.shape[0])
    b -= b.mean()
    x, it, _ = ml.solve(_t(b, cuda), tol=1e-8, maxiter=100, return_info=True)
    x_ref, it_ref = om.solve(b, tol=1e-8, maxiter=100, return_info=True)
    assert it == it_ref and rel(x.cpu().numpy(), x_ref) < 1e-10
    xp, itp, rp = ml.pcg(_t(b, cuda), tol=1e-8, maxiter=200, return_info=True)
    xp_ref, itp_ref = pcg(om.levels[0].A, b, om.aspreconditioner(), tol=1e-8, maxiter=200, return_info=True)
    assert itp == itp_ref and rp < 1e-8 and rel(xp.cpu().numpy(), xp_ref) < 1e-10


def test_one_level_hierarchy_through_pcg_and_batch(cuda):
    """A hierarchy of one level (n <= max_coarse): the cycle is the direct coarse
    solve, so PCG converges in one iteration; in a batch beside a multilevel
    problem both equal their single solves bit for bit."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    A1 = gallery.poisson((6, 6))
    h1 = hierarchy.smoothed_aggregation_solver(A1, max_coarse=100)
    assert len(h1.levels) == 1
    h2 = sa_hierarchy((30, 30))
    mls = [MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda) for h in (h1, h2)]
    Bs = [_t(np.random.default_rng(9 + i).standard_normal((ml.levels[0].A.shape[0], k)), cuda)
          for i, (ml, k) in enumerate(zip(mls, (3, 4)))]
    x1, it1, rr1 = mls[0].pcg(Bs[0], tol=1e-8, maxiter=50, return_info=True)
    assert max(it1) <= 2 and max(rr1) < 1e-8
    Xb, itb, _ = BatchPCG(mls, [3, 4]).pcg(Bs, tol=1e-8, maxiter=50, return_info=True)
    for ml, b, xb, it_ in zip(mls, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=50, return_info=True)
        assert torch.equal(xb, xs) and list(it_) == list(its)


def test_create_refuses_a_declared_width_below_the_widest_slice(cuda):
    """The vector row kernels are specialised on a matrix's declared widest slice, so a
    handle whose descriptor understates it is refused at creation (before any kernel
    runs) with an argument error; the same matrix with its true width works."""
    import torch
    from amjax_amd import MultilevelSolver
    from amjax_amd._lib import AmgError
    h = sa_hierarchy((32, 32))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    d = ml.levels[0].A._desc
    true_w = int(d.max_width)
    d.max_width = true_w - 1
    with pytest.raises((ValueError, AmgError), match="max_width"):
        ml._native()
    d.max_width = true_w
    x = ml.solve(torch.ones(1024, dtype=torch.float64, device=cuda), tol=1e-8, maxiter=50)
    assert torch.isfinite(x).all()
