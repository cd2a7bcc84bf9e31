"""Right-preconditioned BiCGSTAB (jax.scipy.sparse.linalg.bicgstab semantics) on the
GPU against the CPU oracle, whose algorithm is pinned to scipy's bicgstab iterates
(test_oracle.py), on a nonsymmetric upwind convection-diffusion operator."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def convdiff(g, c):
    e = np.ones(g)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    U = sp.diags([-e[:-1], e], [-1, 0])
    I = sp.eye(g)
    return (sp.kron(I, T) + sp.kron(T, I) + c * sp.kron(I, U)).tocsr()


@pytest.fixture(scope="module")
def nonsym():
    from amjax_amd import hierarchy
    A = convdiff(48, 1.5)
    assert abs(A - A.T).max() > 0.5
    return A, hierarchy.smoothed_aggregation_solver(A)


@pytest.mark.parametrize("k", [1, 5])
@pytest.mark.parametrize("cycle", ["V", "W"])
def test_bicgstab_parity(cuda, nonsym, k, cycle):
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, bicgstab
    A, h = nonsym
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(k).standard_normal((A.shape[0], k))
    B[:, -1] *= 1e-3
    X, its, rr = ml.bicgstab(torch.as_tensor(B, device=cuda), tol=1e-10, maxiter=200, cycle=cycle,
                             return_info=True)
    Xr, itr = bicgstab(A, B, M=om.aspreconditioner(cycle), tol=1e-10, maxiter=200, return_info=True)
    assert its == itr
    assert rel(X.cpu().numpy(), Xr) < 1e-10
    res = np.linalg.norm(B - A @ X.cpu().numpy(), axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() < 1e-9 and np.allclose(rr, res, rtol=1e-6, atol=1e-13)


def test_bicgstab_plain_front_end_and_zero_column(cuda, nonsym):
    import torch
    import amjax_amd
    from amjax_amd import MultilevelSolver
    from oracle.solve import bicgstab
    A, h = nonsym
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    B = np.random.default_rng(3).standard_normal((A.shape[0], 3))
    B[:, 1] = 0.0
    Bt = torch.as_tensor(B, device=cuda)
    # unpreconditioned, through the jax-shaped front end
    x, info = amjax_amd.bicgstab(ml.levels[0].A, Bt[:, 0].contiguous(), tol=1e-8, maxiter=2000)
    assert info is None
    xr = bicgstab(A, B[:, 0], tol=1e-8, maxiter=2000)
    assert rel(x.cpu().numpy(), xr) < 1e-9
    X, its, _ = ml.bicgstab(Bt, tol=1e-8, maxiter=200, return_info=True)
    assert its[1] == 0 and np.all(X.cpu().numpy()[:, 1] == 0)
    M = ml.aspreconditioner("V")
    x2, _ = amjax_amd.bicgstab(ml.levels[0].A, Bt[:, 2].contiguous(), tol=1e-8, M=M)
    assert rel(x2.cpu().numpy(), X.cpu().numpy()[:, 2]) < 1e-12


def test_bicgstab_mixed_precision(cuda, nonsym):
    import torch
    from amjax_amd import MultilevelSolver
    A, h = nonsym
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    ml64 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    B = np.random.default_rng(4).standard_normal((A.shape[0], 4))
    Bt = torch.as_tensor(B, device=cuda)
    X64 = ml64.bicgstab(Bt, tol=1e-13, maxiter=300).cpu().numpy()
    X, its, rr = ml32.bicgstab(Bt, tol=1e-10, maxiter=300, return_info=True)
    X = X.cpu().numpy()
    res = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() < 1e-9
    assert (np.linalg.norm(X - X64, axis=0) / np.linalg.norm(X64, axis=0)).max() < 1e-7


@pytest.mark.parametrize("k,restart", [(1, 10), (4, 20), (64, 8)])
def test_gmres_parity(cuda, nonsym, k, restart):
    """Restarted GMRES (jax gmres 'batched' semantics, left preconditioning) against the
    oracle: the same restart counts per column and the same solution to 1e-9."""
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, gmres
    A, h = nonsym
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(40 + k).standard_normal((A.shape[0], k))
    if k > 1:
        B[:, 1] = 0.0
    X, its, rr = ml.gmres(torch.as_tensor(B, device=cuda), tol=1e-10, restart=restart, maxiter=50,
                          return_info=True)
    its, rr = np.atleast_1d(its), np.atleast_1d(rr)
    Xr, itr = gmres(A, B, M=om.aspreconditioner("V"), tol=1e-10, restart=restart, maxiter=50, return_info=True)
    assert list(its) == list(np.atleast_1d(itr))
    X = X.cpu().numpy().reshape(B.shape)
    assert rel(X, Xr) < 1e-9
    live = [c for c in range(k) if np.any(B[:, c])]
    res = np.linalg.norm(B - A @ X, axis=0)[live] / np.linalg.norm(B, axis=0)[live]
    assert res.max() < 1e-8 and np.allclose(rr[live], res, rtol=1e-6, atol=1e-14)
    if k > 1:
        assert np.all(X[:, 1] == 0.0) and its[1] == 0


def test_gmres_mixed_and_plain_front_end(cuda, nonsym):
    """The float32-cycle form reaches a float64 tolerance; amjax_amd.gmres without M
    is unpreconditioned GMRES(m) and matches the oracle's."""
    import torch
    import amjax_amd
    from amjax_amd import MultilevelSolver
    from oracle.solve import gmres
    A, h = nonsym
    b = np.random.default_rng(7).standard_normal(A.shape[0])
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    x, its, rr = ml32.gmres(torch.as_tensor(b, device=cuda), tol=1e-10, restart=20, maxiter=50, return_info=True)
    assert rr < 1e-9 and np.linalg.norm(b - A @ x.cpu().numpy()) / np.linalg.norm(b) < 1e-9
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    x2, _ = amjax_amd.gmres(ml.levels[0].A, torch.as_tensor(b, device=cuda), tol=1e-6, restart=30, maxiter=40)
    xr, kr = gmres(A, b, tol=1e-6, restart=30, maxiter=40, return_info=True)
    assert rel(x2.cpu().numpy(), xr) < 1e-8


def test_gmres_restart_grows_between_calls(cuda, nonsym):
    """A longer restart on a later call reallocates the Krylov basis: its graphs are
    captured again, so the result equals a fresh solver's bit for bit."""
    import torch
    from amjax_amd import MultilevelSolver
    A, h = nonsym
    b = torch.as_tensor(np.random.default_rng(9).standard_normal((A.shape[0], 4)), device=cuda)
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    x5, it5, _ = ml.gmres(b, tol=1e-10, restart=5, maxiter=80, return_info=True)
    x20, it20, _ = ml.gmres(b, tol=1e-10, restart=20, maxiter=80, return_info=True)
    fresh = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    y20, jt20, _ = fresh.gmres(b, tol=1e-10, restart=20, maxiter=80, return_info=True)
    assert list(it20) == list(jt20) and torch.equal(x20, y20)
    assert max(it20) < max(it5)


def test_gmres_edges(cuda):
    """jax gmres edge cases against the oracle: a restart longer than n is cut to n
    (one restart then solves the system), maxiter 0 returns x = 0 with no restarts,
    an atol above tol ||b|| sets the stopping threshold, and a W-cycle preconditioner."""
    import torch
    from amjax_amd import MultilevelSolver, hierarchy
    from oracle.solve import OracleMG, gmres
    A = convdiff(5, 1.5)
    n = A.shape[0]
    h = hierarchy.smoothed_aggregation_solver(A)
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(12).standard_normal((n, 3))
    Bt = torch.as_tensor(B, device=cuda)
    X, its, _ = ml.gmres(Bt, tol=1e-10, restart=32, maxiter=10, return_info=True)
    Xr, itr = gmres(A, B, M=om.aspreconditioner("V"), tol=1e-10, restart=32, maxiter=10, return_info=True)
    assert n < 32 and list(its) == list(itr) == [1, 1, 1]
    assert rel(X.cpu().numpy(), Xr) < 1e-9
    X0, its0, _ = ml.gmres(Bt, tol=1e-10, restart=5, maxiter=0, return_info=True)
    assert torch.count_nonzero(X0) == 0 and list(its0) == [0, 0, 0]
    A2, h2 = convdiff(48, 1.5), None
    h2 = hierarchy.smoothed_aggregation_solver(A2)
    ml2 = MultilevelSolver.from_pyamg(h2, WITHRHO, WITHRHO, device=cuda)
    om2 = OracleMG(h2, WITHRHO, WITHRHO)
    B2 = np.random.default_rng(13).standard_normal((A2.shape[0], 2))
    atol = 1e-4 * np.linalg.norm(B2[:, 0])
    for cyc, tol, at in (("V", 1e-12, atol), ("W", 1e-9, 0.0)):
        X2, it2, _ = ml2.gmres(torch.as_tensor(B2, device=cuda), tol=tol, atol=at, restart=6, maxiter=40,
                               cycle=cyc, return_info=True)
        Xr2, itr2 = gmres(A2, B2, M=om2.aspreconditioner(cyc), tol=tol, atol=at, restart=6, maxiter=40,
                          return_info=True)
        assert list(it2) == list(itr2)
        assert rel(X2.cpu().numpy(), Xr2) < 1e-8


def test_krylov_x0(cuda, nonsym):
    """A start x0 (jax's argument) for cg and bicgstab: the same solution as the oracle
    started from x0 with the threshold from b, x0 = the solution returns it unchanged,
    and a zero column of b with atol = 0 gets x = 0."""
    import torch
    import amjax_amd
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, bicgstab, pcg
    h = sa_hierarchy((40, 40))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    A = om.levels[0].A
    rng = np.random.default_rng(21)
    B = rng.standard_normal((A.shape[0], 3))
    X0 = rng.standard_normal((A.shape[0], 3))
    B[:, 2] = 0.0
    M = ml.aspreconditioner("V")
    X, _ = amjax_amd.cg(ml.levels[0].A, torch.as_tensor(B, device=cuda), x0=torch.as_tensor(X0, device=cuda),
                        tol=1e-6, M=M)
    X = X.cpu().numpy()
    for c in range(2):
        xr = pcg(A, B[:, c], om.aspreconditioner(), tol=1e-6, x0=X0[:, c])
        assert rel(X[:, c], xr) < 1e-9
        assert np.linalg.norm(B[:, c] - A @ X[:, c]) <= 1e-6 * np.linalg.norm(B[:, c])
    assert np.all(X[:, 2] == 0.0)
    xs = pcg(A, B[:, 0], om.aspreconditioner(), tol=1e-13)
    xk, _ = amjax_amd.cg(ml.levels[0].A, torch.as_tensor(B[:, 0], device=cuda), x0=torch.as_tensor(xs, device=cuda),
                         tol=1e-8, M=M)
    assert np.array_equal(xk.cpu().numpy(), xs)
    An, hn = nonsym
    mln = MultilevelSolver.from_pyamg(hn, WITHRHO, WITHRHO, device=cuda)
    omn = OracleMG(hn, WITHRHO, WITHRHO)
    bn = rng.standard_normal(An.shape[0])
    xn0 = rng.standard_normal(An.shape[0])
    xg, _ = amjax_amd.bicgstab(mln.levels[0].A, torch.as_tensor(bn, device=cuda),
                               x0=torch.as_tensor(xn0, device=cuda), tol=1e-8, M=mln.aspreconditioner("V"))
    xr = bicgstab(An, bn, M=omn.aspreconditioner("V"), tol=1e-8, x0=xn0)
    assert rel(xg.cpu().numpy(), xr) < 1e-7
    assert np.linalg.norm(bn - An @ xg.cpu().numpy()) <= 1.0001e-8 * np.linalg.norm(bn)


def test_gmres_x0(cuda, nonsym):
    """gmres with a start x0: the oracle's solution from the same x0 (ptol from b),
    x0 = the solution returns it unchanged, a zero column of b gets x = 0."""
    import torch
    import amjax_amd
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, gmres
    A, h = nonsym
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    rng = np.random.default_rng(31)
    B = rng.standard_normal((A.shape[0], 3))
    X0 = rng.standard_normal((A.shape[0], 3))
    B[:, 2] = 0.0
    M = ml.aspreconditioner("V")
    X, _ = amjax_amd.gmres(ml.levels[0].A, torch.as_tensor(B, device=cuda), x0=torch.as_tensor(X0, device=cuda),
                           tol=1e-9, restart=10, maxiter=50, M=M)
    X = X.cpu().numpy()
    Xr = gmres(A, B[:, :2], M=om.aspreconditioner("V"), tol=1e-9, restart=10, maxiter=50, x0=X0[:, :2])
    assert rel(X[:, :2], Xr) < 1e-8
    assert np.all(X[:, 2] == 0.0)
    xs = gmres(A, B[:, 0], M=om.aspreconditioner("V"), tol=1e-13, restart=20, maxiter=50)
    xk, _ = amjax_amd.gmres(ml.levels[0].A, torch.as_tensor(B[:, 0], device=cuda),
                            x0=torch.as_tensor(xs, device=cuda), tol=1e-8, restart=10, M=M)
    assert np.array_equal(xk.cpu().numpy(), xs)


@pytest.mark.parametrize("k,restart", [(1, 10), (4, 6), (64, 8)])
def test_gmres_incremental(cuda, nonsym, k, restart):
    """solve_method='incremental' against the oracle's: the same restart counts and
    the same solution; it needs no more restarts than 'batched'."""
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, gmres
    A, h = nonsym
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(50 + k).standard_normal((A.shape[0], k))
    if k > 1:
        B[:, 1] = 0.0
    Bt = torch.as_tensor(B, device=cuda)
    X, its, rr = ml.gmres(Bt, tol=1e-10, restart=restart, maxiter=60, return_info=True, solve_method="incremental")
    its = np.atleast_1d(its)
    Xr, itr = gmres(A, B, M=om.aspreconditioner("V"), tol=1e-10, restart=restart, maxiter=60, return_info=True,
                    solve_method="incremental")
    assert list(its) == list(np.atleast_1d(itr))
    assert rel(X.cpu().numpy().reshape(B.shape), Xr) < 1e-9
    _, itb, _ = ml.gmres(Bt, tol=1e-10, restart=restart, maxiter=60, return_info=True)
    assert np.all(its <= np.atleast_1d(itb))
    if k > 1:
        assert np.all(X[:, 1].cpu().numpy() == 0.0) and its[1] == 0


def test_plain_solver_cache_is_bounded(cuda):
    """cg without M builds a device copy of A per operator: only the most recently
    used four are kept (a long-running process passing new operators does not grow)."""
    import torch
    import amjax_amd
    from amjax_amd import gallery, krylov
    krylov._plain.clear()
    b = torch.ones(16 * 16, dtype=torch.float64, device=cuda)
    ops = [gallery.poisson((16, 16)) * (1.0 + i) for i in range(6)]
    for A in ops:
        x, _ = amjax_amd.cg(A, b, tol=1e-10, maxiter=500)
        assert torch.linalg.norm(b - torch.as_tensor(A @ x.cpu().numpy(), device=cuda)) < 1e-8 * 16
    assert len(krylov._plain) == krylov._PLAIN_MAX
    kept = [entry[0] for entry in krylov._plain.values()]
    assert all(a is b for a, b in zip(kept, ops[-krylov._PLAIN_MAX:]))
