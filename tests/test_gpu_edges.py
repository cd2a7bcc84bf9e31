"""Edge shapes: more right-hand sides than one column chunk (k = 100: two 64-lane
chunks on gridDim.y), tiny and two-level hierarchies, row counts below one slice."""

import numpy as np
import pytest

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def test_k100_solve_pcg_bicgstab_batch(cuda):
    from amjax_amd import BatchPCG, MultilevelSolver
    from oracle.solve import OracleMG, bicgstab, pcg
    h = sa_hierarchy((40, 40))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    A = om.levels[0].A
    B = np.random.default_rng(9).standard_normal((1600, 100))
    B[:, 70] = 0.0
    Bt = _t(B, cuda)
    cols = [0, 63, 64, 70, 99]
    X, its, _ = ml.solve(Bt, tol=1e-10, maxiter=200, return_info=True)
    Xr, itr = om.solve(B[:, cols], tol=1e-10, maxiter=200, return_info=True)
    assert [its[c] for c in cols] == itr
    assert rel(X.cpu().numpy()[:, cols], Xr) < 1e-12
    X, its, _ = ml.pcg(Bt, tol=1e-10, maxiter=200, return_info=True)
    Xr, itr = pcg(A, B[:, cols], om.aspreconditioner("V"), tol=1e-10, maxiter=200, return_info=True)
    assert [its[c] for c in cols] == itr
    assert rel(X.cpu().numpy()[:, cols], Xr) < 1e-12
    X, its, _ = ml.bicgstab(Bt, tol=1e-10, maxiter=200, return_info=True)
    Xr, itr = bicgstab(A, B[:, cols], om.aspreconditioner("V"), tol=1e-10, maxiter=200, return_info=True)
    assert [its[c] for c in cols] == itr
    assert rel(X.cpu().numpy()[:, cols], Xr) < 1e-10
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    (Xb,) = BatchPCG([ml32], [100]).pcg([Bt], tol=1e-8, maxiter=200)
    assert rel(Xb.cpu().numpy(), ml32.pcg(Bt, tol=1e-8, maxiter=200).cpu().numpy()) == 0.0


@pytest.mark.parametrize("grid", [(7, 5), (9, 9)])
def test_tiny_problems(cuda, grid):
    """Fewer rows than one 64-row slice, and hierarchies of one or two levels."""
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    from oracle.solve import OracleMG, pcg
    A = gallery.poisson(grid)
    h = hierarchy.smoothed_aggregation_solver(A)
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    n = A.shape[0]
    B = np.random.default_rng(n).standard_normal((n, 3))
    X, its, _ = ml.solve(_t(B, cuda), tol=1e-10, maxiter=100, return_info=True)
    Xr, itr = om.solve(B, tol=1e-10, maxiter=100, return_info=True)
    assert list(its) == list(itr) and rel(X.cpu().numpy(), Xr) < 1e-12
    X = ml.pcg(_t(B, cuda), tol=1e-12, maxiter=100).cpu().numpy()
    assert rel(X, pcg(om.levels[0].A, B, om.aspreconditioner("V"), tol=1e-12, maxiter=100)) < 1e-12


def test_many_block_widths_evict_workspaces(cuda):
    """More distinct k than the handle keeps workspaces for: the least recently
    used workspace (and its captured graphs) is dropped and rebuilt on return."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((24, 24))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    rng = np.random.default_rng(17)
    B = rng.standard_normal((576, 12))
    Xr, itr = om.solve(B, tol=1e-10, maxiter=100, return_info=True)
    for rounds in range(2):
        for k in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12):
            X, its, _ = ml.solve(_t(B[:, :k], cuda), tol=1e-10, maxiter=100, return_info=True)
            assert list(its) == list(itr[:k])
            assert rel(X.cpu().numpy(), Xr[:, :k]) < 1e-12
