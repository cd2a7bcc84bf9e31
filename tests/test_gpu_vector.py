"""Vector mode of the row kernels (16 bytes of a row per lane) against the
one-column-per-lane kernels: the cycle is bit-identical, solves agree to the
rounding of the reordered dot products, and columns that stop at different
iterations (masked Jacobi stores inside one lane's columns) stay correct."""

import numpy as np
import pytest

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _pair(grid, dtype, device):
    from amjax_amd import MultilevelSolver
    h = sa_hierarchy(grid)
    vec = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=dtype, device=device)
    lane = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=dtype, device=device)
    lane.vector = False
    return h, vec, lane


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("grid,k", [((40, 40), 16), ((40, 40), 48), ((12, 12, 12), 20), ((40, 40), 100),
                                    ((30, 30), 64)])
def test_cycle_bitwise(cuda, dtype, grid, k):
    import torch
    h, vec, lane = _pair(grid, dtype, cuda)
    n = h.levels[0].A.shape[0]
    B = torch.as_tensor(np.random.default_rng(k).standard_normal((n, k)), device=cuda)
    M1, M2 = vec.aspreconditioner("V"), lane.aspreconditioner("V")
    Z1, Z2 = M1(B), M2(B)
    assert torch.equal(Z1, Z2)
    # stationary solve: identical cycles; only the residual norms are summed in another order
    X1, it1, _ = vec.solve(B, tol=1e-8, maxiter=60, return_info=True)
    X2, it2, _ = lane.solve(B, tol=1e-8, maxiter=60, return_info=True)
    assert list(it1) == list(it2)
    assert torch.equal(X1, X2)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pcg_close(cuda, dtype):
    import torch
    h, vec, lane = _pair((40, 40), dtype, cuda)
    n = h.levels[0].A.shape[0]
    B = torch.as_tensor(np.random.default_rng(5).standard_normal((n, 32)), device=cuda)
    X1, it1, r1 = vec.pcg(B, tol=1e-10, maxiter=200, return_info=True)
    X2, it2, r2 = lane.pcg(B, tol=1e-10, maxiter=200, return_info=True)
    assert list(it1) == list(it2)
    assert rel(X1.cpu().numpy(), X2.cpu().numpy()) < 1e-12
    assert max(r1) < 1e-10


def test_columns_stop_apart(cuda):
    """Smooth and rough right-hand sides interleaved: the solve masks converged
    columns inside one lane's group of columns."""
    import torch
    from oracle.solve import OracleMG
    h, vec, lane = _pair((40, 40), np.float64, cuda)
    A = h.levels[0].A
    n = A.shape[0]
    rng = np.random.default_rng(11)
    B = rng.standard_normal((n, 16))
    smooth = A @ np.ones(n)
    B[:, 1::3] = smooth[:, None]
    B[:, 6] = 0.0
    Bt = torch.as_tensor(B, device=cuda)
    X1, it1, _ = vec.solve(Bt, tol=1e-9, maxiter=100, return_info=True)
    X2, it2, _ = lane.solve(Bt, tol=1e-9, maxiter=100, return_info=True)
    assert len(set(it1)) > 1
    assert list(it1) == list(it2)
    assert torch.equal(X1, X2)
    om = OracleMG(h, WITHRHO, WITHRHO)
    Xr, itr = om.solve(B, tol=1e-9, maxiter=100, return_info=True)
    assert list(it1) == list(itr)
    assert rel(X1.cpu().numpy(), Xr) < 1e-12


def test_batch_with_and_without_vector_mode(cuda):
    import torch
    from amjax_amd import BatchPCG
    h, vec, lane = _pair((40, 40), np.float32, cuda)
    n = h.levels[0].A.shape[0]
    B = torch.as_tensor(np.random.default_rng(3).standard_normal((n, 16)), device=cuda)
    (Xv,) = BatchPCG([vec], [16]).pcg([B], tol=1e-8, maxiter=100)
    (Xl,) = BatchPCG([lane], [16]).pcg([B], tol=1e-8, maxiter=100)
    assert torch.equal(Xv, vec.pcg(B, tol=1e-8, maxiter=100))
    assert torch.equal(Xl, lane.pcg(B, tol=1e-8, maxiter=100))
    assert rel(Xv.cpu().numpy(), Xl.cpu().numpy()) < 1e-6
