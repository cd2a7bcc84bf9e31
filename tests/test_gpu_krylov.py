"""Right-preconditioned BiCGSTAB (jax.scipy.sparse.linalg.bicgstab semantics) on the
GPU against the CPU oracle, whose algorithm is pinned to scipy's bicgstab iterates
(test_oracle.py), on a nonsymmetric upwind convection-diffusion operator."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel

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
