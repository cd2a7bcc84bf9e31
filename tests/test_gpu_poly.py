"""Polynomial smoothers ('chebyshev', 'richardson': PyAMG's polynomial sweep) as
fused HIP sweeps, against the CPU oracle (whose sweep is pinned by the closed
form of its error propagator in test_oracle.py)."""

import numpy as np
import pytest

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

CHEB = ("chebyshev", {})
CHEB2 = ("chebyshev", {"degree": 2, "iterations": 2})
RICH = ("richardson", {"omega": 1.0, "iterations": 2})
WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


@pytest.mark.parametrize("k", [1, 3, 64])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sell_poly_matches_oracle(cuda, k, dtype):
    from amjax_amd import SparseMatrix, gallery
    from amjax_amd.relaxation import polynomial
    from oracle.solve import chebyshev_polynomial_coefficients, polynomial as opoly
    A = gallery.poisson((37, 29)).astype(dtype)
    coefs = list(-chebyshev_polynomial_coefficients(8.0 / 30, 8.8, 3)[:-1])
    rng = np.random.default_rng(k)
    x = rng.standard_normal((A.shape[0], k)).astype(dtype)
    b = rng.standard_normal((A.shape[0], k)).astype(dtype)
    M = SparseMatrix.from_scipy(A, dtype, cuda)
    tol = 1e-13 if dtype == np.float64 else 1e-5
    for iters in (1, 2, 3):
        for x0 in (x, np.zeros_like(x)):
            y = polynomial(M, _t(x0, cuda), _t(b, cuda), coefs, iterations=iters).cpu().numpy()
            ref = opoly(A.astype(np.float64), x0.astype(np.float64), b.astype(np.float64), coefs, iterations=iters)
            assert rel(y, ref) < tol, (iters, rel(y, ref))
    y = polynomial(M, _t(x, cuda), _t(b, cuda), [0.1]).cpu().numpy()
    assert rel(y, x + 0.1 * (b - A @ x)) < tol


@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("pre,post", [(CHEB, CHEB), (RICH, RICH), (CHEB2, WITHRHO), (None, CHEB)])
def test_cycle_parity_polynomial_smoothers(cuda, cycle, pre, post):
    """float64 iterates after 1, 5 and 10 cycles within 1e-12 of the oracle; the
    solve's iteration count is identical."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, presmoother=pre, postsmoother=post, device=cuda)
    om = OracleMG(h, presmoother=pre, postsmoother=post)
    b = np.random.default_rng(0).standard_normal(4096)
    x_ref = np.zeros(4096)
    x = _t(np.zeros(4096), cuda)
    bt = _t(b, cuda)
    for it in range(1, 11):
        x_ref = om.cycle(x_ref, b, cycle)
        x = ml._cycle(x, bt, cycle)
        if it in (1, 5, 10):
            assert rel(x.cpu().numpy(), x_ref) < 1e-12, (it, rel(x.cpu().numpy(), x_ref))
    xs, its, _ = ml.solve(bt, tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
    xr, itr = om.solve(b, tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
    assert its == itr
    assert rel(xs.cpu().numpy(), xr) < 1e-11


def test_pcg_with_chebyshev_and_transposed_gradient(cuda):
    """PCG with Chebyshev smoothing matches the oracle (count and iterate); the
    fixed-cycle b-gradient equals the transpose of the dense cycle operator."""
    import torch
    from amjax_amd import MultilevelSolver, cycles
    from oracle.solve import OracleMG, pcg
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, CHEB, CHEB, device=cuda)
    om = OracleMG(h, CHEB, CHEB)
    b = np.random.default_rng(1).standard_normal(4096)
    x_ref, it_ref = pcg(om.levels[0].A, b, om.aspreconditioner("V"), tol=1e-10, maxiter=200, return_info=True)
    x, it, _ = ml.pcg(_t(b, cuda), tol=1e-10, maxiter=200, return_info=True)
    assert it == it_ref and rel(x.cpu().numpy(), x_ref) < 1e-11
    hs = sa_hierarchy((12, 12))
    pre, post = ("chebyshev", {"degree": 2}), ("richardson", {"iterations": 1, "omega": 1.3})
    mls = MultilevelSolver.from_pyamg(hs, pre, post, device=cuda)
    oms = OracleMG(hs, pre, post)
    n, m = 144, 3

    def Cm(v):
        x = np.zeros(n)
        for _ in range(m):
            x = oms.cycle(x, v, "V")
        return x

    C = np.stack([Cm(e) for e in np.eye(n)], axis=1)
    assert np.abs(C - C.T).max() > 1e-8   # pre != post: not self-adjoint
    v = np.random.default_rng(7).standard_normal(n)
    bt = _t(np.random.default_rng(8).standard_normal(n), cuda).requires_grad_(True)
    y = cycles(mls, bt, m)
    assert rel(y.detach().cpu().numpy(), C @ bt.detach().cpu().numpy()) < 1e-12
    (g,) = torch.autograd.grad((y * _t(v, cuda)).sum(), bt)
    assert rel(g.cpu().numpy(), C.T @ v) < 1e-12


def test_batch_runs_polynomial_smoothers(cuda):
    """A one-problem batch with Chebyshev smoothing equals the single solve."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver
    h = sa_hierarchy((40, 40))
    ml = MultilevelSolver.from_pyamg(h, CHEB, CHEB, device=cuda)
    b = torch.as_tensor(np.random.default_rng(1).standard_normal((1600, 1)), device=cuda)
    batch = BatchPCG([ml], [1])
    nl, _ = batch.info()
    assert nl > 0
    (xb,) = batch.pcg([b], tol=1e-8, maxiter=100)
    assert torch.equal(xb, ml.pcg(b, tol=1e-8, maxiter=100))
