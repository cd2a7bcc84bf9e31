"""Gradients of the GPU solves (reference: tests/test_multilevel.py:155-320).

* solve(b, A=A) VJP: central finite differences along a random direction
  (rtol 1e-3, the reference's bar) and values-gradient == dense-gradient on the
  pattern (1e-10, the reference's bar);
* PCG implicit adjoint: values-gradient == dense-gradient on the pattern;
* fixed-cycle b-gradient: equals the transpose of the dense m-cycle operator
  built column by column from the oracle (non-symmetric smoothing, so a wrong
  transposition fails), and central differences (1e-6 relative).
"""

import numpy as np
import pytest

from conftest import rel, rs_hierarchy, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev, **kw):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev, **kw)


@pytest.mark.parametrize("cycle", ["V", "W", "F"])
def test_vjp_dense_A_matches_finite_differences(cuda, cycle):
    """tests/test_multilevel.py:155-182: grad wrt dense A vs central FD, eps 1e-4, rtol 1e-3."""
    import torch
    from amjax_amd import MultilevelSolver
    from amjax_amd.gallery import poisson
    A_s = poisson((10, 10))
    h = rs_hierarchy((10, 10))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    b = _t(np.random.default_rng(42).random(100), cuda)
    d = _t(np.random.default_rng(0).standard_normal((100, 100)), cuda)
    A = _t(A_s.toarray(), cuda, dtype=torch.float64).requires_grad_(True)
    f = lambda M: ml.solve(b, A=M, tol=1e-10, maxiter=100, cycle=cycle).sum()
    loss = f(A)
    (g,) = torch.autograd.grad(loss, A)
    assert g.shape == A.shape
    eps = 1e-4
    with torch.no_grad():
        fd = (f(A + eps * d) - f(A - eps * d)) / (2 * eps)
    np.testing.assert_allclose(float((g * d).sum()), float(fd), rtol=1e-3)


def test_vjp_values_equal_dense_on_pattern(cuda):
    """tests/test_multilevel.py:209-247 (1e-10)."""
    import torch
    from amjax_amd import MultilevelSolver, SparseMatrix
    from amjax_amd.gallery import poisson
    A_s = poisson((5, 5))
    h = rs_hierarchy((5, 5))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    b = _t(np.random.default_rng(42).random(25), cuda)
    coo = A_s.tocoo()
    Ad = _t(A_s.toarray(), cuda).requires_grad_(True)
    (gd,) = torch.autograd.grad(ml.solve(b, A=Ad, tol=1e-10, maxiter=100).sum(), Ad)
    base = SparseMatrix.from_scipy(A_s, np.float64, cuda)
    vals = _t(base.host.data, cuda).requires_grad_(True)
    (gv,) = torch.autograd.grad(ml.solve(b, A=base.with_values(vals), tol=1e-10, maxiter=100).sum(), vals)
    H = base.host.tocoo()
    np.testing.assert_allclose(gv.cpu().numpy(), gd.cpu().numpy()[H.row, H.col], rtol=1e-10, atol=1e-10)


def test_vjp_b_gradient_and_multi_rhs(cuda):
    """grad wrt b = lam = solve(v) (amjax/multilevel.py:442); k columns at once."""
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((32, 32))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((1024, 3))
    W = rng.standard_normal((1024, 3))
    bt = _t(B, cuda).requires_grad_(True)
    x = ml.solve(bt, A=ml.levels[0].A, tol=1e-12, maxiter=200)
    (g,) = torch.autograd.grad((x * _t(W, cuda)).sum(), bt)
    lam = om.solve(W, tol=1e-12, maxiter=200)
    assert rel(g.cpu().numpy(), lam) < 1e-10


def test_pcg_values_gradient_equals_dense(cuda):
    """tests/test_multilevel.py:249-290: fixed AMG preconditioner, grad wrt values == dense grad."""
    import torch
    from amjax_amd import MultilevelSolver, SparseMatrix, cg
    from amjax_amd.gallery import poisson
    A_s = poisson((5, 5))
    h = rs_hierarchy((5, 5))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    M = ml.aspreconditioner("V")
    b = _t(np.random.default_rng(42).random(25), cuda)
    base = SparseMatrix.from_scipy(A_s, np.float64, cuda)
    vals = _t(base.host.data, cuda).requires_grad_(True)
    x, _ = cg(base.with_values(vals), b, M=M, tol=1e-12, maxiter=30)
    (gv,) = torch.autograd.grad(x.sum(), vals)
    # dense reference on the full pattern: -lam x^T with lam = A^-1 1
    Ad = A_s.toarray()
    xs = np.linalg.solve(Ad, b.cpu().numpy())
    lam = np.linalg.solve(Ad.T, np.ones(25))
    H = base.host.tocoo()
    np.testing.assert_allclose(gv.cpu().numpy(), -(lam[H.row] * xs[H.col]), rtol=1e-8, atol=1e-10)


def test_fixed_cycle_b_gradient_is_transposed_operator(cuda):
    """x = C_m(b) after m cycles: grad_b = C_m^T v with C_m built densely from the oracle."""
    import torch
    from amjax_amd import MultilevelSolver, cycles
    from oracle.solve import OracleMG
    h = sa_hierarchy((12, 12))
    pre = ("jacobi", {"iterations": 2, "omega": 0.6})
    post = ("jacobi", {"iterations": 1, "omega": 0.9})
    ml = MultilevelSolver.from_pyamg(h, pre, post, device=cuda)
    om = OracleMG(h, pre, post)
    n, m = 144, 4

    def Cm(b):
        x = np.zeros(n)
        for _ in range(m):
            x = om.cycle(x, b, "V")
        return x

    C = np.stack([Cm(e) for e in np.eye(n)], axis=1)
    assert np.abs(C - C.T).max() > 1e-6  # the smoothing is not symmetric
    v = np.random.default_rng(7).standard_normal(n)
    bt = _t(np.random.default_rng(8).standard_normal(n), cuda).requires_grad_(True)
    x = cycles(ml, bt, m)
    assert rel(x.detach().cpu().numpy(), C @ bt.detach().cpu().numpy()) < 1e-12
    (g,) = torch.autograd.grad((x * _t(v, cuda)).sum(), bt)
    assert rel(g.cpu().numpy(), C.T @ v) < 1e-12


@pytest.mark.parametrize("cycle,cpl", [("W", 1), ("F", 0), ("F", 1), ("F", 2)])
def test_fixed_w_and_f_cycle_b_gradient_is_transposed_operator(cuda, cycle, cpl):
    """W and F cycles on a four-level hierarchy with symmetric smoothing.  The
    F-cycle operator is not symmetric there (its coarse correction runs F then
    V), so its b-gradient needs the transposed cycle 'T' (V then F)."""
    import torch
    from amjax_amd import MultilevelSolver, cycles, gallery, hierarchy
    from oracle.solve import OracleMG
    h = hierarchy.smoothed_aggregation_solver(gallery.poisson((40, 40)), max_coarse=10)
    assert len(h.levels) >= 4
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    n, m = 1600, 2

    def Cm(b):
        x = np.zeros(n)
        for _ in range(m):
            x = om.cycle(x, b, cycle, cpl)
        return x

    C = np.stack([Cm(e) for e in np.eye(n)], axis=1)
    if cycle == "F" and cpl > 0:
        assert np.abs(C - C.T).max() > 1e-6 * np.abs(C).max()
    v = np.random.default_rng(17).standard_normal(n)
    bt = _t(np.random.default_rng(18).standard_normal(n), cuda).requires_grad_(True)
    x = cycles(ml, bt, m, cycle=cycle, cycles_per_level=cpl)
    assert rel(x.detach().cpu().numpy(), C @ bt.detach().cpu().numpy()) < 1e-12
    (g,) = torch.autograd.grad((x * _t(v, cuda)).sum(), bt)
    assert rel(g.cpu().numpy(), C.T @ v) < 1e-12


def test_fixed_cycle_gradient_central_differences(cuda):
    """loss = 0.5 ||C_m(b) - x*||^2, 10 V-cycles: directional FD with h = 1e-6."""
    import torch
    from amjax_amd import MultilevelSolver, cycles
    h = sa_hierarchy((32, 32))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    rng = np.random.default_rng(4)
    xs = _t(rng.standard_normal((1024, 2)), cuda)
    b = _t(rng.standard_normal((1024, 2)), cuda).requires_grad_(True)
    loss = lambda bb: 0.5 * ((cycles(ml, bb, 10) - xs) ** 2).sum()
    (g,) = torch.autograd.grad(loss(b), b)
    for s in range(5):
        d = _t(np.random.default_rng(100 + s).standard_normal((1024, 2)), cuda)
        with torch.no_grad():
            fd = (loss(b + 1e-6 * d) - loss(b - 1e-6 * d)) / 2e-6
        np.testing.assert_allclose(float((g * d).sum()), float(fd), rtol=1e-6)


def test_gradient_descent_on_A_decreases_loss(cuda):
    """tests/test_multilevel.py:292-320, with step 1e-5: on the in-repo RS hierarchy a
    step of 1e-4 makes the fixed-hierarchy solve diverge at the second step in the
    oracle as well (loss 8e81), so the GPU trajectory is checked against the oracle's."""
    import scipy.sparse as sp
    import torch
    from amjax_amd import MultilevelSolver
    from amjax_amd.gallery import poisson
    from oracle.solve import OracleMG
    A_s = poisson((10, 10))
    h = rs_hierarchy((10, 10))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    bn = np.random.default_rng(42).random(100)
    b = _t(bn, cuda)
    target = torch.ones(100, dtype=torch.float64, device=cuda)
    A = _t(A_s.toarray(), cuda)
    Ar = A_s.toarray()
    losses, ref = [], []
    for _ in range(21):
        A = A.detach().requires_grad_(True)
        L = ((ml.solve(b, A=A, tol=1e-10, maxiter=100) - target) ** 2).sum()
        losses.append(float(L.detach()))
        (g,) = torch.autograd.grad(L, A)
        A = A - 1e-5 * g
        xr = om.solve(bn, A=sp.csr_matrix(Ar), tol=1e-10, maxiter=100)
        ref.append(float(((xr - 1) ** 2).sum()))
        lam = om.solve(2 * (xr - 1), A=sp.csr_matrix(Ar), tol=1e-10, maxiter=100)
        Ar = Ar + 1e-5 * np.outer(lam, xr)
    np.testing.assert_allclose(losses, ref, rtol=1e-8)
    assert losses[-1] < losses[0]
    assert all(b2 <= a2 for a2, b2 in zip(losses, losses[1:]))


def test_gradients_with_x0(cuda):
    """A start x0 (amjax/multilevel.py:385-393: solve for b - A x0, add x0 back):
    gradients reach b, A's values (sparse and dense A) and x0 through the shift, and
    at convergence equal the closed forms lam = A^-T w, -lam x^T and 0; cg with x0
    has the same values-gradient."""
    import torch
    from amjax_amd import MultilevelSolver, SparseMatrix, cg
    from amjax_amd.gallery import poisson
    A_s = poisson((5, 5))
    h = rs_hierarchy((5, 5))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    rng = np.random.default_rng(8)
    b_np, w_np, x0_np = rng.random(25), rng.standard_normal(25), rng.standard_normal(25)
    Ad = A_s.toarray()
    xs = np.linalg.solve(Ad, b_np)
    lam = np.linalg.solve(Ad.T, w_np)
    base = SparseMatrix.from_scipy(A_s, np.float64, cuda)
    H = base.host.tocoo()
    w = _t(w_np, cuda)
    vals = _t(base.host.data, cuda).requires_grad_(True)
    bt = _t(b_np, cuda).requires_grad_(True)
    x0 = _t(x0_np, cuda).requires_grad_(True)
    x = ml.solve(bt, A=base.with_values(vals), x0=x0, tol=1e-13, maxiter=200)
    gb, gv, gx0 = torch.autograd.grad((x * w).sum(), (bt, vals, x0))
    np.testing.assert_allclose(x.detach().cpu().numpy(), xs, rtol=1e-10)
    np.testing.assert_allclose(gb.cpu().numpy(), lam, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gv.cpu().numpy(), -(lam[H.row] * xs[H.col]), rtol=1e-8, atol=1e-11)
    assert np.abs(gx0.cpu().numpy()).max() < 1e-9 * np.abs(w_np).max()
    A_t = _t(Ad, cuda).requires_grad_(True)
    xd = ml.solve(_t(b_np, cuda), A=A_t, x0=_t(x0_np, cuda), tol=1e-13, maxiter=200)
    (gA,) = torch.autograd.grad((xd * w).sum(), A_t)
    np.testing.assert_allclose(gA.cpu().numpy(), -np.outer(lam, xs), rtol=1e-8, atol=1e-11)
    vals2 = _t(base.host.data, cuda).requires_grad_(True)
    xc, _ = cg(base.with_values(vals2), _t(b_np, cuda), x0=_t(x0_np, cuda), M=ml.aspreconditioner("V"),
               tol=1e-13, maxiter=60)
    (gv2,) = torch.autograd.grad((xc * w).sum(), vals2)
    np.testing.assert_allclose(gv2.cpu().numpy(), -(lam[H.row] * xs[H.col]), rtol=1e-8, atol=1e-11)


def test_transpose_shares_a_self_adjoint_hierarchy(cuda):
    """A symmetric SA hierarchy with R = P^T and alike pre- and post-smoothers is
    its own transpose: transpose() returns the solver (no second upload); different
    pre- and post-smoothers give a separate transposed solver."""
    from amjax_amd import MultilevelSolver
    h = sa_hierarchy((16, 16))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    assert ml.transpose() is ml
    pre = ("jacobi", {"iterations": 2, "omega": 0.6})
    post = ("jacobi", {"iterations": 1, "omega": 0.9})
    ml2 = MultilevelSolver.from_pyamg(h, pre, post, device=cuda)
    mt = ml2.transpose()
    assert mt is not ml2 and mt is ml2.transpose()
