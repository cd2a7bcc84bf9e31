"""GPU parity of the HIP solve phase against the CPU oracle (same hierarchy, same inputs).

Bar: float64 iterates within 1e-12 relative of the oracle after fixed cycle
counts; identical iteration counts to tolerance; float32 paths within the
float32 bounds stated per test.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel, rs_hierarchy, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})
DEFAULT = ("jacobi", {"iterations": 1, "omega": 1.0})


def _t(x, dev, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev, dtype=dtype)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("k", [1, 3, 16, 64, 100])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_matches_scipy(cuda, k, dtype):
    from amjax_amd import SparseMatrix
    rng = np.random.default_rng(k)
    M = sp.random(300, 257, density=0.03, random_state=k, format="csr") + sp.eye(300, 257, format="csr")
    M = M.astype(dtype)
    x = rng.standard_normal((257, k)).astype(dtype)
    A = SparseMatrix.from_scipy(M, dtype, cuda)
    y = _np(A @ _t(x, cuda))
    tol = 1e-13 if dtype == np.float64 else 2e-6
    assert rel(y, M.astype(np.float64) @ x.astype(np.float64)) < tol
    # empty rows and a 1-row matrix
    E = sp.csr_matrix((5, 7), dtype=dtype)
    assert np.all(_np(SparseMatrix.from_scipy(E, dtype, cuda) @ _t(np.ones((7, k), dtype), cuda)) == 0)


def test_jacobi_known_answers(cuda):
    """tests/relaxation/test_relaxation.py:42-80 of the reference, on the GPU."""
    import torch
    from amjax_amd import SparseMatrix, jacobi, inverse_diagonal

    def make(N):
        M = sp.diags([2 * np.ones(N), -np.ones(N - 1), -np.ones(N - 1)], [0, -1, 1], format="csr")
        A = SparseMatrix.from_scipy(M, np.float64, cuda)
        return A, inverse_diagonal(A)

    f = lambda v: torch.as_tensor(np.asarray(v, dtype=np.float64), device=cuda)
    A, D = make(1)
    np.testing.assert_allclose(_np(jacobi(A, f([0.0]), f([0.0]), D)), [0.0])
    A, D = make(3)
    np.testing.assert_allclose(_np(jacobi(A, f([0, 0, 0]), f([0, 1, 2]), D)), [0.0, 0.5, 1.0])
    np.testing.assert_allclose(_np(jacobi(A, f([0, 1, 2]), f([0, 0, 0]), D)), [0.5, 1.0, 0.5])
    A, D = make(1)
    np.testing.assert_allclose(_np(jacobi(A, f([0.0]), f([10.0]), D)), [5.0])
    A, D = make(3)
    np.testing.assert_allclose(_np(jacobi(A, f([0, 1, 2]), f([10, 20, 30]), D)), [5.5, 11.0, 15.5])
    x = _np(jacobi(A, f([0, 1, 2]), f([10, 20, 30]), D, omega=1.0 / 3.0))
    np.testing.assert_allclose(x, 2.0 / 3.0 * np.arange(3) + 1.0 / 3.0 * np.array([5.5, 11.0, 15.5]))


def test_dense_coarse_solvers(cuda):
    """A x = b for pinv / lu / qr / jacobi (tests/test_multilevel.py:16-47)."""
    import scipy.linalg as sla
    import torch
    from amjax_amd import SparseMatrix, coarse_grid_solver, inverse_diagonal, MultilevelSolver
    from amjax_amd.gallery import poisson
    cases = [sp.diags(np.arange(1, 5, dtype=float)).tocsr(), poisson((4,)), poisson((4, 4))]
    for M in cases:
        A = SparseMatrix.from_scipy(M, np.float64, cuda)
        b = torch.arange(M.shape[0], dtype=torch.float64, device=cuda)
        Ad = M.toarray()
        for name, kw in [("jacobi", {"iterations": 500, "omega": 2 / 3}), ("pinv", {}), ("lu", {}), ("qr", {})]:
            lvl = MultilevelSolver.Level()
            lvl.A = A
            lvl.Dinv = inverse_diagonal(A)
            if name == "pinv":
                lvl.A_inv = torch.as_tensor(np.linalg.pinv(Ad), device=cuda)
            elif name == "lu":
                lu, piv = sla.lu_factor(Ad)
                lvl.lu_factor = torch.as_tensor(lu, device=cuda)
                lvl.piv = torch.as_tensor(piv.astype(np.int32), device=cuda)
            elif name == "qr":
                Q, R = np.linalg.qr(Ad)
                lvl.Q, lvl.R_mat = torch.as_tensor(Q, device=cuda), torch.as_tensor(R, device=cuda)
            x = coarse_grid_solver(name, lvl, **kw)(A, torch.zeros_like(b), b)
            np.testing.assert_almost_equal(Ad @ _np(x), _np(b), decimal=3)


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("smoother", [WITHRHO, DEFAULT, ("jacobi", {"iterations": 2, "omega": 0.8}), None])
def test_cycle_parity_fixed_counts(cuda, cycle, smoother, fuse):
    """FP64 iterates after 1, 5 and 10 cycles within 1e-12 of the oracle, with the
    fused sweeps (rounding order changed) and with one kernel per reference operation."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, presmoother=smoother, postsmoother=smoother, device=cuda)
    ml.fuse = fuse
    om = OracleMG(h, presmoother=smoother, postsmoother=smoother)
    b = np.random.default_rng(0).standard_normal(4096)
    x_ref = np.zeros(4096)
    x = _t(np.zeros(4096), cuda)
    bt = _t(b, cuda)
    for it in range(1, 11):
        x_ref = om.cycle(x_ref, b, cycle)
        x = ml._cycle(x, bt, cycle)
        if it in (1, 5, 10):
            assert rel(_np(x), x_ref) < 1e-12, (it, rel(_np(x), x_ref))
    z = ml.aspreconditioner(cycle)(bt)
    assert rel(_np(z), om.aspreconditioner(cycle)(b)) < 1e-12


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("cycle", ["V", "W", "F"])
@pytest.mark.parametrize("coarse", ["pinv", "lu", "qr", "jacobi"])
def test_solve_parity(cuda, cycle, coarse, fuse):
    """Standalone solve: same iteration count and iterates as the oracle."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, relres
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, coarse_solver=coarse, device=cuda)
    ml.fuse = fuse
    om = OracleMG(h, WITHRHO, WITHRHO, coarse_solver=coarse)
    b = np.random.default_rng(0).standard_normal(4096)
    x_ref, it_ref = om.solve(b, tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
    x, it, rr = ml.solve(_t(b, cuda), tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
    assert it == it_ref
    assert rel(_np(x), x_ref) < 1e-12
    assert rr < 1e-8 and abs(rr - relres(om.levels[0].A, b, x_ref)) < 1e-12


@pytest.mark.parametrize("cpl", [0, 2, 3])
def test_f_cycle_cycles_per_level(cuda, cpl):
    """F-cycle with cycles_per_level V-cycles after each F visit
    (amjax/multilevel.py:331-334): fixed-count iterates, the preconditioner and the
    solve against the oracle, whose F recursion is pinned by the dense propagator."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    b = np.random.default_rng(3).standard_normal(4096)
    x_ref, x = np.zeros(4096), _t(np.zeros(4096), cuda)
    for _ in range(4):
        x_ref = om.cycle(x_ref, b, "F", cpl)
        x = ml._cycle(x, _t(b, cuda), "F", cpl)
    assert rel(_np(x), x_ref) < 1e-12
    z = ml.aspreconditioner("F", cpl)(_t(b, cuda))
    assert rel(_np(z), om.aspreconditioner("F", cpl)(b)) < 1e-12
    x_ref, it_ref = om.solve(b, tol=1e-8, maxiter=100, cycle="F", cycles_per_level=cpl, return_info=True)
    x, it, _ = ml.solve(_t(b, cuda), tol=1e-8, maxiter=100, cycle="F", cycles_per_level=cpl, return_info=True)
    assert it == it_ref and rel(_np(x), x_ref) < 1e-12


def test_solve_rs_hierarchy(cuda):
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = rs_hierarchy((20, 20))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    b = np.random.default_rng(1).random(400)
    for cycle in "VWF":
        x_ref, it_ref = om.solve(b, tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
        x, it, _ = ml.solve(_t(b, cuda), tol=1e-8, maxiter=100, cycle=cycle, return_info=True)
        assert it == it_ref and rel(_np(x), x_ref) < 1e-12


@pytest.mark.parametrize("k", [4, 64])
def test_multi_rhs_solve_freezes_columns(cuda, k):
    """k independent solves (vmap semantics): per-column iteration counts and iterates."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((4096, k)) * np.logspace(0, 3, k)[None, :]
    B[:, 0] = 0.0  # a zero column: normb guard, converged at once
    if k > 2:
        B[:, 2] = om.levels[0].A @ np.ones(4096)  # smooth right-hand side
    X_ref, its_ref = om.solve(B, tol=1e-8, maxiter=100, return_info=True)
    X, its, rr = ml.solve(_t(B, cuda), tol=1e-8, maxiter=100, return_info=True)
    assert list(its) == list(its_ref)
    for c in range(k):
        assert rel(_np(X)[:, c], X_ref[:, c]) < 1e-12 or np.all(X_ref[:, c] == 0)


def test_solve_x0_and_maxiter(cuda):
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    rng = np.random.default_rng(2)
    b, x0 = rng.standard_normal(4096), rng.standard_normal(4096)
    x_ref = om.solve(b, x0=x0, tol=1e-10, maxiter=100)
    x = ml.solve(_t(b, cuda), x0=_t(x0, cuda), tol=1e-10, maxiter=100)
    assert rel(_np(x), x_ref) < 1e-12
    # both meet the shifted tolerance ||b - A x|| < tol ||b - A x0|| (amjax/multilevel.py:385-393)
    A = om.levels[0].A
    r0 = np.linalg.norm(b - A @ x0)
    assert np.linalg.norm(b - A @ _np(x)) < 1e-10 * r0
    assert np.linalg.norm(b - A @ x_ref) < 1e-10 * r0
    # forced non-convergence: maxiter bounds the count, tol=0 never stops early
    x, it, _ = ml.solve(_t(b, cuda), tol=0.0, maxiter=3, return_info=True)
    x_ref, it_ref = om.solve(b, tol=0.0, maxiter=3, return_info=True)
    assert it == it_ref == 3 and rel(_np(x), x_ref) < 1e-12
    x, it, _ = ml.solve(_t(b, cuda), tol=1e-8, maxiter=0, return_info=True)
    assert it == 0 and np.all(_np(x) == 0)


@pytest.mark.parametrize("fuse", [True, False])
@pytest.mark.parametrize("cycle", ["V", "W", "F"])
def test_pcg_parity(cuda, cycle, fuse):
    """AMG-preconditioned CG: iteration count and iterate of jax-cg semantics."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg, relres
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    ml.fuse = fuse
    om = OracleMG(h, WITHRHO, WITHRHO)
    A = om.levels[0].A
    b = np.random.default_rng(0).standard_normal(4096)
    x_ref, it_ref = pcg(A, b, om.aspreconditioner(cycle), tol=1e-8, maxiter=500, return_info=True)
    x, it, rr = ml.pcg(_t(b, cuda), tol=1e-8, maxiter=500, cycle=cycle, return_info=True)
    assert it == it_ref
    assert rel(_np(x), x_ref) < 1e-12
    assert abs(rr - relres(A, b, x_ref)) < 1e-12
    # plain CG (M = identity)
    x_ref, it_ref = pcg(A, b, None, tol=1e-8, maxiter=2000, return_info=True)
    x, it, _ = ml.pcg(_t(b, cuda), tol=1e-8, maxiter=2000, precondition=False, return_info=True)
    assert it == it_ref and rel(_np(x), x_ref) < 1e-10


def test_pcg_multi_rhs(cuda):
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(3).standard_normal((4096, 8)) * np.logspace(-2, 2, 8)[None, :]
    X_ref, its_ref = pcg(om.levels[0].A, B, om.aspreconditioner(), tol=1e-8, maxiter=500, return_info=True)
    X, its, _ = ml.pcg(_t(B, cuda), tol=1e-8, maxiter=500, return_info=True)
    assert list(its) == list(its_ref)
    assert rel(_np(X), X_ref) < 1e-12


def test_mixed_precision_reaches_fp64_tolerance(cuda):
    """float32 cycle inside float64 PCG / iterative refinement reaches 1e-10 with
    the error against a float64 oracle converged to 1e-14 stated."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    import torch
    h = sa_hierarchy((64, 64))
    om = OracleMG(h, WITHRHO, WITHRHO)
    A = om.levels[0].A
    b = np.random.default_rng(3).standard_normal((4096, 4))
    X64 = pcg(A, b, om.aspreconditioner(), tol=1e-14, maxiter=500)
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    X, its, rr = ml32.pcg(_t(b, cuda), tol=1e-10, maxiter=500, return_info=True)
    assert X.dtype == torch.float64
    assert max(rr) < 1e-10
    assert rel(_np(X), X64) < 1e-8
    X, its, rr = ml32.solve(_t(b, cuda), tol=1e-10, maxiter=200, return_info=True)
    assert X.dtype == torch.float64
    assert max(rr) < 1e-10 and rel(_np(X), X64) < 1e-8
    # all-float32 solve keeps float32 (tests/test_multilevel.py:402-435)
    x32 = ml32.solve(_t(b[:, 0].astype(np.float32), cuda), tol=1e-4, maxiter=20)
    assert x32.dtype == torch.float32


def test_error_conventions_gpu(cuda):
    from amjax_amd import MultilevelSolver
    h = sa_hierarchy((16, 16))
    ml = MultilevelSolver.from_pyamg(h, device=cuda)
    b = _t(np.ones(256), cuda)
    with pytest.raises(ValueError):
        ml.solve(b, cycle="X")
    with pytest.raises(ValueError):
        ml.aspreconditioner("AMLI")
