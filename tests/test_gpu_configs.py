"""Full-size parity cases of the measured configurations.

* Mixed precision (2-D Poisson 1000x1000 and 3-D Poisson 100^3, N = 10^6, SA,
  k = 1 and k = 32, RHS ~ N(0,1) seed 3): the float32 cycle inside float64 CG,
  the reliable-update CG (float32 vectors), float64 iterative refinement with
  one cycle per step and refinement around a float32-vector CG each reach 1e-10
  on the true float64 residual; the error against a float64 reference
  converged to 1e-14 is printed and bounded.  The float64 reference itself
  matches the CPU oracle on one column.
* Differentiation (2-D variable-coefficient diffusion 512x512, conductance
  exp(g), g a squared-exponential Gaussian random field with length 16 cells,
  sigma 1, seed 1; k = 16; loss 0.5 ||X - X*||^2, X* ~ N(0,1) seed 4): the
  b-gradient through 10 fixed V-cycles and the A-values gradient through the
  converged PCG (1e-10) against central differences along 20 random directions
  with step 1e-6.
"""

import functools

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _relres(A, B, X):
    R = B - A @ X
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


@functools.lru_cache(maxsize=1)
def _poisson_sa(grid):
    from amjax_amd import gallery, hierarchy
    A = gallery.poisson(grid)
    return A, hierarchy.smoothed_aggregation_solver(A)


@pytest.mark.parametrize("k", [1, 32])
@pytest.mark.parametrize("grid", [(1000, 1000), (100, 100, 100)])
def test_config_mixed_precision_1e10(cuda, grid, k):
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    A, h = _poisson_sa(grid)
    N = A.shape[0]
    B = np.random.default_rng(3).standard_normal((N, 32))[:, :k]
    Bt = torch.as_tensor(B, device=cuda)
    ml64 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    X64, it64, rr64 = ml64.pcg(Bt, tol=1e-14, maxiter=1000, return_info=True)
    X64 = X64.cpu().numpy()
    # the recurrence residual reaches 1e-14; the true float64 residual of a 10^6-row
    # system floors near 1e-13 from the rounding of b - A x itself
    assert max(_relres(A, B, X64)) < 1e-12
    # the float64 GPU reference against the CPU oracle on one column
    om = OracleMG(h, WITHRHO, WITHRHO)
    x0 = pcg(om.levels[0].A, B[:, 0], om.aspreconditioner(), tol=1e-14, maxiter=1000)
    assert np.linalg.norm(X64[:, 0] - x0) / np.linalg.norm(x0) < 1e-11
    nrm = np.linalg.norm(X64, axis=0)
    report = {}
    forms = (("mixed-cg", lambda: ml32.pcg(Bt, tol=1e-10, maxiter=1000, return_info=True)),
             ("mixed-ru", lambda: ml32.pcg_ru(Bt, tol=1e-10, maxiter=1000, return_info=True)[:2]),
             ("mixed-ir", lambda: ml32.solve(Bt, tol=1e-10, maxiter=1000, return_info=True)),
             ("mixed-ircg", lambda: ml32.pcg_ir(Bt, tol=1e-10, inner_tol=1e-4, maxiter=20,
                                                return_info=True)[:2]),
             ("fp64-cg", lambda: ml64.pcg(Bt, tol=1e-10, maxiter=1000, return_info=True)))
    for name, fn in forms:
        X, its = fn()[:2]
        assert X.dtype == torch.float64
        X = X.cpu().numpy()
        rr = _relres(A, B, X)
        err = np.linalg.norm(X - X64, axis=0) / nrm
        report[name] = (max(its), float(rr.max()), float(err.max()))
        assert rr.max() < 1e-10, (name, rr.max())
        assert err.max() < 1e-6, (name, err.max())
    print(f"\n{grid} k={k}: " + "; ".join(f"{f} iters {v[0]} relres {v[1]:.2e} err-vs-fp64 {v[2]:.2e}"
                                         for f, v in report.items()), flush=True)


@pytest.fixture(scope="module")
def diffusion512():
    from amjax_amd import gallery, hierarchy
    g = gallery.gaussian_random_field((512, 512), length=16, sigma=1.0, seed=1)
    A = gallery.diffusion_fv(g)
    return A, hierarchy.smoothed_aggregation_solver(A)


def test_config_diffusion_fixed_cycle_b_gradient(cuda, diffusion512):
    import torch
    from amjax_amd import MultilevelSolver, cycles
    A, h = diffusion512
    N = A.shape[0]
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    rng = np.random.default_rng(4)
    Xs = torch.as_tensor(rng.standard_normal((N, 16)), device=cuda)
    B = torch.as_tensor(rng.standard_normal((N, 16)), device=cuda).requires_grad_(True)
    loss = lambda bb: 0.5 * ((cycles(ml, bb, 10) - Xs) ** 2).sum()
    (g,) = torch.autograd.grad(loss(B), B)
    worst = 0.0
    for s in range(20):
        d = torch.as_tensor(np.random.default_rng(1000 + s).standard_normal((N, 16)), device=cuda)
        with torch.no_grad():
            fd = float((loss(B + 1e-6 * d) - loss(B - 1e-6 * d)) / 2e-6)
        ad = float((g * d).sum())
        worst = max(worst, abs(ad - fd) / abs(fd))
    print(f"\ndiffusion 512^2 fixed-cycle b-gradient: worst relative FD mismatch {worst:.2e}", flush=True)
    assert worst < 1e-5


def test_config_diffusion_pcg_values_gradient(cuda, diffusion512):
    import torch
    from amjax_amd import MultilevelSolver, cg
    A, h = diffusion512
    N = A.shape[0]
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    M = ml.aspreconditioner("V")
    base = ml.levels[0].A
    rng = np.random.default_rng(4)
    Xs = torch.as_tensor(rng.standard_normal((N, 16)), device=cuda)
    B = torch.as_tensor(rng.standard_normal((N, 16)), device=cuda)
    v0 = torch.as_tensor(base.host.data.copy(), device=cuda)
    vals = v0.clone().requires_grad_(True)
    X, _ = cg(base.with_values(vals), B, M=M, tol=1e-10, maxiter=500)
    (g,) = torch.autograd.grad(0.5 * ((X - Xs) ** 2).sum(), vals)

    def loss_at(v):
        with torch.no_grad():
            Xv, _ = cg(base.with_values(v), B, M=M, tol=1e-14, maxiter=500)
            return float(0.5 * ((Xv - Xs) ** 2).sum())

    H = sp.csr_matrix(base.host)
    worst = 0.0
    for s in range(20):
        D = H.copy()
        D.data = np.random.default_rng(2000 + s).standard_normal(D.nnz)
        D = ((D + D.T) * 0.5).tocsr()          # symmetric direction on the (symmetric) pattern
        D.sort_indices()
        d = torch.as_tensor(D.data, device=cuda)
        fd = (loss_at(v0 + 1e-6 * d) - loss_at(v0 - 1e-6 * d)) / 2e-6
        ad = float((g * d).sum())
        worst = max(worst, abs(ad - fd) / abs(fd))
    print(f"\ndiffusion 512^2 PCG values gradient: worst relative FD mismatch {worst:.2e}", flush=True)
    assert worst < 1e-4
