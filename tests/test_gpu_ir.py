"""Float64 iterative refinement around a CG in the hierarchy's dtype (amg_pcg_ir)
against the oracle's restatement (oracle.solve.pcg_ir) and against plain PCG."""

import numpy as np
import pytest

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


@pytest.mark.parametrize("k", [1, 3, 16])
def test_one_float64_step_is_pcg(cuda, k):
    """float64 vectors, one step, Hestenes-Stiefel beta: the inner CG is PCG."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    h = sa_hierarchy((40, 40))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    A = om.levels[0].A
    B = np.random.default_rng(k).standard_normal((A.shape[0], k))
    X, its, steps, rr = ml.pcg_ir(_t(B, cuda), tol=1e-9, inner_tol=1e-9, maxiter=1, flexible=False,
                                  return_info=True)
    Xr, itr = pcg(A, B, om.aspreconditioner(), tol=1e-9, return_info=True)
    assert list(its) == list(itr) and list(steps) == [1] * k
    assert rel(X.cpu().numpy(), Xr) < 1e-12


@pytest.mark.parametrize("k", [1, 4, 64])
def test_mixed_refinement_against_oracle(cuda, k):
    """float32 inner CG: the same number of refinement steps as the oracle, the
    inner iteration totals within one per step, and a float64 residual below a
    tolerance far under float32 precision."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg, pcg_ir
    h = sa_hierarchy((40, 40))
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    om64 = OracleMG(h, WITHRHO, WITHRHO)
    om32 = OracleMG(h, WITHRHO, WITHRHO, dtype=np.float32)
    A = om64.levels[0].A
    B = np.random.default_rng(20 + k).standard_normal((A.shape[0], k))
    if k > 1:
        B[:, 1] = 0.0
    X, its, steps, rr = ml32.pcg_ir(_t(B, cuda), tol=1e-11, inner_tol=1e-4, maxiter=10, return_info=True)
    its, steps, rr = np.atleast_1d(its), np.atleast_1d(steps), np.atleast_1d(rr)
    Xr, itr, str_ = pcg_ir(A, B if k > 1 else B[:, 0], om32.aspreconditioner(), tol=1e-11, inner_tol=1e-4,
                           maxiter=10, return_info=True)
    itr, str_ = np.atleast_1d(itr), np.atleast_1d(str_)
    assert list(steps) == list(str_)
    assert np.all(np.abs(np.asarray(its) - np.asarray(itr)) <= np.asarray(steps))
    X = X.cpu().numpy().reshape(B.shape)
    live = [c for c in range(k) if np.any(B[:, c])]
    true = np.linalg.norm(B - A @ X, axis=0)[live] / np.linalg.norm(B, axis=0)[live]
    assert max(true) < 1e-11 and max(rr[live]) < 1e-11
    X64 = pcg(A, B, om64.aspreconditioner(), tol=1e-14).reshape(B.shape)
    assert rel(X, X64) < 1e-9
    if k > 1:
        assert np.all(X[:, 1] == 0.0) and steps[1] == 0


def test_refinement_faster_steps_than_cycles(cuda):
    """On a 2-D Poisson problem the refinement needs 2-4 steps to 1e-10 with an
    inner tolerance of 1e-4, and the summed inner iterations stay within two of
    the float64-vector PCG's iteration count per step."""
    from amjax_amd import MultilevelSolver
    h = sa_hierarchy((64, 64))
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    B = _t(np.random.default_rng(3).standard_normal((4096, 8)), cuda)
    X, its, steps, rr = ml32.pcg_ir(B, tol=1e-10, inner_tol=1e-4, maxiter=10, return_info=True)
    _, itc, _ = ml32.pcg(B, tol=1e-10, maxiter=500, return_info=True)
    assert max(rr) < 1e-10
    assert all(2 <= s <= 4 for s in steps)
    assert all(i <= c + 2 * s for i, c, s in zip(its, itc, steps))


def test_inner_tolerance_changes_between_calls(cuda):
    """The captured step graphs are reused across calls: a new inner tolerance
    (and inner iteration cap) still takes effect."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg_ir
    h = sa_hierarchy((40, 40))
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    om32 = OracleMG(h, WITHRHO, WITHRHO, dtype=np.float32)
    A = sa_hierarchy((40, 40)).levels[0].A
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    seen = []
    for itol, imax in ((1e-2, 100), (1e-5, 100), (1e-2, 100), (1e-5, 3)):
        x, its, steps, rr = ml32.pcg_ir(_t(b, cuda), tol=1e-10, inner_tol=itol, maxiter=30, inner_maxiter=imax,
                                        return_info=True)
        _, itr, str_ = pcg_ir(A, b, om32.aspreconditioner(), tol=1e-10, inner_tol=itol, maxiter=30,
                              inner_maxiter=imax, return_info=True)
        assert steps == str_ and abs(its - itr) <= steps and rr < 1e-10
        seen.append(steps)
    assert seen[0] == seen[2] and seen[1] < seen[0] and seen[3] > seen[1]
