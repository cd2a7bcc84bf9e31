"""CG with reliable updates (amg_pcg_ru) against the oracle.

A float64 handle computes the same sequence as the oracle's float64
restatement (iterates 1e-12, identical iteration and update counts), which
itself equals pcg in float64.  A float32 handle (the mixed form) reaches the
tolerance on the true float64 residual within one iteration of the oracle's
float32 restatement, with its error against a float64 solve converged to
1e-14 stated.
"""

import numpy as np
import pytest

from conftest import rel, sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


@pytest.mark.parametrize("delta", [0.5, 1e-2, 1e-5])
@pytest.mark.parametrize("k", [1, 4])
def test_ru_float64_matches_oracle(cuda, delta, k):
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg_ru
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(21).standard_normal((4096, k)) * np.logspace(-1, 1, k)[None, :]
    X, its, ups, rr = ml.pcg_ru(_t(B, cuda), tol=1e-10, delta=delta, maxiter=300, return_info=True)
    X_ref, its_ref, ups_ref = pcg_ru(om.levels[0].A, B, om.aspreconditioner(), tol=1e-10, delta=delta,
                                     maxiter=300, dtype=np.float64, return_info=True)
    assert list(its) == list(its_ref) and list(ups) == list(ups_ref)
    assert rel(X.cpu().numpy(), X_ref) < 1e-12
    assert max(rr) < 1e-10


@pytest.mark.parametrize("flexible", [False, True])
def test_ru_mixed_reaches_tolerance(cuda, flexible):
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg, pcg_ru
    h = sa_hierarchy((64, 64))
    om = OracleMG(h, WITHRHO, WITHRHO)
    om32 = OracleMG(h, WITHRHO, WITHRHO, dtype=np.float32)
    A = om.levels[0].A
    B = np.random.default_rng(22).standard_normal((4096, 8))
    X64 = pcg(A, B, om.aspreconditioner(), tol=1e-14, maxiter=500)
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    X, its, ups, rr = ml32.pcg_ru(_t(B, cuda), tol=1e-10, maxiter=300, flexible=flexible, return_info=True)
    assert X.dtype == torch.float64
    X = X.cpu().numpy()
    res = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert res.max() < 1e-10 and max(rr) < 1e-10
    M32 = om32.aspreconditioner()
    _, its_ref, ups_ref = pcg_ru(A, B, lambda r: np.asarray(M32(np.asarray(r, np.float32)), np.float32), tol=1e-10,
                                 maxiter=300, dtype=np.float32, flexible=flexible, return_info=True)
    assert all(abs(a - b) <= 1 for a, b in zip(its, its_ref))
    err = (np.linalg.norm(X - X64, axis=0) / np.linalg.norm(X64, axis=0)).max()
    assert err < 1e-8
    print(f"\nreliable-update CG float32 flexible={flexible}: iterations {max(its)}, updates {max(ups)}, "
          f"err vs float64 {err:.1e}", flush=True)


def test_ru_maxiter_folds_the_last_correction(cuda):
    """maxiter = 3: three iterations, the correction folded into x, as float64 pcg."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    h = sa_hierarchy((64, 64))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    b = np.random.default_rng(23).standard_normal(4096)
    x, it, ups, _ = ml.pcg_ru(_t(b, cuda), tol=1e-12, maxiter=3, return_info=True)
    assert it == 3 and ups >= 1
    assert rel(x.cpu().numpy(), pcg(om.levels[0].A, b, om.aspreconditioner(), tol=0.0, maxiter=3)) < 1e-12
    x0, it0, ups0, _ = ml.pcg_ru(_t(b, cuda), tol=1e-8, maxiter=0, return_info=True)
    assert it0 == 0 and ups0 == 0 and np.all(x0.cpu().numpy() == 0)


@pytest.mark.parametrize("cycle,cpl", [("W", 1), ("F", 1), ("F", 2)])
def test_ru_w_and_f_cycles_match_oracle(cuda, cycle, cpl):
    """W- and F-cycle preconditioners: float64 iterates within 1e-12 of the oracle
    with the same iteration and update counts."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg_ru
    h = sa_hierarchy((48, 48))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(24).standard_normal((2304, 2))
    X, its, ups, rr = ml.pcg_ru(_t(B, cuda), tol=1e-10, delta=1e-2, maxiter=200, cycle=cycle,
                                cycles_per_level=cpl, return_info=True)
    M = lambda r: om.cycle(np.zeros_like(r), r, cycle, cpl)
    X_ref, its_ref, ups_ref = pcg_ru(om.levels[0].A, B, M, tol=1e-10, delta=1e-2, maxiter=200,
                                     dtype=np.float64, return_info=True)
    assert list(its) == list(its_ref) and list(ups) == list(ups_ref)
    assert rel(X.cpu().numpy(), X_ref) < 1e-12
    assert max(rr) < 1e-10


def test_ru_k64_block_with_zero_and_scaled_columns(cuda):
    """A k = 64 block (vector kernels) with a zero column and columns scaled over
    eight decades: every column stops on its own true residual; float64 iterates and
    counts equal the oracle's, the zero column returns 0 with no iteration."""
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg_ru
    h = sa_hierarchy((40, 40))
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    om = OracleMG(h, WITHRHO, WITHRHO)
    B = np.random.default_rng(25).standard_normal((1600, 64)) * np.logspace(-4, 4, 64)[None, :]
    B[:, 5] = 0.0
    X, its, ups, rr = ml.pcg_ru(_t(B, cuda), tol=1e-9, maxiter=200, return_info=True)
    X_ref, its_ref, ups_ref = pcg_ru(om.levels[0].A, B, om.aspreconditioner(), tol=1e-9, delta=1e-3,
                                     maxiter=200, dtype=np.float64, return_info=True)
    assert list(its) == list(its_ref) and list(ups) == list(ups_ref)
    assert its[5] == 0 and np.all(X.cpu().numpy()[:, 5] == 0)
    keep = [c for c in range(64) if c != 5]
    assert rel(X.cpu().numpy()[:, keep], X_ref[:, keep]) < 1e-12
    assert max(rr[c] for c in keep) < 1e-9


def test_ru_mixed_k64_block_with_zero_and_scaled_columns(cuda):
    """The float32 form at k = 64 (vector kernels: the update's lanes hold four
    float32 columns each, so columns whose updates fall in different iterations
    share a lane): columns scaled over eight decades and a zero column each stop
    on their own float64 residual, checked afresh as ||b - A x||, with iteration
    counts within one of the oracle's float32 restatement."""
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg_ru
    h = sa_hierarchy((40, 40))
    om32 = OracleMG(h, WITHRHO, WITHRHO, dtype=np.float32)
    A = OracleMG(h, WITHRHO, WITHRHO).levels[0].A
    B = np.random.default_rng(26).standard_normal((1600, 64)) * np.logspace(-4, 4, 64)[None, :]
    B[:, 5] = 0.0
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    X, its, ups, rr = ml32.pcg_ru(_t(B, cuda), tol=1e-9, maxiter=200, return_info=True)
    assert X.dtype == torch.float64
    X = X.cpu().numpy()
    assert its[5] == 0 and np.all(X[:, 5] == 0)
    keep = [c for c in range(64) if c != 5]
    res = np.linalg.norm(B - A @ X, axis=0)[keep] / np.linalg.norm(B, axis=0)[keep]
    assert res.max() < 1e-9 and max(rr[c] for c in keep) < 1e-9
    M32 = om32.aspreconditioner()
    _, its_ref, _ = pcg_ru(A, B[:, keep], lambda r: np.asarray(M32(np.asarray(r, np.float32)), np.float32),
                           tol=1e-9, maxiter=200, dtype=np.float32, return_info=True)
    assert all(abs(its[c] - i) <= 1 for c, i in zip(keep, its_ref))
