"""Batched AMG-PCG over independent problems (one launch per phase for all).

Each problem of a batch computes what MultilevelSolver.pcg computes for it
alone, bit for bit: the batched kernels run the same device bodies on the same
per-problem work decomposition.  The single solves are themselves pinned to
the CPU oracle; one problem of the batch is also checked against the oracle
directly.  Problem sizes are chosen so that row counts are not multiples of
64, hierarchy depths differ and k covers 1, 3 (padded column lanes), 16 and 64.
"""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})
GRIDS = [(37, 29), (20, 20, 20), (70, 53), (13, 11, 9), (90, 90)]
KS = [1, 3, 16, 64, 4]


def _problems(dtype, device):
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    out = []
    for i, (g, k) in enumerate(zip(GRIDS, KS)):
        A = gallery.poisson(g)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=dtype, device=device)
        B = np.random.default_rng(100 + i).standard_normal((A.shape[0], k))
        out.append((A, h, ml, B))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("grid,split", [(0, True), (0, False), (64, True)])
def test_batch_equals_single_solves_bitwise(cuda, dtype, grid, split):
    import torch
    from amjax_amd import BatchPCG
    probs = _problems(dtype, cuda)
    Bs = [torch.as_tensor(B, device=cuda) for _, _, _, B in probs]
    batch = BatchPCG([ml for _, _, ml, _ in probs], KS, grid=grid, split_kp=split)
    Xb, itb, rrb = batch.pcg(Bs, tol=1e-8, maxiter=200, return_info=True)
    nl, nb = batch.info()
    print(f"\nbatch {np.dtype(dtype).name} grid {grid} split {split}: {nl} launches / iteration, "
          f"{nb} virtual blocks", flush=True)
    for (A, _, ml, B), b, xb, it_, rr_ in zip(probs, Bs, Xb, itb, rrb):
        xs, its, rrs = ml.pcg(b, tol=1e-8, maxiter=200, return_info=True)
        assert torch.equal(xb, xs)
        assert it_ == its and rr_ == rrs
        X = xb.cpu().numpy()
        res = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() < 1e-8 * 1.01


def test_batch_against_oracle(cuda):
    import torch
    from amjax_amd import BatchPCG
    from oracle.solve import OracleMG, pcg
    probs = _problems(np.float64, cuda)
    batch = BatchPCG([ml for _, _, ml, _ in probs], KS)
    Xb, itb, _ = batch.pcg([torch.as_tensor(B, device=cuda) for *_, B in probs], tol=1e-10, maxiter=200,
                           return_info=True)
    for (A, h, ml, B), xb, it_ in zip(probs, Xb, itb):
        om = OracleMG(h, WITHRHO, WITHRHO)
        x0 = pcg(om.levels[0].A, B[:, 0], om.aspreconditioner(), tol=1e-10, maxiter=200)
        assert rel(xb.cpu().numpy()[:, 0], x0) < 1e-9


def test_batch_repeated_calls_and_maxiter(cuda):
    """Per-call inputs, parameters and the completion count are reset every call."""
    import torch
    from amjax_amd import BatchPCG
    probs = _problems(np.float32, cuda)
    batch = BatchPCG([ml for _, _, ml, _ in probs], KS)
    B1 = [torch.as_tensor(B, device=cuda) for *_, B in probs]
    B2 = [2.0 * b.flip(0) for b in B1]
    X1 = batch.pcg(B1, tol=1e-8, maxiter=200)
    _, it3, _ = batch.pcg(B2, tol=1e-8, maxiter=3, return_info=True)
    assert all(i == 3 for its in it3 for i in its)
    X2 = batch.pcg(B2, tol=1e-8, maxiter=200)
    X1b = batch.pcg(B1, tol=1e-8, maxiter=200)
    for (_, _, ml, _), b1, b2, x1, x2, x1b in zip(probs, B1, B2, X1, X2, X1b):
        assert torch.equal(x1, x1b)
        assert torch.equal(x2, ml.pcg(b2, tol=1e-8, maxiter=200))


def test_batch_rejects_unsupported(cuda):
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    A = gallery.poisson((30, 30))
    h = hierarchy.smoothed_aggregation_solver(A)
    ml64 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    ml32 = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
    with pytest.raises(ValueError, match="dtype"):
        BatchPCG([ml64, ml32], [1, 1])
    with pytest.raises(ValueError):
        BatchPCG([ml64], [1]).pcg([np.zeros((900, 2))])


def test_autotuned_plans(cuda):
    """Tuned launch shapes change only the work split: k >= 2 results are bitwise
    those of the fixed shape, k = 1 agrees to rounding; batch and single solves
    with tuning on stay bitwise equal."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy, _lib
    probs = []
    for i, (g, k) in enumerate([((120, 120), 1), ((30, 30, 30), 16), ((150, 110), 64), ((26, 24, 22), 4)]):
        A = gallery.poisson(g)
        h = hierarchy.smoothed_aggregation_solver(A)
        ref = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
        tuned = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=cuda)
        tuned.autotune = True
        B = torch.as_tensor(np.random.default_rng(7 + i).standard_normal((A.shape[0], k)), device=cuda)
        probs.append((A, ref, tuned, B, k))
    for A, ref, tuned, B, k in probs:
        x0 = ref.pcg(B, tol=1e-10, maxiter=200)
        x1 = tuned.pcg(B, tol=1e-10, maxiter=200)
        if k >= 2:
            assert torch.equal(x0, x1)
        else:
            assert float((x0 - x1).norm() / x0.norm()) < 1e-9
    rep = _lib.tune_report()
    print("\nautotuner cache:", *rep, sep="\n  ", flush=True)
    assert rep, "no matrix was tuned"
    batch = BatchPCG([p[2] for p in probs], [p[4] for p in probs])
    Xb = batch.pcg([p[3] for p in probs], tol=1e-10, maxiter=200)
    for (A, ref, tuned, B, k), xb in zip(probs, Xb):
        assert torch.equal(xb, tuned.pcg(B, tol=1e-10, maxiter=200))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_polynomial_smoothers(cuda, dtype):
    """Chebyshev and Richardson smoothing (zero-guess first terms, multi-term sweeps,
    one- and three-coefficient polynomials) in a batch beside Jacobi problems: every
    problem equals its single solve bit for bit, and one matches the oracle."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    from oracle.solve import OracleMG, pcg
    specs = [("chebyshev", {}), ("richardson", {"iterations": 2}), WITHRHO, ("chebyshev", {"degree": 2})]
    grids = [(37, 29), (13, 11, 9), (70, 53), (20, 20, 20)]
    ks = [3, 16, 1, 64]
    probs = []
    for i, (g, k, sm) in enumerate(zip(grids, ks, specs)):
        A = gallery.poisson(g)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, sm, sm, dtype=dtype, device=cuda)
        B = np.random.default_rng(300 + i).standard_normal((A.shape[0], k))
        probs.append((A, h, ml, B, sm))
    Bs = [torch.as_tensor(B, device=cuda) for _, _, _, B, _ in probs]
    batch = BatchPCG([p[2] for p in probs], ks)
    Xb, itb, _ = batch.pcg(Bs, tol=1e-8, maxiter=200, return_info=True)
    for (A, h, ml, B, sm), b, xb, it_ in zip(probs, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=200, return_info=True)
        assert torch.equal(xb, xs) and it_ == its
    if dtype == np.float64:
        A, h, ml, B, sm = probs[0]
        om = OracleMG(h, sm, sm)
        Xr = pcg(A, B, om.aspreconditioner(), tol=1e-8, maxiter=200)
        assert rel(Xb[0].cpu().numpy(), Xr) < 1e-10


def test_batch_rebuilds_after_plan_change(cuda):
    """A solver whose native plan changes after the batch was built (new smoothers,
    fuse toggled) gets a new handle, possibly at the old address: the batch must
    rebuild and equal the new single solves bit for bit."""
    import torch
    from amjax_amd import BatchPCG, change_smoothers
    probs = _problems(np.float32, cuda)[:3]
    mls = [ml for _, _, ml, _ in probs]
    Bs = [torch.as_tensor(B, device=cuda) for *_, B in probs]
    batch = BatchPCG(mls, KS[:3])
    batch.pcg(Bs, tol=1e-8, maxiter=200)
    change_smoothers(mls[0], ("jacobi", {"iterations": 2, "omega": 0.6}), ("jacobi", {"iterations": 1, "omega": 0.7}))
    mls[1].fuse = False
    Xb, itb, _ = batch.pcg(Bs, tol=1e-8, maxiter=200, return_info=True)
    for ml, b, xb, it_ in zip(mls, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=200, return_info=True)
        assert torch.equal(xb, xs) and list(it_) == list(its)


def test_batch_default_maxiter_is_per_problem(cuda):
    """maxiter=None gives every problem its own default (10 n_p), as a single pcg."""
    import torch
    from amjax_amd import BatchPCG
    probs = _problems(np.float64, cuda)[:2]
    Bs = [torch.as_tensor(B, device=cuda) for *_, B in probs]
    Xb, itb, _ = BatchPCG([ml for _, _, ml, _ in probs], KS[:2]).pcg(Bs, tol=1e-8, return_info=True)
    for (_, _, ml, _), b, xb, it_ in zip(probs, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, return_info=True)
        assert torch.equal(xb, xs) and list(it_) == list(its)


@pytest.mark.parametrize("cycle,cpl", [("W", 1), ("F", 0), ("F", 1), ("F", 2)])
def test_batch_w_and_f_cycles(cuda, cycle, cpl):
    """W and F cycles in a batch of problems of different depth and k: every
    problem equals its single pcg(cycle=...) solve bit for bit."""
    import torch
    from amjax_amd import BatchPCG
    probs = _problems(np.float32, cuda)
    Bs = [torch.as_tensor(B, device=cuda) for _, _, _, B in probs]
    batch = BatchPCG([ml for _, _, ml, _ in probs], KS, cycle=cycle, cycles_per_level=cpl)
    Xb, itb, _ = batch.pcg(Bs, tol=1e-8, maxiter=200, return_info=True)
    for (A, _, ml, B), b, xb, it_ in zip(probs, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=200, cycle=cycle, cycles_per_level=cpl, return_info=True)
        assert torch.equal(xb, xs) and it_ == its


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_lu_and_qr_coarse_solvers(cuda, dtype):
    """LU, QR (float64 factors) and Jacobi coarse solves in a batch beside a pinv
    problem: every problem equals its single solve bit for bit."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    grids, ks, coarse = [(37, 29), (13, 11, 9), (70, 53), (20, 20)], [3, 16, 1, 4], ["lu", "qr", "pinv", "jacobi"]
    probs = []
    for i, (g, k, cs) in enumerate(zip(grids, ks, coarse)):
        A = gallery.poisson(g)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, coarse_solver=cs, dtype=dtype, device=cuda)
        probs.append((ml, np.random.default_rng(400 + i).standard_normal((A.shape[0], k))))
    Bs = [torch.as_tensor(B, device=cuda) for _, B in probs]
    Xb, itb, _ = BatchPCG([ml for ml, _ in probs], ks).pcg(Bs, tol=1e-8, maxiter=200, return_info=True)
    for (ml, _), b, xb, it_ in zip(probs, Bs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=200, return_info=True)
        assert torch.equal(xb, xs) and it_ == its


@pytest.mark.parametrize("cycle", ["V", "W"])
@pytest.mark.parametrize("iterations", [0, 1, 3])
def test_batch_jacobi_coarse_solver(cuda, iterations, cycle):
    """The Jacobi coarse solver in a batch with odd, one and zero sweeps (zero: the
    coarse correction stays at its zero guess), V and W cycles (the coarse solve
    runs several times at one position): bitwise equal to the single solves."""
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    probs = []
    for i, (g, k) in enumerate([((20, 20), 4), ((13, 11, 9), 1), ((37, 29), 16)]):
        A = gallery.poisson(g)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, coarse_solver="jacobi",
                                         coarse_solver_kwargs={"iterations": iterations}, dtype=np.float32,
                                         device=cuda)
        probs.append((ml, k, torch.as_tensor(np.random.default_rng(500 + i).standard_normal((A.shape[0], k)),
                                             device=cuda)))
    batch = BatchPCG([p[0] for p in probs], [p[1] for p in probs], cycle=cycle)
    Xb, itb, _ = batch.pcg([p[2] for p in probs], tol=1e-8, maxiter=300, return_info=True)
    for (ml, _, b), xb, it_ in zip(probs, Xb, itb):
        xs, its, _ = ml.pcg(b, tol=1e-8, maxiter=300, cycle=cycle, return_info=True)
        assert torch.equal(xb, xs) and list(it_) == list(its)
