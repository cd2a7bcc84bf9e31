"""One-launch coarse tail (AMG_OPT_TAIL).

From the first level l >= 1 with at most AMG_OPT_TAIL rows whose operator
image and level vectors fit the LDS of one workgroup (160 KiB on MI355X), a
cycle's visits of those levels run as one launch that computes the same row
sums and epilogues as the per-level launches: zero-guess V visits always, and
W, F and 'T' visits and nonzero-guess visits when the restrictions fit in the
image too.  Cycles and whole solves with the tail equal those without it bit
for bit (float32 and float64, k in {1, 3, 64}, V, W, F and 'T' cycles from a
zero and a nonzero guess, PCG and the reliable-update CG); the start level
follows the documented rule on hierarchies whose image sizes are known, and
configurations the tail does not cover report -1.
"""

import numpy as np
import pytest

from conftest import sa_hierarchy

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


def _ml(h, dtype, dev, tail=None, smoother=WITHRHO, coarse="pinv"):
    from amjax_amd import MultilevelSolver
    ml = MultilevelSolver.from_pyamg(h, smoother, smoother, coarse_solver=coarse, dtype=dtype, device=dev)
    ml.tail = tail
    return ml


def _cycle(ml, b, x, cycle="V", cpl=1, x_zero=True):
    from amjax_amd import _lib
    code = _lib.load().amg_cycle(ml._native(), cycle.encode(), int(cpl), b.data_ptr(), x.data_ptr(), b.shape[1],
                                 int(x_zero), _lib.stream_ptr(b.device))
    _lib.check(code, "amg_cycle")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tail_start_level(cuda, dtype):
    # 200 x 200: 40000, 6700, 765, 90, 12 and 2 unknowns.  The image from level
    # 1 takes ~1.2 MiB; from level 2 ~151 KiB in float32 (it fits) and ~232 KiB
    # in float64 (it does not); from level 3 ~21 / 32 KiB.
    h = sa_hierarchy((200, 200))
    assert _ml(h, dtype, cuda).tail_level() == (2 if dtype == np.float32 else 3)
    assert _ml(h, dtype, cuda, tail=100).tail_level() == 3
    assert _ml(h, dtype, cuda, tail=10).tail_level() == -1
    assert _ml(h, dtype, cuda, tail=0).tail_level() == -1
    assert _ml(h, dtype, cuda, smoother=("chebyshev", {})).tail_level() == -1
    assert _ml(h, dtype, cuda, coarse="lu").tail_level() == -1
    ml_nf = _ml(h, dtype, cuda)
    ml_nf.fuse = False
    assert ml_nf.tail_level() == -1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("k", [1, 3, 64])
@pytest.mark.parametrize("cycle,x_zero", [("V", True), ("V", False), ("W", True), ("W", False), ("F", True),
                                          ("F", False), ("T", True)])
def test_tail_cycle_bitwise(cuda, dtype, k, cycle, x_zero):
    import torch
    h = sa_hierarchy((200, 200))
    on, off = _ml(h, dtype, cuda), _ml(h, dtype, cuda, tail=0)
    assert on.tail_level() >= 1 and off.tail_level() == -1
    n = h.levels[0].A.shape[0]
    rng = np.random.default_rng(31 + k)
    b = torch.as_tensor(rng.standard_normal((n, k)), device=cuda, dtype=on.dtype)
    x0 = torch.as_tensor(rng.standard_normal((n, k)), device=cuda, dtype=on.dtype)
    outs = []
    for ml in (on, off):
        x = x0.clone()
        _cycle(ml, b, x, cycle, 2, x_zero)
        torch.cuda.synchronize()
        outs.append(x)
    assert torch.isfinite(outs[0]).all()
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("cpl", [0, 1, 3])
def test_tail_f_and_t_cycles_per_level_bitwise(cuda, dtype, cpl):
    """W, F and its transpose 'T' with 0, 1 and 3 V-cycles per level on a deep
    hierarchy (6 levels at 500 x 500: in float32 the tail from level 3 holds the
    restrictions; in float64 they do not fit there, and the W, F and nonzero
    visits take a second tail from level 4)."""
    import torch
    from amjax_amd import gallery, hierarchy
    h = hierarchy.smoothed_aggregation_solver(gallery.poisson((500, 500)))
    on, off = _ml(h, dtype, cuda), _ml(h, dtype, cuda, tail=0)
    assert on.tail_level() == 3
    n = h.levels[0].A.shape[0]
    b = torch.as_tensor(np.random.default_rng(35).standard_normal((n, 2)), device=cuda, dtype=on.dtype)
    for cycle in ("W", "F", "T"):
        outs = []
        for ml in (on, off):
            x = torch.zeros_like(b)
            _cycle(ml, b, x, cycle, cpl, True)
            torch.cuda.synchronize()
            outs.append(x)
        assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tail_pcg_bitwise(cuda, dtype):
    import torch
    h = sa_hierarchy((200, 200))
    on, off = _ml(h, dtype, cuda), _ml(h, dtype, cuda, tail=0)
    n = h.levels[0].A.shape[0]
    B = torch.as_tensor(np.random.default_rng(33).standard_normal((n, 4)), device=cuda)
    for cycle in ("V", "W", "F"):
        X1, it1, r1 = on.pcg(B, tol=1e-8, maxiter=200, cycle=cycle, return_info=True)
        X2, it2, r2 = off.pcg(B, tol=1e-8, maxiter=200, cycle=cycle, return_info=True)
        assert list(it1) == list(it2) and max(r1) < 1e-8
        assert torch.equal(X1, X2)
    x1, i1, r1 = on.solve(B, tol=1e-8, maxiter=100, cycle="W", return_info=True)
    x2, i2, r2 = off.solve(B, tol=1e-8, maxiter=100, cycle="W", return_info=True)
    assert list(i1) == list(i2) and max(r1) < 1e-8
    assert torch.equal(x1, x2)
    if dtype == np.float32:
        X1, it1, u1, r1 = on.pcg_ru(B, tol=1e-10, maxiter=200, return_info=True)
        X2, it2, u2, r2 = off.pcg_ru(B, tol=1e-10, maxiter=200, return_info=True)
        assert list(it1) == list(it2) and list(u1) == list(u2) and max(r1) < 1e-10
        assert torch.equal(X1, X2)


def test_tail_small_hierarchy(cuda):
    """64 x 64 (4096, 704, 80, 9 unknowns): in float32 the tail covers every
    level below 0 (image ~140 KiB); in float64 it starts at level 2."""
    import torch
    h = sa_hierarchy((64, 64))
    for dtype in (np.float32, np.float64):
        on, off = _ml(h, dtype, cuda), _ml(h, dtype, cuda, tail=0)
        assert on.tail_level() == (1 if dtype == np.float32 else 2)
        b = torch.as_tensor(np.random.default_rng(34).standard_normal((4096, 5)), device=cuda, dtype=on.dtype)
        xs = []
        for ml in (on, off):
            x = torch.zeros_like(b)
            _cycle(ml, b, x)
            torch.cuda.synchronize()
            xs.append(x)
        assert torch.equal(xs[0], xs[1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("problem", ["rs500", "sa1d"])
def test_tail_deep_hierarchies_bitwise(cuda, dtype, problem):
    """Deep hierarchies: Ruge-Stueben at 500 x 500 (9 levels, 250000 down to 10
    unknowns, classical interpolation) and SA on a 1-D Poisson problem of 8192
    unknowns (8 levels, 8192 down to 4).  On the 1-D one the tail spans the 5
    levels from level 2, so a zero-guess W-cycle visit needs 139 phases, more
    than one launch's program holds (96), and runs as per-level launches, while
    V visits take the tail; every cycle equals the per-level launches bit for
    bit."""
    import torch
    from amjax_amd import gallery, hierarchy
    if problem == "rs500":
        h = hierarchy.ruge_stuben_solver(gallery.poisson((500, 500)))
    else:
        h = hierarchy.smoothed_aggregation_solver(gallery.poisson((8192,)))
    on, off = _ml(h, dtype, cuda), _ml(h, dtype, cuda, tail=0)
    assert on.tail_level() >= 1 and off.tail_level() == -1
    print(problem, np.dtype(dtype).name, "levels", len(h.levels), "tail from", on.tail_level(), flush=True)
    n = h.levels[0].A.shape[0]
    rng = np.random.default_rng(36)
    b = torch.as_tensor(rng.standard_normal((n, 3)), device=cuda, dtype=on.dtype)
    x0 = torch.as_tensor(rng.standard_normal((n, 3)), device=cuda, dtype=on.dtype)
    for cycle, cpl, x_zero in (("V", 1, True), ("V", 1, False), ("W", 1, True), ("W", 1, False),
                               ("F", 2, True), ("T", 2, True)):
        outs = []
        for ml in (on, off):
            x = x0.clone()
            _cycle(ml, b, x, cycle, cpl, x_zero)
            torch.cuda.synchronize()
            outs.append(x)
        assert torch.isfinite(outs[0]).all(), cycle
        assert torch.equal(outs[0], outs[1]), (cycle, cpl, x_zero)
    B = b.double()
    X1, it1, r1 = on.pcg(B, tol=1e-8, maxiter=300, cycle="W", return_info=True)
    X2, it2, r2 = off.pcg(B, tol=1e-8, maxiter=300, cycle="W", return_info=True)
    assert list(it1) == list(it2) and max(r1) < 1e-8
    assert torch.equal(X1, X2)
