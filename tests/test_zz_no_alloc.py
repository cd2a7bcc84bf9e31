"""The allocation-free cycle amg_cycle_noalloc against amg_cycle, through the C-ABI.

Both entry points run the same kernels; the allocation-free one carves its
level vectors from the library's per-GPU scratch array, which keeps the last
call's contents.  Each case calls it twice in a row, the second time on a
smaller input (another hierarchy, fewer columns), and checks each output
against amg_cycle's bit for bit; requests beyond the scratch size fail with
AMG_E_ARG.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})
CHEB = ("chebyshev", {})


def _solver(grid, dtype, smoother, coarse, dev):
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    h = hierarchy.smoothed_aggregation_solver(gallery.poisson(grid))
    return MultilevelSolver.from_pyamg(h, smoother, smoother, coarse_solver=coarse, dtype=dtype, device=dev)


def _cycle(fn, ml, cycle, cpl, b, x, x_zero):
    from amjax_amd import _lib
    code = getattr(_lib.load(), fn)(ml._native(), cycle.encode(), int(cpl), b.data_ptr(), x.data_ptr(),
                                    b.shape[1], int(x_zero), _lib.stream_ptr(b.device))
    _lib.check(code, fn)


def _pair(ml, k, cycle, cpl, x_zero, seed):
    """(amg_cycle output, amg_cycle_noalloc output) for one random right-hand side block."""
    import torch
    n = ml.levels[0].A.shape[0]
    rng = np.random.default_rng(seed)
    b = torch.as_tensor(rng.standard_normal((n, k)), device=ml.device, dtype=ml.dtype)
    x0 = torch.as_tensor(rng.standard_normal((n, k)), device=ml.device, dtype=ml.dtype)
    ref, out = x0.clone(), x0.clone()
    _cycle("amg_cycle", ml, cycle, cpl, b, ref, x_zero)
    _cycle("amg_cycle_noalloc", ml, cycle, cpl, b, out, x_zero)
    torch.cuda.synchronize()
    return ref, out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("cycle,cpl,x_zero", [("V", 1, 1), ("W", 1, 0), ("F", 2, 1)])
def test_noalloc_cycle_equals_cycle_twice(cuda, dtype, cycle, cpl, x_zero):
    import torch
    big = _solver((96, 80), dtype, WITHRHO, "pinv", cuda)
    small = _solver((40, 30), dtype, WITHRHO, "pinv", cuda)
    for ml, k, seed in ((big, 8, 1), (small, 1, 2)):
        ref, out = _pair(ml, k, cycle, cpl, x_zero, seed)
        assert torch.equal(ref, out), (ml.levels[0].A.shape, k)
        assert torch.isfinite(out).all()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_polynomial_smoother_and_coarse_solvers(cuda, dtype):
    """Polynomial-smoother levels (two more vectors each), LU with float64 factors
    (the coarse work vector) and the Jacobi coarse solver, then a smaller call."""
    import torch
    cases = [((64, 64), CHEB, "lu", 4), ((50, 45), WITHRHO, "jacobi", 3), ((20, 20), WITHRHO, "qr", 1)]
    for i, (grid, sm, coarse, k) in enumerate(cases):
        ml = _solver(grid, dtype, sm, coarse, cuda)
        ref, out = _pair(ml, k, "V", 1, i % 2, 10 + i)
        assert torch.equal(ref, out), (grid, coarse)


def test_noalloc_rejects_requests_beyond_the_scratch(cuda):
    import torch
    from amjax_amd import _lib
    ml = _solver((96, 96), np.float64, WITHRHO, "pinv", cuda)
    n = ml.levels[0].A.shape[0]
    k = 200  # about 0.5 MB of vectors per column: beyond 64 MiB
    b = torch.zeros((n, k), dtype=torch.float64, device=cuda)
    x = torch.zeros_like(b)
    code = _lib.load().amg_cycle_noalloc(ml._native(), b"V", 1, b.data_ptr(), x.data_ptr(), k, 1,
                                         _lib.stream_ptr(cuda))
    assert code == -1
    assert "limit" in _lib.last_error()
    assert _lib.AMG_NOALLOC_SCRATCH_MB == 64
    # the library stays usable
    ref, out = _pair(ml, 2, "V", 1, 1, 3)
    assert torch.equal(ref, out)
