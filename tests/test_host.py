"""CPU tests of the host layer: C-ABI exports, packing, hierarchy builders,
smoother configuration and error conventions (no GPU needed)."""

import ctypes
import os
import re
import types

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT, sa_hierarchy

from amjax_amd import gallery, hierarchy
from amjax_amd._pack import sell_arrays
from amjax_amd.relaxation import smoothing


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "amjax_amd.h")).read()
    return sorted(set(re.findall(r"^int\s+(amg_\w+)\s*\(", text, flags=re.M)))


def test_abi_library_exports_every_declared_symbol():
    from amjax_amd import _lib
    lib = _lib.load()
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    assert sorted(_lib.EXPORTS) == declared
    for name in declared:
        assert hasattr(lib, name), name
    # error path without a device: last_error is callable and returns a string
    assert isinstance(_lib.last_error(), str)


def test_abi_rejects_bad_arguments_without_gpu():
    from amjax_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.amg_create(None, 0, None, 0, None, ctypes.byref(h)) == -1
    assert "null" in _lib.last_error()
    assert lib.amg_destroy(None) == -1
    assert lib.amg_dense_solve(None, 0, None, None, 1, None) == -1
    b = ctypes.c_void_p()
    assert lib.amg_batch_create(None, None, 0, ctypes.byref(b)) == -1
    assert lib.amg_batch_pcg(None, None, None, 1e-8, 0.0, 10, 0, None, None, None) == -1
    assert lib.amg_batch_destroy(None) == -1


def test_setup_library_loads():
    from amjax_amd.hierarchy import _lib as setup_lib
    lib = setup_lib()
    for name in ("amg_setup_standard_aggregation", "amg_setup_gauss_seidel", "amg_setup_rs_cf_splitting"):
        assert hasattr(lib, name)


def _emulate_sell(sptr, col, val, n, x):
    y = np.zeros((n,) + x.shape[1:])
    for row in range(n):
        s = row >> 6
        for p in range(sptr[s] + (row & 63), sptr[s + 1], 64):
            if col[p] >= 0:
                y[row] += val[p] * x[col[p]]
    return y


@pytest.mark.parametrize("shape,density", [((300, 257), 0.03), ((64, 64), 0.2), ((1, 5), 1.0), ((70, 3), 0.5)])
def test_sell_packing_roundtrip(shape, density):
    M = sp.random(*shape, density=density, random_state=3, format="csr")
    sptr, col, val, nnz = sell_arrays(M)
    assert nnz == M.nnz
    assert sptr[0] == 0 and sptr[-1] == col.size and np.all(np.diff(sptr) % 64 == 0)
    assert np.all((col >= -1) & (col < shape[1]))
    assert np.all(val[col < 0] == 0)
    x = np.random.default_rng(0).standard_normal((shape[1], 2))
    np.testing.assert_allclose(_emulate_sell(sptr, col, val, shape[0], x), M @ x, rtol=1e-13, atol=1e-13)


def test_poisson_matches_stencil():
    A = gallery.poisson((500, 500))
    assert A.shape == (250000, 250000) and A.nnz == 1248000
    A3 = gallery.poisson((100, 100, 100))
    assert A3.nnz == 6940000 and A3.diagonal().max() == 6
    A1 = gallery.poisson((4,)).toarray()
    np.testing.assert_array_equal(A1, [[2, -1, 0, 0], [-1, 2, -1, 0], [0, -1, 2, -1], [0, 0, -1, 2]])
    A2 = gallery.poisson((3, 4)).toarray()
    # lexicographic ordering, last index fastest: (i, j) -> 4 i + j
    assert A2[5, 4] == -1 and A2[5, 6] == -1 and A2[5, 1] == -1 and A2[5, 9] == -1 and A2[5, 5] == 4


def test_sa_hierarchy_properties():
    h = sa_hierarchy((64, 64))
    assert len(h.levels) >= 3 and h.levels[-1].A.shape[0] <= 10
    for i, l in enumerate(h.levels[:-1]):
        n, nc = l.A.shape[0], h.levels[i + 1].A.shape[0]
        assert l.P.shape == (n, nc) and l.R.shape == (nc, n)
        assert abs(l.R - l.P.T).max() == 0
        Ac = h.levels[i + 1].A
        assert abs(Ac - Ac.T).max() < 1e-12 * abs(Ac).max()
        assert np.all(np.linalg.eigvalsh(Ac.toarray()) > 0) if nc < 800 else True


def test_standard_aggregation_covers_connected_nodes():
    A = gallery.poisson((30, 30))
    AggOp, roots = hierarchy.standard_aggregation(hierarchy.symmetric_strength(A))
    assert AggOp.shape[0] == 900
    assert np.all(np.diff(AggOp.indptr) == 1)  # every node in exactly one aggregate
    assert np.all(np.asarray(AggOp.sum(axis=0)).ravel() >= 1)
    agg = AggOp.indices
    assert np.all(agg[roots] == np.arange(AggOp.shape[1]))


def test_rs_splitting_and_interpolation():
    A = gallery.poisson((20, 20))
    h = hierarchy.ruge_stuben_solver(A)
    assert len(h.levels) >= 3
    P = h.levels[0].P
    # interpolation preserves constants away from the boundary rows: rows sum to <= 1
    rs = np.asarray(P.sum(axis=1)).ravel()
    assert np.all(rs <= 1 + 1e-12) and np.all(rs > 0)


def test_hierarchy_save_load_roundtrip(tmp_path):
    h = sa_hierarchy((16, 16))
    p = tmp_path / "h.npz"
    h.save(p)
    g = hierarchy.Hierarchy.load(p)
    assert len(g.levels) == len(h.levels)
    for a, b in zip(h.levels, g.levels):
        assert abs(a.A - b.A).max() == 0
        if hasattr(a, "P"):
            assert abs(a.P - b.P).max() == 0 and abs(a.R - b.R).max() == 0


def _host_ml(grid=(20, 20)):
    """A MultilevelSolver-shaped object with host levels, for smoother configuration."""
    from oracle.solve import inverse_diagonal
    h = sa_hierarchy(grid)
    levels = []
    for l in h.levels:
        lv = types.SimpleNamespace(A=l.A, Dinv=inverse_diagonal(l.A))
        levels.append(lv)
    return types.SimpleNamespace(levels=levels, symmetric_smoothing=False)


def test_change_smoothers_symmetry_flag_and_baking():
    """tests/relaxation/test_smoothing.py:40-62 (flag), smoothing.py:59-82 (baking)."""
    for method in [("jacobi", {"iterations": 1, "withrho": True}), ("jacobi", {"iterations": 2, "withrho": True}),
                   None]:
        ml = _host_ml()
        smoothing.change_smoothers(ml, method, method)
        assert ml.symmetric_smoothing
    for pre, post in [(("jacobi", {"iterations": 2}), ("jacobi", {"iterations": 1})), ("jacobi", None),
                      (None, "jacobi")]:
        ml = _host_ml()
        smoothing.change_smoothers(ml, pre, post)
        assert not ml.symmetric_smoothing
    ml = _host_ml()
    smoothing.change_smoothers(ml, ("jacobi", {"withrho": True}), ("jacobi", {"withrho": True}))
    lv = ml.levels[0]
    name, kw = lv._presmoother_spec
    assert name == "jacobi" and kw["withrho"] is False
    assert kw["omega"] == pytest.approx(1.0 / lv._spectral_radius)
    assert lv.presmoother._omega == lv.postsmoother._omega
    # per-level list, last entry repeated
    ml = _host_ml()
    smoothing.change_smoothers(ml, [("jacobi", {"iterations": 3}), None], "jacobi")
    assert ml.levels[0].presmoother._iterations == 3
    assert all(l.presmoother.__name__ == "none" for l in ml.levels[1:-1])


def test_smoother_errors():
    ml = _host_ml()
    for name in smoothing._NOT_IMPLEMENTED_SMOOTHERS:
        with pytest.raises(NotImplementedError):
            smoothing.change_smoothers(ml, (name, {}), "jacobi")
    with pytest.raises(ValueError):
        smoothing.change_smoothers(ml, "bogus", "jacobi")
    with pytest.raises(ValueError):
        smoothing.change_smoothers(ml, 3, "jacobi")
    with pytest.raises(ValueError):
        smoothing.change_smoothers(ml, [], "jacobi")
    lv = types.SimpleNamespace()
    with pytest.raises(AttributeError):
        smoothing.rebuild_smoother(lv)


def test_from_pyamg_validates_before_upload():
    """Errors are raised at configuration time, before any device work
    (tests/test_multilevel.py:335-368)."""
    from amjax_amd import MultilevelSolver
    from amjax_amd.multilevel import _NOT_IMPLEMENTED_SOLVERS, coarse_grid_solver
    h = sa_hierarchy((10, 10))
    for name in _NOT_IMPLEMENTED_SOLVERS:
        with pytest.raises(NotImplementedError):
            MultilevelSolver.from_pyamg(h, coarse_solver=name)
    for name in smoothing._NOT_IMPLEMENTED_SMOOTHERS:
        with pytest.raises(NotImplementedError):
            MultilevelSolver.from_pyamg(h, presmoother=(name, {}))
    with pytest.raises(ValueError):
        coarse_grid_solver("bad_solver", MultilevelSolver.Level())


def test_cycle_complexity_of_solver_object():
    """MultilevelSolver.cycle_complexity on the reference's hierarchy (100, 25, 9, 4 nnz)."""
    from amjax_amd import MultilevelSolver

    class FakeA:
        def __init__(self, n):
            self.nnz = n * n
            self.shape = (n, n)

    levels = []
    for n in (10, 5, 3, 2):
        lv = MultilevelSolver.Level()
        lv.A = FakeA(n)
        lv.P = object()
        lv.R = object()
        levels.append(lv)
    dummy = lambda A, x, b: x
    assert MultilevelSolver(levels[:1], dummy).cycle_complexity("V") == 1.0
    assert MultilevelSolver(levels[:2], dummy).cycle_complexity("W") == 2.25
    mg = MultilevelSolver(levels[:3], dummy)
    assert (mg.cycle_complexity("V"), mg.cycle_complexity("W"), mg.cycle_complexity("F")) == (2.59, 3.18, 3.18)
    mg = MultilevelSolver(levels, dummy)
    assert (mg.cycle_complexity("V"), mg.cycle_complexity("W"), mg.cycle_complexity("F")) == (2.72, 3.88, 3.66)
    with pytest.raises(NotImplementedError):
        mg.cycle_complexity("AMLI")
    assert mg.operator_complexity() == pytest.approx(138 / 100)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says why (no GPU needed)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
