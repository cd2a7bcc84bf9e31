"""Committed golden vectors of the 64x64 configuration (tests/golden/make_golden.py).

CPU: the oracle reproduces them (regression of the checker itself).
GPU: the HIP path matches them to the float64 parity bar (1e-12 relative,
identical iteration counts).
"""

import os

import numpy as np
import pytest

from conftest import rel

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})
DEFAULT = ("jacobi", {"iterations": 1, "omega": 1.0})


@pytest.fixture(scope="module")
def golden():
    from amjax_amd import hierarchy
    g = dict(np.load(os.path.join(HERE, "poisson64.npz"), allow_pickle=False))
    h = hierarchy.Hierarchy.load(os.path.join(HERE, "poisson64_hierarchy.npz"))
    return g, h


def test_fixture_config(golden):
    g, h = golden
    assert h.levels[0].A.shape == (4096, 4096) and h.levels[0].A.nnz == 20224
    np.testing.assert_array_equal(g["b"], np.random.default_rng(0).standard_normal(4096))


def test_oracle_reproduces_golden(golden):
    from oracle.solve import OracleMG, pcg
    g, h = golden
    b = g["b"]
    for sname, spec in (("withrho", WITHRHO), ("default", DEFAULT)):
        om = OracleMG(h, spec, spec)
        for cyc in ("VWF" if sname == "withrho" else "V"):
            x = np.zeros(4096)
            for it in range(1, 11):
                x = om.cycle(x, b, cyc)
                if it in (1, 10):
                    assert rel(x, g[f"cyc{cyc}_{sname}_{it}"]) < 1e-14
    om = OracleMG(h, WITHRHO, WITHRHO)
    for cyc in "VWF":
        x, it = om.solve(b, tol=1e-8, maxiter=100, cycle=cyc, return_info=True)
        assert it == int(g[f"solve_{cyc}_iters"]) and rel(x, g[f"solve_{cyc}"]) < 1e-14
    x, it = pcg(om.levels[0].A, b, om.aspreconditioner(), tol=1e-8, maxiter=500, return_info=True)
    assert it == int(g["pcg_V_iters"]) and rel(x, g["pcg_V"]) < 1e-14


@pytest.mark.gpu
def test_gpu_matches_golden(golden, cuda):
    import torch
    from amjax_amd import MultilevelSolver
    g, h = golden
    bt = torch.as_tensor(g["b"], device=cuda)
    for sname, spec in (("withrho", WITHRHO), ("default", DEFAULT)):
        ml = MultilevelSolver.from_pyamg(h, spec, spec, device=cuda)
        for cyc in ("VWF" if sname == "withrho" else "V"):
            x = torch.zeros(4096, dtype=torch.float64, device=cuda)
            for it in range(1, 11):
                x = ml._cycle(x, bt, cyc)
                if it in (1, 10):
                    assert rel(x.cpu().numpy(), g[f"cyc{cyc}_{sname}_{it}"]) < 1e-12
    ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, device=cuda)
    for cyc in "VWF":
        x, it, _ = ml.solve(bt, tol=1e-8, maxiter=100, cycle=cyc, return_info=True)
        assert it == int(g[f"solve_{cyc}_iters"]) and rel(x.cpu().numpy(), g[f"solve_{cyc}"]) < 1e-12
    mld = MultilevelSolver.from_pyamg(h, DEFAULT, DEFAULT, device=cuda)
    x, it, _ = mld.solve(bt, tol=1e-8, maxiter=300, return_info=True)
    assert it == int(g["solve_default_V_iters"]) and rel(x.cpu().numpy(), g["solve_default_V"]) < 1e-12
    x, it, _ = ml.pcg(bt, tol=1e-8, maxiter=500, return_info=True)
    assert it == int(g["pcg_V_iters"]) and rel(x.cpu().numpy(), g["pcg_V"]) < 1e-12
