"""Writes tests/golden/poisson64.npz: the 64x64 Poisson configuration and oracle outputs.

Configuration: A = five-point Poisson 64x64 (N = 4096, nnz = 20,224), in-repo
smoothed-aggregation hierarchy (stored in the fixture so it stays fixed),
b = default_rng(0).standard_normal(4096), pinv coarse solve.
Outputs (oracle, float64):
  cyc{V,W,F}_withrho_{1,10}, cycV_default_{1,10}: iterate after that many cycles from 0
  solve_{V,W,F}: x and iteration count of solve(tol=1e-8, maxiter=100), withrho
  solve_default_V: same with the reference's default smoother (omega=1, no rho)
  pcg_V: x and iteration count of AMG-PCG (tol=1e-8, maxiter=500), withrho
Run: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})
DEFAULT = ("jacobi", {"iterations": 1, "omega": 1.0})


def main():
    from amjax_amd import gallery, hierarchy
    from oracle.solve import OracleMG, pcg

    h = hierarchy.smoothed_aggregation_solver(gallery.poisson((64, 64)))
    hp = os.path.join(HERE, "poisson64_hierarchy.npz")
    h.save(hp)
    h = hierarchy.Hierarchy.load(hp)
    b = np.random.default_rng(0).standard_normal(4096)
    out = {"b": b}
    for sname, spec in (("withrho", WITHRHO), ("default", DEFAULT)):
        om = OracleMG(h, spec, spec)
        for cyc in ("VWF" if sname == "withrho" else "V"):
            x = np.zeros(4096)
            for it in range(1, 11):
                x = om.cycle(x, b, cyc)
                if it in (1, 10):
                    out[f"cyc{cyc}_{sname}_{it}"] = x.copy()
    om = OracleMG(h, WITHRHO, WITHRHO)
    for cyc in "VWF":
        x, it = om.solve(b, tol=1e-8, maxiter=100, cycle=cyc, return_info=True)
        out[f"solve_{cyc}"] = x
        out[f"solve_{cyc}_iters"] = np.array(it)
    omd = OracleMG(h, DEFAULT, DEFAULT)
    x, it = omd.solve(b, tol=1e-8, maxiter=300, return_info=True)
    out["solve_default_V"] = x
    out["solve_default_V_iters"] = np.array(it)
    x, it = pcg(om.levels[0].A, b, om.aspreconditioner(), tol=1e-8, maxiter=500, return_info=True)
    out["pcg_V"] = x
    out["pcg_V_iters"] = np.array(it)
    np.savez_compressed(os.path.join(HERE, "poisson64.npz"), **out)
    print({k: (int(v) if v.ndim == 0 else v.shape) for k, v in out.items() if "iters" in k})


if __name__ == "__main__":
    main()
