"""The uneven batch at full size: 40 independent SPD problems from seed 2.

Half 2-D five-point Poisson n x n with n ~ U[64, 768], half 3-D seven-point
Poisson n^3 with n ~ U[16, 80], each with its own smoothed-aggregation
hierarchy (withrho Jacobi, pinv coarse solve) and k ~ U{1, 4, 16, 64}
right-hand sides ~ N(0, 1), all drawn in order from default_rng(2)
(``bench.batch_problems``).  Float32 cycle inside float64 PCG to 1e-8.

Bars: the batch (plain and persistent grid) equals the per-problem loop of
``MultilevelSolver.pcg`` bit for bit, with the same iteration counts; every
column's true float64 relative residual is < 1e-8; the smallest problem
matches the CPU oracle (float64 single solve at 1e-10, and the mixed solve
against the oracle's float32-cycle PCG).
"""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

WITHRHO = ("jacobi", {"iterations": 1, "withrho": True})


@pytest.fixture(scope="module")
def batch40():
    import torch
    import bench
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    dev = torch.device("cuda", 0)
    probs, rng = bench.batch_problems()
    out = []
    for grid, k in probs:
        A = gallery.poisson(grid)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, WITHRHO, WITHRHO, dtype=np.float32, device=dev)
        ml.autotune = True
        B = rng.standard_normal((A.shape[0], k))
        out.append({"grid": grid, "k": k, "A": A, "h": h, "ml": ml, "B": B,
                    "Bt": torch.as_tensor(B, device=dev)})
    return out


@pytest.mark.parametrize("grid", [0, 2048])
def test_batch40_equals_loop_and_converges(cuda, batch40, grid):
    import torch
    from amjax_amd import BatchPCG
    assert len(batch40) == 40
    batch = BatchPCG([p["ml"] for p in batch40], [p["k"] for p in batch40], grid=grid)
    Xb, itb, rrb = batch.pcg([p["Bt"] for p in batch40], tol=1e-8, maxiter=500, return_info=True)
    worst, its_all = 0.0, []
    for p, xb, it_, rr_ in zip(batch40, Xb, itb, rrb):
        xs, its, rrs = p["ml"].pcg(p["Bt"], tol=1e-8, maxiter=500, return_info=True)
        assert torch.equal(xb, xs), p["grid"]
        assert list(it_) == list(its) and list(rr_) == list(rrs)
        X = xb.cpu().numpy()
        res = np.linalg.norm(p["B"] - p["A"] @ X, axis=0) / np.linalg.norm(p["B"], axis=0)
        assert res.max() < 1e-8, (p["grid"], res.max())
        worst = max(worst, float(res.max()))
        its_all += list(it_)
    nl, nb = batch.info()
    print(f"\nbatch40 grid {grid}: iterations {min(its_all)}..{max(its_all)}, worst relres {worst:.2e}, "
          f"{nl} launches / iteration", flush=True)


def test_batch40_smallest_problem_against_oracle(cuda, batch40):
    import torch
    from amjax_amd import MultilevelSolver
    from oracle.solve import OracleMG, pcg
    p = min(batch40, key=lambda q: q["A"].shape[0])
    b = p["B"][:, 0]
    bt = torch.as_tensor(np.ascontiguousarray(b), device=cuda)
    om64 = OracleMG(p["h"], WITHRHO, WITHRHO)
    A = om64.levels[0].A
    ml64 = MultilevelSolver.from_pyamg(p["h"], WITHRHO, WITHRHO, device=cuda)
    x, it, _ = ml64.pcg(bt, tol=1e-10, maxiter=500, return_info=True)
    x_ref, it_ref = pcg(A, b, om64.aspreconditioner(), tol=1e-10, maxiter=500, return_info=True)
    assert it == it_ref and rel(x.cpu().numpy(), x_ref) < 1e-10
    # the mixed form: float32 cycle (oracle in float32) inside float64 CG
    om32 = OracleMG(p["h"], WITHRHO, WITHRHO, dtype=np.float32)
    M32 = om32.aspreconditioner()

    def M(r):
        return np.asarray(M32(r.astype(np.float32)), dtype=np.float64)

    xm_ref, itm_ref = pcg(A, b, M, tol=1e-8, maxiter=500, return_info=True)
    xm, itm, _ = p["ml"].pcg(bt, tol=1e-8, maxiter=500, return_info=True)
    assert abs(itm - itm_ref) <= 1
    assert rel(xm.cpu().numpy(), xm_ref) < 1e-6
