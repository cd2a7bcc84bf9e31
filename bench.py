#!/usr/bin/env python
"""Headline benchmark: ms to 1e-8 relative FP64 residual on 2-D Poisson 500x500.

Workload (one "step" = one complete solve of all k right-hand sides):
  A = five-point Poisson 500x500 (N = 250,000, nnz = 1,248,000), float64;
  smoothed-aggregation hierarchy (in-repo builder, PyAMG defaults), V-cycle,
  1 pre + 1 post damped Jacobi with omega = 1/rho(D^-1 A), pinv coarse solve;
  B ~ N(0, 1) from numpy default_rng(0), shape (N, k), k = 64, resident in HBM;
  AMG-preconditioned CG to ||b - A x|| <= 1e-8 ||b|| (true float64 residual) per
  column, mixed precision: CG with reliable updates (float32 cycle and CG
  vectors, float64 dot products, the float64 iterate and true residual
  refreshed whenever the recurrence residual has fallen 1e3-fold).  Also
  reported: the plain float64 path, the float32 cycle inside float64-vector CG,
  k = 1, the standalone V-cycle solve, float64 iterative refinement around a
  float32-vector CG ("ircg", inner tolerance 1e-4), and every mixed result's
  error against a float64 solve converged to 1e-14.

Prints one JSON line.  Single GPU only.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _ir_info(out):
    """(x, inner iterations, true relative residuals) of a pcg_ir return_info result."""
    x, its, _, rr = out
    return x, its, rr


def _ru_info(out):
    """(x, CG iterations, true relative residuals) of a pcg_ru return_info result."""
    x, its, _, rr = out
    return x, its, rr


def reference_rhs_counts(h, dev, tol, log):
    """The reference benchmark's right-hand side (benchmarks/benchmark.ipynb cell 3 and 10:
    np.random.seed(42), then np.random.rand(N), uniform on [0, 1)) on the 500x500 SA
    hierarchy and on the Ruge-Stueben hierarchy of the same matrix, float64, withrho
    Jacobi, pinv: V-cycle solve and PCG iteration counts and relative residuals at `tol`
    (the published summary lists n_iter 40 for SA and 13 for Ruge-Stueben,
    benchmarks/latest_summary.json:27,80).  The notebook draws further vectors from the
    same stream for earlier configurations, so this is the first draw only."""
    import numpy as np
    import torch
    from amjax_amd import MultilevelSolver, hierarchy
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    np.random.seed(42)
    b = torch.as_tensor(np.random.rand(h.levels[0].A.shape[0]), device=dev)
    out = {}
    for name, hh in (("sa", h), ("rs", hierarchy.ruge_stuben_solver(h.levels[0].A))):
        ml = MultilevelSolver.from_pyamg(hh, spec, spec, coarse_solver="pinv", device=dev)
        _, it, rr = ml.solve(b, tol=tol, maxiter=250, return_info=True)
        _, itp, rrp = ml.pcg(b, tol=tol, maxiter=500, return_info=True)
        out[name] = {"levels": len(hh.levels), "vcycles": int(it), "vcycle_relres": float(rr),
                     "pcg_iters": int(itp), "pcg_relres": float(rrp)}
        log(f"reference RHS {name}: {len(hh.levels)} levels, V-cycle solve {it} cycles ({rr:.2e}), "
            f"PCG {itp} iterations ({rrp:.2e})")
    out["published_n_iter"] = {"sa": 40, "rs": 13}
    return out


def batch_problems(seed=2, count=40):
    """The uneven batch: problem i is 2-D five-point Poisson n x n with n ~ U[64, 768]
    (even i) or 3-D seven-point Poisson n^3 with n ~ U[16, 80] (odd i), k ~ U{1, 4, 16, 64}
    right-hand sides ~ N(0, 1); all draws in order from numpy default_rng(seed)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2 == 0:
            n = int(rng.integers(64, 769))
            grid = (n, n)
        else:
            n = int(rng.integers(16, 81))
            grid = (n, n, n)
        k = int(rng.choice([1, 4, 16, 64]))
        out.append((grid, k))
    return out, rng


def batch_workload(dev, tol, log, grid_cap=0, split_kp=True, autotune=True, reps=3):
    """40 independent SPD problems, each with its own SA hierarchy (Jacobi, pinv), float32
    cycle inside float64 PCG to `tol`: a per-problem loop of MultilevelSolver.pcg calls
    against one BatchPCG call (one launch per phase across all problems)."""
    import numpy as np
    import torch
    from amjax_amd import BatchPCG, MultilevelSolver, gallery, hierarchy
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    probs, rng = batch_problems()
    t0 = time.perf_counter()
    mls, Bs, ks = [], [], []
    for grid, k in probs:
        A = gallery.poisson(grid)
        h = hierarchy.smoothed_aggregation_solver(A)
        ml = MultilevelSolver.from_pyamg(h, spec, spec, coarse_solver="pinv", dtype=np.float32, device=dev)
        ml.autotune = autotune
        mls.append(ml)
        Bs.append(torch.as_tensor(rng.standard_normal((A.shape[0], k)), device=dev))
        ks.append(k)
    N = [ml.levels[0].A.shape[0] for ml in mls]
    log(f"batch: 40 problems, N {min(N)}..{max(N)} (sum {sum(N)}), sum k*N {sum(n * k for n, k in zip(N, ks))}, "
        f"setup {time.perf_counter() - t0:.1f}s")
    maxit = 500

    def loop_once(record=None):
        outs = []
        for ml, b in zip(mls, Bs):
            t = time.perf_counter()
            outs.append(ml.pcg(b, tol=tol, maxiter=maxit, return_info=True))
            if record is not None:
                record.append((time.perf_counter() - t) * 1e3)
        return outs

    batch = BatchPCG(mls, ks, grid=grid_cap, split_kp=split_kp, autotune=autotune)

    def batch_once():
        return batch.pcg(Bs, tol=tol, maxiter=maxit, return_info=True)

    loop_once()
    batch_once()
    torch.cuda.synchronize()
    loop_ms, per = [], []
    for _ in range(reps):
        rec = []
        t = time.perf_counter()
        lo = loop_once(rec)
        torch.cuda.synchronize()
        loop_ms.append((time.perf_counter() - t) * 1e3)
        per.append(rec)
    batch_ms = []
    for _ in range(reps):
        t = time.perf_counter()
        bo = batch_once()
        torch.cuda.synchronize()
        batch_ms.append((time.perf_counter() - t) * 1e3)
    best = int(np.argmin(loop_ms))
    solve_ms = np.array(per[best])
    done_ms = np.cumsum(solve_ms)  # completion time of each problem when all 40 arrive together
    identical = all(torch.equal(x, xb) for (x, _, _), xb in zip(lo, bo[0]))
    iters = [max(it) for _, it, _ in lo]
    worst = max(max(rr) for _, _, rr in lo)
    nl, nb = batch.info()
    res = {
        "problems": len(mls), "N_range": [min(N), max(N)], "k_counts": {str(k): ks.count(k) for k in (1, 4, 16, 64)},
        "iters_range": [min(iters), max(iters)], "max_relres": worst,
        "loop_ms": round(min(loop_ms), 3), "batch_ms": round(min(batch_ms), 3),
        "loop_solves_per_s": round(len(mls) / min(loop_ms) * 1e3, 1),
        "batch_solves_per_s": round(len(mls) / min(batch_ms) * 1e3, 1),
        "speedup": round(min(loop_ms) / min(batch_ms), 3),
        "loop_solve_ms_p50_p99": [round(float(np.percentile(solve_ms, 50)), 3),
                                  round(float(np.percentile(solve_ms, 99)), 3)],
        "loop_completion_ms_p50_p99": [round(float(np.percentile(done_ms, 50)), 3),
                                       round(float(np.percentile(done_ms, 99)), 3)],
        "batch_completion_ms_p50_p99": [round(min(batch_ms), 3), round(min(batch_ms), 3)],
        "batch_launches_per_iteration": nl, "batch_virtual_blocks_per_iteration": nb,
        "batch_equals_loop_bitwise": bool(identical),
    }
    log(f"batch: loop {res['loop_ms']} ms, batch {res['batch_ms']} ms (x{res['speedup']}), "
        f"iters {res['iters_range']}, identical {identical}")
    del batch
    return res


def config3_workload(dev, log, reps=3):
    """Mixed precision at N = 10^6: 2-D Poisson 1000x1000 and 3-D Poisson 100^3, SA,
    V-cycle, withrho Jacobi, pinv; B ~ N(0,1) from default_rng(3), k = 1 and 32; solved
    to 1e-10 relative float64 residual by float64 PCG, mixed PCG (float32 cycle), mixed
    iterative refinement (one float32 cycle per step) and float64 refinement around a
    float32-vector CG (inner tolerance 1e-4); error against a float64 PCG solve
    converged to 1e-14."""
    import numpy as np
    import torch
    from amjax_amd import MultilevelSolver, gallery, hierarchy
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    out = {}
    for grid in ((1000, 1000), (100, 100, 100)):
        A = gallery.poisson(grid)
        h = hierarchy.smoothed_aggregation_solver(A)
        N = A.shape[0]
        ml64 = MultilevelSolver.from_pyamg(h, spec, spec, device=dev)
        ml32 = MultilevelSolver.from_pyamg(h, spec, spec, dtype=np.float32, device=dev)
        ml64.autotune = ml32.autotune = True
        Bfull = torch.as_tensor(np.random.default_rng(3).standard_normal((N, 32)), device=dev)
        name = "x".join(map(str, grid))
        for k in (1, 32):
            B = Bfull[:, :k].contiguous()
            Xref = ml64.pcg(B, tol=1e-14, maxiter=1000)
            nref = torch.linalg.norm(Xref, dim=0)
            for label, fn in (("fp64_pcg", lambda: ml64.pcg(B, tol=1e-10, maxiter=1000, return_info=True)),
                              ("mixed_pcg", lambda: ml32.pcg(B, tol=1e-10, maxiter=1000, return_info=True)),
                              ("mixed_ru", lambda: _ru_info(ml32.pcg_ru(B, tol=1e-10, maxiter=1000, return_info=True))),
                              ("mixed_ir", lambda: ml32.solve(B, tol=1e-10, maxiter=1000, return_info=True)),
                              ("mixed_ircg", lambda: _ir_info(ml32.pcg_ir(B, tol=1e-10, inner_tol=1e-4, maxiter=20,
                                                                          return_info=True)))):
                fn()
                torch.cuda.synchronize()
                ts = []
                for _ in range(reps):
                    t = time.perf_counter()
                    X, its, rr = fn()
                    torch.cuda.synchronize()
                    ts.append((time.perf_counter() - t) * 1e3)
                its = its if isinstance(its, list) else [its]
                rr = rr if isinstance(rr, list) else [rr]
                err = float((torch.linalg.norm(X - Xref, dim=0) / nref).max())
                out[f"{name}_k{k}_{label}"] = {"ms": round(min(ts), 3), "iters": max(its),
                                               "max_relres": max(rr), "err_vs_fp64": err}
                log(f"config3 {name} k={k} {label}: {min(ts):.2f} ms, iters {max(its)}, "
                    f"relres {max(rr):.1e}, err vs fp64 {err:.1e}")
        del ml64, ml32
        torch.cuda.empty_cache()
    return out


def config4_workload(dev, log, reps=3):
    """Differentiation, 512x512 variable-coefficient diffusion (conductance exp(g), g a
    squared-exponential GRF, length 16 cells, sigma 1, seed 1), SA, k = 16, loss
    0.5 ||X - X*||^2 with X* ~ N(0,1) (seed 4): forward + backward through 10 fixed
    V-cycles (b-gradient) and through PCG to 1e-10 (A-values gradient)."""
    import numpy as np
    import torch
    from amjax_amd import MultilevelSolver, cg, cycles, gallery, hierarchy
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    g = gallery.gaussian_random_field((512, 512), length=16, sigma=1.0, seed=1)
    A = gallery.diffusion_fv(g)
    h = hierarchy.smoothed_aggregation_solver(A)
    N = A.shape[0]
    ml = MultilevelSolver.from_pyamg(h, spec, spec, device=dev)
    ml.autotune = True
    rng = np.random.default_rng(4)
    Xs = torch.as_tensor(rng.standard_normal((N, 16)), device=dev)
    B0 = torch.as_tensor(rng.standard_normal((N, 16)), device=dev)
    M = ml.aspreconditioner("V")
    base = ml.levels[0].A
    v0 = torch.as_tensor(base.host.data.copy(), device=dev)

    def fixed_cycles():
        B = B0.clone().requires_grad_(True)
        (gb,) = torch.autograd.grad(0.5 * ((cycles(ml, B, 10) - Xs) ** 2).sum(), B)
        return gb

    def pcg_values():
        vals = v0.clone().requires_grad_(True)
        X, _ = cg(base.with_values(vals), B0, M=M, tol=1e-10, maxiter=500)
        (gv,) = torch.autograd.grad(0.5 * ((X - Xs) ** 2).sum(), vals)
        return gv

    out = {}
    for label, fn in (("b_grad_10_vcycles", fixed_cycles), ("values_grad_pcg_1e10", pcg_values)):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t) * 1e3)
        out[label] = {"ms_forward_backward": round(min(ts), 3)}
        log(f"config4 {label}: {min(ts):.2f} ms (forward + backward, k = 16)")
    return out


def config2_variants(dev, tol, log, autotune=True, steps=5):
    """The other cases of the 2-D Poisson configurations, at k = 1 (float64 unless
    named mixed): W and F cycles of the 500x500 SA hierarchy (standalone solve and
    PCG), the Ruge-Stueben hierarchy of the same matrix (V-cycle solve and PCG, float64
    and mixed), and the 64x64 case with the reference's default Jacobi (omega 1, no
    withrho) and with withrho (standalone solve and PCG).  B ~ N(0,1) from
    default_rng(0)."""
    import numpy as np
    import torch
    from amjax_amd import MultilevelSolver, gallery, hierarchy

    def timed(fn, info):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        _, it_, rr_ = info()
        return {"ms": round(ms, 4), "iters": int(it_), "relres": float(rr_)}

    withrho = ("jacobi", {"iterations": 1, "withrho": True})
    out = {}
    A = gallery.poisson((500, 500))
    b = torch.as_tensor(np.random.default_rng(0).standard_normal((A.shape[0], 1)), device=dev)[:, 0].contiguous()
    sa = hierarchy.smoothed_aggregation_solver(A)
    ml = MultilevelSolver.from_pyamg(sa, withrho, withrho, device=dev)
    ml.autotune = autotune
    for cyc in ("W", "F"):
        out[f"sa500_{cyc}_solve"] = timed(lambda: ml.solve(b, tol=tol, maxiter=250, cycle=cyc),
                                          lambda: ml.solve(b, tol=tol, maxiter=250, cycle=cyc, return_info=True))
        out[f"sa500_{cyc}_pcg"] = timed(lambda: ml.pcg(b, tol=tol, maxiter=500, cycle=cyc),
                                        lambda: ml.pcg(b, tol=tol, maxiter=500, cycle=cyc, return_info=True))
    rs = hierarchy.ruge_stuben_solver(A)
    for label, dt in (("fp64", np.float64), ("mixed", np.float32)):
        mr = MultilevelSolver.from_pyamg(rs, withrho, withrho, dtype=dt, device=dev)
        mr.autotune = autotune
        out[f"rs500_V_solve_{label}"] = timed(lambda: mr.solve(b, tol=tol, maxiter=250),
                                              lambda: mr.solve(b, tol=tol, maxiter=250, return_info=True))
        out[f"rs500_pcg_{label}"] = timed(lambda: mr.pcg(b, tol=tol, maxiter=500),
                                          lambda: mr.pcg(b, tol=tol, maxiter=500, return_info=True))
    A64 = gallery.poisson((64, 64))
    h64 = hierarchy.smoothed_aggregation_solver(A64)
    b64 = torch.as_tensor(np.random.default_rng(0).standard_normal(A64.shape[0]), device=dev)
    for label, spec in (("default_jacobi", ("jacobi", {"iterations": 1, "omega": 1.0})), ("withrho", withrho)):
        m64 = MultilevelSolver.from_pyamg(h64, spec, spec, device=dev)
        out[f"sa64_V_solve_{label}"] = timed(lambda: m64.solve(b64, tol=tol, maxiter=100),
                                             lambda: m64.solve(b64, tol=tol, maxiter=100, return_info=True))
        out[f"sa64_pcg_{label}"] = timed(lambda: m64.pcg(b64, tol=tol, maxiter=100),
                                         lambda: m64.pcg(b64, tol=tol, maxiter=100, return_info=True))
    for key, v in out.items():
        log(f"config2/1 {key}: {v['ms']:.3f} ms, {v['iters']} iterations, relres {v['relres']:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=500)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-cols", type=int, default=64, help="columns of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 40-problem uneven batch")
    ap.add_argument("--no-configs", action="store_true", help="skip the N = 10^6 and gradient configurations")
    ap.add_argument("--batch-grid", type=int, default=0, help="persistent grid of the batched launches (0: off)")
    ap.add_argument("--autotune", type=int, default=1, help="row kernels use the autotuner's launch shapes")
    ap.add_argument("--form", choices=("ru", "pcg", "fp64"), default="ru",
                    help="timed solve: mixed CG with reliable updates (the headline), float32 cycle inside "
                         "float64 CG, or float64 PCG")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the multigrid solve of one problem runs on one device "
              "(multi-GPU sharding is not built)", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    import torch
    from amjax_amd import MultilevelSolver, gallery, hierarchy, _lib

    dev = torch.device("cuda", 0)
    n = args.grid
    N = n * n
    k = args.k
    t0 = time.perf_counter()
    A = gallery.poisson((n, n))
    h = hierarchy.smoothed_aggregation_solver(A)
    log(f"hierarchy {[l.A.shape[0] for l in h.levels]} built in {time.perf_counter() - t0:.2f}s")
    spec = ("jacobi", {"iterations": 1, "withrho": True})
    ml64 = MultilevelSolver.from_pyamg(h, spec, spec, coarse_solver="pinv", dtype=np.float64, device=dev)
    ml32 = MultilevelSolver.from_pyamg(h, spec, spec, coarse_solver="pinv", dtype=np.float32, device=dev)
    ml64.autotune = ml32.autotune = bool(args.autotune)
    Bh = np.random.default_rng(0).standard_normal((N, k))
    B = torch.as_tensor(Bh, device=dev)
    b1 = B[:, :1].contiguous()
    tol = args.tol
    maxit = 500

    def run_timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    # headline: mixed-precision CG with reliable updates, k RHS (--form: another solve form)
    if args.form == "ru":
        head = lambda: ml32.pcg_ru(B, tol=tol, maxiter=maxit)
        info = lambda: ml32.pcg_ru(B, tol=tol, maxiter=maxit, return_info=True)
    else:
        mlh = ml32 if args.form == "pcg" else ml64
        head = lambda: mlh.pcg(B, tol=tol, maxiter=maxit)
        info = lambda: (lambda o: (o[0], o[1], [0] * len(o[1]), o[2]))(mlh.pcg(B, tol=tol, maxiter=maxit,
                                                                                return_info=True))
    log(f"warmup + timed headline ({args.form})")
    ms = run_timed(head, args.steps, args.warmup)
    Xh, its, ups, rr = info()
    log(f"{args.form} k={k}: {ms:.3f} ms, iters {min(its)}-{max(its)}, updates {max(ups)}, "
        f"max relres {max(rr):.2e}")
    # error of the mixed results against a float64 PCG solve converged to 1e-14
    Xref = ml64.pcg(B, tol=1e-14, maxiter=2 * maxit)
    nref = torch.linalg.norm(Xref, dim=0)

    def err_vs_fp64(X):
        X = X.reshape(Xref.shape[0], -1)
        c = X.shape[1]
        return float((torch.linalg.norm(X - Xref[:, :c], dim=0) / nref[:c]).max())

    head_err = err_vs_fp64(Xh)
    del Xh

    extra = {}
    if not args.no_extra:
        steps2 = max(1, min(args.steps, 5))
        cases = [
            ("pcg_fp64_k%d" % k, lambda: ml64.pcg(B, tol=tol, maxiter=maxit),
             lambda: ml64.pcg(B, tol=tol, maxiter=maxit, return_info=True)),
            # float32 cycle inside float64 CG (float64 CG vectors)
            ("pcg_mixed_k%d" % k, lambda: ml32.pcg(B, tol=tol, maxiter=maxit),
             lambda: ml32.pcg(B, tol=tol, maxiter=maxit, return_info=True)),
            ("ru_mixed_k1", lambda: ml32.pcg_ru(b1, tol=tol, maxiter=maxit),
             lambda: _ru_info(ml32.pcg_ru(b1, tol=tol, maxiter=maxit, return_info=True))),
            ("pcg_mixed_k1", lambda: ml32.pcg(b1, tol=tol, maxiter=maxit),
             lambda: ml32.pcg(b1, tol=tol, maxiter=maxit, return_info=True)),
            ("pcg_fp64_k1", lambda: ml64.pcg(b1, tol=tol, maxiter=maxit),
             lambda: ml64.pcg(b1, tol=tol, maxiter=maxit, return_info=True)),
            ("vcycle_fp64_k1", lambda: ml64.solve(b1, tol=tol, maxiter=250),
             lambda: ml64.solve(b1, tol=tol, maxiter=250, return_info=True)),
            ("vcycle_mixed_k1", lambda: ml32.solve(b1, tol=tol, maxiter=250),
             lambda: ml32.solve(b1, tol=tol, maxiter=250, return_info=True)),
            ("vcycle_fp64_k%d" % k, lambda: ml64.solve(B, tol=tol, maxiter=250),
             lambda: ml64.solve(B, tol=tol, maxiter=250, return_info=True)),
            ("vcycle_mixed_k%d" % k, lambda: ml32.solve(B, tol=tol, maxiter=250),
             lambda: ml32.solve(B, tol=tol, maxiter=250, return_info=True)),
            # float64 refinement around a float32-vector CG (inner tolerance 1e-4);
            # iterations: inner CG iterations summed over the refinement steps
            ("ircg_mixed_k%d" % k, lambda: ml32.pcg_ir(B, tol=tol, inner_tol=1e-4, maxiter=20),
             lambda: _ir_info(ml32.pcg_ir(B, tol=tol, inner_tol=1e-4, maxiter=20, return_info=True))),
            ("ircg_mixed_k1", lambda: ml32.pcg_ir(b1, tol=tol, inner_tol=1e-4, maxiter=20),
             lambda: _ir_info(ml32.pcg_ir(b1, tol=tol, inner_tol=1e-4, maxiter=20, return_info=True))),
        ]
        for name, fn, info in cases:
            t = run_timed(fn, steps2, 1)
            X_, it_, rr_ = info()
            it_ = it_ if isinstance(it_, list) else [it_]
            rr_ = rr_ if isinstance(rr_, list) else [rr_]
            extra[name] = {"ms": round(t, 4), "iters": [min(it_), max(it_)], "max_relres": max(rr_)}
            if "mixed" in name:
                extra[name]["err_vs_fp64"] = err_vs_fp64(X_)
            del X_
            log(f"{name}: {t:.3f} ms iters {min(it_)}-{max(it_)} relres {max(rr_):.2e}"
                + (f" err vs fp64 {extra[name]['err_vs_fp64']:.1e}" if "mixed" in name else ""))
        fp = extra["pcg_fp64_k%d" % k]["ms"]
        extra["mixed_speedup_ru_k%d" % k] = round(fp / ms, 3)
        extra["mixed_speedup_pcg_k%d" % k] = round(fp / extra["pcg_mixed_k%d" % k]["ms"], 3)
        extra["mixed_speedup_ircg_k%d" % k] = round(fp / extra["ircg_mixed_k%d" % k]["ms"], 3)
        extra["mixed_speedup_ru_k1"] = round(extra["pcg_fp64_k1"]["ms"] / extra["ru_mixed_k1"]["ms"], 3)
        extra["reference_rhs_500"] = reference_rhs_counts(h, dev, tol, log)

    if not args.no_batch:
        extra["batch40"] = batch_workload(dev, tol, log, grid_cap=args.batch_grid, autotune=bool(args.autotune))
        if args.batch_grid == 0:
            # the persistent form: every launch a grid of 8 blocks per CU looping over its work
            cus = torch.cuda.get_device_properties(dev).multi_processor_count
            extra["batch40_persistent"] = batch_workload(dev, tol, log, grid_cap=8 * cus,
                                                         autotune=bool(args.autotune))
    if not args.no_configs:
        extra["poisson_variants_k1"] = config2_variants(dev, tol, log, autotune=bool(args.autotune))
        extra["mixed_precision_1M"] = config3_workload(dev, log)
        extra["gradients_512"] = config4_workload(dev, log)

    # per-kernel achieved bandwidth (HIP events on the launch stream, 50 launches each);
    # algorithmic bytes per launch follow layout-independent formulas (matrix entries
    # and indices once, each vector read or written once)
    lib = _lib.load()
    stream = _lib.stream_ptr(dev)
    reps = 50
    a0 = h.levels[0].A.nnz
    n1 = h.levels[1].A.shape[0]
    p0 = h.levels[0].P.nnz
    ml32._native()
    d0 = ml32._derived[0][1].nnz

    def time_kernel(ml, kernel, level=0):
        t = ctypes.c_float()
        _lib.check(lib.amg_time_kernel(ml._native(), kernel, level, k, reps, ctypes.byref(t), stream),
                   "amg_time_kernel")
        return t.value * 1e-3  # s

    kernels = {}

    def record(name, seconds, nbytes):
        kernels[name] = {"us": round(seconds * 1e6, 2), "bytes": int(nbytes),
                         "GBps": round(nbytes / seconds / 1e9, 1), "frac": round(nbytes / seconds / 1e9 / HBM_PEAK_GBS, 4)}

    # one 128,000,000-byte device copy: the known-size transfer that calibrates the
    # FETCH_SIZE / WRITE_SIZE counters when this script runs under rocprofv3 --pmc
    cal = torch.empty(32_000_000, dtype=torch.float32, device=dev)
    torch.empty_like(cal).copy_(cal)
    del cal
    # the streaming ceiling: a 256 MiB device-to-device copy (512 MiB moved, twice the
    # Infinity Cache, so it streams from and to HBM), the runtime's own copy kernel
    src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(20):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    record("device_copy_256MiB", e0.elapsed_time(e1) * 1e-3 / 20, 2 * src.numel() * 4)
    del src, dst
    COLD = 256  # AMG_TIME_COLD: 512 MB read before each timed launch (inputs from HBM)
    ru_bytes = 6 * N * k * 4                        # p, Ap, xh, rh read; xh, rh written (float32)
    upd_bytes = 6 * N * k * 8 + N * k * 4           # p, Ap, x, r read; x, r write (f64); r as f32 write

    def in_solve(ml, kernel, reps_):
        t = ctypes.c_float()
        _lib.check(lib.amg_time_kernel(ml._native(), kernel, 0, k, reps_, ctypes.byref(t), stream), "amg_time_kernel")
        return t.value * 1e-3

    # the dominant kernel of the headline solve: the reliable-update CG's update, timed with
    # events around each of its launches inside an eager solve (hot: as the solve runs;
    # cold: inputs from HBM), and alone (50 back-to-back launches, and cache-cold)
    record("ru_update_f32_k%d" % k, in_solve(ml32, 9, 10), ru_bytes)
    record("ru_update_f32_cold_k%d" % k, in_solve(ml32, 9 | COLD, 10), ru_bytes)
    record("ru_update_f32_alone_k%d" % k, time_kernel(ml32, 11), ru_bytes)
    record("ru_update_f32_alone_cold_k%d" % k, time_kernel(ml32, 11 | COLD), ru_bytes)
    # the float64 CG update of the float64-vector PCG forms
    record("pcg_update_f64outer_k%d" % k, in_solve(ml32, 8, 10), upd_bytes)
    record("pcg_update_f64outer_alone_k%d" % k, time_kernel(ml32, 5), upd_bytes)
    record("pcg_update_f64outer_cold_k%d" % k, time_kernel(ml32, 5 | COLD), upd_bytes)
    spmv64 = a0 * 12 + (N + 1) * 4 + 2 * N * k * 8
    jac32 = a0 * 8 + (N + 1) * 4 + N * 4 + 3 * N * k * 4
    down32 = d0 * 8 + (n1 + 1) * 4 + N * k * 4 + n1 * k * 4
    prol32 = p0 * 8 + (N + 1) * 4 + n1 * k * 4 + N * 4 + 2 * N * k * 4
    for name, ml, kern, nbytes in (("spmv_A0_f64", ml64, 0, spmv64), ("jacobi_A0_f32", ml32, 1, jac32),
                                   ("down_D0_f32", ml32, 6, down32), ("prolong_x1_P0_f32", ml32, 7, prol32)):
        record(f"{name}_k{k}", time_kernel(ml, kern), nbytes)
        record(f"{name}_cold_k{k}", time_kernel(ml, kern | COLD), nbytes)
    dom = kernels["ru_update_f32_k%d" % k]
    cold = kernels["ru_update_f32_cold_k%d" % k]
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_ru_update.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tr = json.load(f)
        if tr.get("k") == k and tr.get("n") == N:
            traffic = tr.get("hbm_bytes_per_launch")
    log("kernels: " + ", ".join(f"{n_}: {v['us']}us {v['GBps']}GB/s" for n_, v in kernels.items()))

    cpu = None
    if not args.no_cpu:
        from oracle.solve import OracleMG, pcg
        om = OracleMG(h, spec, spec)
        cols = max(1, min(args.cpu_cols, k))
        M = om.aspreconditioner()
        t = time.perf_counter()
        pcg(om.levels[0].A, Bh[:, :cols], M, tol=tol, maxiter=maxit)
        dt = time.perf_counter() - t
        cpu = {"value": round(dt * 1e3 * k / cols, 1), "unit": f"ms per {k}-RHS solve",
               "cores": 1, "kind": "port",
               "sample": f"numpy/scipy oracle, float64 AMG-PCG on {cols} of the {k} columns "
                         f"(scaled by {k}/{cols}), host {os.cpu_count()} cpus, scipy SpMV single-threaded"}
        log(f"cpu baseline: {cpu['value']} ms ({dt:.1f}s sample)")

    forms = {"ru": "mixed FP32 CG with FP64 reliable updates", "pcg": "mixed FP32-cycle/FP64-PCG",
             "fp64": "FP64 PCG"}
    out = {
        "metric": f"ms to 1e-8 rel. residual, 250k Poisson k=64, {forms[args.form]}",
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": round(ms / 295.15, 5) if (n == 500 and k == 64) else None,
        "dtype": {"ru": "f32 cycle and CG vectors / f64 dots and reliable updates",
                  "pcg": "f32 cycle / f64 CG", "fp64": "f64"}[args.form],
        "data": "synthetic: B ~ N(0,1) default_rng(0)",
        "config": {"workload": f"2D Poisson {n}x{n} (N={N}), SA/V-cycle/Jacobi(withrho)/pinv, "
                               f"AMG-preconditioned CG to {tol:g} true float64 relative residual, k={k} RHS, "
                               + {"ru": "float32 cycle and CG vectors, float64 dot products and reliable updates",
                                  "pcg": "float32 cycle inside float64 CG",
                                  "fp64": "float64 throughout"}[args.form],
                   "iters": [min(its), max(its)], "reliable_updates": max(ups), "max_relres": max(rr),
                   "err_vs_fp64": head_err},
        "roofline": {"bound": "hbm", "achieved": dom["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac"], "traffic": traffic,
                     "kernel": f"k_ru_update_v<float,{min(k, 64)}> (reliable-update CG update xh += alpha p, "
                               f"rh -= alpha Ap, rh.rh; float32, finest level), {dom['bytes']} algorithmic "
                               f"bytes/launch, mean of its launches inside the headline solve",
                     "us_per_launch": dom["us"],
                     "achieved_cold": cold["GBps"], "frac_cold": cold["frac"], "us_per_launch_cold": cold["us"],
                     "which_is_hbm": "achieved_cold: 512 MB are read before each timed launch, so the "
                                     "launch reads its inputs from HBM; achieved (in the solve) is helped by "
                                     "the 256 MB Infinity Cache holding p and Ap, written just before"},
        "spmv_fp64_gbs": kernels["spmv_A0_f64_k%d" % k]["GBps"],
        "kernels": kernels,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
