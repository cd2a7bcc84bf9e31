"""bench.py end to end on a small problem: one JSON line with the contract's keys.

The headline workload itself (500 x 500, k = 64) is what the round-end benchmark
runs; this keeps the script's code paths (headline, float64 reference, kernel
table with its cache-cold timings, roofline) exercised on a 64 x 64 problem.
"""

import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_small_problem_prints_the_contract_line(cuda, capsys, monkeypatch):
    sys.path.insert(0, ROOT)
    import bench
    monkeypatch.setattr(sys, "argv", ["bench.py", "--grid", "64", "--k", "16", "--steps", "2", "--warmup", "1",
                                      "--no-extra", "--no-batch", "--no-configs", "--no-cpu"])
    bench.main()
    line = capsys.readouterr().out.strip().splitlines()[-1]
    out = json.loads(line)
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 2 and out["higher_is_better"] is False
    assert out["value"] > 0 and out["vs_baseline"] is None
    assert out["config"]["max_relres"] < 1e-8 and out["config"]["err_vs_fp64"] < 1e-6
    roof = out["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert 0 < roof["frac"] < 1.5 and roof["achieved"] > 0 and roof["achieved_cold"] > 0
