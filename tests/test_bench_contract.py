"""The bench.py contract: one JSON line with the metric, throughput, roofline and CPU
baseline; one GPU only.  The GPU test runs a short bench of the 101^3 cube."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout.strip() == ""


def test_cpu_baseline_leg_reports_a_bounded_sample():
    import bench
    from ela3d_amd import workloads
    cfg, model = workloads.homogeneous_cube(nstep=4, n=24)
    b = bench.cpu_baseline(cfg, model, budget_s=0.5)
    assert set(b) >= {"value", "unit", "cores", "kind", "sample"}
    assert b["value"] > 0 and b["unit"] == "Mcell-updates/s" and b["kind"] == "port"
    assert 1 <= b["cores"] <= b["nproc"] and "steps" in b["sample"]
    assert b["cores"] <= b["cores_available"]


@pytest.mark.gpu
def test_bench_json_line(gpu):
    p = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--workload", "homogeneous_cube",
                        "--steps", "20", "--warmup", "2", "--no-cpu", "--no-tune", "--shape", "128,2,1",
                        "--accuracy-steps", "40", "--serving-shots", "0"],
                       cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 2 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["unit"] == "Mcell-updates/s" and d["dtype"] == "f32"
    assert np.isclose(d["value"], d["config"]["cells"] / (d["ms_per_step"] * 1e-3) / 1e6, rtol=1e-3)
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert 0 < r["frac"] < 1 and np.isclose(r["frac"], r["achieved"] / r["peak"], rtol=1e-3)
    assert "workload" in d["config"] and d["cpu_baseline"] is None
    acc = d["config"]["accuracy"]
    assert acc["nstep"] == 40 and acc["seismogram_rel_l2_vs_fp64"] <= 1e-4 and acc["within_bar"]
    assert d["config"]["serving"] is None


@pytest.mark.gpu
def test_serving_summary_modes(gpu):
    """The shortened serving leg of the bench line: every mode runs every shot."""
    import bench
    out = bench.serving_summary(nshots=6, nstep_range=(20, 40), edges=(32, 40), max_active=3)
    for mode in ("batched", "graph_one_at_a_time", "plain_loop"):
        m = out[mode]
        assert m["failed"] == 0 and m["shots_per_s"] > 0, mode
        assert m["poisson_p99_s"] >= m["poisson_p50_s"] > 0, mode
    assert out["speedup_vs_plain_loop"] > 0
