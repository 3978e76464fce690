"""bench.py's command-line contract: one JSON line with the metric, the
roofline and CPU-baseline objects; runs on more than one GPU are refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""


@pytest.mark.gpu
def test_bench_json_line_on_a_small_graph():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--log-n", "14", "--degree", "8",
                          "--steps", "3", "--warmup", "1", "--no-extras", "--cpu-seconds", "0.5"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1 and d["dtype"] == "f32"
    assert d["value"] > 0 and d["ms_per_step"] > 0
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert d["context"]["gflops_step"] > 0 and d["context"]["gather_no_reuse_GBps"] > 0
    c = d["cpu_baseline"]
    assert c["kind"] == "port" and c["value"] > 0 and c["cores"] >= 1
    assert d["config"]["workload"].startswith("rmat14_deg8_community_part0of8")


def test_precision_schedule_table_is_complete():
    """The committed schedule table of the precision section names a full
    plan tuning for every (dim, storage dtype) the section runs, and the
    workload fields bench.py checks before using it."""
    import cola_amd as cola
    with open(os.path.join(ROOT, "profiles", "plan_rows_precision_rmat22_tuning.json")) as f:
        d = json.load(f)
    assert d["order"] == "community" and d["row_order"] == 1
    assert (d["log_n"], d["degree"], d["nparts"], d["rank"]) == (22, 32, 8, 0)
    assert d["command"].startswith("python -m cola_amd.tune_precision")
    for dim in ("64", "128", "256"):
        for name in ("f32", "bf16", "f16", "fp8_e4m3"):
            t = d["tunings"][dim][name]
            assert sorted(t) == sorted(cola.Plan.TUNING_FIELDS), (dim, name)


@pytest.mark.gpu
def test_tune_precision_writes_a_table(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    out = tmp_path / "table.json"
    run = subprocess.run([sys.executable, "-m", "cola_amd.tune_precision", "--log-n", "12", "--degree", "8",
                          "--nparts", "2", "--dims", "32", "--dtypes", "f32", "bf16", "--autotunes", "2",
                          "--rounds", "2", "--reps", "2", "--out", str(out)],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert run.returncode == 0, run.stderr[-2000:]
    import cola_amd as cola
    d = json.loads(out.read_text())
    assert (d["log_n"], d["nparts"], d["rank"], d["order"]) == (12, 2, 0, "community")
    for name in ("f32", "bf16"):
        assert sorted(d["tunings"]["32"][name]) == sorted(cola.Plan.TUNING_FIELDS)
        s = d["step_ms"]["32"][name]
        assert 1 <= s["candidates"] <= 2 and len(s["median_ms"]) == s["candidates"] and min(s["median_ms"]) > 0
