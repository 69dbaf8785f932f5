"""bench.py's output contract: one JSON line with the metric, the roofline
and CPU-baseline objects; one GPU only; the PMC traffic figure attached only
to the workload it was measured on."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_traffic_only_for_the_measured_workload():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.pmc_traffic("f8", 204800, 2560, 4) > 0
    assert bench.pmc_traffic("f16", 204800, 2560, 2) > 0
    assert bench.pmc_traffic("f8", 10240, 256, 4) is None
    assert bench.pmc_traffic("f8", 204800, 2560, 2) is None
    assert bench.pmc_traffic("f16", 204800, 2560, 4) is None


@pytest.mark.gpu
def test_small_run_prints_the_contract_line():
    from tests.gpu_util import torch_cuda
    torch_cuda()
    r = subprocess.run([sys.executable, BENCH, "--n", "10240", "--b", "256", "--steps", "1", "--warmup", "1",
                        "--cpu-n", "1024", "--cpu-b", "256"], capture_output=True, text=True, timeout=300)
    print(r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    line = json.loads(r.stdout.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in line, key
    assert line["n_gpus"] == 1 and line["steps"] == 1 and line["warmup"] == 1 and line["value"] > 0
    assert line["valid"] and line["harness"] < 16
    assert "workload" in line["config"] and line["config"]["n"] == 10240
    roof = line["roofline"]
    assert roof["bound"] == "mfma" and roof["unit"] == "TFLOP/s" and roof["peak"] == 5000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    assert roof["traffic"] is None  # measured at n = 204800 only
    cpu = line["cpu_baseline"]
    assert cpu["value"] > 0 and cpu["kind"] in ("port", "reference") and cpu["cores"] >= 1
