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


def test_sq_summary_derives_clock_busy_and_waits(tmp_path, capsys):
    """--summarize-sq on a synthetic counter file: one FP8 trailing-update
    dispatch of 1 ms at 2 GHz on 8 XCDs with the matrix pipes of its 1024
    SIMDs half busy, one wave per SIMD, a quarter of the wave cycles waiting
    (an eighth on LDS)."""
    sys.path.insert(0, ROOT)
    import bench
    cyc = 2.0e6  # 1 ms at 2 GHz
    vals = {"GRBM_GUI_ACTIVE": 8 * cyc, "SQ_BUSY_CYCLES": 8 * cyc, "SQ_WAVE_CYCLES": 1024 * cyc / 4,
            "SQ_VALU_MFMA_BUSY_CYCLES": 0.5 * 1024 * cyc, "SQ_WAIT_ANY": 0.25 * 1024 * cyc / 4,
            "SQ_WAIT_INST_LDS": 0.125 * 1024 * cyc / 4, "SQ_ACTIVE_INST_LDS": 0.0, "SQ_LDS_BANK_CONFLICT": 0.0}
    path = tmp_path / "sq.csv"
    with open(path, "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp\n")
        for k, v in vals.items():
            f.write(f'7,"void k_gemm_f8<false, 4>(long)",{k},{v},1000000,2000000\n')
        f.write('8,"k_other(long)",GRBM_GUI_ACTIVE,5,0,10\n')
    bench.summarize_sq(str(path))
    out = capsys.readouterr().out
    assert "k_gemm_f8<false: 1 launches, 1.0 ms" in out
    assert "clock 2.000 GHz; MFMA busy 0.500; waves/SIMD 1.00; waiting 0.250 (LDS 0.125)" in out
    assert "k_gemm_f16<false" not in out
