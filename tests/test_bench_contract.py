"""bench.py's output contract: one JSON line with the fields a driver reads,
and a clear refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.gpu
def test_small_run_prints_one_contract_line():
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--qubits", "20", "--steps", "2",
                        "--warmup", "1", "--no-cpu", "--no-density", "--no-trotter", "--no-swap",
                        "--fp32-qubits", "22"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [x for x in r.stdout.splitlines() if x.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["value"] > 0
    assert d["unit"] == "GB/s" and d["dtype"] == "f64" and d["higher_is_better"] is True
    assert d["config"]["qubits"] == 20 and "workload" in d["config"]
    rl = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rl, k
    assert rl["bound"] == "hbm" and abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-3
    f = d["fused_fp32"]
    assert f["fidelity_error"] < 1e-6 and f["speedup_vs_fp64_per_gate"] > 0
    big = d["fused_fp32_large"]
    assert big["qubits"] == 22 and big["fidelity_error"] < 1e-6
