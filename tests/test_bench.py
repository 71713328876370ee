"""bench.py's output contract: one JSON line with the metric, the roofline of
the dominant kernel and the CPU baseline; more than one GPU is refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and p.stdout == ""


@pytest.mark.gpu
def test_bench_json_line():
    print("bench --config 1 --steps 3 --warmup 1", flush=True)
    p = subprocess.run([sys.executable, BENCH, "--config", "1", "--steps", "3", "--warmup", "1", "--cpu-seconds", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    lines = [l for l in p.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["unit"] == "set-samples/s" and d["data"] == "synthetic" and "workload" in d["config"]
    # AUTO runs a 2-way search this small (2.7e8 set-samples) on the VALU
    # kernel, whose 32 x 32 units spread it over more workgroups
    assert d["dtype"] == "u32" and d["config"]["path"] == "fused_valu"
    r = d["roofline"]
    assert r["bound"] == "valu" and r["unit"] == "T lane-ops/s" and r["peak"] > 0 and r["achieved"] > 0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9 and "traffic" in r
    c = d["cpu_baseline"]
    assert c["value"] > 0 and c["cores"] >= 1 and c["kind"] in ("port", "reference") and c["sample"]
    assert d["best"]["idx"] and d["best"]["fp32_max_rel_err"] <= 1e-5
    # the plain round loop, counted both ways
    assert set(d["plain_path"]) == {"mfma_count", "valu_count"}
