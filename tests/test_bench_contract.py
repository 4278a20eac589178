"""bench.py's output contract: one JSON line with the metric, the roofline of
the dominant kernel and the CPU baseline; a multi-GPU request exits with a
message (the solver runs on one device)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def run_bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def test_more_than_one_gpu_exits_with_message():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.gpu
def test_headline_line_has_the_contract_fields():
    # a small instance of the headline workload: random 4-regular graph,
    # 8-bit with stochastic Y rounding (headline), 10-, 12- and 8-bit and
    # packed dSB against plain FP32 dSB, CPU baseline on a 1 s sample
    # --steps 150: a run executes whole window periods, so 200 steps run and
    # the line reports 200 (the timed region's real size)
    r = run_bench("--nodes-log2", "14", "--agents", "128", "--steps", "150", "--warmup", "10",
                  "--cpu-seconds", "1", "--instances", "64", "--seq-sample", "4")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config", "roofline",
              "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 200 and d["warmup"] == 10
    assert d["config"]["steps_requested"] == 150
    assert d["value"] > 0 and d["unit"] == "agent-steps/s" and d["higher_is_better"] is True
    assert "workload" in d["config"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    assert "traffic" in rf  # null away from the profiled size
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("port", "reference")
    assert cb["sample"]
    # headline: the 8-bit form with stochastic Y rounding; packed and plain
    # dSB follow the same trajectory (identical best cut), the narrow forms
    # their own
    assert d["config"]["form"] == "q8s" and d["packed"]["best_cut_rel_diff_vs_plain"] == 0.0
    assert abs(d["best_cut_rel_diff_vs_plain"]) < 0.01
    for f in ("q10", "q12", "q8"):
        assert d[f]["form"] == f and abs(d[f]["best_cut_rel_diff_vs_plain"]) < 0.01
    assert d["energy_check"]["max_abs_err_cut"] == 0.0
    bf = d["bf16_bsb"]
    assert bf["form"] == "bf16" and bf["plain_bsb"]["form"] == "f32" and bf["value"] > 0
    assert abs(bf["best_cut_rel_diff_vs_plain_bsb"]) < 0.01
    sv = d["serving"]
    assert sv["alone_vs_mix_identical"] == "4/4" and sv["closed_loop_solves_per_s"] > 0
    gr = d["graph"]
    assert gr["use_graph"] == 2 and gr["steps_executed"] == 200 and gr["value"] > 0
    assert gr["identical_to_direct"] is True
