"""The bench.py output contract: one JSON line with the metric, the value
and its unit, the run shape, the roofline of the dominant kernel and the
CPU baseline; more than one GPU is refused."""
import io
import json
import os
import subprocess
import sys
from contextlib import redirect_stdout

import pytest

from conftest import ROOT


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr and r.stdout == ""


@pytest.mark.gpu
def test_bench_json_line(monkeypatch):
    sys.path.insert(0, ROOT)
    import bench
    monkeypatch.setattr(sys, "argv", ["bench.py", "--config", "c4", "--steps", "1", "--warmup",
                                      "0", "--cpu-seconds", "1", "--simp-iters", "2",
                                      "--simp-fp32ir-iters", "1"])
    out = io.StringIO()
    with redirect_stdout(out):
        bench.main()
    lines = [ln for ln in out.getvalue().splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 0
    assert d["value"] > 0 and d["higher_is_better"] is True and d["dtype"] == "f64"
    assert d["vs_baseline"] is None and d["data"] == "synthetic"
    assert "workload" in d["config"] and d["config"]["true_rel_res"] <= 1e-8
    rl = d["roofline"]
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("reference", "port")
    assert d["mixed_true_rel_res"] <= 1e-8
    assert rl["byte_model"] and rl["csr_bytes_per_launch"] > rl["bytes_per_launch"]
    assert cb["host"]["nproc"] >= 1 and cb["host"]["affinity"] >= 1
    st = rl["stream_context"]
    assert 0 < st["copy_gbs"] < 1.2 * rl["peak"]
    assert abs(st["spmv_frac_of_copy"] - rl["achieved"] / st["copy_gbs"]) < 1e-2
    # the mixed-precision speedup split into the preconditioner's gain and
    # the precision's, and the FP32-CG + FP64-refinement scheme beside them
    dec = d["decomposition"]
    assert abs(dec["preconditioner_factor"] * dec["precision_factor"]
               - d["mixed_speedup_vs_fp64"]) <= 0.01 * d["mixed_speedup_vs_fp64"]
    assert dec["cheb_fcg_fp64_precond"]["true_rel_res"] <= 1e-8
    assert d["fp32ir"]["seconds"] == d["time_to_1e8_fp32ir_s"] > 0
    assert set(d["simp_decomposition"]["ms_per_iter"]) == {
        "jacobi_pcg_fp64", "cheb_fcg_fp64_precond", "cheb_fcg_fp16_precond"}
    assert d["simp_fp32ir"]["iterations"] == 1
