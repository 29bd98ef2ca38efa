"""bench.py's contract, checked on the CPU: it refuses several GPUs with a
JSON message and status 2, keeps its documented defaults and workloads, and its
per-kernel algorithmic bytes add up to the step's roofline bytes.  Nothing
here touches a GPU."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_several_gpus_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_defaults_and_workloads(monkeypatch):
    b = _bench()
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = b.parse()
    assert (a.gpus, a.steps, a.warmup, a.cpu_steps) == (1, 50, 5, 6)
    assert a.workload == b.DEFAULT_WORKLOAD == "retau550_768x384x384"
    assert b.WORKLOADS["retau550_768x384x384"][:3] == (768, 384, 384)
    assert b.WORKLOADS["retau1000_1536x768x512"][:3] == (1536, 768, 512)


def test_step_bytes_add_up():
    b = _bench()
    u = b.ROUTINE_BYTES_FP64
    assert sum(u.values()) == b.STEP_BYTES_FP64 == 312
    k = b.BYTES_PER_POINT["fp64"]
    for n in ("rk2", "calc_rhs", "fft_x_fwd", "fft_y_fwd", "solve_trid", "fft_y_bwd", "fft_x_bwd",
              "correct_vel", "mean_partial", "force_add"):
        assert k[n] == u[n], n
    # inside the step stage 1 also reads p (the previous step's projection is
    # applied on load); the fused stage 2 + calcRHS reads u, v, w once, and p
    assert k["rk1"] == u["rk1"] + 8
    assert k["rk2_rhs"] == u["rk2"] + u["calc_rhs"] - 24 + 8
    # FP32 fields halve every field access; FP32-mixed stage 2 also reads p
    f = b.BYTES_PER_POINT["fp32_mixed"]
    for n in ("rk1", "calc_rhs", "fft_x_fwd", "fft_y_fwd", "solve_trid", "fft_y_bwd", "fft_x_bwd",
              "correct_vel", "mean_partial", "force_add"):
        assert 2 * f[n] == k[n], n
    assert f["rk2"] == k["rk2"] // 2 + 4
    # the fused FP32-mixed stages read u, v, w, p and write u, v, w
    assert f["rk12"] == 4 * 4 + 3 * 4
    assert b.STAT_BYTES == {"fp64": 208, "fp32_mixed": 192}
