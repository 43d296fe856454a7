"""bench.py command line on the CPU: it measures one GPU and refuses more with a message
before touching a device."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "--gpus 1" in (p.stdout + p.stderr)


def _bench_module():
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_driver_invocation_reports_the_headline_fields():
    """Under the driver's `--steps 20 --warmup 5` both forms are reported to the target: the
    record assembled from stand-in leg results has every headline field filled and
    steps x ms_per_step covering exactly the K timed sweeps."""
    bench = _bench_module()
    args = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    assert args.steps == 20 and args.warmup == 5 and not args.sweeps_only
    assert args.form == "both"
    plain = {"time_to_target_s": 36.0, "wall_s": 36.0, "sweeps": 157, "sweeps_to_target": 157,
             "setup_ms": 130.0, "ms_per_sweep": 227.0, "residual_true_final": 9.9e-4,
             "rmse_vs_truth_in_mask": 0.06, "timed_ms": [227.0 + i for i in range(20)]}
    learned = {"time_to_target_s": 2.6, "wall_s": 2.6, "sweeps": 10,
               "rmse_vs_truth_in_mask": 0.03, "residual_true_final": 9.5e-4}
    cpu = {"value": 1.0e5, "cores": 16, "unit": "s", "kind": "port", "sample": "s"}
    rec = bench.headline_record(args, plain, learned, dict(learned), {"frac": 0.2}, {}, {}, cpu,
                                True, 3936)
    for key in ("value", "speedup_vs_plain", "rmse_ratio_vs_plain", "ms_per_step"):
        assert rec[key] is not None, key
    assert rec["cpu_baseline"]["value"] is not None
    assert rec["plain_fp32"]["time_to_target_s"] is not None
    assert rec["steps"] == 20 and rec["warmup"] == 5
    assert abs(rec["ms_per_step"] - (227.0 + 9.5)) < 1e-9
    assert abs(rec["speedup_vs_plain"] - 36.0 / 2.6) < 1e-12
    assert rec["value"] == 2.6 and rec["higher_is_better"] is False
    assert "timed_ms" not in rec["plain_fp32"]
