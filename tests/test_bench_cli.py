"""bench.py's command-line contract and its link to the committed counter
profiles (CPU only: no GPU work is started)."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_bad_step_counts_are_refused_before_any_gpu_work():
    for args in (["--steps", "0"], ["--timing-steps", "0"], ["--warmup", "-1"]):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
        assert ">=" in r.stderr, args


def test_committed_profiles_match_the_default_tuning():
    """The bench line's roofline.traffic and VALU estimate come from these
    records; they apply only when the run's tuning equals the recorded one, and
    the autotuner keeps the default unless another setting is clearly faster."""
    bench = _bench()
    default = {"cells_per_h": 2, "prune": 1, "block": 128, "blocks_per_xcd": 0}
    for workload in ("dam1m", "dam8m"):
        rec = bench.profile_record(workload, "force_integrate", default)
        assert rec is not None, workload
        assert rec["traffic_bytes_per_launch"] > 0
        assert 0.0 < rec["valu_issue_estimate"] < 2.0
        assert os.path.exists(os.path.join(ROOT, rec["source"]))
    assert bench.profile_record("dam1m", "force_integrate", dict(default, block=256)) is None


def test_byte_models_add_up():
    bench = _bench()
    assert bench.STEP_BYTES == sum(bench.ALGO_BYTES[k] for k in ("density", "force", "integrate")) == 128
    assert bench.ALGO_BYTES["force_integrate"] == 68
    assert bench.MLP_FLOPS_PER_PAIR == 9728
    with open(os.path.join(ROOT, "profiles", "traffic_dam1m.json")) as f:
        assert json.load(f)["tuning"]["block"] == 128
