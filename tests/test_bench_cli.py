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


def test_committed_profile_records_are_keyed_and_sourced():
    """The bench line's roofline.traffic and counter-derived utilisations come
    from profiles/traffic_<workload>_<tuning tag>.json; a record applies only to the
    kernel, settled state and tuning it was measured at, and names the
    committed rocprofv3 summary it was computed from."""
    import glob
    bench = _bench()
    recs = sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")))
    assert recs
    for path in recs:
        rec = json.load(open(path))
        for key in ("kernel", "workload", "settle", "tuning", "traffic_bytes_per_launch", "source",
                    "rocprof_avg_ms_timed_launches", "valu_busy"):
            assert key in rec, (path, key)
        assert os.path.basename(path) == f"traffic_{rec['workload']}_{bench.profile_tag(rec['tuning'])}.json"
        assert rec["traffic_bytes_per_launch"] > 0
        assert 0.0 < rec["valu_busy"] <= 1.0
        assert os.path.exists(os.path.join(ROOT, rec["source"]))
        assert bench.profile_record(rec["workload"], rec["kernel"], rec["tuning"], rec["settle"]) == rec
        other = dict(rec["tuning"], block=384)
        assert bench.profile_record(rec["workload"], rec["kernel"], other, rec["settle"]) is None
        assert bench.profile_record(rec["workload"], rec["kernel"], rec["tuning"], rec["settle"] + 1) is None


def test_byte_models_add_up():
    bench = _bench()
    assert bench.STEP_BYTES == sum(bench.ALGO_BYTES[k] for k in ("density", "force", "integrate")) == 128
    assert bench.ALGO_BYTES["force_integrate"] == 68
    assert bench.MLP_FLOPS_PER_PAIR == 9728
