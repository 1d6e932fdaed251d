"""bench.py's contract pieces that need no GPU: the one-GPU refusal and the
HBM traffic figure it reads from the committed counter summary."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU" in r.stderr


def test_traffic_comes_from_the_committed_counter_summary():
    t = _bench().traffic_from_profiles()
    assert t is not None and 1e5 < t < 1e8   # bytes per step launch
