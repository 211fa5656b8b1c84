"""Host side of bench.py's contract (no GPU): one GPU only, and the roofline traffic comes from the
committed rocprofv3 counter lines of every closest-hit variant."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_profiled_traffic_of_every_variant():
    b = _bench()
    rays = 16777216
    vals = [b.profiled_traffic(v, rays) for v in range(5)]
    assert all(isinstance(v, int) and 1e9 < v < 1e10 for v in vals), vals
    assert b.profiled_traffic(2, rays + 1) is None and b.profiled_traffic(99, rays) is None
