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
    vals = [b.profiled_traffic(v, rays) for v in range(8)]
    assert all(isinstance(v, int) and 1e9 < v < 1e10 for v in vals), vals
    assert b.profiled_traffic(2, rays + 1) is None and b.profiled_traffic(99, rays) is None


def test_host_cpu_and_trimmed_mean():
    b = _bench()
    h = b.host_cpu()
    assert 1 <= h["threads"] <= len(os.sched_getaffinity(0)) and h["nproc"] >= 1 and h["cpu_model"]
    assert b.trimmed_mean([5.0, 1.0, 2.0, 3.0, 100.0]) == 10.0 / 3.0  # min and max dropped
    assert b.trimmed_mean([2.0, 4.0]) == 3.0


def test_fixture_loader_checks_the_pin():
    """load_fixture returns the converged image only when its SHA-256 matches the manifest; RMSE against
    the fixture of the fixture's own mean is zero."""
    import json

    b = _bench()
    man = os.path.join(ROOT, "tests", "golden", "converged.json")
    if not os.path.exists(man):
        return
    for name in json.load(open(man)):
        fx = b.load_fixture(name)
        assert fx is not None, name
        assert fx["mean"].shape == fx["var"].shape and fx["mean"].shape[2] == 3
        assert b.rmse(fx["mean"], fx) == 0.0
    assert b.load_fixture("no_such_fixture") is None
