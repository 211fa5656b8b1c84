"""The C-ABI library loads and exports every function declared in include/*.h; host-only entry
points behave as documented. No GPU compute here."""
import ctypes as C
import os
import re

import pytest

import prtgpu as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\**\s+\**(prt_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = P.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(P.EXPORTS) == names


def test_version_and_last_error():
    assert P.lib().prt_version() >= 10000
    rc = P.lib().prt_samples_for_pass(None, 1, None)
    assert rc == 1
    assert b"NULL" in P.lib().prt_last_error()


def test_pass_schedule_matches_reference():
    # progressive_test.go:8-34
    p = P.sampling_params(400, 225, 5, 5, 50, max_passes=7, initial_samples=1)
    assert [P.get_samples_for_pass(p, k) for k in range(1, 8)] == [1, 9, 17, 25, 33, 41, 50]
    p1 = P.sampling_params(400, 225, 5, 5, 64, max_passes=1)
    assert P.get_samples_for_pass(p1, 1) == 64
    with pytest.raises(P.PrtError):
        P.get_samples_for_pass(p, 0)


def test_workload_bad_arguments_raise():
    with pytest.raises(P.PrtError):
        P.icosphere(0)


@pytest.mark.skipif(P.device_count() > 0, reason="a GPU is present")
def test_scene_create_without_device_fails_cleanly():
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    with pytest.raises(P.PrtError) as e:
        P.Scene(s)
    assert e.value.code == 4
