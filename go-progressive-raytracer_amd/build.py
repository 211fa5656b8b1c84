"""Builds libprt.so (host C++ + gfx950 HIP kernels) in-tree with hipcc.

Usage: python build.py [--jobs N]. Every call recompiles all sources.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libprt.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["api.cpp", "bvh_build.cpp", "workload.cpp", "ply.cpp", "pbrt.cpp", "server.cpp", "sse.cpp", "scene.hip", "pt.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PRT_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-c", *COMMON]
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += [os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    return src, r.returncode, errs, obj


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    objs = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, errs, obj in ex.map(_compile, SOURCES):
            if rc != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n{errs}")
            if verbose and errs.strip():
                print(errs, file=sys.stderr)
            objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp, "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
