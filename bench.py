"""Benchmark: closest-hit throughput (Mrays/s) of the HIP traversal on one MI355X.

Workload (one "step"): 16,777,216 rays (origins uniform in the mesh box scaled 2x, directions
uniform on the sphere, tmin 0.001, tmax inf) against a seeded displaced geodesic icosphere of
1,310,720 triangles (an icosahedron subdivided 8 times), binned-SAH 4-wide BVH, leaf <= 4.
Rays, BVH and hit buffer are resident in HBM before timing starts.

Prints one JSON line (see the keys below). --gpus above 1 is not supported (one GPU).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "go-progressive-raytracer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--frequency", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (this build runs on a single "
              "MI355X; sharding across GPUs is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import devmem as D
    import prtgpu as P
    import scenes

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if a.calibrate_copy_mib:
        nb = a.calibrate_copy_mib << 20
        src = D.DeviceBuffer.from_array(np.ones(nb // 4, np.float32))
        dst = D.DeviceBuffer(nb)
        for _ in range(2):
            D.copy_d2d(dst, src, nb)
        D.synchronize()
        log(f"copied {nb} bytes device-to-device twice")
        return

    t0 = time.time()
    desc, info = scenes.mesh_scene(frequency=a.frequency)
    v = info["vertices"]
    lo, hi = v.min(0), v.max(0)
    rays = P.random_rays(lo, hi, a.rays, scale=2.0, seed=0)
    log(f"workload generated in {time.time() - t0:.1f}s ({info['faces'].shape[0]} triangles, {a.rays} rays)")
    t0 = time.time()
    scene = P.Scene(desc, device=0)
    si = scene.info()
    log(f"scene built in {time.time() - t0:.1f}s: {si.bvh_nodes} nodes, depth {si.bvh_depth}")

    # inputs resident in HBM before timing: rays (32 B each), hit buffer (16 B each)
    d_rays = D.DeviceBuffer.from_array(rays)
    d_hits = D.DeviceBuffer(a.rays * 16)
    stream = D.Stream()

    def step():
        scene.intersect_closest_device(d_rays.ptr, a.rays, d_hits.ptr, stream.handle)

    for _ in range(a.warmup):
        step()
    D.synchronize()
    evs = [(D.Event(), D.Event()) for _ in range(a.steps)]
    D.synchronize()
    t_start = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    D.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = [e0.elapsed_ms(e1) for e0, e1 in evs]
    avg_ms = float(np.mean(kernel_ms))
    value = a.rays * a.steps / wall / 1e6

    # parity spot check of the timed output against a fresh host call
    hits = d_hits.to_array(P.HIT_DTYPE, a.rays)
    sub = rays[: 1 << 16]
    ref = scene.intersect_closest(sub)
    assert np.array_equal(hits[: 1 << 16]["prim_id"], ref["prim_id"])

    # algorithmic bytes per ray: 32 ray + 16 hit + 32 per box test + 36 per triangle test,
    # traversal counts from the same kernel's counting variant on a fixed 65,536-ray sample
    nodes, boxes, prims = scene.traversal_counts(rays[: 1 << 16])
    bytes_per_ray = 48.0 + 32.0 * boxes.mean() + 36.0 * prims.mean()
    achieved = bytes_per_ray * a.rays / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_ray": round(bytes_per_ray, 1), "box_tests_per_ray": round(float(boxes.mean()), 2),
                "tri_tests_per_ray": round(float(prims.mean()), 2),
                "kernel_ms": round(avg_ms, 3)}

    cpu = None
    if not a.no_cpu:
        import oracle as O

        cores = min(os.cpu_count() or 1, 16)
        os_ = O.OracleScene(desc)
        done, t_c = 0, 0.0
        chunk = 1 << 20
        while done < a.rays and t_c < a.cpu_seconds:
            c0 = time.perf_counter()
            os_.intersect(rays[done:done + chunk], nthreads=cores)
            t_c += time.perf_counter() - c0
            done += min(chunk, a.rays - done)
        cpu = {"value": round(done / t_c / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"first {done} of the {a.rays} rays through the float64 oracle (reference midpoint BVH, "
                         f"leaf 8, Moller-Trumbore), {cores} threads, {t_c:.1f}s"}

    out = {
        "metric": "Mrays/s closest-hit",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"closest-hit microbench: displaced icosphere {info['faces'].shape[0]} triangles "
                               f"(frequency {a.frequency}), {a.rays} rays uniform in 2x AABB, dirs uniform on S2, "
                               f"binned-SAH 4-wide BVH leaf<=4",
                   "triangles": int(info["faces"].shape[0]), "rays": a.rays, "bvh_nodes": int(si.bvh_nodes),
                   "bvh_depth": int(si.bvh_depth)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
