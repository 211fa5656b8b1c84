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
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracing throughput section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (time-to-first-pass) section")
    ap.add_argument("--pt-config", choices=["c0", "c2", "c3", "c2bdpt"], default=None,
                    help="run only this render configuration in the path-tracing section")
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-bdpt", type=int, default=1, help="1: half of the served requests use BDPT")
    ap.add_argument("--serve-cpu-requests", type=int, default=8,
                    help="serving: time the oracle's first pass of the first N requests on the host cores")
    ap.add_argument("--serve-window-s", type=float, default=60.0,
                    help="cancel requests still running this long after the first arrival (the offered load "
                         "exceeds one GPU, so the backlog is cut once every first pass is in)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (the renderer runs on a single "
              "MI355X here)", file=sys.stderr)
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
    variant = int(scene.info().trace_variant)
    traffic = profiled_traffic(variant, a.rays)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
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

    secondary = {}
    if not a.no_pt:
        secondary.update(pt_section(a, log))
    if not a.no_serve:
        secondary["serving"] = serve_section(a, log)

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
                   "bvh_depth": int(si.bvh_depth), "trace_variant": int(scene.info().trace_variant)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


def profiled_traffic(variant, rays):
    """HBM bytes per launch of this traversal variant at this ray count, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (calibrated against a copy kernel); None if none was profiled."""
    import re

    path = os.path.join(ROOT, "profiles", "closest_hit_c1_icosphere1.31M_16Mrays.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    for m in re.finditer(r"traffic: variant=(\d+) rays=(\d+) fetch_bytes=(\d+) write_bytes=(\d+)", text):
        if int(m.group(1)) == variant and int(m.group(2)) == rays:
            return int(m.group(3)) + int(m.group(4))
    return None


def pt_section(a, log):
    """Path-tracing samples/s on the three render configurations, each set next to the oracle
    (the float64 restatement of the reference estimator) timed on a bounded crop with all host
    threads. Both draw the same counter-RNG stream per (pixel, sample), so at equal samples per
    pixel they have the same expected RMSE: samples/s is the equal-RMSE rate."""
    import numpy as np

    import prtgpu as P
    import scenes

    PT, BD = P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT
    # (name, scene, integrator, spp override, oracle spp of the bounded CPU sample)
    configs = [
        ("C0 default scene 400x225, 64 spp, depth 50 (pinhole)", lambda: scenes.default_scene(), PT, None, 256),
        ("C2 metre Cornell boxes 512x512, 1024 spp, depth 8", lambda: scenes.cornell_metre(), PT, None, 64),
        ("C3 1.8M-triangle glass icosphere + HDR sky 1024x1024, 256 spp, depth 16",
         lambda: scenes.mesh_sky_scene(), PT, None, 8),
        ("C2-BDPT metre Cornell boxes 512x512, bidirectional, 64 spp, depth 8", lambda: scenes.cornell_metre(), BD, 64,
         16),
    ]
    out = []
    cores = min(os.cpu_count() or 1, 16)
    tags = ["c0", "c2", "c3", "c2bdpt"]
    for tag, (name, make, integ, spp_override, c_spp) in zip(tags, configs):
        if a.pt_config and tag != a.pt_config:
            continue
        desc, cfg = make()
        w, h = desc.image_size
        spp = spp_override or cfg["spp"]
        scene = P.Scene(desc, device=0)
        params = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=1, seed=1,
                                   integrator=integ)
        pr = P.ProgressiveRaytracer(scene, params)
        pr.render_samples(1)  # warm-up (kernels, first-touch)
        pr.reset()
        t0 = time.perf_counter()
        st = pr.render_samples(spp)
        wall = time.perf_counter() - t0
        gpu_sps = w * h * spp / wall
        # first-pass latency: one sample per pixel on a fresh film (reference pass 1, progressive.go:104-107)
        pr2 = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp,
                                                              max_passes=7, seed=2, integrator=integ))
        t0 = time.perf_counter()
        pr2.render_pass(1)
        ttfp_ms = (time.perf_counter() - t0) * 1e3
        pr2.close()
        entry = {"config": name, "samples_per_s": round(gpu_sps, 1), "render_s": round(wall, 3),
                 "device_ms": round(st.device_ms, 2), "first_pass_ms": round(ttfp_ms, 2)}
        if not a.no_cpu:
            import oracle as O

            os_ = O.OracleScene(desc)
            # bounded sample: the whole image at a few samples per pixel (~2-6 s of CPU work)
            c0 = time.perf_counter()
            os_.render(params, c_spp, 0, nthreads=cores, splats=integ == BD)
            ct = time.perf_counter() - c0
            cpu_sps = w * h * c_spp / ct
            entry["cpu_samples_per_s"] = round(cpu_sps, 1)
            entry["cpu_sample"] = f"full {w}x{h} image, {c_spp} spp, {cores} threads, {ct:.2f}s"
            entry["speedup_equal_rmse"] = round(gpu_sps / cpu_sps, 1)
        log(f"{name}: {entry}")
        out.append(entry)
        pr.close()
        scene.close()
    return {"pt": out, "cores": cores}


def serve_section(a, log):
    """Poisson arrivals (rate 20/s, seed 0) of progressive path-tracing requests mixing the three
    render scenes x {256x256, 512x512, 1280x720, 1920x1080} x spp log-uniform in [16, 1024] x {PT, BDPT},
    7 passes, through the device-resident serving loop (HBM budget 192 GiB of the 288 GiB, admission
    bounded by that budget). The offered load (~4 G samples/s) is several times what one GPU renders,
    so requests still running after the window are cancelled once every first pass is in.
    Time to first pass = submit -> pass 1 (1 spp) readable, per request, and from admission (film and
    device state set up, the reference's post-setup definition, render.go:75,285) -> pass 1; time to
    target = submit -> the last pass (the requested samples per pixel, so the RMSE the request asked
    for); samples/s counts completed passes. The CPU reference first pass (1 spp of the request's scene,
    size and integrator through the oracle on the host cores, scene build excluded) is timed
    sequentially for the first --serve-cpu-requests requests."""
    import numpy as np

    import prtgpu as P
    import scenes

    rng = np.random.default_rng(0)
    made = [scenes.default_scene(), scenes.cornell_metre(), scenes.mesh_sky_scene()]
    sc = [(P.Scene(d), cfg) for d, cfg in made]
    sizes = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    n = a.serve_requests
    arrivals = np.cumsum(rng.exponential(1 / 20.0, n))
    picks = rng.integers(0, 3, n)
    res = rng.integers(0, 4, n)
    spps = np.exp(rng.uniform(np.log(16), np.log(1024), n)).astype(int)
    integs = rng.integers(0, 2, n) if a.serve_bdpt else np.zeros(n, int)
    srv = P.Server(hbm_budget_bytes=192 << 30, max_active=1024)
    ids, meta = [], {}
    t0 = time.perf_counter()
    ttfp, done_t, finished, pix_samples, admitted = {}, {}, set(), {}, {}
    k = 0
    cancelled = False
    while len(finished) < n:
        now = time.perf_counter() - t0
        while k < n and arrivals[k] <= now:
            scene, cfg = sc[picks[k]]
            w, h = sizes[res[k]]
            p = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], int(spps[k]), max_passes=7, seed=k,
                                  integrator=int(integs[k]))
            rid = srv.submit(scene, p)
            meta[rid] = (k, w * h)
            k += 1
        if not cancelled and k == n and len(ttfp) == n and now > a.serve_window_s:
            for rid in meta:
                if rid not in finished:
                    srv.cancel(rid)
            cancelled = True
        for e in srv.poll(timeout_ms=1):
            if e.type == P.EVENT_ADMITTED:
                admitted[e.request_id] = e.t_ms
            elif e.type == P.EVENT_PASS_COMPLETE:
                pix_samples[e.request_id] = meta[e.request_id][1] * e.target_samples
                if e.pass_number == 1:
                    ttfp[e.request_id] = e.t_ms
            elif e.type in (P.EVENT_DONE, P.EVENT_ERROR, P.EVENT_CANCELLED):
                if e.type == P.EVENT_DONE:
                    done_t[e.request_id] = e.t_ms
                finished.add(e.request_id)
                srv.release(e.request_id)
    makespan = time.perf_counter() - t0
    srv.close()
    t = np.array(list(ttfp.values()))
    tp = np.array([ttfp[r] - admitted.get(r, 0.0) for r in ttfp])
    d = np.array(list(done_t.values()))
    out = {"requests": n, "arrival_rate_per_s": 20, "ttfp_ms_p50": round(float(np.percentile(t, 50)), 2),
           "ttfp_ms_p99": round(float(np.percentile(t, 99)), 2), "ttfp_ms_max": round(float(t.max()), 2),
           "ttfp_post_setup_ms_p50": round(float(np.percentile(tp, 50)), 2),
           "ttfp_post_setup_ms_p99": round(float(np.percentile(tp, 99)), 2),
           "completed": len(d), "cancelled_after_window": n - len(d),
           "time_to_target_s_p50": round(float(np.percentile(d, 50)) / 1e3, 3) if len(d) else None,
           "time_to_target_s_p99": round(float(np.percentile(d, 99)) / 1e3, 3) if len(d) else None,
           "makespan_s": round(makespan, 2),
           "samples_per_s": round(sum(pix_samples.values()) / makespan, 1),
           "integrators": "path tracing and BDPT, 1:1" if a.serve_bdpt else "path tracing only"}
    kcpu = 0 if a.no_cpu else min(a.serve_cpu_requests, n)
    if kcpu > 0:
        import oracle as O

        cores = min(os.cpu_count() or 1, 16)
        factories = [scenes.default_scene, scenes.cornell_metre, scenes.mesh_sky_scene]
        by_k = {meta[r][0]: r for r in meta}
        cpu_ms, gpu_ms = [], []
        for k in range(kcpu):
            desc, cfg = factories[picks[k]]()
            w, h = sizes[res[k]]
            desc.camera.width = w
            desc.camera.aspect_ratio = w / h
            os_ = O.OracleScene(desc)
            p = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], int(spps[k]), max_passes=7, seed=k,
                                  integrator=int(integs[k]))
            c0 = time.perf_counter()
            os_.render(p, 1, 0, nthreads=cores, splats=int(integs[k]) == P.INTEGRATOR_BDPT)
            cpu_ms.append((time.perf_counter() - c0) * 1e3)
            gpu_ms.append(ttfp[by_k[k]])
        out["cpu_first_pass"] = {"requests": kcpu, "cores": cores, "kind": "port",
                                 "cpu_ms_p50": round(float(np.percentile(cpu_ms, 50)), 1),
                                 "cpu_ms_max": round(float(max(cpu_ms)), 1),
                                 "gpu_ttfp_ms_p50_same_requests": round(float(np.percentile(gpu_ms, 50)), 2),
                                 "speedup_p50": round(float(np.percentile(np.array(cpu_ms) / np.array(gpu_ms), 50)), 1)}
    log(f"serving: {out}")
    return out


if __name__ == "__main__":
    main()
