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
    ap.add_argument("--cpu-repeats", type=int, default=5,
                    help="repeats of each CPU baseline sample (trimmed mean: min and max dropped)")
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
                "achieved_basis": "algorithmic bytes (32 B ray + 16 B hit + 32 B per box test + 36 B per "
                                  "triangle test) per launch / kernel time; most node reads hit in L2",
                "bytes_per_ray": round(bytes_per_ray, 1), "box_tests_per_ray": round(float(boxes.mean()), 2),
                "tri_tests_per_ray": round(float(prims.mean()), 2),
                "kernel_ms": round(avg_ms, 3)}
    if traffic:
        hbm = traffic / (avg_ms * 1e-3) / 1e9
        roofline.update({"traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel variant at this "
                                           "ray count, calibrated against a copy kernel (profiles/)",
                         "hbm_measured_GBps": round(hbm, 1), "hbm_measured_frac": round(hbm / HBM_PEAK_GBS, 4),
                         "limiter": "dependent node fetches at 6 waves/SIMD (80 VGPRs) with the VALU ~68 % busy "
                                    "(a 4-wide node step is ~125 VALU instructions, a leaf step ~100; leaves "
                                    "gated to 2/8 of a wave); "
                                    f"measured HBM traffic runs at {hbm / 1e3:.2f} TB/s"})

    cpu = None
    if not a.no_cpu:
        import oracle as O

        host = host_cpu()
        cores = host["threads"]
        os_ = O.OracleScene(desc)
        res = {}

        def run_all():
            res["out"] = os_.intersect(rays, nthreads=cores)

        reps = [timed(run_all) for _ in range(a.cpu_repeats)]
        n1 = min(a.rays, 1 << 20)
        t1 = timed(lambda: os_.intersect(rays[:n1], nthreads=1))
        tm = trimmed_mean(reps)
        cpu = {"value": round(a.rays / tm / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"all {a.rays} rays through the float64 oracle (reference midpoint BVH, leaf 8, "
                         f"Moller-Trumbore) on {cores} threads, trimmed mean of {len(reps)} runs "
                         f"({', '.join(f'{t:.2f}' for t in reps)} s)",
               "one_thread_value": round(n1 / t1 / 1e6, 3), "one_thread_sample": f"first {n1} rays, {t1:.2f}s",
               **{k: v for k, v in host.items() if k != "threads"},
               "parity": hit_parity(hits, res["out"][0], res["out"][1])}

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


def hit_parity(hits, ref_t, ref_p):
    """The timed GPU hits of every ray against the float64 oracle's: primitive-id match rate, mismatches whose
    two hits lie at the same distance (1e-4 relative: ties at shared edges and vertices), and the relative
    hit-t error where both hit the same primitive."""
    import numpy as np

    gp, gt = hits["prim_id"].astype(np.int64), hits["t"].astype(np.float64)
    match = gp == ref_p
    both = (gp >= 0) & (ref_p >= 0)
    with np.errstate(invalid="ignore"):  # misses hold t = inf on both sides
        tie = both & ~match & (np.abs(gt - ref_t) <= 1e-4 * np.maximum(1.0, np.abs(ref_t)))
    same = match & both
    rel = np.abs(gt[same] - ref_t[same]) / np.maximum(np.abs(ref_t[same]), 1e-30)
    return {"rays": int(len(gp)), "prim_id_match": round(float(match.mean()), 7),
            "mismatches": int((~match).sum()), "mismatches_at_equal_t": int(tie.sum()),
            "hit_t_rel_err_max": float(rel.max()) if rel.size else 0.0,
            "hit_t_rel_err_mean": float(rel.mean()) if rel.size else 0.0,
            "reference": "float64 oracle (Moller-Trumbore over a midpoint BVH) on the same rays"}


def host_cpu():
    """The host the CPU baselines run on: all threads this process may use (affinity mask, bounded by a
    cgroup CPU quota if one is set), the CPU model and the machine's logical CPU count."""
    threads = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
            threads = max(1, min(threads, int(quota)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": threads, "cpu_model": model, "nproc": os.cpu_count(), "cgroup_cpu_quota": quota}


def timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def trimmed_mean(xs):
    """Mean without the fastest and slowest run (benchmark.sh:158-175 drops min and max)."""
    xs = sorted(xs)
    return sum(xs[1:-1]) / (len(xs) - 2) if len(xs) > 2 else sum(xs) / len(xs)


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


def load_fixture(name):
    """Converged 65,536-spp oracle image of a configuration (tests/golden/make_converged.py): per-pixel mean
    and one-sample variance over its pixel rectangle, or None if it has not been generated."""
    import hashlib

    import numpy as np

    gold = os.path.join(ROOT, "tests", "golden")
    try:
        man = json.load(open(os.path.join(gold, "converged.json")))[name]
        path = os.path.join(gold, "converged", name + ".npz")
        if hashlib.sha256(open(path, "rb").read()).hexdigest() != man["sha256"]:
            return None
        z = np.load(path)
        return {"mean": z["mean"].astype(np.float64), "var": z["var"].astype(np.float64), **man}
    except (OSError, KeyError, ValueError):
        return None


def rmse(img, fx):
    import numpy as np

    return float(np.sqrt(np.mean((np.asarray(img, np.float64) - fx["mean"]) ** 2)))


def pt_section(a, log):
    """Path tracing and BDPT on the render configurations: GPU samples/s at full size; per-kernel device
    time and (path tracing) traversal counts of the trace kernel from prt_film_probe, giving rays/s and
    the algorithmic-byte fraction of the HBM roofline; time to the first pass; and, against the
    configuration's converged 65,536-spp oracle image, the measured RMSE of the GPU image and of the
    oracle's CPU image (the float64 restatement of the reference estimator, timed on all host threads),
    from which the speedup at equal RMSE follows:
        CPU spp needed for the GPU's RMSE  n_c' = n_c (RMSE_c^2 - e^2) / (RMSE_g^2 - e^2),
        e^2 = the fixture's own error (mean variance / 65,536),
        speedup = (n_c' x pixels / CPU samples/s) / (n_g x pixels / GPU samples/s)."""
    import numpy as np

    import prtgpu as P
    import scenes

    PT, BD = P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT
    # (tag, name, scene, integrator, spp override, fixture, CPU samples per pixel of one timed repeat: the
    # configuration's own, so that GPU and CPU errors are compared at equal samples per pixel and the
    # rare high-variance paths of heavy-tailed scenes enter both)
    configs = [
        ("c0", "C0 default scene 400x225, 64 spp, depth 50 (pinhole)", lambda **k: scenes.default_scene(**k), PT,
         None, "c0_default_400x225", 64),
        ("c2", "C2 metre Cornell boxes 512x512, 1024 spp, depth 8", lambda **k: scenes.cornell_metre(**k), PT, None,
         "c2_cornell_metre_128", 1024),
        ("c3", "C3 1.8M-triangle glass icosphere + HDR sky 1024x1024, 256 spp, depth 16",
         lambda **k: scenes.mesh_sky_scene(**k), PT, None, "c3_mesh_sky_crop128", 256),
        ("c2bdpt", "C2-BDPT metre Cornell boxes 512x512, bidirectional, 64 spp, depth 8",
         lambda **k: scenes.cornell_metre(**k), BD, 64, "c2_bdpt_cornell_metre_128", 64),
    ]
    out = []
    host = host_cpu()
    cores = host["threads"]
    for tag, name, make, integ, spp_override, fx_name, c_spp in configs:
        if a.pt_config and tag != a.pt_config:
            continue
        desc, cfg = make()
        w, h = desc.image_size
        spp = spp_override or cfg["spp"]
        scene = P.Scene(desc, device=0)
        params = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=1, seed=1,
                                   integrator=integ)
        pr = P.ProgressiveRaytracer(scene, params)
        # warm-up (kernels, first touch) until the scene's autotuner has settled: it times launches inside a
        # chunk other than the film's first and a later chunk reads them
        tuned = lambda: (scene.info().pt_trace_leaf_gate if integ == PT else scene.info().bdpt_trace_leaf_gate) >= 0
        for _ in range(4):
            pr.render_samples(spp)
            pr.reset()
            if tuned():
                break
        t0 = time.perf_counter()
        st = pr.render_samples(spp)
        wall = time.perf_counter() - t0
        gpu_sps = w * h * spp / wall
        img = pr.read()[0]
        # per-kernel device time, then traversal counts, of the same render (renders are deterministic)
        pr.reset()
        pr.probe(P.PROBE_TIME)
        pr.render_samples(spp)
        kt = pr.probe(P.PROBE_COUNT if integ == PT else 0)
        kernels = {k: round(v, 2) for k, v in kt.items() if k.endswith("_ms") and v > 0}
        entry = {"config": name, "samples_per_s": round(gpu_sps, 1), "render_s": round(wall, 3),
                 "device_ms": round(st.device_ms, 2), "kernel_ms": kernels}
        info = scene.info()  # the choices the per-scene autotuners made on the first large chunk

        def traversal(g):  # 0-8: BVH with a leaf gate of g/8 of a wave; 9: flat (every primitive tested)
            return "flat" if g == 9 else ("bvh" if g >= 0 else "untuned")

        gate = int(info.pt_trace_leaf_gate if integ == PT else info.bdpt_trace_leaf_gate)
        entry["autotune"] = {"traversal": traversal(gate), "leaf_gate_eighths": gate if 0 <= gate <= 8 else None}
        if integ == PT:
            entry["autotune"]["trace_waves_per_simd"] = int(info.pt_trace_waves)
        if integ == PT:
            pr.reset()
            pr.render_samples(spp)
            c = pr.probe(0)
            rays = c["extension_rays"] + c["shadow_rays"]
            boxes = c["extension_box_tests"] + c["shadow_box_tests"]
            prims = c["extension_prim_tests"] + c["shadow_prim_tests"]
            flat = traversal(gate) == "flat"
            # a flat traversal reads each primitive record once per 64-ray wave batch (the same address in
            # every lane), a BVH traversal per ray
            tb = 48.0 * rays + 32.0 * boxes + (36.0 / 64.0 if flat else 36.0) * prims
            tr_s = kt["trace_ms"] * 1e-3
            entry["trace"] = {"rays": int(rays), "extension_rays": int(c["extension_rays"]),
                              "shadow_rays": int(c["shadow_rays"]), "launches": int(c["trace_launches"]),
                              "rays_per_launch": round(rays / max(1, c["trace_launches"]), 1),
                              "grays_per_s": round(rays / tr_s / 1e9, 3),
                              "box_tests_per_ray": round(boxes / rays, 2), "prim_tests_per_ray": round(prims / rays, 2)}
            entry["roofline"] = {"kernel": "k_trace", "bound": "hbm", "achieved": round(tb / tr_s / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(tb / tr_s / 1e9 / HBM_PEAK_GBS, 4),
                                 "achieved_basis": "algorithmic bytes of all trace launches of the render (48 B per "
                                                   "ray + 32 B per box test + 36 B per primitive test, per 64-ray "
                                                   "wave batch in a flat traversal) / their summed device time"}
            if flat:
                entry["roofline"]["limiter"] = ("flat traversal: every primitive tested by every lane in a wave-uniform "
                                                "loop (records read once per wave), so the VALU, not HBM, bounds it")
        # first-pass latency: one sample per pixel on a fresh film (reference pass 1, progressive.go:104-107),
        # the median over five fresh films (the film is created before the clock starts)
        fps = []
        for k in range(5):
            pr2 = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"],
                                                                  spp, max_passes=7, seed=2 + k, integrator=integ))
            t0 = time.perf_counter()
            pr2.render_pass(1)
            fps.append((time.perf_counter() - t0) * 1e3)
            pr2.close()
        entry["first_pass_ms"] = round(float(np.median(fps)), 2)
        entry["first_pass_ms_min_max"] = [round(min(fps), 2), round(max(fps), 2)]
        fx = load_fixture(fx_name)
        if fx is not None:
            # mean squared error against the converged image over independent GPU renders at the
            # configuration's spp (seeds 1..k): single images of heavy-tailed scenes scatter widely
            x0, y0, x1, y1 = fx["rect"]
            same = [w, h] == fx["image_size"]
            if same:
                sc2, d2 = scene, desc
            else:  # the fixture is of a smaller image of the same scene
                d2 = make(width=fx["image_size"][0])[0]
                sc2 = P.Scene(d2, device=0)
            fw, fh = d2.image_size
            k_seeds = 4 if fw * fh * spp > (1 << 26) else 8
            mses = []
            for seed in range(1, k_seeds + 1):
                if same and seed == 1:
                    g_img = img[y0:y1, x0:x1]
                else:
                    f2 = P.ProgressiveRaytracer(sc2, P.sampling_params(fw, fh, cfg["max_depth"], cfg["rr_min_bounces"],
                                                                       spp, max_passes=1, seed=seed, integrator=integ))
                    f2.render_samples(spp)
                    g_img = f2.read()[0][y0:y1, x0:x1]
                    f2.close()
                mses.append(rmse(g_img, fx) ** 2)
            if not same:
                sc2.close()
            n_fix = fx["fixture_spp"]
            e2 = float(fx["var"].mean()) / n_fix
            mse_g = float(np.mean(mses))
            entry["rmse"] = {"fixture": fx_name, "fixture_spp": n_fix, "pixels": int((x1 - x0) * (y1 - y0)),
                             "gpu": round(mse_g ** 0.5, 6), "gpu_spp": spp, "gpu_renders": k_seeds,
                             "gpu_per_render": [round(m ** 0.5, 6) for m in mses],
                             "gpu_expected_unbiased": round(float(np.sqrt(fx["var"].mean() * (1 / spp + 1 / n_fix))), 6)}
        if not a.no_cpu:
            import oracle as O

            if fx is not None:
                dc = desc if [w, h] == fx["image_size"] else make(width=fx["image_size"][0])[0]
                rect = tuple(fx["rect"])
            else:  # no converged image: a bounded sample, the centre 128x128 pixels of the image
                cw, ch = min(w, 128), min(h, 128)
                dc, rect = desc, ((w - cw) // 2, (h - ch) // 2, (w + cw) // 2, (h + ch) // 2)
            os_ = O.OracleScene(dc)
            rw, rh = rect[2] - rect[0], rect[3] - rect[1]
            tile = 64  # the reference's tile size, halved until every host thread has tiles to render
            while tile > 8 and ((rw + tile - 1) // tile) * ((rh + tile - 1) // tile) < 4 * cores:
                tile //= 2
            pc = P.sampling_params(dc.image_size[0], dc.image_size[1], cfg["max_depth"], cfg["rr_min_bounces"],
                                   c_spp, max_passes=1, seed=7, integrator=integ, tile_size=tile)
            npx = rw * rh
            reps, cmses = [], []
            for k in range(a.cpu_repeats):
                t0 = time.perf_counter()
                r = os_.render(pc, c_spp, k * c_spp, rect=rect, nthreads=cores, splats=integ == BD)
                reps.append(time.perf_counter() - t0)
                part = r[0]
                if integ == BD:
                    x0, y0, x1, y1 = rect
                    part = part + r[3][y0:y1, x0:x1]
                if fx is not None:
                    cmses.append(rmse(part / c_spp, fx) ** 2)
            spp1 = max(1, c_spp // cores)
            t1 = timed(lambda: os_.render(pc, spp1, 0, rect=rect, nthreads=1, splats=integ == BD))
            cpu_sps = npx * c_spp / trimmed_mean(reps)
            entry["cpu_samples_per_s"] = round(cpu_sps, 1)
            entry["cpu_one_thread_samples_per_s"] = round(npx * spp1 / t1, 1)
            entry["cpu_sample"] = (f"{rw}x{rh} pixels of the {dc.image_size[0]}x{dc.image_size[1]} image in "
                                   f"{tile}x{tile} tiles, {c_spp} spp per run on {cores} threads, trimmed mean of "
                                   f"{len(reps)} runs ({', '.join(f'{t:.2f}' for t in reps)} s); one thread: "
                                   f"{spp1} spp, {t1:.2f} s")
            if dc is desc and rect != (0, 0, w, h):
                # the converged image is a crop: its pixels (the glass mesh) cost more than the image's average,
                # so throughput is timed on the whole image, as the GPU's is; the crop sets the RMSE per spp
                pf = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], 1, max_passes=1, seed=7,
                                       integrator=integ, tile_size=64)
                wreps = [timed(lambda: os_.render(pf, 1, k, nthreads=cores, splats=integ == BD)) for k in range(3)]
                cpu_sps = w * h / trimmed_mean(wreps)
                entry["cpu_crop_samples_per_s"] = entry["cpu_samples_per_s"]
                entry["cpu_samples_per_s"] = round(cpu_sps, 1)
                entry["cpu_sample"] += (f"; throughput on the whole {w}x{h} image at 1 spp on {cores} threads, trimmed "
                                        f"mean of 3 runs ({', '.join(f'{t:.2f}' for t in wreps)} s)")
            entry["speedup_equal_spp"] = round(gpu_sps / cpu_sps, 1)
            if fx is not None:
                mse_c = float(np.mean(cmses))
                entry["rmse"].update({"cpu": round(mse_c ** 0.5, 6), "cpu_spp": c_spp, "cpu_renders": len(cmses),
                                      "cpu_per_render": [round(m ** 0.5, 6) for m in cmses]})
                if mse_g > e2 and mse_c > e2:
                    need = c_spp * (mse_c - e2) / (mse_g - e2)
                    entry["rmse"]["cpu_spp_for_gpu_rmse"] = round(need, 1)
                    t_cpu = need * npx / cpu_sps
                    t_gpu = spp * npx / gpu_sps
                    entry["speedup_equal_rmse"] = round(t_cpu / t_gpu, 1)
        log(f"{name}: {entry}")
        out.append(entry)
        pr.close()
        scene.close()
    return {"pt": out, "cpu_host": host}


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
    names = ["default", "cornell_metre", "mesh_sky"]
    slow = sorted(ttfp, key=lambda r: -ttfp[r])[:5]
    out["ttfp_slowest"] = [{"scene": names[picks[meta[r][0]]], "size": "x".join(map(str, sizes[res[meta[r][0]]])),
                            "integrator": "bdpt" if integs[meta[r][0]] else "pt", "ttfp_ms": round(ttfp[r], 2),
                            "post_setup_ms": round(ttfp[r] - admitted.get(r, 0.0), 2)} for r in slow]
    kcpu = 0 if a.no_cpu else min(a.serve_cpu_requests, n)
    if kcpu > 0:
        import oracle as O

        cores = host_cpu()["threads"]
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
