"""Cylinders and cones on the HIP path against the float64 oracle (pkg/geometry/cylinder.go,
cone.go): closest/any hit on random rays, path-tracing and BDPT parity on the reference's
cylinder and cone test scenes, the part code in prt_hit.u, and NewCone's argument checks.

Bars as for the other scenes: closest hit >= 99.99 % primitive-id match; path tracing >= 99 % of
per-sample radiances within 1e-3 relative (floor 1e-3); BDPT >= 98 % camera-side and >= 97 % of
splat pixels within 1e-2 relative (floor 0.1), splat totals within 1 %.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))


def _rays(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, P.RAY_DTYPE)
    r["org"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    r["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmin"] = 0.001
    r["tmax"] = np.inf
    return r


def _base():
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    return s


def test_closest_hit_cylinders_and_cones():
    s = _base()
    m = s.lambertian((0.5, 0.5, 0.5))
    rng = np.random.default_rng(8)
    for i in range(60):
        base = rng.uniform(-2, 2, 3)
        axis = rng.normal(size=3) if rng.uniform() < 0.7 else np.eye(3)[rng.integers(3)]
        axis /= np.linalg.norm(axis)
        top = base + axis * rng.uniform(0.3, 1.5)
        if i % 2 == 0:
            s.cylinder(tuple(base), tuple(top), float(rng.uniform(0.1, 0.5)), bool(rng.integers(2)), m)
        else:
            rb = float(rng.uniform(0.15, 0.6))
            rt = 0.0 if rng.uniform() < 0.5 else float(rng.uniform(0.02, rb * 0.8))
            s.cone(tuple(base), rb, tuple(top), rt, bool(rng.integers(2)), m)
    scene, os_ = P.Scene(s), O.OracleScene(s)
    rays = _rays([-3, -3, -3], [3, 3, 3], 200000, 31)
    gpu = scene.intersect_closest(rays)
    ref_t, ref_p, _, _ = os_.intersect(rays)
    match = gpu["prim_id"] == ref_p
    both = (gpu["prim_id"] >= 0) & (ref_p >= 0)
    with np.errstate(invalid="ignore"):
        tie = both & (np.abs(gpu["t"] - ref_t) <= 1e-4 * np.maximum(1.0, ref_t))
    print(f"axial: hits {(ref_p >= 0).sum()}, match {match.mean():.6f}, unexplained {(~match & ~tie).sum()}")
    assert (ref_p >= 0).sum() > 50000
    assert match.mean() >= 0.9999 and (~match & ~tie).sum() <= 2e-5 * len(rays)
    hm = match & (ref_p >= 0)
    assert np.all(np.abs(gpu["t"][hm] - ref_t[hm]) <= 2e-4 * np.maximum(1.0, ref_t[hm]))
    assert (scene.intersect_any(rays).astype(bool) == (ref_p >= 0)).mean() >= 0.9999


@pytest.mark.parametrize("sec,kind", [("cylinder_hit", 4), ("cone_hit", 5)])
def test_reference_cases_and_part_codes(sec, kind):
    """Every ray of the reference's cylinder/cone tests, one scene per case; prt_hit.u holds the part."""
    for case in G[sec]["cases"]:
        s = _base()
        m = s.lambertian((0.5, 0.5, 0.5))
        g = case["g"]
        if kind == 4:
            s.cylinder(g[0:3], g[3:6], g[6], bool(g[7]), m)
        else:
            s.cone(g[0:3], g[3], g[4:7], g[7], bool(g[8]), m)
        r = np.zeros(1, P.RAY_DTYPE)
        r["org"], r["dir"], r["tmin"], r["tmax"] = case["o"], case["d"], case["tmin"], case["tmax"]
        h = P.Scene(s).intersect_closest(r)[0]
        assert (h["prim_id"] == 0) == case["hit"], case["name"]
        if case["hit"] and "p" in case:
            p = np.array(case["o"]) + h["t"] * np.array(case["d"])
            assert np.allclose(p, case["p"], atol=1e-5), case["name"]
            axis = np.array(g[3:6] if kind == 4 else g[4:7], float) - np.array(g[0:3], float)
            axis /= np.linalg.norm(axis)
            on_cap = abs(np.dot(p - np.array(g[0:3]), axis)) < 1e-5 or \
                abs(np.dot(p - np.array(g[3:6] if kind == 4 else g[4:7]), axis)) < 1e-5
            assert (h["u"] != 0) == on_cap, case["name"]


@pytest.mark.parametrize("case", G["axial_boxes"]["cone_invalid"], ids=lambda c: c["name"])
def test_cone_arguments_rejected(case):
    s = _base()
    s.cone((0, 0, 0), case["base_radius"], case["top"], case["top_radius"], False, s.lambertian((0.5, 0.5, 0.5)))
    with pytest.raises(P.PrtError):
        P.Scene(s)


def _params(desc, cfg, spp, seed, integrator):
    w, h = desc.image_size
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, seed=seed, integrator=integrator)


def _agree(g, r, rel, floor):
    d = np.abs(g.astype(np.float64) - r)
    return np.all(d <= rel * np.maximum(np.abs(r), floor), axis=2).mean()


def _first_sample(desc, cfg, seed, integrator):
    w, h = desc.image_size
    pr = P.ProgressiveRaytracer(P.Scene(desc), _params(desc, cfg, 1, seed, integrator))
    pr.render_samples(1)
    mean, _, _, cnt = pr.read()
    assert np.all(cnt == 1)
    sp = pr.splats() if integrator == P.INTEGRATOR_BDPT else np.zeros((h, w, 3))
    ys, xs = np.mgrid[0:h, 0:w]
    pix = np.stack([xs.ravel(), ys.ravel()], 1)
    out, rsp = O.OracleScene(desc).render_samples(_params(desc, cfg, 1, seed, integrator), pix, np.zeros(len(pix)),
                                                  splats=True)
    return mean.astype(np.float64) - sp, out.reshape(h, w, 3), sp, rsp


def _cones_separated():
    """The cone scene with its two coincident cap pairs (the glass cone's base on the blue frustum's
    top cap, a glass cone's base on the ground) 1e-4 apart."""
    s, cfg = scenes.cone_scene(width=128)
    s.cones[3].base_center[1] += 1e-4
    s.cones[5].base_center[1] += 1e-4
    s.cones[5].top_center[1] += 1e-4
    return s, cfg


SCENES = {"cylinders": lambda: scenes.cylinder_scene(width=128), "cones": lambda: scenes.cone_scene(width=128),
          "cones_separated": _cones_separated}
# The reference cone scene puts cap discs exactly on other surfaces: which of two surfaces at the
# same distance a ray takes depends on traversal order and rounding, and the float64 oracle itself
# moves 2.5 % of its pixels (and its mean by 1.5 %) when one cone is lowered by 1e-6. Its bar is
# set to that sensitivity; with the caps 1e-4 apart the usual bar holds.
BARS = {"cones": (0.94, 0.04)}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_path_tracing_parity(name):
    s, cfg = SCENES[name]()
    g, r, _, _ = _first_sample(s, cfg, 4, P.INTEGRATOR_PATH)
    frac = _agree(g, r, 1e-3, 1.0)
    bar, mean_tol = BARS.get(name, (0.99, 0.02))
    print(f"{name} PT: agreement {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f}")
    assert frac >= bar
    assert abs(g.mean() - r.mean()) <= mean_tol * max(r.mean(), 1e-3) + 1e-3


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bdpt_parity(name):
    s, cfg = SCENES[name]()
    g, r, gs, rs = _first_sample(s, cfg, 6, P.INTEGRATOR_BDPT)
    frac = _agree(g, r, 1e-3, 1.0)
    sfrac = _agree(gs, rs, 1e-2, 1e-1)
    bar, mean_tol = BARS.get(name, (0.98, 0.02))
    print(f"{name} BDPT: camera-side {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f}; "
          f"splat total gpu {gs.sum():.5f} oracle {rs.sum():.5f}; splat pixels {sfrac:.5f}")
    assert frac >= bar
    assert abs(g.mean() - r.mean()) <= mean_tol * max(r.mean(), 1e-3) + 1e-3
    assert abs(gs.sum() - rs.sum()) <= 0.01 * abs(rs.sum()) + 1e-3
    assert sfrac >= 0.97
