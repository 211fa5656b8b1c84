"""Pins the float64 oracle against the reference's own test vectors (tests/golden/reference_cases.json)
and against closed forms. CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))


def hit(kind, geom, o, d, tmin, tmax):
    out = np.zeros(10)
    h = O.lib().orc_shape_hit(kind, O._p(O.vec(*geom)), O._p(O.vec(*o)), O._p(O.vec(*d)), tmin, tmax, O._p(out))
    return h == 1, out


@pytest.mark.parametrize("case", G["triangle_hit"]["cases"], ids=lambda c: c["name"])
def test_triangle_hit(case):
    g = G["triangle_hit"]
    h, out = hit(0, g["triangle"], case["o"], case["d"], case["tmin"], case["tmax"])
    assert h == case["hit"]
    if h:
        assert abs(out[0] - case["t"]) <= g["tol"]
        p = np.array(case["o"]) + out[0] * np.array(case["d"])
        assert np.allclose(out[1:4], p, atol=1e-6)


@pytest.mark.parametrize("case", G["sphere_hit"]["cases"], ids=lambda c: c["name"])
def test_sphere_hit(case):
    g = G["sphere_hit"]
    h, out = hit(1, g["sphere"], case["o"], case["d"], case["tmin"], case["tmax"])
    assert h == case["hit"]
    if "t" in case:
        assert abs(out[0] - case["t"]) <= g["tol"]
    if "front" in case:
        assert bool(out[7]) == case["front"]
        assert np.allclose(out[4:7], case["n"], atol=g["tol"])
    if "p" in case:
        assert np.allclose(out[1:4], case["p"], atol=g["tol"])


def test_negative_radius_sphere_flips_normal():
    # outward normal (p - c) / r points inward for r < 0 (sphere.go:63; default_scene.go:70)
    h, out = hit(1, [0, 0, 0, -1.0], [0, 0, 2], [0, 0, -1], 0.001, 1000.0)
    assert h and abs(out[0] - 1.0) < 1e-12
    # outward normal = (0,0,-1) points along the ray: a back face, face-forwarded to (0,0,1)
    assert bool(out[7]) is False
    assert np.allclose(out[4:7], [0, 0, 1])


@pytest.mark.parametrize("case", G["quad_hit"]["cases"], ids=lambda c: c["name"])
def test_quad_hit(case):
    g = G["quad_hit"]
    h, out = hit(2, g["quad"], case["o"], case["d"], case["tmin"], case["tmax"])
    assert h == case["hit"]
    if "t" in case:
        assert abs(out[0] - case["t"]) <= g["tol"]
    if "p" in case:
        assert np.allclose(out[1:4], case["p"], atol=g["tol"])


def test_quad_bbox_slab():
    # axis-aligned quads get a +-0.001 box along the normal (quad.go:44-77)
    out = np.zeros(6)
    O.lib().orc_quad_bbox(O._p(O.vec(0, 2, 0, 1, 0, 0, 0, 0, 1)), O._p(out))
    assert np.allclose(out, [0, 2 - 0.001, 0, 1, 2 + 0.001, 1])


def test_aabb_parallel_axis():
    box = O.vec(0, 0, 0, 1, 1, 1)
    # |d| < 1e-8 on an axis: reject when the origin is outside that slab (aabb.go:65-70)
    assert O.lib().orc_aabb_hit(O._p(box), O._p(O.vec(0.5, 2, -1)), O._p(O.vec(0, 0, 1)), 0, 1e9) == 0
    assert O.lib().orc_aabb_hit(O._p(box), O._p(O.vec(0.5, 0.5, -1)), O._p(O.vec(0, 0, 1)), 0, 1e9) == 1


def test_world_radius_unit_sphere():
    g = G["world_radius"]
    s = P.SceneDesc()
    m = s.lambertian((0.5, 0.5, 0.5))
    s.sphere(g["unit_sphere_center"], g["unit_sphere_radius"], m)
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 16, 1.0, 40)
    c, r, n = O.OracleScene(s).world()
    assert np.allclose(c, g["expected_center"], atol=g["tol"])
    assert abs(r - g["expected_radius"]) <= g["tol"]


@pytest.mark.parametrize("case", G["power_heuristic"]["cases"])
def test_power_heuristic(case):
    nf, f, ng, gg, exp = case
    assert abs(O.lib().orc_power_heuristic(nf, f, ng, gg) - exp) <= G["power_heuristic"]["tol"]


def test_schlick():
    g = G["schlick"]
    r0 = O.lib().orc_reflectance(1.0, g["ratio"])
    r90 = O.lib().orc_reflectance(0.0, g["ratio"])
    r45 = O.lib().orc_reflectance(g["cos45"], g["ratio"])
    lo, hi = g["normal_incidence_range"]
    assert lo <= r0 <= hi and r90 >= g["grazing_min"] and r0 < r45 <= g["cos45_max"] and r45 < r90
    # closed form at normal incidence: ((1-n)/(1+n))^2
    assert abs(r0 - ((1 - g["ratio"]) / (1 + g["ratio"])) ** 2) < 1e-15


def _glass_scene():
    s = P.SceneDesc()
    glass = s.dielectric(1.5)
    lam = s.lambertian((0.5, 0.5, 0.5))
    metal = s.metal((0.8, 0.6, 0.2), 0.0)
    layered = s.layered(glass, lam)
    s.sphere((0, 0, 0), 1, lam)
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 16, 1.0, 40)
    return O.OracleScene(s), dict(glass=glass, lam=lam, metal=metal, layered=layered)


def _brdf(os_, m, i, o, n, front, mode):
    out = np.zeros(3)
    O.lib().orc_material_brdf(os_.h, m, O._p(O.vec(*i)), O._p(O.vec(*o)), O._p(O.vec(*n)), front, mode, O._p(out))
    return out


def test_dielectric_transport_modes():
    g = G["dielectric_transport"]
    os_, m = _glass_scene()
    n = np.array(g["normal"])
    wi = np.array(g["incoming"])
    # refracted direction air->glass (refractVector, dielectric.go:155-161)
    eta = 1 / g["ior"]
    ct = min(-wi @ n, 1.0)
    perp = (wi + n * ct) * eta
    wo = perp - n * math.sqrt(abs(1 - perp @ perp))
    rad = _brdf(os_, m["glass"], wi, wo, n, 1, 0)
    imp = _brdf(os_, m["glass"], wi, wo, n, 1, 1)
    assert rad[0] > 0 and imp[0] > 0
    assert abs(rad[0] / imp[0] - g["radiance_over_importance_front"]) < g["tol"]
    # reflection: same value in both modes
    refl = wi - 2 * (wi @ n) * n
    assert np.allclose(_brdf(os_, m["glass"], wi, refl, n, 1, 0), _brdf(os_, m["glass"], wi, refl, n, 1, 1))
    assert _brdf(os_, m["glass"], wi, refl, n, 1, 0)[0] > 0


def test_lambertian_pdf_and_brdf():
    os_, m = _glass_scene()
    n = np.array([0, 1.0, 0])
    delta = __import__("ctypes").c_int32()
    for th in np.linspace(0.05, 1.5, 7):
        o = np.array([math.sin(th), math.cos(th), 0])
        p = O.lib().orc_material_pdf(os_.h, m["lam"], O._p(O.vec(0, -1, 0)), O._p(O.vec(*o)), O._p(O.vec(*n)),
                                     __import__("ctypes").byref(delta))
        assert abs(p - math.cos(th) / math.pi) < G["lambertian_pdf"]["tol"] and delta.value == 0
    assert np.allclose(_brdf(os_, m["lam"], [0, -1, 0], [0, 1, 0], n, 1, 0), 0.5 / math.pi)
    assert np.allclose(_brdf(os_, m["lam"], [0, -1, 0], [0, -1, 0], n, 1, 0), 0)
    # metal and dielectric are delta; layered never reports delta (layered.go:90-104)
    p = O.lib().orc_material_pdf(os_.h, m["metal"], O._p(O.vec(0, -1, 0)), O._p(O.vec(0, 1, 0)), O._p(O.vec(*n)),
                                 __import__("ctypes").byref(delta))
    assert p == 0 and delta.value == 1
    O.lib().orc_material_pdf(os_.h, m["layered"], O._p(O.vec(0, -1, 0)), O._p(O.vec(0, 1, 0)), O._p(O.vec(*n)),
                             __import__("ctypes").byref(delta))
    assert delta.value == 0


def _light_scene():
    s = P.SceneDesc()
    g = G["quad_light_pdf"]
    s.add_quad_light(g["corner"], g["u"], g["v"], (1, 1, 1))
    gs = G["sphere_light_pdf"]
    s.add_sphere_light(gs["center"], gs["radius"], (1, 1, 1))
    s.add_uniform_infinite_light((1, 1, 1))
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 16, 1.0, 40)
    return O.OracleScene(s)


def test_light_pdfs():
    os_ = _light_scene()
    g = G["quad_light_pdf"]
    p = O.lib().orc_light_pdf(os_.h, 0, O._p(O.vec(*g["point"])), O._p(O.vec(0, 0, 1)), O._p(O.vec(*g["direction"])))
    assert abs(p - g["expected"]) < g["tol"]
    for c in G["sphere_light_pdf"]["cases"]:
        p = O.lib().orc_light_pdf(os_.h, 1, O._p(O.vec(*c["point"])), O._p(O.vec(0, 1, 0)), O._p(O.vec(*c["direction"])))
        assert (p > 0) == c["positive"]
        if "expected" in c:
            assert abs(p - c["expected"]) < G["sphere_light_pdf"]["tol"]
    # outside the sphere: cone pdf 1 / (2 pi (1 - cos theta_max)) (sphere_light.go:158-164)
    p = O.lib().orc_light_pdf(os_.h, 1, O._p(O.vec(3, 0, 0)), O._p(O.vec(0, 1, 0)), O._p(O.vec(-1, 0, 0)))
    ctm = math.sqrt(1 - (1 / 3) ** 2)
    assert abs(p - 1 / (2 * math.pi * (1 - ctm))) < 1e-12
    for d, exp in G["infinite_light_pdf"]["cases"]:
        p = O.lib().orc_light_pdf(os_.h, 2, O._p(O.vec(0, 0, 0)), O._p(O.vec(*G["infinite_light_pdf"]["normal"])),
                                  O._p(O.vec(*d)))
        assert abs(p - exp) < G["infinite_light_pdf"]["tol"]


def test_light_sample_pdf_consistency():
    # Sample's pdf equals PDF() of the sampled direction (quad_light_test.go:467-520)
    os_ = _light_scene()
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = rng.random(2)
        out = np.zeros(8)
        p = O.vec(0.1, 0.2, 1.5)
        O.lib().orc_light_sample(os_.h, 0, O._p(p), O._p(O.vec(0, 0, -1)), O._p(O.vec(*s)), O._p(out))
        pdf2 = O.lib().orc_light_pdf(os_.h, 0, O._p(p), O._p(O.vec(0, 0, -1)), O._p(O.vec(*out[:3])))
        assert abs(out[7] - pdf2) < 1e-9 * max(1, pdf2)
        p3 = O.vec(3.0, 0.5, 0.2)
        O.lib().orc_light_sample(os_.h, 1, O._p(p3), O._p(O.vec(0, 1, 0)), O._p(O.vec(*s)), O._p(out))
        pdf3 = O.lib().orc_light_pdf(os_.h, 1, O._p(p3), O._p(O.vec(0, 1, 0)), O._p(O.vec(*out[:3])))
        assert abs(out[7] - pdf3) < 1e-9 * pdf3


def test_pass_schedule():
    g = G["pass_schedule"]
    got = [O.lib().orc_samples_for_pass(g["initial"], g["max_spp"], g["max_passes"], k) for k in range(1, 8)]
    assert got == g["expected"]


def test_cosine_hemisphere_warp():
    g = G["cosine_hemisphere_mean_cos"]
    rng = np.random.default_rng(0)
    n = O.vec(0.3, 0.8, -0.52)
    n = n / np.linalg.norm(n)
    cs = []
    for _ in range(4000):
        out = np.zeros(3)
        O.lib().orc_sample_cosine_hemisphere(O._p(n), O._p(O.vec(*rng.random(2))), O._p(out))
        assert abs(np.linalg.norm(out) - 1) < 1e-12
        cs.append(out @ n)
    cs = np.array(cs)
    assert cs.min() >= 0
    assert abs(cs.mean() - g["expected"]) < g["tol"]
    # closed form: E[cos] = 2/3 for a cosine-weighted hemisphere
    assert abs(cs.mean() - 2 / 3) < 0.02


def test_disk_and_ball_warps():
    out = np.zeros(3)
    O.lib().orc_sample_point_in_unit_disk(O._p(O.vec(0.5, 0.5)), O._p(out))
    assert np.all(out == 0)
    O.lib().orc_sample_point_in_unit_disk(O._p(O.vec(1.0, 0.5)), O._p(out))
    assert np.allclose(out, [1, 0, 0])
    O.lib().orc_sample_point_in_unit_sphere(O._p(O.vec(0.125, 0.0, 1.0)), O._p(out))
    assert np.allclose(out, [0, 0, 0.5])  # r = cbrt(1/8), cos(theta) = 1


def test_counter_rng_uniformity():
    xs = np.array([O.lib().orc_draw(0, p, s, k) for p in range(40) for s in range(25) for k in range(20)])
    assert xs.min() >= 0 and xs.max() < 1
    assert abs(xs.mean() - 0.5) < 0.01 and abs(xs.var() - 1 / 12) < 0.003
    # consecutive draws uncorrelated
    assert abs(np.corrcoef(xs[:-1], xs[1:])[0, 1]) < 0.03


def _pt_params(desc, cfg, seed=0):
    w, h = desc.image_size
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], cfg["spp"], seed=seed)


def test_uniform_light_every_pixel_exact():
    # no geometry + uniform light of radiance 1: every sample is exactly 1
    s, cfg = scenes.empty_uniform_scene(width=16)
    os_ = O.OracleScene(s)
    sm, ls, lq = os_.render(_pt_params(s, cfg), 3)
    assert np.all(sm == 3.0)


def test_miss_returns_gradient():
    g = G["miss_gradient"]
    s = P.SceneDesc()
    s.add_gradient_infinite_light(g["top"], g["bottom"])
    s.set_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 8, 1.0, 40)
    os_ = O.OracleScene(s)
    out = np.zeros(3)
    O.lib().orc_ray_color(os_.h, __import__("ctypes").byref(P.sampling_params(8, 8, 5, 5, 1)), O._p(O.vec(0, 0, 0)), O._p(O.vec(*g["direction"])),
                          7, O._p(out))
    assert np.allclose(out, g["expected"], atol=g["tol"])


def test_depth_zero_is_black():
    s, cfg = scenes.empty_uniform_scene(width=8)
    os_ = O.OracleScene(s)
    out = np.zeros(3)
    O.lib().orc_ray_color(os_.h, __import__("ctypes").byref(P.sampling_params(8, 8, 0, 0, 1)), O._p(O.vec(0, 0, 0)), O._p(O.vec(0, 0, -1)), 1,
                          O._p(out))
    assert np.all(out == 0)


def test_white_furnace_sphere():
    """Lambertian sphere of albedo a under a uniform sky of radiance 1: with one infinite light
    the light and BSDF pdfs are both cos/pi, so the reference weights direct and indirect by
    1/2 each and every sample that hits the sphere returns exactly a (convex: every bounce
    escapes); sky pixels return 1."""
    s, cfg = scenes.furnace_scene(albedo=0.5, width=24)
    os_ = O.OracleScene(s)
    sm, _, _ = os_.render(_pt_params(s, cfg), 1, sample_offset=5)
    v = sm
    hitpix = np.abs(v - 0.5).max(axis=2) < 1e-9
    skypix = np.abs(v - 1.0).max(axis=2) < 1e-12
    assert np.all(hitpix | skypix)
    assert hitpix.sum() > 20 and skypix.sum() > 20


def test_tile_grid_count():
    g = G["tile_grid"]
    tiles = math.ceil(g["width"] / g["tile"]) * math.ceil(g["height"] / g["tile"])
    assert tiles == g["tiles"]


def test_default_scene_image_height():
    s, _ = scenes.default_scene()
    assert s.image_size == (400, 225)
    assert O.lib().orc_camera_height(O.OracleScene(s).h) == 225
