"""Pins the oracle's disc shape, disc light, disc spot light and point spot light (pkg/geometry/disc.go,
pkg/lights/disc_light.go, disc_spot_light.go, point_spot_light.go) against the reference's own test
vectors (tests/golden/reference_cases.json) and against closed forms. CPU only."""
import ctypes as C
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


def disc_hit(geom, o, d, tmin, tmax):
    out = np.zeros(10)
    h = O.lib().orc_shape_hit(3, O._p(O.vec(*geom)), O._p(O.vec(*o)), O._p(O.vec(*d)), tmin, tmax, O._p(out))
    return h == 1, out


def disc_frame(geom):
    out = np.zeros(12)
    O.lib().orc_disc_frame(O._p(O.vec(*geom)), O._p(out))
    return out[0:3], out[3:6], out[6:9], out[9:12]


def light_sample(os_, li, p, s, n=(0, 1, 0)):
    out = np.zeros(8)
    O.lib().orc_light_sample(os_.h, li, O._p(O.vec(*p)), O._p(O.vec(*n)), O._p(O.vec(*s)), O._p(out))
    return out[0:3], out[3], out[4:7], out[7]


def light_pdf(os_, li, p, d, n=(0, 1, 0)):
    return O.lib().orc_light_pdf(os_.h, li, O._p(O.vec(*p)), O._p(O.vec(*n)), O._p(O.vec(*d)))


def sample_emission(os_, li, sp, sd):
    out = np.zeros(14)
    O.lib().orc_light_sample_emission(os_.h, li, O._p(O.vec(*sp)), O._p(O.vec(*sd)), O._p(out))
    return dict(p=out[0:3], n=out[3:6], dir=out[6:9], emission=out[9:12], area_pdf=out[12], dir_pdf=out[13])


def pdf_le(os_, li, p, d):
    out = np.zeros(2)
    O.lib().orc_light_pdf_le(os_.h, li, O._p(O.vec(*p)), O._p(O.vec(*d)), O._p(out))
    return out[0], out[1]


def one_light_scene(add):
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    add(s)
    return O.OracleScene(s)


def stratified(n):
    g = (np.arange(n) + 0.5) / n
    a, b = np.meshgrid(g, g, indexing="ij")
    return np.stack([a.ravel(), b.ravel()], 1)


# ---- disc geometry (disc_test.go) ----

@pytest.mark.parametrize("case", G["disc_hit"]["cases"], ids=lambda c: c["name"])
def test_disc_hit(case):
    g = G["disc_hit"]
    h, out = disc_hit(g["disc"], case["o"], case["d"], case["tmin"], case["tmax"])
    assert h == case["hit"]
    if h:
        assert abs(out[0] - case["t"]) <= g["tol"]
        assert np.linalg.norm(out[1:4] - np.array(g["disc"][0:3])) <= g["disc"][6] + g["tol"]
        # SetFaceNormal: the stored normal faces the incoming ray
        assert np.dot(out[4:7], case["d"]) < 0


def test_disc_bbox_and_frame():
    g = G["disc_bbox"]
    _, _, lo, hi = disc_frame(g["disc"])
    assert np.allclose(lo, g["expected_min"], atol=g["tol"]) and np.allclose(hi, g["expected_max"], atol=g["tol"])
    f = G["disc_frame"]
    for n in f["normals"]:
        nn = np.array(n, float) / np.linalg.norm(n)
        right, up, _, _ = disc_frame([0, 0, 0, *n, 1.0])
        for a, b in ((right, nn), (up, nn), (right, up)):
            assert abs(np.dot(a, b)) <= f["tol"]
        assert abs(np.linalg.norm(right) - 1) <= f["tol"] and abs(np.linalg.norm(up) - 1) <= f["tol"]
        # right-handed frame: right x up = n (up = n x right)
        assert np.allclose(np.cross(right, up), nn, atol=1e-12)


def test_disc_uv_maps_rim_to_unit_square():
    # uv = ((cp.right / r + 1) / 2, (cp.up / r + 1) / 2), disc.go:76-80
    right, up, _, _ = disc_frame([0, 0, 0, 0, 1, 0, 2.0])
    p = 1.5 * right
    h, out = disc_hit([0, 0, 0, 0, 1, 0, 2.0], p + np.array([0, 1, 0]), [0, -1, 0], 0.001, 10)
    assert h and np.allclose(out[8:10], [(1.5 / 2 + 1) / 2, 0.5], atol=1e-12)


def test_world_bounds_include_discs():
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    m = s.lambertian((0.5, 0.5, 0.5))
    s.disc((1, 2, 3), (0, 1, 0), 2.0, m)
    c, r, n = O.OracleScene(s).world()
    assert n == 1 and np.allclose(c, [1, 2, 3]) and abs(r - math.sqrt(8)) < 1e-12


# ---- disc light (disc_light_test.go) ----

@pytest.fixture(scope="module")
def disc_light():
    g = G["disc_light_pdf"]
    return one_light_scene(lambda s: s.add_disc_light(g["center"], g["normal"], g["radius"], g["emission"]))


@pytest.mark.parametrize("case", G["disc_light_pdf"]["cases"], ids=lambda c: c["name"])
def test_disc_light_pdf(disc_light, case):
    d = np.array(case["d"], float)
    pdf = light_pdf(disc_light, 0, G["disc_light_pdf"]["point"], d / np.linalg.norm(d))
    assert (pdf > 0) == case["positive"]
    if not case["positive"]:
        assert pdf == 0


def test_disc_light_sample_pdf_consistent(disc_light):
    # disc_light_test.go:110-133: the sample's pdf equals PDF() along the sampled direction
    for s in stratified(9):
        d, dist, e, pdf = light_sample(disc_light, 0, (0, 0, 0), s)
        assert pdf > 0 and np.allclose(e, G["disc_light_pdf"]["emission"])
        assert abs(light_pdf(disc_light, 0, (0, 0, 0), d) - pdf) <= 1e-9 * pdf


def test_disc_light_irradiance_closed_form():
    """E[L cos / pdf] over light samples = irradiance of a Lambertian disc of radiance L, radius R at
    height h on its axis: pi L R^2 / (h^2 + R^2). Catches a dropped cos or dist^2 in the area to
    solid-angle conversion and a non-uniform point distribution."""
    h, R, L = 1.5, 0.8, 3.0
    os_ = one_light_scene(lambda s: s.add_disc_light((0, h, 0), (0, -1, 0), R, (L, L, L)))
    acc = 0.0
    ss = stratified(200)
    for s in ss:
        d, dist, e, pdf = light_sample(os_, 0, (0, 0, 0), s)
        acc += e[0] * d[1] / pdf
    est = acc / len(ss)
    exact = math.pi * L * R * R / (h * h + R * R)
    assert abs(est - exact) <= 2e-4 * exact


@pytest.mark.parametrize("case", G["disc_light_emission_pdf"]["cases"], ids=lambda c: c["name"])
def test_disc_light_pdf_le(case):
    g = G["disc_light_emission_pdf"]
    os_ = one_light_scene(lambda s: s.add_disc_light(g["center"], g["normal"], g["radius"], g["emission"]))
    pos, dirp = pdf_le(os_, 0, case["p"], case["d"])
    if case["positive"]:
        assert abs(pos - 1 / (math.pi * g["radius"] ** 2)) <= g["tol"]
        assert abs(dirp - np.dot(case["d"], g["normal"]) / math.pi) <= g["tol"]
    else:
        assert pos == 0


def test_disc_light_sample_emission():
    # disc_light_test.go:206-273: point on the disc, cosine-weighted direction about the normal
    os_ = one_light_scene(lambda s: s.add_disc_light((0, 0, 0), (0, 1, 0), 1.0, (5, 5, 5)))
    for sp, sd in zip(stratified(4), stratified(4)[::-1]):
        es = sample_emission(os_, 0, sp, sd)
        assert np.linalg.norm(es["p"]) <= 1 + 1e-12 and abs(es["p"][1]) < 1e-12
        ct = es["dir"][1]
        assert ct > 0 and abs(es["area_pdf"] - 1 / math.pi) < 1e-12 and abs(es["dir_pdf"] - ct / math.pi) < 1e-12
        assert np.allclose(es["emission"], 5) and np.allclose(es["n"], [0, 1, 0])


# ---- spot lights (disc_spot_light_test.go, point_spot_light_test.go) ----

@pytest.mark.parametrize("case", G["spot_falloff"]["cases"], ids=lambda c: c["name"])
def test_spot_falloff(case):
    """The falloff curve through a point spot light's Sample at unit distance: emission = E * falloff."""
    g = G["spot_falloff"]
    os_ = one_light_scene(lambda s: s.add_point_spot_light((0, 0, 0), (0, -1, 0), (1, 1, 1), g["cone_angle"], g["cone_delta"]))
    a = math.radians(case["angle_deg"])
    p = (math.sin(a), -math.cos(a), 0)
    _, dist, e, pdf = light_sample(os_, 0, p, (0.5, 0.5))
    assert abs(dist - 1) < 1e-12 and pdf == 1.0
    if case["expected"] == "quartic":
        ct, cf = math.cos(math.radians(g["cone_angle"])), math.cos(math.radians(g["cone_angle"] - g["cone_delta"]))
        x = (math.cos(a) - ct) / (cf - ct)
        want = x ** 4
        assert 0 < want < 1
    else:
        want = case["expected"]
    assert abs(e[0] - want) <= g["tol"]


@pytest.mark.parametrize("case", G["disc_spot_sample"]["cases"], ids=lambda c: c["name"])
def test_disc_spot_sample(case):
    g = G["disc_spot_sample"]
    os_ = one_light_scene(lambda s: s.add_spot_light(g["from"], g["to"], g["emission"], g["cone_angle"], g["cone_delta"], g["radius"]))
    for s in stratified(3):
        d, dist, e, pdf = light_sample(os_, 0, case["p"], s)
        spt = np.array(case["p"]) + d * dist
        assert np.linalg.norm(spt - np.array(g["from"])) <= g["radius"] + 1e-9
        assert (np.any(e != 0)) == case["emits"]
        if case["emits"]:
            assert pdf > 0


def test_point_spot_light():
    g = G["point_spot"]
    w = g["within_cone"]
    os_ = one_light_scene(lambda s: s.add_point_spot_light(g["from"], g["to"], w["emission"], w["cone_angle"], w["cone_delta"]))
    d, dist, e, pdf = light_sample(os_, 0, w["p"], (0.3, 0.7))
    assert np.allclose(d, w["direction"], atol=w["tol"]) and abs(dist - w["distance"]) <= w["tol"]
    assert abs(np.linalg.norm(e) - w["emission_length"]) <= w["tol"] and pdf == 1.0
    o = g["outside_cone"]
    os2 = one_light_scene(lambda s: s.add_point_spot_light(g["from"], g["to"], o["emission"], o["cone_angle"], o["cone_delta"]))
    _, _, e2, _ = light_sample(os2, 0, o["p"], (0.3, 0.7))
    assert np.linalg.norm(e2) <= o["max_emission_length"]
    q = g["pdf"]
    os3 = one_light_scene(lambda s: s.add_point_spot_light(g["from"], g["to"], q["emission"], q["cone_angle"], q["cone_delta"]))
    to_light = np.array(g["from"], float) - q["p"]
    to_light /= np.linalg.norm(to_light)
    for case in q["cases"]:
        d = to_light if case["d"] == "to_light" else (-to_light if case["d"] == "away" else np.array(case["d"], float))
        assert light_pdf(os3, 0, q["p"], d) == case["expected"]


def test_point_spot_cone_emission_sampling():
    """SampleEmission draws directions uniformly in the cone: cos(theta) uniform on [cos_total, 1],
    direction pdf 1 / (2 pi (1 - cos_total)), emission E * falloff; PDF_Le agrees at the light."""
    cone, delta = 40.0, 10.0
    os_ = one_light_scene(lambda s: s.add_point_spot_light((1, 2, 3), (1, -5, 3), (2, 2, 2), cone, delta))
    ct, cf = math.cos(math.radians(cone)), math.cos(math.radians(cone - delta))
    cos = []
    for sd in stratified(40):
        es = sample_emission(os_, 0, (0.5, 0.5), sd)
        c = -es["dir"][1]
        cos.append(c)
        assert c >= ct - 1e-12 and np.allclose(es["p"], [1, 2, 3])
        assert abs(es["dir_pdf"] - 1 / (2 * math.pi * (1 - ct))) < 1e-12 and es["area_pdf"] == 1.0
        fall = 1.0 if c >= cf else ((c - ct) / (cf - ct)) ** 4
        assert abs(es["emission"][0] - 2 * fall) < 1e-12
        pos, dirp = pdf_le(os_, 0, es["p"], es["dir"])
        assert pos == 1.0 and abs(dirp - es["dir_pdf"]) < 1e-12
    # the first sample dimension of a 40x40 grid takes 40 stratum midpoints: cos(theta) lands on
    # the midpoints of 40 equal strata of [cos_total, 1]
    levels = np.unique(np.round(cos, 12))
    assert len(levels) == 40
    assert np.allclose(levels, ct + (1 - ct) * (np.arange(40) + 0.5) / 40, atol=1e-12)
    assert pdf_le(os_, 0, (1, 2.01, 3), (0, -1, 0)) == (0.0, 0.0)


def test_point_light_direct_lighting_closed_form():
    """One bounce on a Lambertian floor under a point spot light (max depth 1: no indirect term).
    PathTracingIntegrator's direct lighting weighs the delta light's pdf 1 against the BSDF pdf
    cos/pi with the power heuristic (path_tracing.go:116-157), so the estimate is exactly
    (a/pi) * E * falloff / d^2 * cos * 1 / (1 + (cos/pi)^2)."""
    a, E, h = 0.5, 8.0, 4.0
    cone, delta = 60.0, 10.0
    s = P.SceneDesc()
    s.set_camera((0, 3, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    floor = s.lambertian((a, a, a))
    s.quad((-50, 0, -50), (0, 0, 100), (100, 0, 0), floor)
    s.add_point_spot_light((0, h, 0), (0, 0, 0), (E, E, E), cone, delta)
    os_ = O.OracleScene(s)
    ct, cf = math.cos(math.radians(cone)), math.cos(math.radians(cone - delta))
    for x in (0.0, 1.0, 4.0 * math.tan(math.radians(55.0)), 4.0 * math.tan(math.radians(65.0))):
        out = np.zeros(3)
        O.lib().orc_ray_color(os_.h, C.byref(P.sampling_params(8, 8, 1, 10, 1)), O._p(O.vec(x, 1, 0)),
                              O._p(O.vec(0, -1, 0)), 11, O._p(out))
        d = math.hypot(x, h)
        c = h / d
        fall = 0.0 if c < ct else (1.0 if c >= cf else ((c - ct) / (cf - ct)) ** 4)
        want = a / math.pi * E * fall / (d * d) * c / (1 + (c / math.pi) ** 2)
        assert abs(out[0] - want) <= 1e-12 * max(want, 1e-300), (x, out[0], want)


def test_delta_light_is_never_hit_by_bsdf_rays():
    """A point spot light has no surface: pixels that see only the sky under it return the sky."""
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    s.add_point_spot_light((0, 0, 0), (0, 0, 5), (100, 100, 100), 90.0, 0.0)
    s.add_uniform_infinite_light((0.25, 0.25, 0.25))
    os_ = O.OracleScene(s)
    sm, _, _ = os_.render(P.sampling_params(8, 8, 4, 10, 2), 2)
    assert np.allclose(sm, 0.5)


def test_scene_constructors_build():
    for fn in (lambda: scenes.cornell_scene(width=16, light="point"), lambda: scenes.cornell_scene(width=16, light="sphere"),
               lambda: scenes.caustic_glass_scene(width=16), lambda: scenes.caustic_glass_scene(width=16, light="point"),
               lambda: scenes.lights_scene(width=16)):
        s, cfg = fn()
        c, r, n = O.OracleScene(s).world()
        assert n >= 0 and np.all(np.isfinite(c))
