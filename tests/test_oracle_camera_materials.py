"""Pins of the oracle's camera (GetRay, CalculateRayPDFs, SampleCameraFromPoint, EvaluateRayImportance,
MapRayToPixel), material scatter (metal, dielectric, layered), SampleOnUnitSphere and the environment-map
light, against the reference's own test expectations and against closed forms.

Reference tests followed: pkg/geometry/camera_test.go:11-244, pkg/geometry/camera_splat_test.go:12-314,
pkg/material/metal_test.go:36-148, pkg/material/dielectric_test.go:11-128,
pkg/material/layered_test.go:11-161. Where a reference test draws from math/rand, the oracle's counter
stream (keys 0, 1, 2, ...) stands in for the seeds; the expectations do not depend on the values drawn.
The environment map has no reference counterpart: it is pinned by closed forms (its density integrates
to one, sampling follows the texel weights, a constant map lights a white furnace exactly as a uniform
infinite light does in expectation).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes


def _vec(*x):
    return np.ascontiguousarray(np.array(x, dtype=np.float64).reshape(-1))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _camera_scene(center, look_at, width, aspect, vfov, aperture=0.0, focus=0.0):
    s = P.SceneDesc()
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    s.set_camera(center=center, look_at=look_at, up=(0, 1, 0), width=width, aspect_ratio=aspect, vfov=vfov,
                 aperture=aperture, focus_distance=focus)
    return O.OracleScene(s)


def _info(os_):
    out = np.zeros(11)
    O.lib().orc_camera_info(os_.h, _p(out))
    return {"forward": out[0:3], "lens_area": out[3], "plane_area": out[4], "cos_total": out[5], "vw": out[6],
            "vh": out[7], "focus": out[8], "width": int(out[9]), "height": int(out[10])}


def _pdfs(os_, o, d):
    pos, dr = C.c_double(), C.c_double()
    px, py = C.c_int32(), C.c_int32()
    O.lib().orc_camera_ray_pdfs(os_.h, _p(_vec(*o)), _p(_vec(*d)), C.byref(pos), C.byref(dr), C.byref(px),
                                C.byref(py))
    return pos.value, dr.value, px.value, py.value


def _get_ray(os_, i, j, lens, jit):
    o, d = np.zeros(3), np.zeros(3)
    O.lib().orc_camera_get_ray(os_.h, i, j, _p(_vec(*lens)), _p(_vec(*jit)), _p(o), _p(d))
    return o, d


def _importance(os_, o, d):
    return O.lib().orc_camera_importance(os_.h, _p(_vec(*o)), _p(_vec(*d)))


def _sample_from_point(os_, ref, s):
    out = np.zeros(8)
    ok = O.lib().orc_camera_sample_from_point(os_.h, _p(_vec(*ref)), _p(_vec(*s)), _p(out))
    return bool(ok), out[0:3], out[3:6], out[6], out[7]


# ---- camera_test.go ----------------------------------------------------------------------------------
def _cam45():
    return _camera_scene((0, 0, 0), (0, 0, -1), 400, 1.0, 45.0)


def test_camera_forward():  # TestCameraGetCameraForward
    assert np.allclose(_info(_cam45())["forward"], (0, 0, -1), atol=1e-6)


def test_ray_pdfs_closed_form():
    """TestCameraCalculateRayPDFs_AreaPDF/_DirectionPDF: both pdfs positive for rays in view; closed forms
    pdfPos = 1/lensArea (pinhole: 1) and pdfDir = 1/(A cos^3) with A = (2 tan(vfov/2))^2 * aspect
    (camera.go:194-216). The reference test's "straight above angled" holds because its angled ray
    (0.5, 0, -1) is outside the 45-degree view (pdf 0); within the view 1/cos^3 grows off axis."""
    os_ = _cam45()
    A = (2 * math.tan(math.radians(22.5))) ** 2
    assert _info(os_)["plane_area"] == pytest.approx(A, rel=1e-14)
    pos, d_straight, px, py = _pdfs(os_, (0, 0, 0), (0, 0, -1))
    assert pos == 1.0 and d_straight == pytest.approx(1.0 / A, rel=1e-14)
    assert (px, py) == (200, 200)
    # the reference's angled ray (0.5, 0, -1) leaves the 45-degree field of view: both pdfs are 0
    assert _pdfs(os_, (0, 0, 0), _unit((0.5, 0, -1)))[:2] == (0.0, 0.0)
    ang = _unit((0.3, 0, -1))
    pos2, d_ang, _, _ = _pdfs(os_, (0, 0, 0), ang)
    assert pos2 == 1.0 and d_ang > d_straight > 0
    assert d_ang == pytest.approx(1.0 / (A * (-ang[2]) ** 3), rel=1e-12)


def test_ray_pdfs_backward_ray_is_zero():  # TestCameraCalculateRayPDFs_InvalidRay
    pos, dr, _, _ = _pdfs(_cam45(), (0, 0, 0), (0, 0, 1))
    assert pos == 0 and dr == 0


def test_ray_pdfs_cornell_pinhole_scale():
    """TestCameraCalculateRayPDFs_CornellBoxRealistic / _ScaleConsistency: the pinhole area pdf is 1.0,
    so its ratio to the 130x105 quad light's area pdf equals the light's area (within 1.0)."""
    os_ = _camera_scene((278, 278, -800), (278, 278, 0), 400, 1.0, 40.0)
    pos, dr, _, _ = _pdfs(os_, (278, 278, -800), (0, 0, 1))
    assert abs(pos - 1.0) <= 1e-6 and dr > 0
    assert 1e-10 < pos <= 1.0
    assert abs(pos / (1.0 / (130.0 * 105.0)) - 130.0 * 105.0) <= 1.0


def test_generated_rays_have_positive_pdfs():  # TestCameraCalculateRayPDFs_ConsistencyWithGeneration
    os_ = _camera_scene((0, 0, 0), (0, 0, -1), 100, 1.0, 45.0)
    rng = np.random.default_rng(42)
    for _ in range(200):
        i, j = rng.integers(0, 100, 2)
        o, d = _get_ray(os_, int(i), int(j), rng.random(2), rng.random(2))
        pos, dr, px, py = _pdfs(os_, o, d)
        assert d @ (0, 0, -1) > 0 and pos > 0 and dr > 0
        assert (px, py) == (i, j)  # pinhole: the ray maps back to its own pixel


def test_get_ray_closed_form():
    """GetRay (camera.go:173-190): the direction through pixel (i, j) with jitter s is the unit vector to
    pixel00 + du (i + s.x - 0.5) + dv (j + s.y - 0.5); pixel (0, 0) is the upper-left corner as seen
    through u = w x up, v = w x u."""
    os_ = _camera_scene((0, 0, 0), (0, 0, -1), 400, 2.0, 90.0)  # 400x200, viewport 4 x 2 at focus 1
    inf = _info(os_)
    assert (inf["width"], inf["height"]) == (400, 200)
    assert inf["vh"] == pytest.approx(2.0) and inf["vw"] == pytest.approx(4.0)
    for i, j, s in [(0, 0, (0.5, 0.5)), (399, 199, (0.5, 0.5)), (200, 100, (0.0, 0.0)), (17, 150, (0.25, 0.9))]:
        _, d = _get_ray(os_, i, j, (0.5, 0.5), s)
        # u = w x up = (0,0,1) x (0,1,0) = (-1,0,0); v = w x u = (0,-1,0)
        x = -(-2.0 + 4.0 * (i + s[0]) / 400)
        y = -(-1.0 + 2.0 * (j + s[1]) / 200)
        assert np.allclose(d, _unit((x, y, -1.0)), atol=1e-14)


def test_thin_lens_origin_on_the_lens_disk():
    """A lens camera's ray origins lie on the lens disk (radius aperture/2, in the u-v plane through the
    centre) and every ray of a pixel passes through the same point of the focus plane (focus 2)."""
    os_ = _camera_scene((0, 0, 0), (0, 0, -1), 64, 1.0, 40.0, aperture=0.4, focus=2.0)
    rng = np.random.default_rng(3)
    pts = []
    for _ in range(64):
        o, d = _get_ray(os_, 10, 20, rng.random(2), (0.5, 0.5))
        assert abs(o[2]) < 1e-15 and np.hypot(o[0], o[1]) <= 0.2 + 1e-12
        t = -2.0 / d[2]
        pts.append(o + t * d)
    pts = np.array(pts)
    assert np.abs(pts - pts[0]).max() < 1e-12
    assert _info(os_)["lens_area"] == pytest.approx(math.pi * 0.04, rel=1e-14)


# ---- camera_splat_test.go ------------------------------------------------------------------------------
def _splat_camera(aperture):
    return _camera_scene((0, 0, 0), (0, 0, -1), 800, 16.0 / 9.0, 90.0, aperture=aperture, focus=1.0)


def test_sample_camera_from_point():
    """TestSampleCameraFromPoint: the ray points toward the reference point (dot >= 0.8), pdf and weight
    positive; closed form pdf = dist^2 / (|cos| lensArea) (camera.go:244-250)."""
    os_ = _splat_camera(0.1)
    ref = np.array([0.5, 0.3, -2.0])
    for s in [(0.37, 0.81), (0.5, 0.5), (0.02, 0.97)]:
        ok, o, d, we, pdf = _sample_from_point(os_, ref, s)
        assert ok
        exp = _unit(ref - o)
        assert d @ exp >= 0.8 and np.allclose(d, exp, atol=1e-14)
        assert pdf > 0 and we > 0
        dist = np.linalg.norm(ref - o)
        assert pdf == pytest.approx(dist * dist / (abs(d[2]) * math.pi * 0.05 ** 2), rel=1e-12)
        assert np.hypot(o[0], o[1]) <= 0.05 + 1e-15 and o[2] == 0


def test_sample_camera_from_point_behind_camera():  # TestSampleCameraFromPointBehindCamera
    ok, *_ = _sample_from_point(_splat_camera(0.1), (0, 0, 1.0), (0.3, 0.6))
    assert not ok


def test_map_ray_to_pixel_round_trip():
    """TestMapRayToPixel: centre and corner rays map back to their pixels (the reference allows 2 and 5
    pixels; for a pinhole the map is exact up to the jitter)."""
    os_ = _splat_camera(0.0)
    assert (_info(os_)["width"], _info(os_)["height"]) == (800, 450)
    rng = np.random.default_rng(42)
    for px, py in [(400, 225), (0, 0), (799, 0), (0, 449), (799, 449)]:
        o, d = _get_ray(os_, px, py, rng.random(2), rng.random(2))
        _, _, x, y = _pdfs(os_, o, d)
        assert (x, y) == (px, py)


def test_map_ray_to_pixel_out_of_bounds():  # TestMapRayToPixelOutOfBounds
    os_ = _splat_camera(0.0)
    assert _pdfs(os_, (0, 0, 0), (0, 0, 1))[2:] == (-1, -1)
    assert _pdfs(os_, (0, 0, 0), _unit((10, 10, -1)))[2:] == (-1, -1)


def test_importance_we():
    """TestWeFunction: We = 1/(A lensArea cos^4): positive on axis, the 45-degree ray's importance is
    4x the centre's (tolerance 0.01), zero outside the field of view and backwards."""
    os_ = _splat_camera(0.1)
    inf = _info(os_)
    c = _importance(os_, (0, 0, 0), (0, 0, -1))
    assert c > 0 and c == pytest.approx(1.0 / (inf["plane_area"] * inf["lens_area"]), rel=1e-14)
    e = _importance(os_, (0, 0, 0), _unit((0.7, 0, -0.7)))
    assert e > c and abs(e / c - 4.0) <= 0.01
    assert _importance(os_, (0, 0, 0), _unit((5, 0, -1))) == 0
    assert _importance(os_, (0, 0, 0), (0, 0, 1)) == 0


def test_camera_sampling_consistency():  # TestCameraSamplingConsistency
    os_ = _splat_camera(0.0)
    ref = np.array([0.2, 0.1, -1.0])
    ok, o, d, _, _ = _sample_from_point(os_, ref, (0.4, 0.6))
    assert ok
    _, _, x, y = _pdfs(os_, o, d)
    assert x >= 0
    _, d2 = _get_ray(os_, x, y, (0.5, 0.5), (0.3, 0.7))
    assert d2 @ _unit(ref) >= 0.95


# ---- material scatter ----------------------------------------------------------------------------------
def _material_scene(build):
    s = P.SceneDesc()
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    s.set_camera(center=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=40)
    mats = build(s)
    return O.OracleScene(s), mats


def _scatter(os_, m, ro, rd, hp, hn, front, key):
    out = np.zeros(7)
    ok = O.lib().orc_material_scatter(os_.h, m, _p(_vec(*ro)), _p(_vec(*rd)), _p(_vec(*hp)), _p(_vec(*hn)),
                                      1 if front else 0, key, _p(out))
    return bool(ok), out[0:3], out[3:6], out[6]


def test_metal_perfect_reflection():
    """TestMetal_PerfectReflection: (0,-1,-1)/sqrt2 off normal (0,0,1) reflects to (0,-1,1)/sqrt2 (1e-10),
    attenuation = albedo, pdf 0 (specular)."""
    os_, m = _material_scene(lambda s: s.metal((0.9, 0.9, 0.9), 0.0))
    for key in range(8):
        ok, d, att, pdf = _scatter(os_, m, (0, 1, 1), _unit((0, -1, -1)), (0, 0, 0), (0, 0, 1), True, key)
        assert ok and np.linalg.norm(_unit(d) - _unit((0, -1, 1))) <= 1e-10
        assert np.array_equal(att, (0.9, 0.9, 0.9)) and pdf == 0


def test_metal_fuzzy_reflection():
    """TestMetal_FuzzyReflection: fuzz 0.5, normal incidence: every sample scatters above the surface and
    the directions vary; the reflected direction is mirror + fuzz * (point in the unit ball), left
    unnormalised (metal.go:44-48), so |d - (0,0,1)| <= 0.5."""
    os_, m = _material_scene(lambda s: s.metal((0.8, 0.8, 0.8), 0.5))
    dirs = []
    for key in range(10):
        ok, d, att, pdf = _scatter(os_, m, (0, 0, 1), (0, 0, -1), (0, 0, 0), (0, 0, 1), True, key)
        assert ok and _unit(d)[2] > 0 and pdf == 0
        assert np.linalg.norm(d - (0, 0, 1)) <= 0.5 + 1e-12
        dirs.append(_unit(d))
    assert max(np.linalg.norm(x - dirs[0]) for x in dirs) > 1e-10


def test_metal_absorption_at_grazing_angle():
    """TestMetal_ScatterAbsorption: fuzz 1 at a grazing angle absorbs some rays (direction below the
    surface) and scatters others."""
    os_, m = _material_scene(lambda s: s.metal((0.8, 0.8, 0.8), 1.0))
    res = [_scatter(os_, m, (-1, 0, 0.01), _unit((1, 0, -0.01)), (0, 0, 0), (0, 0, 1), True, k) for k in range(1000)]
    n_ok = sum(r[0] for r in res)
    assert 0 < n_ok < 1000
    for ok, d, _, _ in res:  # absorbed exactly when the direction is not above the surface
        assert ok == (d[2] > 0)


def _glass():
    return _material_scene(lambda s: s.dielectric(1.5))


def test_dielectric_basic_behaviour():
    """TestDielectricBasicBehavior: always scatters, attenuation exactly white, pdf 0, and both reflection
    and refraction occur at 45 degrees. Closed forms: refraction by Snell's law (sin t = sin 45 / 1.5), the
    mirror direction otherwise, and reflection with the Schlick probability
    R = r0 + (1 - r0)(1 - cos)^5, r0 = ((1 - 1/1.5) / (1 + 1/1.5))^2 (dielectric.go:20-99)."""
    os_, m = _glass()
    d_in = _unit((1, -1, 0))
    st = math.sin(math.pi / 4) / 1.5
    refr = np.array([st, -math.sqrt(1 - st * st), 0.0])
    refl = _unit((1, 1, 0))
    n, n_refl = 8000, 0
    for key in range(n):
        ok, d, att, pdf = _scatter(os_, m, (0, 1, 0), d_in, (0, 0, 0), (0, 1, 0), True, key)
        assert ok and np.array_equal(att, (1.0, 1.0, 1.0)) and pdf == 0
        du = _unit(d)
        if du[1] > -0.5:
            assert np.allclose(du, refl, atol=1e-12)
            n_refl += 1
        else:
            assert np.allclose(du, refr, atol=1e-12)
    r0 = ((1 - 1 / 1.5) / (1 + 1 / 1.5)) ** 2
    R = r0 + (1 - r0) * (1 - math.cos(math.pi / 4)) ** 5
    assert 0 < n_refl < n
    assert abs(n_refl - n * R) <= 5 * math.sqrt(n * R * (1 - R))


def test_dielectric_total_internal_reflection():
    """TestDielectricTotalInternalReflection: leaving glass at a shallow angle (back face) reflects every
    time, going up with the x component preserved (1e-10)."""
    os_, m = _glass()
    d_in = _unit((1, -0.1, 0))
    ct = -d_in @ (0, 1, 0)
    assert 1.5 * math.sqrt(1 - ct * ct) > 1.0
    for key in range(10):
        ok, d, _, _ = _scatter(os_, m, (0, 0, 0), d_in, (0, 0, 0), (0, 1, 0), False, key)
        assert ok and d[1] > 0 and abs(d[0] - d_in[0]) <= 1e-10


def test_layered_basic_behaviour():
    """TestLayeredBasicBehavior: red Lambertian over blue Lambertian scatters with both attenuation
    components positive and none above 0.5. The outer Lambertian always scatters outward, so the result
    is the outer layer's own: attenuation albedo/pi, pdf cos/pi (layered.go:24-60)."""
    os_, m = _material_scene(lambda s: s.layered(s.lambertian((0.8, 0.1, 0.1)), s.lambertian((0.1, 0.1, 0.8))))
    for key in range(64):
        ok, d, att, pdf = _scatter(os_, m, (0, 1, 0), (0, -1, 0), (0, 0, 0), (0, 1, 0), True, key)
        assert ok and att[0] > 0 and att[2] > 0 and att.max() <= 0.5
        assert np.allclose(att, np.array([0.8, 0.1, 0.1]) / math.pi, rtol=1e-14)
        assert pdf == pytest.approx(_unit(d)[1] / math.pi, rel=1e-12)


def test_layered_outward_scattering():
    """TestLayeredOutwardScattering: a mirror outer layer reflects the 45-degree ray upward with the outer
    attenuation only (0.9 within 0.1; exactly 0.9 here)."""
    os_, m = _material_scene(lambda s: s.layered(s.metal((0.9, 0.9, 0.9), 0.0), s.lambertian((0.8, 0.1, 0.1))))
    ok, d, att, _ = _scatter(os_, m, (-1, 1, 0), _unit((1, -1, 0)), (0, 0, 0), (0, 1, 0), True, 0)
    assert ok and d[1] > 0 and np.array_equal(att, (0.9, 0.9, 0.9))


def test_layered_inward_scattering():
    """TestLayeredInwardScattering: glass over red Lambertian at normal incidence: when the glass refracts
    inward, the inner layer scatters and the attenuation is red-tinted (exactly 1 x red/pi); otherwise the
    glass reflects with attenuation 1. Reflection happens with the Schlick probability r0 = 0.04."""
    os_, m = _material_scene(lambda s: s.layered(s.dielectric(1.5), s.lambertian((0.8, 0.1, 0.1))))
    n, n_in = 4000, 0
    for key in range(n):
        ok, d, att, _ = _scatter(os_, m, (0, 1, 0), (0, -1, 0), (0, 0, 0), (0, 1, 0), True, key)
        assert ok
        if att[0] < 0.9 or att[1] < 0.9 or att[2] < 0.9:
            n_in += 1
            assert att[0] > att[1] and att[0] > att[2]
            assert np.allclose(att, np.array([0.8, 0.1, 0.1]) / math.pi, rtol=1e-14)
        else:
            assert np.array_equal(att, (1.0, 1.0, 1.0)) and np.allclose(_unit(d), (0, 1, 0), atol=1e-12)
    r0 = 0.04
    assert n_in > 0 and abs((n - n_in) - n * r0) <= 5 * math.sqrt(n * r0 * (1 - r0))


# ---- SampleOnUnitSphere (core/sampling.go:97-104) -------------------------------------------------------
def test_sample_on_unit_sphere_closed_forms():
    """Unit length; on a stratified 256x256 grid the moments are those of the uniform sphere:
    E[x] = E[y] = E[z] = 0, E[x^2] = E[y^2] = E[z^2] = 1/3, E[xy] = 0; and z = 1 - 2 u exactly."""
    n = 256
    g = (np.arange(n) + 0.5) / n
    pts = np.zeros((n * n, 3))
    k = 0
    out = np.zeros(3)
    for u in g:
        for v in g:
            O.lib().orc_sample_on_unit_sphere(_p(_vec(u, v)), _p(out))
            pts[k] = out
            k += 1
    assert np.allclose(np.linalg.norm(pts, axis=1), 1.0, atol=1e-14)
    assert np.allclose(pts.mean(0), 0.0, atol=2e-3)
    assert np.allclose((pts ** 2).mean(0), 1.0 / 3.0, atol=2e-3)
    assert abs((pts[:, 0] * pts[:, 1]).mean()) < 1e-3
    assert np.allclose(pts[::n, 2], 1 - 2 * g, atol=1e-15)


# ---- environment-map light (no reference counterpart) ---------------------------------------------------
def _env_scene(rgb):
    s = P.SceneDesc()
    s.add_envmap_light(np.ascontiguousarray(rgb, np.float32))
    s.set_camera(center=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=40)
    return s


def _env_pdf(os_, d):
    return O.lib().orc_light_pdf(os_.h, 0, _p(_vec(0, 0, 0)), _p(_vec(0, 1, 0)), _p(_vec(*d)))


def _env_sample(os_, s):
    out = np.zeros(8)
    O.lib().orc_light_sample(os_.h, 0, _p(_vec(0, 0, 0)), _p(_vec(0, 1, 0)), _p(_vec(*s)), _p(out))
    return out[0:3], out[4:7], out[7]


def _random_env(h, w, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((h, w, 3)) ** 3 * 5 + 0.01).astype(np.float32)


def test_env_pdf_integrates_to_one():
    """The solid-angle density of the environment light integrates to 1 over the sphere (midpoint rule on
    a grid 4x finer than the map; exact up to the O(dtheta^2) of the per-row density)."""
    rgb = _random_env(16, 32, 1)
    os_ = O.OracleScene(_env_scene(rgb))
    nt, nphi = 16 * 8, 32 * 8
    th = (np.arange(nt) + 0.5) / nt * math.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * math.pi
    tot = 0.0
    for t in th:
        st = math.sin(t)
        row = sum(_env_pdf(os_, (st * math.cos(p), math.cos(t), st * math.sin(p))) for p in ph)
        tot += row * st * (math.pi / nt) * (2 * math.pi / nphi)
    assert tot == pytest.approx(1.0, abs=2e-3)


def test_env_sampling_follows_texel_weights():
    """Stratified samples land in each texel with probability luminance x sin(row centre) / total (the
    sampling distribution of the map), to the resolution of the strata; the returned pdf equals the
    light's PDF of the returned direction and the emission is the texel the direction falls in."""
    h, w = 4, 8
    rgb = _random_env(h, w, 2)
    os_ = O.OracleScene(_env_scene(rgb))
    lum = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    wgt = lum * np.sin((np.arange(h)[:, None] + 0.5) / h * math.pi)
    want = wgt / wgt.sum()
    n = 200
    g = (np.arange(n) + 0.5) / n
    hist = np.zeros((h, w))
    for a in g:
        for b in g:
            d, e, pdf = _env_sample(os_, (a, b))
            theta = math.acos(max(-1.0, min(1.0, d[1])))
            phi = math.atan2(d[2], d[0]) % (2 * math.pi)
            y, x = min(int(theta / math.pi * h), h - 1), min(int(phi / (2 * math.pi) * w), w - 1)
            hist[y, x] += 1
            assert pdf == pytest.approx(_env_pdf(os_, d), rel=1e-12)
            assert np.allclose(e, rgb[y, x], rtol=1e-6)
    assert np.abs(hist / n ** 2 - want).max() <= 2.0 / n


def test_env_map_orientation():
    """Row 0 is the +Y pole; column 0 starts at +X and columns advance toward +Z (phi = atan2(z, x)),
    read back as the radiance of miss rays (an empty scene returns the infinite lights' emission), and a
    direction sampled from texel (y, x) carries that texel's radiance."""
    h, w = 4, 8
    rgb = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        for x in range(w):
            rgb[y, x] = (y + 1, x + 1, 0.5)
    os_ = O.OracleScene(_env_scene(rgb))
    p = P.sampling_params(8, 8, 4, 5, 1)
    # 4 rows of 45 degrees of polar angle from +Y, 8 columns of 45 degrees of azimuth from +X toward +Z
    cases = [((0, 1, 0), (0, 0)), ((0, -1, 0), (3, 0)), ((1, 0.2, 0.2), (1, 0)), ((0.2, 0.2, 1), (1, 1)),
             ((-0.2, 0.2, 1), (1, 2)), ((-1, -0.2, 0.2), (2, 3)), ((0.2, -0.2, -1), (2, 6)), ((1, 0.9, -0.05), (1, 7))]
    for d, (y, x) in cases:
        out = np.zeros(3)
        O.lib().orc_ray_color(os_.h, C.byref(p), _p(_vec(0, 0, 0)), _p(_vec(*_unit(d))), 7, _p(out))
        assert np.allclose(out, (y + 1, x + 1, 0.5)), (d, out)
    for a in (np.arange(2 * h) + 0.5) / (2 * h):
        for b in (np.arange(2 * w) + 0.5) / (2 * w):
            d, e, _ = _env_sample(os_, (a, b))
            y = int(math.acos(max(-1.0, min(1.0, d[1]))) / (math.pi / 4))
            x = int((math.atan2(d[2], d[0]) % (2 * math.pi)) / (math.pi / 4))
            assert np.allclose(e, (y + 1, x + 1, 0.5))


def test_constant_env_map_furnace_matches_uniform_light():
    """A Lambertian sphere (albedo 0.5) under a constant environment map of radiance 1: every sample of
    the sphere has expectation exactly the albedo (the reference estimator's two MIS weights sum to one
    per direction and every bounce escapes the convex sphere), as under a uniform infinite light of
    radiance 1, where each sample is exactly 0.5 (both densities are cos/pi). Sky pixels are exactly 1."""
    s_env, _ = scenes.furnace_scene(0.5, width=16)
    s_uni, cfg = scenes.furnace_scene(0.5, width=16)
    s_env.lights = []
    s_env.add_envmap_light(np.ones((16, 32, 3), np.float32))
    os_env, os_uni = O.OracleScene(s_env), O.OracleScene(s_uni)
    p = P.sampling_params(16, 16, cfg["max_depth"], cfg["rr_min_bounces"], 1, seed=4)
    pix = np.array([[8, 8]] * 20000, np.int32)  # the sphere covers the centre pixel
    smp = np.arange(20000, dtype=np.int32)
    uni = os_uni.render_samples(p, pix, smp)
    env = os_env.render_samples(p, pix, smp)
    assert np.allclose(uni, 0.5, atol=1e-12)
    m, se = env.mean(0), env.std(0) / math.sqrt(len(env))
    assert np.all(np.abs(m - 0.5) <= 4 * se + 1e-3), (m, se)
    corner = os_env.render_samples(p, np.array([[0, 0]] * 16, np.int32), np.arange(16, dtype=np.int32))
    assert np.allclose(corner, 1.0, atol=1e-12)
