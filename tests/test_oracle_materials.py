"""Pins the oracle's image textures, hit UVs and Mix material (pkg/material/image_texture.go, mix.go,
the UV of each shape's Hit) against the reference's own test vectors and closed forms. CPU only."""
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


def tex_eval(pixels, u, v):
    a = np.ascontiguousarray(pixels, dtype=np.float32)
    out = np.zeros(3)
    O.lib().orc_texture_eval(O._p(a), a.shape[1], a.shape[0], u, v, O._p(out))
    return out


def test_image_texture_reference_cases():
    g = G["image_texture"]
    c = g["checker_2x2"]
    for case in c["cases"]:
        assert np.array_equal(tex_eval(c["pixels"], *case["uv"]), case["rgb"])
    w = g["wrap_1x1"]
    for uv in w["uvs"]:
        assert np.array_equal(tex_eval(w["pixels"], *uv), w["rgb"])
    ramp = np.array([[[(y * 4 + x) / 15.0] * 3 for x in range(4)] for y in range(4)])
    for case in g["ramp_4x4"]["cases"]:
        assert np.allclose(tex_eval(ramp, *case["uv"]), case["rgb"], atol=1e-7)


def test_texture_lookup_wraps_by_truncation():
    """u - int(u) truncates toward zero, so -0.25 wraps to 0.75 and -1.25 to 0.75; v is flipped."""
    ramp = np.array([[[x, y, 0] for x in range(4)] for y in range(4)], float)
    assert np.array_equal(tex_eval(ramp, -0.25, 0.9), [3, 0, 0])
    assert np.array_equal(tex_eval(ramp, -1.25, -0.1), [3, 0, 0])  # v: -0.1 -> 0.9 -> row int(0.1*4) = 0
    assert np.array_equal(tex_eval(ramp, 0.999999, 0.0), [3, 3, 0])  # v = 0 -> row int(1.0*4) clamped to 3


def test_procedural_textures_match_reference_definitions():
    ck = P.checkerboard_texture(8, 6, 2, (1, 1, 1), (0, 0, 0))
    assert ck.shape == (6, 8, 3) and np.all(ck[0, 0] == 1) and np.all(ck[0, 2] == 0) and np.all(ck[2, 2] == 1)
    uvd = P.uv_debug_texture(5, 3)
    assert np.allclose(uvd[2, 4], [1, 1, 0]) and np.allclose(uvd[1, 2], [0.5, 0.5, 0])
    gr = P.gradient_texture(2, 3, (1, 0, 0), (0, 1, 0))
    assert np.allclose(gr[0, 1], [1, 0, 0]) and np.allclose(gr[1, 0], [0.5, 0.5, 0]) and np.allclose(gr[2, 0], [0, 1, 0])


def shape_uv(kind, g, o, d):
    out = np.zeros(10)
    h = O.lib().orc_shape_hit(kind, O._p(O.vec(*g)), O._p(O.vec(*o)), O._p(O.vec(*d)), 0.001, 1000.0, O._p(out))
    assert h == 1
    return out[8:10]


@pytest.mark.parametrize("kind,g,o,d,uv", [
    # sphere (sphere.go:64-68): v = acos(-n.y)/pi, u = (atan2(-n.z, n.x) + pi)/(2 pi)
    (1, [0, 0, 0, 1.0], [0, 5, 0], [0, -1, 0], [0.5, 1.0]),
    (1, [0, 0, 0, 1.0], [5, 0, 0], [-1, 0, 0], [0.5, 0.5]),
    (1, [0, 0, 0, 1.0], [0, 0, 5], [0, 0, -1], [0.25, 0.5]),
    (1, [0, 0, 0, 1.0], [0, -5, 0], [0, 1, 0], [0.5, 0.0]),
    # quad (quad.go:165): (alpha, beta)
    (2, [0, 0, 0, 2, 0, 0, 0, 0, 4], [0.5, 1, 1], [0, -1, 0], [0.25, 0.25]),
    # disc (disc.go:76-80): right = (0,0,1), up = (1,0,0) for normal +y
    (3, [0, 0, 0, 0, 1, 0, 2.0], [1, 1, 0.5], [0, -1, 0], [0.625, 0.75]),
    # cylinder body: tangent (0,0,-1), bitangent (-1,0,0) for axis +y; caps: (0,0,-+1), (-1,0,0)
    (4, [0, 0, 0, 0, 2, 0, 1.0, 1], [2, 0.5, 0], [-1, 0, 0], [0.25, 0.25]),
    (4, [0, 0, 0, 0, 2, 0, 1.0, 1], [0.5, 3, 0], [0, -1, 0], [0.5, 0.25]),
    (4, [0, 0, 0, 0, 2, 0, 1.0, 1], [0.5, -1, 0], [0, 1, 0], [0.5, 0.25]),
    # cone body at half height
    (5, [0, 0, 0, 1.0, 0, 2, 0, 0.0, 0], [2, 1, 0], [-1, 0, 0], [0.25, 0.5]),
])
def test_hit_uv_closed_forms(kind, g, o, d, uv):
    assert np.allclose(shape_uv(kind, g, o, d), uv, atol=1e-12)


def test_triangle_vertex_uvs_interpolate():
    s = P.SceneDesc()
    s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    m = s.lambertian((0.5, 0.5, 0.5))
    s.mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float), np.array([[0, 1, 2]]), m,
           uvs=np.array([[0.1, 0.2], [0.9, 0.2], [0.1, 0.8]]))
    r = np.zeros(1, P.RAY_DTYPE)
    r["org"], r["dir"], r["tmin"], r["tmax"] = (0.25, 0.25, -1), (0, 0, 1), 0.001, 10
    t, prim, u, v = O.OracleScene(s).intersect(r)
    # weights 1 - u - v = 0.5, u = 0.25 (vertex 1), v = 0.25 (vertex 2)
    assert prim[0] == 0 and abs(u[0] - 0.3) < 1e-7 and abs(v[0] - 0.35) < 1e-7


def _plane_furnace(texture_rgb):
    s = P.SceneDesc()
    s.set_camera((0, 3, 0.01), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
    tex = s.texture(texture_rgb)
    s.quad((-1, 0, -1), (2, 0, 0), (0, 0, 2), s.lambertian((0, 0, 0), texture=tex))
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    return s


def test_textured_plane_under_uniform_sky_returns_the_texel():
    """A single Lambertian plane under a uniform sky of radiance 1 returns exactly its albedo (direct
    and indirect weighted 1/2 each, every bounce escapes), so the radiance of a ray is the texel at
    the hit's (alpha, beta)."""
    W, H = 8, 4
    s = _plane_furnace(P.uv_debug_texture(W, H))
    os_ = O.OracleScene(s)
    rng = np.random.default_rng(3)
    for _ in range(20):
        x, z = rng.uniform(-0.99, 0.99, 2)
        out = np.zeros(3)
        O.lib().orc_ray_color(os_.h, C.byref(P.sampling_params(8, 8, 4, 10, 1)), O._p(O.vec(x, 1, z)),
                              O._p(O.vec(0, -1, 0)), 5, O._p(out))
        alpha, beta = (x + 1) / 2, (z + 1) / 2
        px, py = min(int(alpha * W), W - 1), min(int((1 - beta) * H), H - 1)
        assert np.allclose(out, [px / (W - 1), py / (H - 1), 0], atol=1e-12), (x, z, out)


def test_mix_furnace_two_values():
    """Mix(Lambertian a1, Lambertian a2, r) on a sphere under a uniform sky: the BRDF and pdf are the
    r-weighted averages (abar/pi, cos/pi) while the scattered direction's weight uses the picked
    component, so a sample is exactly (abar + a_k)/2 with k = 2 for a fraction r of the samples."""
    a1, a2, r = 0.2, 0.8, 0.3
    abar = (1 - r) * a1 + r * a2
    s = P.SceneDesc()
    s.set_camera((0, 0, 4), (0, 0, 0), (0, 1, 0), 32, 1.0, 30)
    m = s.mix(s.lambertian((a1, a1, a1)), s.lambertian((a2, a2, a2)), r)
    s.sphere((0, 0, 0), 1.0, m)
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    os_ = O.OracleScene(s)
    sm, _, _ = os_.render(P.sampling_params(32, 32, 5, 10, 1), 1)
    v = sm[..., 0]
    on = v != 1.0
    lo, hi = (abar + a1) / 2, (abar + a2) / 2
    assert np.all(np.isclose(v[on], lo, atol=1e-12) | np.isclose(v[on], hi, atol=1e-12))
    frac = np.isclose(v[on], hi, atol=1e-12).mean()
    assert on.sum() > 300 and abs(frac - r) < 0.07, frac


def test_scene_constructors_with_textures_and_mix_build():
    for fn in (lambda: scenes.texture_scene(width=32), lambda: scenes.sphere_grid_scene(4, "mixed", width=32),
               lambda: scenes.sphere_grid_scene(4, "glossy", width=32)):
        s, cfg = fn()
        c, r, n = O.OracleScene(s).world()
        assert n > 0 and np.all(np.isfinite(c))
