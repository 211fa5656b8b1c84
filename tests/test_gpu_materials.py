"""Image textures, hit UVs and the Mix material on the HIP path against the float64 oracle
(pkg/material/image_texture.go, procedural_textures.go, mix.go; the UVs of every shape's Hit).

Bars as for the other scenes: path tracing >= 99 % of per-sample radiances within 1e-3 relative
(floor 1e-3); BDPT >= 98 % camera-side and >= 97 % of splat pixels within 1e-2 relative
(floor 0.1), splat totals within 1 %. A float32 UV that lands on the other side of a texel edge
than the float64 one changes a sample, so textured scenes also count such samples here.
"""
import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu


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


def _textured_mesh():
    """Two triangles with tiled vertex UVs (0..3, wrapping) under a checkerboard, plus a Mix sphere."""
    s = P.SceneDesc()
    s.set_camera((0, 2, 4), (0, 0, 0), (0, 1, 0), 96, 1.0, 50)
    tex = s.texture(P.checkerboard_texture(64, 64, 8, (0.9, 0.8, 0.2), (0.1, 0.2, 0.7)))
    m = s.lambertian((0, 0, 0), texture=tex)
    v = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2]], float)
    f = np.array([[0, 1, 2], [0, 2, 3]])
    uv = np.array([[0, 0], [3, 0], [3, 3], [0, 3]], float)
    s.mesh(v, f, m, uvs=uv)
    s.sphere((0, 0.6, 0), 0.6, s.mix(s.lambertian((0.8, 0.3, 0.3)), s.metal((0.9, 0.9, 0.9), 0.2), 0.4))
    tex2 = s.texture(P.uv_debug_texture(32, 32))
    s.sphere((1.3, 0.4, 0.5), 0.4, s.metal((0, 0, 0), 0.3, texture=tex2))
    s.add_sphere_light((2, 4, 2), 0.8, (12, 12, 12))
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    return s, dict(max_depth=8, rr_min_bounces=4)


SCENES = {
    "texture": lambda: scenes.texture_scene(width=128),
    "spheregrid_mixed": lambda: scenes.sphere_grid_scene(6, "mixed", width=128),
    "spheregrid_glossy": lambda: scenes.sphere_grid_scene(6, "glossy", width=128),
    "textured_mesh": _textured_mesh,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_path_tracing_parity(name):
    s, cfg = SCENES[name]()
    g, r, _, _ = _first_sample(s, cfg, 2, P.INTEGRATOR_PATH)
    frac = _agree(g, r, 1e-3, 1.0)
    print(f"{name} PT: agreement {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f}")
    assert frac >= 0.99
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3


@pytest.mark.parametrize("name", sorted(SCENES))
def test_bdpt_parity(name):
    s, cfg = SCENES[name]()
    g, r, gs, rs = _first_sample(s, cfg, 3, P.INTEGRATOR_BDPT)
    frac = _agree(g, r, 1e-3, 1.0)
    sfrac = _agree(gs, rs, 1e-2, 1e-1)
    print(f"{name} BDPT: camera-side {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f}; "
          f"splat total gpu {gs.sum():.5f} oracle {rs.sum():.5f}; splat pixels {sfrac:.5f}")
    assert frac >= 0.98
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3
    assert abs(gs.sum() - rs.sum()) <= 0.01 * abs(rs.sum()) + 1e-3
    assert sfrac >= 0.97


def test_mix_furnace_two_values():
    """The oracle's closed form on the device: every sphere sample is (abar + a_k)/2."""
    a1, a2, r = 0.2, 0.8, 0.3
    abar = (1 - r) * a1 + r * a2
    s = P.SceneDesc()
    s.set_camera((0, 0, 4), (0, 0, 0), (0, 1, 0), 64, 1.0, 30)
    s.sphere((0, 0, 0), 1.0, s.mix(s.lambertian((a1, a1, a1)), s.lambertian((a2, a2, a2)), r))
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    pr = P.ProgressiveRaytracer(P.Scene(s), P.sampling_params(64, 64, 5, 10, 1))
    pr.render_samples(1)
    v = pr.read()[0][..., 0].astype(np.float64)
    on = v != 1.0
    lo, hi = (abar + a1) / 2, (abar + a2) / 2
    ok = np.isclose(v[on], lo, atol=2e-6) | np.isclose(v[on], hi, atol=2e-6)
    assert on.sum() > 1000 and ok.mean() >= 0.999
    assert abs(np.isclose(v[on], hi, atol=2e-6).mean() - r) < 0.03


def test_invalid_materials_rejected():
    def base():
        s = P.SceneDesc()
        s.set_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 8, 1.0, 40)
        return s
    s = base()
    s.sphere((0, 0, 0), 1, s.lambertian((1, 1, 1), texture=2))
    s.texture(np.zeros((2, 2, 3)))
    with pytest.raises(P.PrtError):
        P.Scene(s)
    s = base()
    inner = s.layered(s.dielectric(1.5), s.lambertian((0.5, 0.5, 0.5)))
    s.sphere((0, 0, 0), 1, s.mix(inner, s.lambertian((1, 1, 1)), 0.5))
    with pytest.raises(P.PrtError):
        P.Scene(s)
    s = base()
    mx = s.mix(s.lambertian((1, 1, 1)), s.metal((1, 1, 1), 0.0), 0.5)
    s.sphere((0, 0, 0), 1, s.layered(s.dielectric(1.5), mx))
    with pytest.raises(P.PrtError):
        P.Scene(s)
