"""Per-sample parity at the benchmarked scene sizes: the 1.8M-triangle glass icosphere under the HDR sky
(frequency 300, the full mesh) for path tracing and BDPT, BDPT on the metre Cornell box at depth 8 (the
benchmarked BDPT configuration), and a weighted light sampler with non-uniform weights
(pkg/lights/weighted_light_sampler.go:82-96) on the device.

Both sides draw the same counter-RNG values per (pixel, sample, bounce, dimension), so single samples
are compared one to one (bars as in test_gpu_pt.py / test_gpu_bdpt.py: >= 99 % (PT) / 98 % (BDPT camera
side) of per-sample radiances within 1e-3 relative, floor 1e-3; BDPT splat totals within 1 % and >= 97 %
of splat pixels within 1e-2 relative, floor 0.1).
"""
import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu


def _params(desc, cfg, integrator, seed):
    w, h = desc.image_size
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], 1, seed=seed, integrator=integrator)


def _agree(g, r, rel=1e-3, floor=1.0):
    d = np.abs(g.astype(np.float64) - r)
    return np.all(d <= rel * np.maximum(np.abs(r), floor), axis=2).mean()


def _first_sample(desc, cfg, integrator, seed, scene=None, os_=None):
    """(gpu camera-side radiance, gpu splats, oracle radiance, oracle splats) of sample 0 of every pixel."""
    scene = scene or P.Scene(desc)
    p = _params(desc, cfg, integrator, seed)
    pr = P.ProgressiveRaytracer(scene, p)
    pr.render_samples(1)
    mean, _, _, cnt = pr.read()
    assert np.all(cnt == 1)
    bd = integrator == P.INTEGRATOR_BDPT
    sp = pr.splats() if bd else np.zeros_like(mean)
    pr.close()
    w, h = desc.image_size
    os_ = os_ or O.OracleScene(desc)
    ys, xs = np.mgrid[0:h, 0:w]
    pix = np.stack([xs.ravel(), ys.ravel()], 1)
    if bd:
        r, rs = os_.render_samples(p, pix, np.zeros(len(pix)), splats=True)
    else:
        r, rs = os_.render_samples(p, pix, np.zeros(len(pix))), None
    return mean.astype(np.float64) - sp, sp, r.reshape(h, w, 3), rs


def _check(name, g, gs, r, rs, bar):
    frac = _agree(g, r)
    print(f"{name}: per-sample agreement {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f}")
    assert frac >= bar
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3
    if rs is not None:
        tg, tr = gs.sum(), rs.sum()
        fs = _agree(gs, rs, 1e-2, 1e-1)
        print(f"{name}: splat total gpu {tg:.6f} oracle {tr:.6f}; splat pixels agreeing {fs:.5f}")
        assert abs(tg - tr) <= 0.01 * abs(tr) + 1e-3
        assert fs >= 0.97


@pytest.fixture(scope="module")
def mesh_sky():
    """The benchmark's 1.8M-triangle scene (20 x 300^2 faces) rendered at 128x128; one device scene and
    one oracle scene shared by the tests."""
    desc, cfg = scenes.mesh_sky_scene(width=128)
    scene = P.Scene(desc)
    assert scene.info().num_triangles == 1_800_000
    return desc, cfg, scene, O.OracleScene(desc)


def test_full_mesh_sky_path_tracing(mesh_sky):
    desc, cfg, scene, os_ = mesh_sky
    g, gs, r, rs = _first_sample(desc, cfg, P.INTEGRATOR_PATH, 3, scene, os_)
    _check("1.8M-triangle mesh + sky, PT 128x128", g, gs, r, rs, 0.99)


def test_full_mesh_sky_bdpt(mesh_sky):
    desc, cfg, scene, os_ = mesh_sky
    g, gs, r, rs = _first_sample(desc, cfg, P.INTEGRATOR_BDPT, 4, scene, os_)
    _check("1.8M-triangle mesh + sky, BDPT 128x128", g, gs, r, rs, 0.98)


def test_cornell_metre_bdpt_depth8():
    desc, cfg = scenes.cornell_metre(width=128)
    assert cfg["max_depth"] == 8
    g, gs, r, rs = _first_sample(desc, cfg, P.INTEGRATOR_BDPT, 6)
    _check("metre Cornell BDPT depth 8, 128x128", g, gs, r, rs, 0.98)


@pytest.mark.parametrize("integrator", [P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT])
def test_nonuniform_light_weights(integrator):
    """The default scene's sphere light and sky sampled with weights 3:1 (NewWeightedLightSampler); the
    selection CDF walk, the pdf scaling by the selection probability and the MIS light-pdf sum all use
    the weights."""
    desc, cfg = scenes.default_scene(width=128)
    desc.light_weights = [3.0, 1.0]
    g, gs, r, rs = _first_sample(desc, cfg, integrator, 8)
    _check(f"default scene, light weights 3:1, integrator {integrator}", g, gs, r, rs,
           0.99 if integrator == P.INTEGRATOR_PATH else 0.98)
    # the weights change the image's samples (against the uniform sampler on the same stream)
    desc_u, _ = scenes.default_scene(width=128)
    gu, _, _, _ = _first_sample(desc_u, cfg, integrator, 8)
    assert _agree(g, gu) < 0.9


def _film(scene, desc, cfg, spp, seed):
    w, h = desc.image_size
    pr = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp,
                                                         seed=seed))
    pr.render_samples(spp)
    out = pr.read()
    pr.close()
    return out


def _mesh_under_sky():
    """Triangle-only scene: the 1.31M-triangle displaced icosphere under a uniform infinite light (shadow
    rays towards the light are any-hit traversals of the mesh)."""
    desc, _ = scenes.mesh_scene(frequency=256, width=96)
    desc.add_uniform_infinite_light((0.8, 0.9, 1.0))
    return desc, dict(max_depth=8, rr_min_bounces=3)


@pytest.mark.parametrize("case", ["mesh_sky_mixed", "mesh_triangles_only"])
def test_long_ray_parking_leaves_films_unchanged(case, mesh_sky, monkeypatch):
    """Rays still traversing after the step budget (PRT_PT_PARK_STEPS, PRT_PT_PARK_STEPS_SMALL for launches
    with fewer rays than the grid's lanes; scenes of >= 2^16 primitives park after 128 / 64 steps by
    default) are finished by a whole wave each (coop_trace); the closest hit is the
    least (t, primitive index) pair whichever traversal finds it, so films are bit-identical with parking
    off (0), at the default, and at budgets of 24 and 4 steps that park most rays (closest and any hit)."""
    if case == "mesh_sky_mixed":
        desc, cfg, scene, _ = mesh_sky
    else:
        desc, cfg = _mesh_under_sky()
        scene = P.Scene(desc)
    films = {}
    for budget in ("0", None, "24", "4"):
        for var in ("PRT_PT_PARK_STEPS", "PRT_PT_PARK_STEPS_SMALL"):
            if budget is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, budget)
        films[budget] = _film(scene, desc, cfg, 4, seed=11)
    base = films["0"]
    assert base[0].mean() > 0
    for b, f in films.items():
        for x, y in zip(base, f):
            assert np.array_equal(x, y), f"film with park budget {b} differs from the film without parking"


def test_long_rays_and_tail_in_adaptive_passes(mesh_sky, monkeypatch):
    """Adaptive films (rounds over an active-pixel list, tile_renderer.go:50-101) run the same launches:
    with long-ray parking at a 4-step budget and the tail kernel forced from bounce 1, the films after three
    progressive passes equal the all-wavefront render without parking, sample counts included."""
    desc, cfg, scene, _ = mesh_sky
    w, h = desc.image_size

    def render():
        p = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], 16, max_passes=4, seed=13,
                              adaptive_min_samples=0.25, adaptive_threshold=0.05)
        pr = P.ProgressiveRaytracer(scene, p)
        for k in range(1, 4):
            pr.render_pass(k)
        out = pr.read()
        pr.close()
        return out

    for var in ("PRT_PT_PARK_STEPS", "PRT_PT_PARK_STEPS_SMALL", "PRT_PT_TAIL_PATHS"):
        monkeypatch.setenv(var, "0")
    base = render()
    monkeypatch.setenv("PRT_PT_PARK_STEPS", "4")
    monkeypatch.setenv("PRT_PT_PARK_STEPS_SMALL", "4")
    monkeypatch.setenv("PRT_PT_TAIL_PATHS", str(1 << 40))
    render()  # a non-adaptive chunk has not recorded the profile: record it with a plain film
    pr = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], 2, seed=1))
    pr.render_samples(2)
    pr.close()
    other = render()
    assert base[3].min() >= 1 and base[3].max() <= 16
    for x, y in zip(base, other):
        assert np.array_equal(x, y)
