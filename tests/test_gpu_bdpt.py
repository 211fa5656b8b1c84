"""Bidirectional path-tracing parity of the HIP megakernel against the float64 oracle restatement of
BDPTIntegrator (pkg/integrator/bdpt.go, bdpt_mis.go) and the splat path of pkg/renderer
(SplatQueue + processSplats, progressive.go:370-390).

Both sides draw the same counter-RNG values per (pixel, sample, subpath bounce, strategy), so the
camera-side radiance of single samples is compared one to one, and the light-tracing splats of a
whole 1-spp image are compared pixel by pixel. Float32 vs float64 paths agree except where a
float32 decision (Fresnel choice, edge hit, visibility at a grazing connection) flips a branch.
Bars: >= 98 % of per-sample radiances within 1e-3 relative (per channel, floor 1e-3 absolute);
>= 97 % of splat pixels within 1e-2 relative (floor 1e-3) and the splat total within 1 %;
closed-form scenes exact to float32 rounding.
"""
import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu


def params_for(desc, cfg, spp, seed=0, max_passes=1):
    w, h = desc.image_size
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=max_passes, seed=seed,
                             integrator=P.INTEGRATOR_BDPT)


def gpu_first_sample(desc, cfg, seed=0):
    """(per-pixel radiance of sample 0 without splats, splat image) of a 1-spp BDPT render."""
    scene = P.Scene(desc)
    pr = P.ProgressiveRaytracer(scene, params_for(desc, cfg, 1, seed))
    pr.render_samples(1)
    mean, ls, lq, cnt = pr.read()
    sp = pr.splats()
    assert np.all(cnt == 1)
    return mean.astype(np.float64) - sp, sp


def oracle_first_sample(desc, cfg, seed=0):
    w, h = desc.image_size
    os_ = O.OracleScene(desc)
    ys, xs = np.mgrid[0:h, 0:w]
    pix = np.stack([xs.ravel(), ys.ravel()], 1)
    out, sp = os_.render_samples(params_for(desc, cfg, 1, seed), pix, np.zeros(len(pix)), splats=True)
    return out.reshape(h, w, 3), sp


def agree_fraction(g, r, rel=1e-3, floor=1.0):
    d = np.abs(g.astype(np.float64) - r)
    tol = rel * np.maximum(np.abs(r), floor)
    return np.all(d <= tol, axis=2).mean()


def small_cornell(width=64):
    s, cfg = scenes.cornell_scene(width=width)
    return s, dict(cfg, max_depth=5)


SCENES = {
    "cornell": lambda: small_cornell(64),
    "default": lambda: scenes.default_scene(width=96),
    "mesh_sky": lambda: scenes.mesh_sky_scene(frequency=8, width=64),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_per_sample_parity(name):
    s, cfg = SCENES[name]()
    g, gs = gpu_first_sample(s, cfg, seed=5)
    r, rs = oracle_first_sample(s, cfg, seed=5)
    frac = agree_fraction(g, r)
    print(f"{name}: camera-side agreement {frac:.5f}; mean gpu {g.mean():.5f} oracle {r.mean():.5f}")
    assert frac >= 0.98
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3
    tot_g, tot_r = gs.sum(), rs.sum()
    print(f"{name}: splat total gpu {tot_g:.5f} oracle {tot_r:.5f}; "
          f"pixels agreeing {agree_fraction(gs, rs, 1e-2, 1e-1):.5f}")
    assert abs(tot_g - tot_r) <= 0.01 * abs(tot_r) + 1e-3
    assert agree_fraction(gs, rs, 1e-2, 1e-1) >= 0.97


def test_uniform_light_every_pixel_exact():
    """No geometry, uniform infinite light of radiance 1: every pixel is exactly 1 and nothing splats."""
    s, cfg = scenes.empty_uniform_scene(width=32)
    scene = P.Scene(s)
    pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4))
    pr.render_samples(4)
    mean, ls, lq, cnt = pr.read()
    assert np.all(pr.splats() == 0)
    assert np.all(mean == 1.0) and np.all(cnt == 4)


def test_deterministic_and_reset():
    s, cfg = small_cornell(48)
    scene = P.Scene(s)
    pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4))
    pr.render_samples(4)
    a = pr.read()[0]
    sa = pr.splats()
    pr.reset()
    pr.render_samples(2)
    pr.render_samples(2)
    b = pr.read()[0]
    sb = pr.splats()
    # each sample is added with its own compensated step in sample order and splats are exact fixed-point
    # sums: 4 samples at once and 2 + 2 give the same film
    assert np.array_equal(sa, sb)
    assert np.array_equal(a, b)
    pr2 = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4))
    pr2.render_samples(4)
    assert np.array_equal(pr2.read()[0], a) and np.array_equal(pr2.splats(), sa)


def test_bdpt_matches_pt_mean():
    """BDPT and PT estimate the same image: mean luminance within 3 % on the Cornell box
    (progressive_integration_test.go compares the two integrators with 5-15 % bounds)."""
    s, cfg = scenes.cornell_scene(width=64)
    scene = P.Scene(s)
    w, h = s.image_size
    pt = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], 1024,
                                                         seed=1))
    pt.render_samples(1024)
    bd = P.ProgressiveRaytracer(scene, params_for(s, cfg, 256, seed=2))
    bd.render_samples(256)
    lum = lambda im: (0.2126 * im[..., 0] + 0.7152 * im[..., 1] + 0.0722 * im[..., 2]).mean()
    lp, lb = lum(pt.read()[0]), lum(bd.read()[0])
    print(f"PT {lp:.5f} BDPT {lb:.5f}")
    assert abs(lb - lp) <= 0.03 * lp


def test_statistical_mean_vs_oracle():
    """Image mean over 16 spp agrees with the oracle's 16-spp image mean (same samples)."""
    s, cfg = small_cornell(32)
    scene = P.Scene(s)
    pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 16, seed=9))
    pr.render_samples(16)
    g = pr.read()[0].astype(np.float64)
    os_ = O.OracleScene(s)
    sm, ls, lq, sp = os_.render(params_for(s, cfg, 16, seed=9), 16, splats=True)
    r = (sm + sp) / 16
    print(f"gpu {g.mean():.5f} oracle {r.mean():.5f}")
    assert abs(g.mean() - r.mean()) <= 0.01 * r.mean()


def test_depth_limit_rejected():
    s, cfg = small_cornell(16)
    scene = P.Scene(s)
    p = params_for(s, dict(cfg, max_depth=65), 1)
    with pytest.raises(P.PrtError):
        P.ProgressiveRaytracer(scene, p)


PAIR_SCENES = {
    "cornell_metre_d8": lambda: (scenes.cornell_metre(width=64)[0], dict(scenes.cornell_metre(width=64)[1],
                                                                          max_depth=8)),
    "default_d12": lambda: (scenes.default_scene(width=64)[0], dict(scenes.default_scene(width=64)[1], max_depth=12)),
    "mesh_sky_d16": lambda: (scenes.mesh_sky_scene(frequency=8, width=48)[0],
                             dict(scenes.mesh_sky_scene(frequency=8, width=48)[1], max_depth=16)),
    "texture_d6": lambda: (scenes.texture_scene(width=64)[0], dict(scenes.texture_scene(width=64)[1], max_depth=6)),
}


@pytest.mark.parametrize("name", sorted(PAIR_SCENES))
def test_pair_queues_match_per_sample_connections(name, monkeypatch):
    """The general connections (s, t >= 2) evaluated through per-pair queues give the same film as the
    per-sample loop over the pairs (PRT_BDPT_PAIR_QUEUES=0): the same contributions, MIS weights and visibility
    rays, added in the same order. The two kernels are compiled separately and may contract multiply-adds
    differently, so a weighted contribution can differ in its last bits: pixels agree to 2e-6 relative (a
    dropped, doubled or misweighted pair moves a pixel by far more). Splats do not pass through the pair
    kernels and stay identical. Scenes with specular vertices (metal, glass), a mesh BVH with an environment
    map, textures, and depths 6 to 16."""
    s, cfg = PAIR_SCENES[name]()
    scene = P.Scene(s)
    films = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PRT_BDPT_PAIR_QUEUES", flag)
        pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 8, seed=3))
        pr.render_samples(8)
        films.append((pr.read()[0], pr.splats()))
        del pr
    (a, sa), (b, sb) = films
    print(f"{name}: mean {a.mean():.6f} / {b.mean():.6f}")
    assert np.array_equal(sa, sb)
    d = np.abs(a.astype(np.float64) - b)
    rel = d / np.maximum(np.abs(b.astype(np.float64)), 1e-6)
    print(f"{name}: max relative difference {rel.max():.3e}, pixels identical {np.mean(np.all(a == b, axis=2)):.4f}")
    assert rel.max() <= 2e-6
