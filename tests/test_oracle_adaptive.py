"""Pins the oracle's adaptive sampling (adaptiveSamplePixelWithSplats / shouldStopSampling,
pkg/renderer/tile_renderer.go:50-101) and its pass loop (RenderPass, progressive.go:126-223).

* constant radiance stops at exactly max(1, int(target * AdaptiveMinSamples)) samples
  (TestTileRendererAdaptiveSampling, tile_renderer_test.go:134-173, with a constant-colour integrator);
* AdaptiveMinSamples = 1.0 reduces to exact samples per pixel (identical sums to orc_render);
* an independent Python replay of the reference loop (sample index = the pixel's count, stop test before
  every sample, targets from getSamplesForPass, state carried across passes) over single-sample
  radiances from orc_render_samples gives the same counts and sums.
CPU only."""
import math

import numpy as np

import oracle as O
import prtgpu as P
import scenes


def params(desc, cfg, spp, passes, amin, thr, seed=0, integrator=P.INTEGRATOR_PATH):
    w, h = desc.image_size
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=passes, seed=seed,
                             adaptive_min_samples=amin, adaptive_threshold=thr, integrator=integrator)


def test_constant_radiance_stops_at_min_samples():
    s, cfg = scenes.empty_uniform_scene(width=8)
    os_ = O.OracleScene(s)
    film = os_.render_passes(params(s, cfg, 100, 1, 0.1, 0.001), 1, 1)
    assert np.all(film["count"] == 10)
    assert np.all(film["sum"] == 10.0)


def test_black_pixels_stop_at_min_samples():
    """No lights: every sample is black, mean <= 1e-8 and variance 0 < 1e-6 -> stop at the minimum."""
    s = P.SceneDesc()
    s.sphere((0, 0, -1), 0.5, s.lambertian((0.5, 0.5, 0.5)))
    s.set_camera(center=(0, 0, 0), look_at=(0, 0, -1), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=45.0)
    os_ = O.OracleScene(s)
    film = os_.render_passes(params(s, dict(max_depth=4, rr_min_bounces=4), 64, 1, 0.15, 0.01), 1, 1)
    assert np.all(film["count"] == 9) and np.all(film["sum"] == 0.0)


def test_exact_spp_when_min_samples_is_one():
    s, cfg = scenes.cornell_scene(width=24)
    os_ = O.OracleScene(s)
    p = params(s, cfg, 17, 3, 1.0, 0.01, seed=4)
    film = os_.render_passes(p, 1, 3)
    assert np.all(film["count"] == 17)
    ref, _, _ = os_.render(p, 17, 0)
    assert np.allclose(film["sum"], ref, rtol=1e-12, atol=1e-12)


def _replay(os_, desc, p, passes):
    """The reference loop over per-sample radiances (Python, float64)."""
    W, H = desc.image_size
    n = W * H
    cnt = np.zeros(n, np.int64)
    s_rgb = np.zeros((n, 3))
    ls = np.zeros(n)
    lq = np.zeros(n)
    pix = np.stack(np.meshgrid(np.arange(W), np.arange(H)), -1).reshape(-1, 2)
    cache = {}

    def sample(k, idx):
        key = (k, idx)
        if key not in cache:
            # fetch a block of sample indices for this pixel at once
            idxs = np.arange(idx, idx + 16)
            vals = os_.render_samples(p, np.repeat(pix[k:k + 1], len(idxs), 0), idxs)
            for i2, v in zip(idxs, vals):
                cache[(k, int(i2))] = v
        return cache[key]

    for ps in range(1, passes + 1):
        target = P.get_samples_for_pass(p, ps)
        min_s = max(1, int(target * p.adaptive_min_samples))
        for k in range(n):
            while cnt[k] < target:
                if cnt[k] >= min_s:
                    mean, msq = ls[k] / cnt[k], lq[k] / cnt[k]
                    var = max(0.0, msq - mean * mean)
                    stop = var < 1e-6 if mean <= 1e-8 else math.sqrt(var) / mean < p.adaptive_threshold
                    if stop:
                        break
                c = sample(k, int(cnt[k]))
                s_rgb[k] += c
                l = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
                ls[k] += l
                lq[k] += l * l
                cnt[k] += 1
    return cnt.reshape(H, W), s_rgb.reshape(H, W, 3)


def test_matches_python_replay_of_the_reference_loop():
    s, cfg = scenes.cornell_scene(width=12)
    cfg = dict(cfg, max_depth=4)
    os_ = O.OracleScene(s)
    p = params(s, cfg, 40, 4, 0.2, 0.5, seed=11)
    film = os_.render_passes(p, 1, 4)
    cnt, srgb = _replay(os_, s, p, 4)
    print("counts", np.bincount(cnt.ravel()))
    assert len(np.unique(cnt)) > 2  # the stop actually varies between pixels
    assert np.array_equal(film["count"], cnt)
    assert np.allclose(film["sum"], srgb, rtol=1e-12, atol=1e-12)
