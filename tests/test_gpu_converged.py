"""Image-level parity against converged references: the GPU image of each benchmark configuration, at
the configuration's sample count, per pixel within 3 sigma of the 65,536-spp float64 oracle image
(tests/golden/make_converged.py; the method of pkg/renderer/progressive_integration_test.go:29-552, with
PixelStats' per-pixel moments, pkg/renderer/stats.go:20-55).

Per pixel and channel, z = (gpu - fixture mean) / sqrt(var (1/n + 1/N) + (1e-4 mean)^2 + 1e-12), with var
the fixture's per-sample variance, n the GPU's and N = 65,536 the fixture's sample count (independent
seeds). Bars:
* the fraction of |z| > 3 is at most 1.25 x the larger of the normal tail (0.27 %) and the fraction an
  independent oracle render at n samples gives (heavy-tailed pixels such as caustics exceed the normal
  tail at low sample counts; measured per configuration when the fixture was made), plus 4 binomial
  standard errors;
* the mean of z clipped to +-10 over the image is within 0.02 plus 3 standard deviations of the calibration
  renders' values of their average (a dropped or misweighted estimator term shifts it by many times that);
* the spread of the clipped z is at most 1.2 x the calibration's plus 0.05.
z is clipped for these two because a pixel that catches a rare bright path (a sun caustic through the glass
mesh) at a few hundred samples has a z in the hundreds, which alone would set the unclipped mean and spread.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["c0_default_400x225", "c2_cornell_metre_128", "c3_mesh_sky_crop128", "c2_bdpt_cornell_metre_128"]


def _fixture(name):
    try:
        man = json.load(open(os.path.join(GOLD, "converged.json")))[name]
    except (OSError, KeyError):
        pytest.skip(f"converged fixture {name} not generated")
    path = os.path.join(GOLD, "converged", name + ".npz")
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == man["sha256"], "fixture does not match its pin"
    z = np.load(path)
    return man, z["mean"].astype(np.float64), z["var"].astype(np.float64)


@pytest.mark.parametrize("name", NAMES)
def test_gpu_image_within_3_sigma_of_converged(name):
    man, mean, var = _fixture(name)
    desc, cfg = getattr(scenes, man["scene"])(**man["scene_args"])
    w, h = desc.image_size
    assert [w, h] == man["image_size"]
    n = man["config_spp"]
    integ = P.INTEGRATOR_BDPT if man["integrator"] == "bdpt" else P.INTEGRATOR_PATH
    p = P.sampling_params(w, h, man["max_depth"], man["rr_min_bounces"], n, seed=1, integrator=integ)
    pr = P.ProgressiveRaytracer(P.Scene(desc), p)
    pr.render_samples(n)
    img, _, _, cnt = pr.read()
    assert np.all(cnt == n)
    x0, y0, x1, y1 = man["rect"]
    g = img[y0:y1, x0:x1].astype(np.float64)
    sd = np.sqrt(var * (1.0 / n + 1.0 / man["fixture_spp"]) + (1e-4 * np.abs(mean)) ** 2 + 1e-12)
    z = (g - mean) / sd
    frac = float((np.abs(z) > 3).mean())
    cal = man["calibration"]
    p0 = max(0.0027, max(c["frac_abs_z_gt_3"] for c in cal))
    allow = 1.25 * p0 + 4 * math.sqrt(p0 * (1 - p0) / z.size)
    zc = np.clip(z, -10.0, 10.0)
    std_allow = 1.2 * max(c["std_zc"] for c in cal) + 0.05
    mz = [c["mean_zc"] for c in cal]
    mz_mid, mz_sd = float(np.mean(mz)), float(np.std(mz, ddof=1))
    print(f"{name}: |z|>3 fraction {frac:.5f} (allowed {allow:.5f}; oracle at {n} spp "
          f"{[round(c['frac_abs_z_gt_3'], 5) for c in cal]}), clipped z mean {zc.mean():+.4f} (allowed "
          f"{mz_mid:+.4f} +- {0.02 + 3 * mz_sd:.4f}), std {zc.std():.3f} (allowed {std_allow:.3f}); unclipped "
          f"z mean {z.mean():+.4f} std {z.std():.3f}; RMSE {np.sqrt(((g - mean) ** 2).mean()):.5f}, "
          f"image mean gpu {g.mean():.5f} fixture {mean.mean():.5f}")
    assert frac <= allow
    assert abs(float(zc.mean()) - mz_mid) <= 0.02 + 3 * mz_sd
    assert float(zc.std()) <= std_allow
