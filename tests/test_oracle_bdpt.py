"""Pins the oracle's BDPT (restatement of pkg/integrator/bdpt.go, bdpt_mis.go).

* MIS weights of the reference's hand-built paths, expected values from bdpt_mis_test.go:144-430
  (fixtures: createTestCameraPath / createTestLightPath / createSampledLightVertex /
  createSimpleTestScene / createSceneWithLightsAndWeights, bdpt_test.go:280-377,724-868,
  bdpt_mis_test.go:965-985).
* BDPT and PT estimate the same image (PT vs BDPT luminance, progressive_integration_test.go).
CPU only."""
import math

import numpy as np
import pytest

import oracle as O
import prtgpu as P
import scenes

PI = math.pi


def simple_scene():
    """createSimpleTestScene: white sphere + 1x1 quad light at y=2; camera at the origin."""
    s = P.SceneDesc()
    mats = {"white": s.lambertian((0.7, 0.7, 0.7)), "glass": s.dielectric(1.5), "emissive": s.emissive((5, 5, 5)),
            "metal": s.metal((0.9, 0.9, 0.9), 0.0)}
    s.sphere((0, 0, -1), 0.5, mats["white"])
    s.add_quad_light((-0.5, 2.0, -0.5), (1, 0, 0), (0, 0, 1), (5, 5, 5))
    s.set_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 100, 1.0, 45.0)
    return s, mats


def weighted_scene():
    """createSceneWithLightsAndWeights with two 1x1 quad lights, weights 0.2 / 0.8."""
    s = P.SceneDesc()
    white = s.lambertian((0.7, 0.7, 0.7))
    s.sphere((0, 0, -5), 0.5, white)
    s.add_quad_light((-1.5, 2.0, -1.5), (1, 0, 0), (0, 0, 1), (2, 2, 2))
    s.add_quad_light((-0.5, 2.0, -1.5), (1, 0, 0), (0, 0, 1), (5, 5, 5))
    s.light_weights = [0.2, 0.8]
    s.set_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 100, 1.0, 45.0)
    return s


def light_pool():
    """Light objects carried by test vertices: createTestAreaLight (4x4 at y=2) and the 1x1 'hit' lights."""
    s = P.SceneDesc()
    s.add_quad_light((-2, 2, -3), (4, 0, 0), (0, 0, 4), (5, 5, 5))       # 0: createTestAreaLight
    s.add_quad_light((-0.5, 2.0, -1.5), (1, 0, 0), (0, 0, 1), (4, 4, 4))  # 1: hitLight of PathTracing_s0t3
    s.add_quad_light((-0.5, 2.0, -1.5), (1, 0, 0), (0, 0, 1), (5, 5, 5))  # 2: hitLight of the weighted case
    s.set_camera((0, 0, 0), (0, 0, -1), (0, 1, 0), 100, 1.0, 45.0)
    return s


def vtx(p, n, mat=-1, light_ref=-1, light_index=0, flags=0, fwd=0.0, rev=0.0):
    return (p, n, mat, light_ref, light_index, flags, fwd, rev)


def camera_path(kinds, positions, mats, hit_light=None, hit_index=0):
    """createTestCameraPathWithLight (bdpt_test.go:728-785)."""
    vs = [vtx(positions[0], (0, 0, 1), flags=O.V_CAMERA, fwd=1.0)]
    for i, k in enumerate(kinds, start=1):
        flags = 0
        if k in ("glass", "metal"):
            flags |= O.V_SPECULAR
        lref = -1
        if k == "emissive":
            flags |= O.V_LIGHT
            if i == len(positions) - 1 and hit_light is not None:
                lref = hit_light
        vs.append(vtx(positions[i], (0, 1, 0), mats[k], lref, hit_index if lref >= 0 else 0, flags, 1 / PI, 1 / PI))
    return np.array(vs, dtype=O.ORC_VERTEX)


def light_path(kinds, positions, mats):
    """createTestLightPath (bdpt_test.go:788-838): vertex 0 on createTestAreaLight, pdfPos from PDF_Le."""
    p0 = positions[0]
    on_light = abs(p0[1] - 2.0) < 1e-12 and -2 <= p0[0] <= 2 and -3 <= p0[2] <= 1
    vs = [vtx(p0, (0, -1, 0), mats[kinds[0]], 0, 0, O.V_LIGHT, (1 / 16.0) if on_light else 0.0, 0.0)]
    for i in range(1, len(positions)):
        k = kinds[i]
        flags = O.V_SPECULAR if k in ("glass", "metal") else 0
        vs.append(vtx(positions[i], (0, 1, 0), mats[k], -1, 0, flags, 1 / PI, 1 / PI))
    return np.array(vs, dtype=O.ORC_VERTEX)


def sampled_light(mats):
    """createSampledLightVertex (bdpt_mis_test.go:965-985)."""
    return np.array([vtx((0, 2, -1), (0, -1, 0), mats["emissive"], 0, 0, O.V_LIGHT, 1.0, 0.0)], dtype=O.ORC_VERTEX)


CASES = [
    # name, camera kinds, camera positions, light kinds, light positions, sampled?, s, t, expected
    ("TrivialCase_s0t2", ["white"], [(0, 0, 0), (0, 0, -1)], [], [], False, 0, 2, 1.0),
    ("DirectLighting_s1t2", ["white"], [(0, 0, 0), (0, 0, -1)], ["emissive"], [(0, 2, -1)], True, 1, 2, 0.752119),
    ("DirectLighting_s1t3", ["white", "white"], [(0, 0, 0), (0, 0, -1), (-1, 0, -1)], ["emissive"], [(-1, 2, -1)],
     True, 1, 3, 0.583572),
    ("IndirectLighting_s2t2", ["white"], [(0, 0, 0), (0, 0, -1)], ["emissive", "white"], [(0, 2, -1), (0, 1, -1)],
     False, 2, 2, 0.097708),
    ("ComplexPath_s2t3", ["white", "white"], [(0, 0, 0), (0, 0, -1), (1, 0, -1)], ["emissive", "white"],
     [(1, 3, -1), (1, 2, -1)], False, 2, 3, 0.069731),
    ("PathTracing_s0t4", ["white", "white", "white"], [(0, 0, 0), (0, 0, -1), (1, 0, -1), (2, 0, -1)], [], [],
     False, 0, 4, 0.041875),
    ("MultiBounceGlass_s1t3", ["glass", "white"], [(0, 0, 0), (0, 0, -1), (0, 0, -2)], ["emissive"], [(0, 2, -2)],
     True, 1, 3, 0.951539),
    ("SpecularMirrorPath_s1t3", ["metal", "white"], [(0, 0, 0), (0, 0, -1), (1, 0, -1)], ["emissive"],
     [(1, 2, -1)], True, 1, 3, 0.951539),
    ("DirectSpecularLighting_s1t2", ["metal"], [(0, 0, 0), (0, 0, -1)], ["emissive"], [(0, 2, -1)], True, 1, 2,
     0.444444),
    ("ComplexSpecularPath_s2t2", ["metal"], [(0, 0, 0), (0, 0, -1)], ["emissive", "metal"],
     [(0, 3, -1), (0, 2, -1)], False, 2, 2, 0.095925),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_mis_golden(case):
    name, ck, cpos, lk, lpos, samp, s, t, expected = case
    desc, mats = simple_scene()
    sc = O.OracleScene(desc)
    pool = O.OracleScene(light_pool())
    cp = camera_path(ck, cpos, mats)
    lp = light_path(lk, lpos, mats) if lk else np.zeros(0, O.ORC_VERTEX)
    sv = sampled_light(mats) if samp else None
    w = O.bdpt_mis(sc, pool, cp, lp, sv, s, t)
    assert 0 <= w <= 1
    assert abs(w - expected) <= 1e-3, f"{name}: {w} vs {expected}"


def test_mis_golden_hits_light():
    """PathTracing_s0t3_HitsLight: camera path ends on a 1x1 quad light carried by the vertex."""
    desc, mats = simple_scene()
    sc = O.OracleScene(desc)
    pool = O.OracleScene(light_pool())
    cp = camera_path(["white", "emissive"], [(0, 0, 0), (0, 0, -1), (0, 2, -1)], mats, hit_light=1, hit_index=0)
    w = O.bdpt_mis(sc, pool, cp, np.zeros(0, O.ORC_VERTEX), None, 0, 3)
    assert abs(w - 0.202964) <= 1e-3


def test_mis_golden_weighted_sampler():
    """WeightedLightSampler_PathTracing_s0t3: the hit light is index 1 of a 0.2/0.8 weighted sampler."""
    d = weighted_scene()
    sc = O.OracleScene(d)
    pool = O.OracleScene(light_pool())
    mats = {"white": 0, "emissive": 1}
    # materials of the weighted scene: 0 white, 1/2 the two light emissives; the path's emissive vertex
    # only needs a material for IsOnSurface, any emissive index works
    cp = camera_path(["white", "emissive"], [(0, 0, 0), (0, 0, -1), (0, 2, -1)], mats, hit_light=2, hit_index=1)
    w = O.bdpt_mis(sc, pool, cp, np.zeros(0, O.ORC_VERTEX), None, 0, 3)
    assert abs(w - 0.241453) <= 1e-3


def _lum_image(os_, desc, cfg, integrator, spp, seed):
    w, h = desc.image_size
    p = P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, seed=seed, integrator=integrator)
    if integrator == P.INTEGRATOR_BDPT:
        s, ls, lq, sp = os_.render(p, spp, splats=True)
        return (s + sp) / spp
    s, ls, lq = os_.render(p, spp)
    return s / spp


def test_bdpt_matches_pt_on_cornell():
    """Average luminance of BDPT and PT agree (progressive_integration_test.go:29-552 uses 5-15 %)."""
    desc, cfg = scenes.cornell_scene(width=32)
    os_ = O.OracleScene(desc)
    pt = _lum_image(os_, desc, cfg, P.INTEGRATOR_PATH, 256, 1)
    bd = _lum_image(os_, desc, cfg, P.INTEGRATOR_BDPT, 64, 2)
    lum = lambda im: (0.2126 * im[..., 0] + 0.7152 * im[..., 1] + 0.0722 * im[..., 2]).mean()
    print(f"PT {lum(pt):.4f} BDPT {lum(bd):.4f}")
    assert abs(lum(bd) - lum(pt)) <= 0.05 * lum(pt)


def test_bdpt_uniform_light_empty_scene():
    """No geometry, uniform infinite light of radiance 1: the s=0 strategy sees the background; light
    paths never reach the camera, so every pixel is exactly 1."""
    desc, cfg = scenes.empty_uniform_scene(width=16)
    os_ = O.OracleScene(desc)
    im = _lum_image(os_, desc, cfg, P.INTEGRATOR_BDPT, 2, 3)
    assert np.allclose(im, 1.0, atol=1e-12)
