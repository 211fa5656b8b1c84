"""Scenes read by the PBRT loader render on the HIP path as in the oracle: the reference's Cornell PBRT
files and a scene with a PBRT point light (a sphere light that is not a shape), PT and BDPT.
Bars as for the other scenes (per-sample agreement, splat totals)."""
import os

import numpy as np
import pytest

import oracle as O
import prtgpu as P

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

POINT_SCENE = '''LookAt 0 1 6  0 0.5 0  0 1 0
Camera "perspective" "float fov" 35
Film "rgb" "integer xresolution" 96 "integer yresolution" 72
WorldBegin
Material "diffuse" "rgb reflectance" [0.6 0.6 0.6]
Shape "bilinearPatch" "point3 P00" [-4 0 -4] "point3 P01" [-4 0 4] "point3 P10" [4 0 -4] "point3 P11" [4 0 4]
AttributeBegin
  Material "dielectric" "float eta" 1.5
  Shape "sphere" "float radius" 0.6 "point3 center" [-0.8 0.6 0]
AttributeEnd
AttributeBegin
  Material "conductor" "rgb eta" [0.9 0.7 0.4] "float roughness" 0.2
  Shape "box" "point3 center" [0.9 0.5 0] "point3 size" [0.4 0.5 0.4] "point3 rotation" [0 0.5 0]
AttributeEnd
LightSource "point" "rgb I" [6 6 6] "point3 from" [0 2.5 1]
LightSource "infinite-gradient" "rgb topColor" [0.2 0.3 0.5] "rgb bottomColor" [0.1 0.1 0.1]
WorldEnd'''


def _scene(name):
    if name == "point_light":
        return P.parse_pbrt(POINT_SCENE)
    s = P.load_pbrt(os.path.join(HERE, "golden", "pbrt", "scenes", name + ".pbrt"))
    s.camera.width = 64  # a smaller film of the same camera
    return s


def _first_sample(desc, seed, integrator):
    w, h = desc.image_size
    cfg = desc.sampling if hasattr(desc, "sampling") else {}
    p = P.sampling_params(w, h, cfg.get("max_depth", 5), cfg.get("rr_min_bounces", 16), 1, seed=seed,
                          integrator=integrator)
    pr = P.ProgressiveRaytracer(P.Scene(desc), p)
    pr.render_samples(1)
    mean, _, _, cnt = pr.read()
    assert np.all(cnt == 1)
    sp = pr.splats() if integrator == P.INTEGRATOR_BDPT else np.zeros((h, w, 3))
    ys, xs = np.mgrid[0:h, 0:w]
    pix = np.stack([xs.ravel(), ys.ravel()], 1)
    out, rsp = O.OracleScene(desc).render_samples(p, pix, np.zeros(len(pix)), splats=True)
    return mean.astype(np.float64) - sp, out.reshape(h, w, 3), sp, rsp


def _agree(g, r, rel, floor):
    return np.all(np.abs(g - r) <= rel * np.maximum(np.abs(r), floor), axis=2).mean()


@pytest.mark.parametrize("name", ["cornell-boxes", "cornell", "point_light"])
@pytest.mark.parametrize("integrator", [P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT])
def test_pbrt_scene_parity(name, integrator):
    s = _scene(name)
    g, r, gs, rs = _first_sample(s, 3, integrator)
    frac = _agree(g, r, 1e-3, 1.0)
    print(f"{name} integrator {integrator}: agreement {frac:.5f}; mean gpu {g.mean():.6f} oracle {r.mean():.6f};"
          f" splats gpu {gs.sum():.4f} oracle {rs.sum():.4f}")
    assert r.mean() > 0
    assert frac >= (0.99 if integrator == P.INTEGRATOR_PATH else 0.98)
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3
    if integrator == P.INTEGRATOR_BDPT:
        assert abs(gs.sum() - rs.sum()) <= 0.01 * abs(rs.sum()) + 1e-3
        assert _agree(gs, rs, 1e-2, 1e-1) >= 0.97
