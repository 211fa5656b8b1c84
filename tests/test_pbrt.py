"""The PBRT loader (include/prt_loaders.h; loaders.ParsePBRT + scene.NewPBRTScene) against the
reference's own PBRT test inputs and expectations (tests/golden/pbrt_cases.json) and the scene files
its tests load (tests/golden/pbrt/scenes/). Host code only, plus the oracle for the converted scene."""
import json
import os

import numpy as np
import pytest

import oracle as O
import prtgpu as P

HERE = os.path.dirname(os.path.abspath(__file__))
G = json.load(open(os.path.join(HERE, "golden", "pbrt_cases.json")))
LIGHT = {"quad": P.LIGHT_QUAD, "sphere": P.LIGHT_SPHERE, "uniform": P.LIGHT_UNIFORM_INFINITE,
         "gradient": P.LIGHT_GRADIENT_INFINITE, "gradient_top": P.LIGHT_GRADIENT_INFINITE}


def arr(ptr, n):
    return [ptr[i] for i in range(n)]


def check(scene, e):
    d = scene.desc
    mats = arr(d.materials, d.num_materials)
    spheres = arr(d.spheres, d.num_spheres)
    lights = arr(d.lights, d.num_lights)
    for k in ("width", "height"):
        if k in e:
            assert scene.sampling[k] == e[k]
    if "vfov" in e:
        assert scene.camera.vfov == e["vfov"]
    if "center" in e:
        assert list(scene.camera.center) == e["center"]
    if "spp" in e:
        assert scene.sampling["spp"] == e["spp"]
    if "shapes" in e:
        assert scene.reference_shapes == e["shapes"]
    if "quads" in e:
        assert d.num_quads == e["quads"]
    if "spheres" in e:
        assert [(s.radius, list(mats[s.material].albedo)) for s in spheres] == [(r, a) for r, a in e["spheres"]]
        assert all(mats[s.material].type == P.MAT_LAMBERTIAN for s in spheres)
    if "metal_spheres" in e:
        got = [(s.radius, list(mats[s.material].albedo), mats[s.material].fuzz) for s in spheres
               if mats[s.material].type == P.MAT_METAL]
        assert got == [(r, a, f) for r, a, f in e["metal_spheres"]]
    if "dielectric_spheres" in e:
        assert [(s.radius, mats[s.material].ior) for s in spheres] == [(r, i) for r, i in e["dielectric_spheres"]]
    if "emissive_spheres" in e:
        assert [(s.radius, list(mats[s.material].emission)) for s in spheres] == [(r, em) for r, em in e["emissive_spheres"]]
        assert all(mats[s.material].type == P.MAT_EMISSIVE for s in spheres)
    if "sphere_centers" in e:
        assert [list(s.center) for s in spheres] == e["sphere_centers"]
    if "lights" in e:
        assert [L.type for L in lights] == [LIGHT[t] for t, _ in e["lights"]]
        for L, (t, v) in zip(lights, e["lights"]):
            if v is None:
                continue
            if t == "quad":  # the emitting quad's material carries L
                assert list(mats[d.quads[L.shape].material].emission) == v
            elif t == "gradient_top":
                assert list(L.top) == v and list(L.bottom) == [1, 1, 1]
            else:
                assert list(L.emission) == v
    if "point_lights" in e:
        pts = [(list(L.from_), list(L.emission), L.radius) for L in lights if L.type == P.LIGHT_SPHERE and L.shape == -1]
        assert pts == [(f, i, 0.1) for f, i in e["point_lights"]]
    if "boxes" in e:
        assert [(list(b.center), list(b.size), list(b.rotation)) for b in arr(d.boxes, d.num_boxes)] == \
            [tuple(x) for x in e["boxes"]]
    if "metal_default" in e:
        m = [m for m in mats if m.type == P.MAT_METAL][0]
        assert list(m.albedo) == e["metal_default"][0] and m.fuzz == e["metal_default"][1]
    if "meshes" in e:
        assert [(m.num_vertices, m.num_faces) for m in arr(d.meshes, d.num_meshes)] == [tuple(x) for x in e["meshes"]]


@pytest.mark.parametrize("case", G["cases"], ids=lambda c: c["name"])
def test_reference_pbrt_cases(case):
    e = case["expect"]
    if "error" in e:
        with pytest.raises(P.PrtError) as ei:
            P.parse_pbrt(case["text"])
        assert e["error"] in str(ei.value)
    else:
        check(P.parse_pbrt(case["text"]), e)


@pytest.mark.parametrize("f", G["files"], ids=lambda f: f["file"])
def test_reference_scene_files(f):
    s = P.load_pbrt(os.path.join(HERE, "golden", "pbrt", "scenes", f["file"]))
    check(s, {k: ([[t, None] for t in v] if k == "lights" else v) for k, v in f.items()
              if k not in ("file", "source", "boxes")})
    if "boxes" in f:
        assert s.desc.num_boxes == f["boxes"]


def test_load_path_rules():
    """validateFilePath (pbrt.go:440-480): a .pbrt file under scenes/ or the temporary directory."""
    for bad in ("", "model.pbrt", "scenes/model.txt", "scenes/" + "x" * 600 + ".pbrt"):
        with pytest.raises(P.PrtError):
            P.load_pbrt(bad)
    with pytest.raises(P.PrtError):  # allowed path, missing file
        P.load_pbrt("scenes/does-not-exist.pbrt")


def test_point_light_is_a_sphere_light_without_a_shape():
    """A PBRT point light becomes NewSphereLight(from, 0.1, emissive(I)) that is not added to the
    shapes: camera rays never hit it, next-event estimation samples it."""
    s = P.parse_pbrt('''LookAt 0 0 5  0 0 0  0 1 0
Camera "perspective" "float fov" 30
Film "rgb" "integer xresolution" 24 "integer yresolution" 24
WorldBegin
Material "diffuse" "rgb reflectance" [0.5 0.5 0.5]
Shape "bilinearPatch" "point3 P00" [-3 -1 -3] "point3 P01" [-3 -1 3] "point3 P10" [3 -1 -3] "point3 P11" [3 -1 3]
LightSource "point" "rgb I" [40 40 40] "point3 from" [0 0 0]
WorldEnd''')
    os_ = O.OracleScene(s)
    _, _, n = os_.world()
    assert n == 1  # only the floor is geometry
    sm, _, _ = os_.render(P.sampling_params(24, 24, 3, 16, 1), 1)
    assert sm.max() > 0  # the floor is lit through next-event estimation
    centre = sm[11:13, 11:13]  # looking straight at the light's position: no emitter is hit there
    assert np.all(centre < 40)
