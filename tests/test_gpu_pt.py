"""Path-tracing parity of the HIP wavefront integrator against the float64 oracle restatement of
PathTracingIntegrator (pkg/integrator/path_tracing.go) and the film of pkg/renderer.

Both sides draw the same counter-RNG values per (pixel, sample, bounce, dimension), so single
samples are compared one to one: float32 vs float64 paths agree except where a float32
decision (Russian roulette threshold, Fresnel choice, edge hit) flips a branch.
Bar: >= 99 % of per-sample radiances within 1e-3 relative (per channel, floor 1e-3 absolute);
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
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=max_passes, seed=seed)


def gpu_first_sample(desc, cfg, seed=0):
    scene = P.Scene(desc)
    pr = P.ProgressiveRaytracer(scene, params_for(desc, cfg, 1, seed))
    pr.render_samples(1)
    mean, ls, lq, cnt = pr.read()
    assert np.all(cnt == 1)
    return mean


def oracle_samples(desc, cfg, sample=0, seed=0):
    w, h = desc.image_size
    os_ = O.OracleScene(desc)
    ys, xs = np.mgrid[0:h, 0:w]
    pix = np.stack([xs.ravel(), ys.ravel()], 1)
    out = os_.render_samples(params_for(desc, cfg, 1, seed), pix, np.full(len(pix), sample))
    return out.reshape(h, w, 3)


def agree_fraction(g, r, rel=1e-3):
    d = np.abs(g.astype(np.float64) - r)
    tol = rel * np.maximum(np.abs(r), 1.0)
    return np.all(d <= tol, axis=2).mean()


def test_uniform_light_every_pixel_exact():
    s, cfg = scenes.empty_uniform_scene(width=32)
    scene = P.Scene(s)
    pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4))
    pr.render_samples(4)
    mean, ls, lq, cnt = pr.read()
    assert np.all(mean == 1.0) and np.all(cnt == 4)
    assert np.all(pr.image()[..., :3] == 255)


def test_white_furnace():
    s, cfg = scenes.furnace_scene(albedo=0.5, width=32)
    m = gpu_first_sample(s, cfg)
    on = np.abs(m - 0.5).max(axis=2) < 2e-5
    sky = m.max(axis=2) == 1.0
    assert np.all(on | sky) and on.sum() > 50 and sky.sum() > 50


@pytest.mark.parametrize("name", ["default", "cornell_metre", "cornell"])
def test_per_sample_parity(name):
    if name == "default":
        s, cfg = scenes.default_scene(width=160)
    elif name == "cornell_metre":
        s, cfg = scenes.cornell_metre(width=96)
    else:
        s, cfg = scenes.cornell_scene(width=96)
    g = gpu_first_sample(s, cfg, seed=3)
    r = oracle_samples(s, cfg, seed=3)
    frac = agree_fraction(g, r)
    print(f"{name}: per-sample agreement {frac:.5f}; mean gpu {g.mean():.5f} oracle {r.mean():.5f}")
    assert frac >= 0.99
    assert abs(g.mean() - r.mean()) <= 0.02 * max(r.mean(), 1e-3) + 1e-3


def test_per_sample_parity_mesh_envmap():
    # same code path as the 1.8M-triangle sky scene, with a small mesh so the oracle is quick
    s, cfg = scenes.mesh_sky_scene(frequency=24, width=64)
    g = gpu_first_sample(s, cfg, seed=1)
    r = oracle_samples(s, cfg, seed=1)
    frac = agree_fraction(g, r)
    print(f"mesh+envmap: per-sample agreement {frac:.5f}")
    assert frac >= 0.99


def test_deterministic_and_pass_schedule():
    s, cfg = scenes.default_scene(width=80)
    scene = P.Scene(s)
    p = params_for(s, cfg, 50, seed=9, max_passes=7)
    pr = P.ProgressiveRaytracer(scene, p)
    targets = []
    for k in range(1, 8):
        img, st = pr.render_pass(k)
        targets.append(st.target_samples)
        assert st.average_samples == st.target_samples
    assert targets == [1, 9, 17, 25, 33, 41, 50]
    a = pr.read()
    pr2 = P.ProgressiveRaytracer(scene, p)
    for k in range(1, 8):
        pr2.render_pass(k)
    b = pr2.read()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)  # bit-reproducible
    # RGBA8 image = uint8(255 * clamp(sqrt(mean))) (progressive.go:482-495)
    mean = a[0].astype(np.float64)
    exp = (255 * np.clip(np.sqrt(mean), 0, 1)).astype(np.uint8)
    assert np.array_equal(pr.image()[..., :3], exp)


def test_statistical_mean_vs_oracle():
    """GPU image at 256 spp vs an independent oracle image at 1024 spp (different seeds): the
    per-pixel luminance difference is within 4 standard errors for >= 98 % of pixels and the
    image means agree within 2 %."""
    s, cfg = scenes.cornell_metre(width=32)
    scene = P.Scene(s)
    pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 256, seed=11))
    pr.render_samples(256)
    mean, ls, lq, cnt = pr.read()
    n = cnt.astype(np.float64)
    lum_g = ls / n
    var_g = np.maximum(lq / n - lum_g ** 2, 0) / n
    os_ = O.OracleScene(s)
    sm, ls_r, lq_r = os_.render(params_for(s, cfg, 1024, seed=12), 1024)
    lum_r = ls_r / 1024
    var_r = np.maximum(lq_r / 1024 - lum_r ** 2, 0) / 1024
    z = np.abs(lum_g - lum_r) / np.sqrt(var_g + var_r + 1e-12)
    print(f"|z|>4 fraction {(z > 4).mean():.4f}; mean lum gpu {lum_g.mean():.5f} oracle {lum_r.mean():.5f}")
    assert (z > 4).mean() <= 0.02
    assert abs(lum_g.mean() - lum_r.mean()) <= 0.02 * lum_r.mean()


def test_mesh_options_rotation_and_custom_normals():
    """Mesh rotation about a centre and per-triangle custom normals (TriangleMeshOptions) render the same
    samples on the device as in the oracle; an emissive triangle with a flipped custom normal is dark."""
    import math

    v, f = P.icosphere(6, radius=0.8, center=(0.2, 0.9, 0.0), amplitude=0.05, seed=1)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n[::3] *= -1.0  # every third triangle's normal flipped
    s = P.SceneDesc()
    glass = s.dielectric(1.5)
    s.mesh(v, f, glass, normals=n, rotation=(0.3, -0.9, 0.2), center=(0.2, 0.9, 0.0))
    ground = s.lambertian((0.5, 0.5, 0.5))
    s.quad((-5, 0, -5), (10, 0, 0), (0, 0, 10), ground)
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    s.set_camera(center=(0, 1.2, 3.5), look_at=(0, 0.8, 0), up=(0, 1, 0), width=64, aspect_ratio=1.0, vfov=40.0)
    cfg = dict(max_depth=8, rr_min_bounces=4)
    g = gpu_first_sample(s, cfg, seed=2)
    r = oracle_samples(s, cfg, seed=2)
    frac = agree_fraction(g, r)
    print(f"mesh options: per-sample agreement {frac:.5f}")
    assert frac >= 0.99
    for normals, expect in ((None, 2.0), ([[0, 0, -1.0]], 0.0)):
        e = P.SceneDesc()
        em = e.emissive((2.0, 2.0, 2.0))
        e.mesh(np.array([[-3, -3, 0], [3, -3, 0], [0, 3, 0]], float), [[0, 1, 2]], em, normals=normals)
        e.set_camera(center=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=20.0)
        m = gpu_first_sample(e, dict(max_depth=4, rr_min_bounces=4))
        assert np.allclose(m[4, 4], expect)


@pytest.mark.parametrize("integrator", [P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT])
def test_rgba8_image_is_the_gamma_of_the_mean(integrator):
    """prt_film_read_rgba8 (converted on the device) equals uint8(255 * min(1, sqrt(mean))) computed in
    float64 from the film's mean (BDPT: with its splats), for the whole image and a rectangle."""
    s, cfg = scenes.cornell_scene(width=72)
    w, h = s.image_size
    p = P.sampling_params(w, h, 5, 16, 8, seed=4, max_passes=3, integrator=integrator)
    pr = P.ProgressiveRaytracer(P.Scene(s), p)
    pr.render_pass(2)
    mean = pr.read()[0].astype(np.float64)
    exp = (255 * np.where(mean > 0, np.minimum(1.0, np.sqrt(np.maximum(mean, 0))), 0)).astype(np.uint8)
    img = pr.image()
    assert np.array_equal(img[..., :3], exp) and np.all(img[..., 3] == 255)
    sub = pr.image(5, 7, 41, 30)
    assert sub.shape == (23, 36, 4) and np.array_equal(sub[..., :3], exp[7:30, 5:41])
    assert pr.image(3, 3, 3, 9).shape == (6, 0, 4)


def test_probe_counts_and_times():
    """prt_film_probe: the counting trace kernel leaves films bit-identical; at depth 1 every camera sample
    traces exactly one extension ray; each node visit is four box tests; per-kernel times are positive;
    a probe call clears what it returns."""
    s, cfg = scenes.default_scene(width=96)
    scene = P.Scene(s)
    w, h = s.image_size
    plain = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4, seed=2))
    plain.render_samples(4)
    probed = P.ProgressiveRaytracer(scene, params_for(s, cfg, 4, seed=2))
    probed.probe(P.PROBE_COUNT | P.PROBE_TIME)
    probed.render_samples(4)
    st = probed.probe(0)
    assert np.array_equal(plain.read()[0], probed.read()[0])
    assert st["extension_rays"] >= w * h * 4 and st["shadow_rays"] > 0
    assert st["extension_box_tests"] == 4 * st["extension_node_visits"] > 0
    assert st["shadow_box_tests"] == 4 * st["shadow_node_visits"] > 0
    assert st["trace_launches"] == cfg["max_depth"] + 1
    assert st["trace_ms"] > 0 and st["shade_ms"] > 0 and st["camera_ms"] > 0 and st["film_ms"] > 0
    assert probed.probe(0)["extension_rays"] == 0
    one = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, 1, cfg["rr_min_bounces"], 3, seed=2))
    one.probe(P.PROBE_COUNT)
    one.render_samples(3)
    assert one.probe(0)["extension_rays"] == w * h * 3


def test_film_independent_of_chunking():
    """8 spp rendered in one call, as 4 + 4, and as eight single samples give bit-identical films (every
    sample enters the compensated sums on its own, in sample order)."""
    s, cfg = scenes.cornell_metre(width=64)
    scene = P.Scene(s)
    films = []
    for split in ([8], [4, 4], [1] * 8):
        pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 8, seed=4))
        for k in split:
            pr.render_samples(k)
        films.append(pr.read())
        pr.close()
    for other in films[1:]:
        for x, y in zip(films[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("integrator", ["pt", "bdpt"])
def test_leaf_gate_leaves_films_unchanged(integrator, monkeypatch):
    """The leaf gate of the render kernels' traversals (pinned per scene with PRT_TRACE_LEAF_GATE, else
    picked by the autotuner) only reschedules lanes: films, splats included, are bit-identical for every
    gate, and so is the flat traversal (PRT_TRACE_LEAF_GATE=9: every primitive tested without the BVH,
    scenes of <= 64 primitives). The autotuner times launches inside a chunk of >= 2^20 (path tracing) or
    2^18 (BDPT) paths that is not a one-sample first chunk, and a later chunk reads the timings: rendering
    in three calls tunes the scene during the film, and a second film then renders with the tuned gate."""
    s, cfg = scenes.cornell_metre(width=128)
    w, h = s.image_size
    integ = P.INTEGRATOR_BDPT if integrator == "bdpt" else P.INTEGRATOR_PATH
    per_call = 16 if integrator == "bdpt" else 64  # 128^2 x per_call = 2^18 / 2^20 paths per chunk

    def render(scene):
        pr = P.ProgressiveRaytracer(scene, P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"],
                                                             3 * per_call, seed=3, integrator=integ))
        for _ in range(3):
            pr.render_samples(per_call)
        out = pr.read() + (pr.splats(),)
        pr.close()
        return out

    def tuned_gate(scene):
        i = scene.info()
        return i.bdpt_trace_leaf_gate if integrator == "bdpt" else i.pt_trace_leaf_gate

    films = []
    for g in ("0", "4", "8", "9"):  # 9: flat traversal (the metre Cornell box has 18 primitives)
        monkeypatch.setenv("PRT_TRACE_LEAF_GATE", g)
        scene = P.Scene(s)
        assert scene.info().pt_trace_leaf_gate == int(g) and scene.info().bdpt_trace_leaf_gate == int(g)
        films.append(render(scene))
        assert tuned_gate(scene) == int(g)  # a pinned gate is kept
    monkeypatch.delenv("PRT_TRACE_LEAF_GATE")
    scene = P.Scene(s)
    assert tuned_gate(scene) == -1
    films.append(render(scene))  # the second call starts the tuning, the third reads it
    assert tuned_gate(scene) in (0, 4, 8, 9)
    if integrator == "pt":
        assert scene.info().pt_trace_waves > 0
    films.append(render(scene))  # renders with the tuned gate from its first chunk
    for other in films[1:]:
        for x, y in zip(films[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("name", ["default", "cornell_metre", "mesh_under_sky"])
def test_tail_kernel_leaves_films_unchanged(name, monkeypatch):
    """From the first bounce at which a chunk is expected to trace few paths (PRT_PT_TAIL_PATHS, default
    2^16, read from the survival profile that earlier chunks of the scene record on the device), one launch
    (k_tail) runs every remaining path to its end: shading, its shadow ray and its next ray bounce after
    bounce. A path's radiance terms keep the wavefront order, so films are bit-identical with the tail
    switched off, forced from bounce 1, and from an intermediate bounce. The first film of each scene
    records the profile; the second uses it."""
    if name == "default":
        s, cfg = scenes.default_scene(width=96)
    elif name == "cornell_metre":
        s, cfg = scenes.cornell_metre(width=64)
    else:  # triangles only, under a uniform sky
        s, _ = scenes.mesh_scene(frequency=24, width=64)
        s.add_uniform_infinite_light((0.8, 0.9, 1.0))
        cfg = dict(max_depth=12, rr_min_bounces=3)
    films = []
    for limit in ("0", str(1 << 40), "2000"):
        monkeypatch.setenv("PRT_PT_TAIL_PATHS", limit)
        scene = P.Scene(s)
        for _ in range(2):
            pr = P.ProgressiveRaytracer(scene, params_for(s, cfg, 8, seed=6))
            pr.render_samples(8)
            films.append(pr.read())
            pr.close()
    assert films[0][0].mean() > 0
    for other in films[1:]:
        for x, y in zip(films[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("integrator", [P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT])
def test_films_of_one_scene_one_after_another(integrator):
    """A scene's leaf-gate tuning jobs and survival-profile copies run on the stream of the film that
    started them; a film that is destroyed takes its stream with it. Films of one scene created, rendered
    and destroyed one after another (each large enough to start a tuning job) all render without error,
    and the same film renders the same image before and after the scene's tuning settled."""
    s, cfg = scenes.cornell_metre(width=512)
    scene = P.Scene(s)
    p = P.sampling_params(1024, 1024, cfg["max_depth"], cfg["rr_min_bounces"], 2, max_passes=1, seed=3,
                          integrator=integrator)
    imgs = []
    for _ in range(5):
        pr = P.ProgressiveRaytracer(scene, p)
        pr.render_samples(2)
        imgs.append(pr.read()[0])
        pr.close()
    for img in imgs[1:]:
        assert np.array_equal(img, imgs[0])
