"""Scene constructors of the benchmark configurations, mirroring pkg/scene.

Each returns (SceneDesc, dict of sampling settings). Coordinates, materials and lights follow
the reference constructors cited in each docstring.
"""
import math

import numpy as np

import prtgpu as P


def default_scene(pinhole=True, width=400):
    """NewDefaultScene (pkg/scene/default_scene.go:11-95).

    pinhole=True forces aperture 0: MergeCameraConfig only applies non-zero overrides
    (pkg/geometry/camera.go:341), so the reference cannot zero it through an override.
    """
    s = P.SceneDesc()
    s.set_camera(center=(0, 0.75, 2), look_at=(0, 0.5, -1), up=(0, 1, 0), width=width,
                 aspect_ratio=16.0 / 9.0, vfov=40.0, aperture=0.0 if pinhole else 0.05, focus_distance=0.0)
    green = s.lambertian((0.8 * 0.6, 0.8 * 0.6, 0.0))
    blue = s.lambertian((0.1, 0.2, 0.5))
    red = s.lambertian((0.65, 0.25, 0.2))
    silver = s.metal((0.8, 0.8, 0.8), 0.0)
    gold = s.metal((0.8, 0.6, 0.2), 0.3)
    glass = s.dielectric(1.5)
    coated = s.layered(glass, red)
    # AddSphereLight first (default_scene.go:78-82), then the shapes in the reference order
    s.add_sphere_light((30, 30.5, 15), 10, (15.0, 14.0, 13.0))
    s.sphere((0, 0.5, -1), 0.5, coated)
    s.sphere((-1, 0.5, -1), 0.5, silver)
    s.sphere((1, 0.5, -1), 0.5, gold)
    ground_quad(s, (0, 0, 0), 10000.0, green)
    s.sphere((0.5, 0.25, -0.5), 0.25, glass)
    s.sphere((-0.5, 0.25, -0.5), 0.25, glass)
    s.sphere((-0.5, 0.25, -0.5), -0.24, glass)
    s.sphere((-0.5, 0.25, -0.5), 0.20, blue)
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    cfg = dict(max_depth=50, rr_min_bounces=20, spp=64)
    return s, cfg


def ground_quad(s, center, size, material):
    """NewGroundQuad (pkg/scene/scene.go:33-42)."""
    corner = (center[0] - size / 2, center[1], center[2] - size / 2)
    return s.quad(corner, (size, 0, 0), (0, 0, size), material)


def cornell_scene(scale=1.0, width=400, light="reference", tall_angle_deg=-20.0, short_angle_deg=18.0):
    """NewCornellScene(CornellBoxes, CornellQuadLight) (pkg/scene/cornell.go:32-257), uniformly
    scaled by `scale`. light="reference": 130x105 quad, emission (18,15,8)*2.5;
    light="metre": radiance 15 on the same (scaled) quad; light="point": CornellPointLight, a
    180-degree point spot light of intensity (18,15,8)*20000 at the quad's centre; light="sphere":
    CornellSphereLight, a radius-52.5 sphere light of emission (18,15,8)*2.05 (cornell.go:150-190)."""
    k = scale
    s = P.SceneDesc()
    s.set_camera(center=(278 * k, 278 * k, -800 * k), look_at=(278 * k, 278 * k, 0), up=(0, 1, 0), width=width,
                 aspect_ratio=1.0, vfov=40.0, aperture=0.0, focus_distance=0.0)
    white = s.lambertian((0.73, 0.73, 0.73))
    red = s.lambertian((0.65, 0.05, 0.05))
    green = s.lambertian((0.12, 0.45, 0.15))
    s.quad((0, 0, 0), (556 * k, 0, 0), (0, 0, 556 * k), white)          # floor
    s.quad((0, 556 * k, 0), (556 * k, 0, 0), (0, 0, 556 * k), white)    # ceiling
    s.quad((0, 0, 556 * k), (556 * k, 0, 0), (0, 556 * k, 0), white)    # back wall
    s.quad((0, 0, 0), (0, 0, 556 * k), (0, 556 * k, 0), red)            # left wall
    s.quad((556 * k, 0, 0), (0, 0, 556 * k), (0, 556 * k, 0), green)    # right wall
    centre = (278.0 * k, (556 - 0.001) * k, 279.5 * k)
    if light == "point":
        # the reference passes (0,-1,0) as the spot's target point; a 180-degree cone lights every
        # direction, so only the sampled cone frame depends on it
        s.add_point_spot_light(centre, (0, -1, 0), (18.0 * 20000, 15.0 * 20000, 8.0 * 20000), 180.0, 0.0)
    elif light == "sphere":
        s.add_sphere_light(centre, 52.5 * k, (18.0 * 2.05, 15.0 * 2.05, 8.0 * 2.05))
    else:
        emission = (18.0 * 2.5, 15.0 * 2.5, 8.0 * 2.5) if light == "reference" else (15.0, 15.0, 15.0)
        s.add_quad_light((213 * k, (556 - 0.001) * k, 227 * k), (130 * k, 0, 0), (0, 0, 105 * k), emission)
    white_box = s.lambertian((0.73, 0.73, 0.73))
    mirror = s.metal((0.9, 0.9, 0.9), 0.0)
    s.box((370 * k, 82.5 * k, 169 * k), (82.5 * k, 82.5 * k, 82.5 * k), (0, short_angle_deg * math.pi / 180, 0),
          white_box)
    s.box((185 * k, 165 * k, 351 * k), (82.5 * k, 165 * k, 82.5 * k), (0, tall_angle_deg * math.pi / 180, 0), mirror)
    cfg = dict(max_depth=5, rr_min_bounces=16, spp=150)
    return s, cfg


def cornell_metre(width=512):
    """Metre-scale Cornell box with two boxes rotated +-18 degrees, radiance-15 quad light,
    512x512, PT depth 8, 1024 spp (x0.001 of cornell.go's geometry)."""
    s, _ = cornell_scene(scale=0.001, width=width, light="metre", tall_angle_deg=-18.0, short_angle_deg=18.0)
    return s, dict(max_depth=8, rr_min_bounces=16, spp=1024)


def mesh_scene(frequency=256, seed=0, width=512):
    """Closest-hit workload: one displaced icosphere (20*f^2 triangles) at the origin."""
    v, f = P.icosphere(frequency, radius=1.0, center=(0, 0, 0), amplitude=0.1, seed=seed)
    s = P.SceneDesc()
    m = s.lambertian((0.7, 0.7, 0.7))
    s.mesh(v, f, m)
    s.set_camera(center=(0, 0, 4), look_at=(0, 0, 0), up=(0, 1, 0), width=width, aspect_ratio=1.0, vfov=40.0)
    return s, dict(vertices=v, faces=f)


def mesh_sky_scene(frequency=300, seed=0, width=1024, sky=None):
    """1.8M-triangle displaced glass icosphere (IOR 1.5) over a Lambertian ground, lit by a
    1024x512 HDR sky environment light; 1024x1024, PT depth 16, 256 spp."""
    v, f = P.icosphere(frequency, radius=1.0, center=(0, 1.15, 0), amplitude=0.1, seed=seed)
    s = P.SceneDesc()
    glass = s.dielectric(1.5)
    ground = s.lambertian((0.5, 0.5, 0.5))
    s.mesh(v, f, glass)
    ground_quad(s, (0, 0, 0), 20.0, ground)
    if sky is None:
        sky = P.hdr_sky(1024, 512, seed=seed, sun_peak=1.0e4)
    s.add_envmap_light(sky)
    s.set_camera(center=(0, 1.6, 4.5), look_at=(0, 1.0, 0), up=(0, 1, 0), width=width, aspect_ratio=1.0,
                 vfov=40.0)
    return s, dict(max_depth=16, rr_min_bounces=5, spp=256)


def empty_uniform_scene(width=32, emission=(1.0, 1.0, 1.0)):
    """No geometry, uniform infinite light: every pixel equals the emission exactly
    (progressive_integration_test.go:44-70)."""
    s = P.SceneDesc()
    s.add_uniform_infinite_light(emission)
    s.set_camera(center=(0, 0, 0), look_at=(0, 0, -1), up=(0, 1, 0), width=width, aspect_ratio=1.0, vfov=45.0)
    return s, dict(max_depth=5, rr_min_bounces=5, spp=4)


def furnace_scene(albedo=0.5, width=32):
    """White furnace: a Lambertian sphere of albedo a under a uniform sky of radiance 1. With a
    single infinite light the light and BSDF pdfs are both cos/pi, so the reference estimator
    weights direct and indirect lighting by 1/2 each; every bounce off the convex sphere escapes,
    so each sample on the sphere is exactly a and each sky sample exactly 1."""
    s = P.SceneDesc()
    m = s.lambertian((albedo, albedo, albedo))
    s.sphere((0, 0, -3), 1.0, m)
    s.add_uniform_infinite_light((1.0, 1.0, 1.0))
    s.set_camera(center=(0, 0, 0), look_at=(0, 0, -1), up=(0, 1, 0), width=width, aspect_ratio=1.0, vfov=60.0)
    return s, dict(max_depth=50, rr_min_bounces=50, spp=16)


def dragon_scene(ply_path=None, finish="gold", width=900):
    """NewDragonScene (pkg/scene/dragon.go:14-219): Z-up PBRT camera, sphere light + gradient sky, a
    10000-unit ground quad at z = -40, and the PLY mesh rotated -53 degrees about Y around the origin
    (the reference's exact rotation). Without a PLY file the scene has no mesh (loadMesh = false)."""
    s = P.SceneDesc()
    s.set_camera(center=(277, -240, 250), look_at=(0, 60, -30), up=(0, 0, 1), width=width, aspect_ratio=1.0,
                 vfov=33.0, aperture=0.0, focus_distance=0.0)
    s.add_sphere_light((0, 200, 800), 300.0, (15.0 * 0.25, 14.0 * 0.25, 12.0 * 0.25))
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (0.0, 0.0, 0.0))
    ground = s.lambertian((0.6, 0.6, 0.6))
    s.quad((-5000, -5000, -40), (10000, 0, 0), (0, 10000, 0), ground)
    if ply_path is not None:
        finishes = {
            "plastic": lambda: s.layered(s.dielectric(1.4), s.lambertian((0.7, 0.5, 0.8))),
            "matte": lambda: s.lambertian((0.75, 0.65, 0.55)),
            "mirror": lambda: s.metal((0.9, 0.9, 0.9), 0.0),
            "glass": lambda: s.dielectric(1.5),
            "copper": lambda: s.metal((0.8, 0.4, 0.2), 0.1),
        }
        mat = finishes.get(finish, lambda: s.metal((0.7, 0.5, 0.2), 0.02))()
        d = P.load_ply(ply_path)
        rot = (0.0, -53.0 * 3.14159265359 / 180.0, 0.0)
        # the reference passes per-vertex PLY normals as per-triangle normals, which NewTriangleMesh only
        # accepts when the counts agree (triangle_mesh.go:42-45); otherwise the mesh uses its own normals
        n = d.get("normals")
        if n is not None and len(n) != len(d["faces"]):
            raise ValueError("PLY normals are per vertex; NewTriangleMesh needs one normal per triangle")
        s.mesh(d["vertices"], d["faces"], mat, normals=n, rotation=rot, center=(0.0, 0.0, 0.0))
    return s, dict(max_depth=50, rr_min_bounces=15, spp=200, adaptive_min_samples=0.15, adaptive_threshold=0.01)


def caustic_glass_scene(mesh_dir=None, light="spot", width=525):
    """NewCausticGlassScene (pkg/scene/caustic_glass.go:17-188): the glass.pbrt camera, a spot light
    from (0,5,9) to (-5,2.75,0) of intensity (139.81,118.64,105.39) with a 30-degree cone and 5-degree
    falloff (light="spot": disc spot light of radius 0.7; light="point": point spot light), a
    uniform 0.1 sky, and, when `mesh_dir` holds them, mesh_00001.ply (glass, ior 1.25) and
    mesh_00002.ply (Lambertian 0.64)."""
    import os
    s = P.SceneDesc()
    s.set_camera(center=(-5.5, 7, -5.5), look_at=(-4.75, 2.25, 0), up=(0, 1, 0), width=width,
                 aspect_ratio=525.0 / 750.0, vfov=30.0 * 1.5, aperture=0.0, focus_distance=0.0)
    frm, to = (0, 5, 9), (-5, 2.75, 0)
    intensity = (139.8113403320, 118.6366500854, 105.3887557983)
    if light == "point":
        s.add_point_spot_light(frm, to, intensity, 30.0, 5.0)
    else:
        s.add_spot_light(frm, to, intensity, 30.0, 5.0, 0.7)
    s.add_uniform_infinite_light((0.1, 0.1, 0.1))
    if mesh_dir:
        for name, mat in (("mesh_00001.ply", s.dielectric(1.25)), ("mesh_00002.ply", s.lambertian((0.6399999857,) * 3))):
            path = os.path.join(mesh_dir, name)
            if os.path.exists(path):
                ply = P.load_ply(path)
                s.mesh(ply.vertices, ply.faces, mat)
    cfg = dict(max_depth=20, rr_min_bounces=16, spp=8192, adaptive_min_samples=0.2, adaptive_threshold=0.005)
    return s, cfg


def lights_scene(width=96):
    """A metre-scale Cornell box lit by every finite light type: a disc light under the ceiling, a
    disc spot light and a point spot light aimed at the floor from the upper corners, and a
    Lambertian disc and a glass sphere on the floor."""
    s = P.SceneDesc()
    s.set_camera(center=(0.278, 0.278, -0.8), look_at=(0.278, 0.278, 0), up=(0, 1, 0), width=width,
                 aspect_ratio=1.0, vfov=40.0, aperture=0.0, focus_distance=0.0)
    white = s.lambertian((0.73, 0.73, 0.73))
    red = s.lambertian((0.65, 0.05, 0.05))
    green = s.lambertian((0.12, 0.45, 0.15))
    k = 0.001
    s.quad((0, 0, 0), (556 * k, 0, 0), (0, 0, 556 * k), white)
    s.quad((0, 556 * k, 0), (556 * k, 0, 0), (0, 0, 556 * k), white)
    s.quad((0, 0, 556 * k), (556 * k, 0, 0), (0, 556 * k, 0), white)
    s.quad((0, 0, 0), (0, 0, 556 * k), (0, 556 * k, 0), red)
    s.quad((556 * k, 0, 0), (0, 0, 556 * k), (0, 556 * k, 0), green)
    s.add_disc_light((0.278, 0.5559, 0.2795), (0, -1, 0), 0.06, (12.0, 11.0, 9.0))
    s.add_spot_light((0.05, 0.5, 0.05), (0.3, 0.0, 0.35), (4.0, 6.0, 9.0), 35.0, 8.0, 0.02)
    s.add_point_spot_light((0.5, 0.5, 0.1), (0.2, 0.0, 0.4), (0.08, 0.06, 0.03), 40.0, 10.0)
    blue = s.lambertian((0.2, 0.3, 0.7))
    s.disc((0.15, 0.004, 0.3), (0, 1, 0), 0.09, blue)
    s.sphere((0.38, 0.1, 0.3), 0.1, s.dielectric(1.5))
    cfg = dict(max_depth=6, rr_min_bounces=3, spp=16)
    return s, cfg


def _primitive_test_camera(s, width):
    s.set_camera(center=(0, 1.5, 4), look_at=(0, 1, 0), up=(0, 1, 0), width=width, aspect_ratio=16.0 / 9.0,
                 vfov=50.0, aperture=0.0, focus_distance=0.0)


def cylinder_scene(width=400):
    """NewCylinderTestScene (pkg/scene/cylinder_scene.go:9-122): an open gold tube, two capped
    Lambertian cylinders and a capped glass cylinder on a ground quad, a sphere light and a blue
    gradient sky; 200 spp, depth 50, Russian roulette from bounce 20."""
    s = P.SceneDesc()
    _primitive_test_camera(s, width)
    gray, red, blue = s.lambertian((0.5, 0.5, 0.5)), s.lambertian((0.8, 0.2, 0.2)), s.lambertian((0.2, 0.2, 0.8))
    gold, glass = s.metal((0.8, 0.6, 0.2), 0.1), s.dielectric(1.5)
    ground_quad(s, (0, 0, 0), 10000.0, gray)
    s.cylinder((-0.3, 1.0, -1.5), (0, 1.2, 2.0), 0.35, False, gold)
    s.cylinder((-2.5, 0.3, 0), (-1.5, 0.3, 0), 0.3, True, blue)
    s.cylinder((1.8, 0, 0), (1.8, 2, 0), 0.5, True, red)
    s.cylinder((0.5, 0, 1), (0.5, 0.6, 1), 0.2, True, glass)
    s.add_sphere_light((3, 5, 3), 1.5, (10.0, 10.0, 10.0))
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    cfg = dict(max_depth=50, rr_min_bounces=20, spp=200, adaptive_min_samples=0.15, adaptive_threshold=0.01)
    return s, cfg


def cone_scene(width=400):
    """NewConeTestScene (pkg/scene/cone_scene.go:9-160): pointed, frustum, tilted and glass cones
    (capped and open) on a ground quad, a sphere light and a blue gradient sky; 200 spp, depth 50."""
    s = P.SceneDesc()
    _primitive_test_camera(s, width)
    gray, red = s.lambertian((0.5, 0.5, 0.5)), s.lambertian((0.8, 0.2, 0.2))
    blue, green = s.lambertian((0.2, 0.2, 0.8)), s.lambertian((0.2, 0.8, 0.2))
    gold, glass = s.metal((0.8, 0.6, 0.2), 0.1), s.dielectric(1.5)
    ground_quad(s, (0, 0, 0), 10000.0, gray)
    s.cone((0, 0, 0), 0.5, (0, 2, 0), 0.0, True, red)
    s.cone((-2, 0.8, -0.8), 0.5, (-2, 0.2, 0.5), 0.2, True, gold)
    s.cone((2, 0, 0), 0.8, (2, 0.6, 0), 0.5, True, blue)
    s.cone((2, 0.6, 0), 0.5, (2, 1.8, 0), 0.0, True, glass)
    s.cone((-1.5, 0, -0.5), 0.4, (-1.2, 1.2, -0.3), 0.15, False, green)
    s.cone((-0.8, 0, 1.2), 0.3, (-0.8, 0.8, 1.2), 0.0, True, glass)
    s.cone((0.8, 0.2, 1.2), 0.3, (0.8, 1.0, 1.2), 0.0, True, glass)
    s.add_sphere_light((3, 5, 3), 1.5, (10.0, 10.0, 10.0))
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    cfg = dict(max_depth=50, rr_min_bounces=20, spp=200, adaptive_min_samples=0.15, adaptive_threshold=0.01)
    return s, cfg


def texture_scene(width=800):
    """NewTextureTestScene (pkg/scene/texture_test_scene.go:11-161): a checkerboard sphere, a gradient
    cylinder, a UV-debug frustum, a brick-checker box, a checkerboard disc, a gradient quad and a
    UV-debug triangle on a brick-checker ground; a sphere light and a dim gradient sky; 100 spp,
    depth 10, Russian roulette from bounce 5."""
    s = P.SceneDesc()
    s.set_camera(center=(0, 2, 10), look_at=(0, 1, 0), up=(0, 1, 0), width=width, aspect_ratio=16.0 / 9.0,
                 vfov=50.0, aperture=0.0, focus_distance=0.0)
    checker = s.texture(P.checkerboard_texture(256, 256, 32, (0.9, 0.9, 0.9), (0.2, 0.2, 0.8)))
    gradient = s.texture(P.gradient_texture(256, 256, (1.0, 0.2, 0.2), (0.2, 1.0, 0.2)))
    uvdebug = s.texture(P.uv_debug_texture(256, 256))
    brick = s.texture(P.checkerboard_texture(512, 512, 16, (0.7, 0.3, 0.1), (0.5, 0.2, 0.05)))
    checker_m = s.lambertian((0, 0, 0), texture=checker)
    gradient_m = s.lambertian((0, 0, 0), texture=gradient)
    uvdebug_m = s.lambertian((0, 0, 0), texture=uvdebug)
    brick_m = s.lambertian((0, 0, 0), texture=brick)
    s.sphere((-6, 1, 0), 1.0, checker_m)
    s.cylinder((-4, 0, 0), (-4, 2, 0), 0.6, True, gradient_m)
    s.cone((-2, 0, 0), 0.8, (-2, 2, 0), 0.2, True, uvdebug_m)
    s.box((0, 1, 0), (0.8, 0.8, 0.8), (0, 0, 0), brick_m)
    s.disc((2, 1.2, 0), (0, 0, 1), 0.9, checker_m)
    s.quad((3.5, 0, 0.2), (1.5, 0, -0.3), (0, 2, 0), gradient_m)
    s.mesh(np.array([[5.5, 0, 0], [7, 0, 0], [6.25, 2, 0]], float), np.array([[0, 1, 2]]), uvdebug_m)
    s.quad((-10, 0, -5), (20, 0, 0), (0, 0, 15), brick_m)
    s.add_sphere_light((0, 8, 5), 2.0, (20, 20, 20))
    s.add_gradient_infinite_light((0.3, 0.4, 0.6), (0.2, 0.2, 0.2))
    cfg = dict(max_depth=10, rr_min_bounces=5, spp=100, adaptive_min_samples=0.15, adaptive_threshold=0.01)
    return s, cfg


def oklch_to_rgb(l, c, h):
    """OKLCh -> linear sRGB through OKLab (Ottosson's published matrices), clamped to [0, 1]; the
    sphere grid's palette (spheregrid.go:11-38)."""
    a, b = c * math.cos(math.radians(h)), c * math.sin(math.radians(h))
    lms = np.array([l + 0.3963377774 * a + 0.2158037573 * b,
                    l - 0.1055613458 * a - 0.0638541728 * b,
                    l - 0.0894841775 * a - 1.2914855480 * b]) ** 3
    m = np.array([[4.0767416621, -3.3077115913, 0.2309699292],
                  [-1.2684380046, 2.6097574011, -0.3413193965],
                  [-0.0041960863, -0.7034186147, 1.7076147010]])
    return tuple(float(x) for x in np.clip(m @ lms, 0.0, 1.0))


def sphere_grid_scene(grid_size=10, finish="metallic", width=800):
    """NewSphereGridScene (pkg/scene/spheregrid.go:41-160): grid_size^2 spheres over a 9x9 area,
    OKLCh hues; finish = matte | glossy (glass over Lambertian) | mirror | glass | mixed (50/50
    Lambertian/metal Mix) | metallic; a sphere light, a gradient sky, f/0.02 lens; 100 spp, depth 40."""
    s = P.SceneDesc()
    s.set_camera(center=(4.5, 6, 18), look_at=(4.5, 0.8, 4.5), up=(0, 1, 0), width=width, aspect_ratio=16.0 / 9.0,
                 vfov=40.0, aperture=0.02, focus_distance=0.0)
    s.add_sphere_light((20, 25, 20), 8, (8.0, 7.5, 6.5))
    ground_quad(s, (0, 0, 0), 10000.0, s.lambertian((0.5, 0.5, 0.5)))
    area = 9.0
    spacing = area / (grid_size - 1)
    radius = max(0.02, min(0.35, spacing * 0.35))
    for i in range(grid_size):
        for j in range(grid_size):
            x = i * spacing - area / 2.0 + 4.5
            z = j * spacing - area / 2.0 + 4.5
            hue = math.fmod(i / (grid_size - 1) * 330.0 + j / (grid_size - 1) * 120.0, 360.0)
            chroma = 0.05 + 0.8 * 0.2 + 0.2 * 0.2 * (j / (grid_size - 1))
            color = oklch_to_rgb(0.65, chroma, hue)
            if finish == "matte":
                m = s.lambertian(color)
            elif finish == "glossy":
                m = s.layered(s.dielectric(1.5), s.lambertian(color))
            elif finish == "mirror":
                m = s.metal((0.9, 0.9, 0.9), 0.0)
            elif finish == "glass":
                m = s.dielectric(1.5)
            elif finish == "mixed":
                m = s.mix(s.lambertian(color), s.metal(color, 0.1), 0.5)
            else:
                m = s.metal(color, 0.05)
            s.sphere((x, radius, z), radius, m)
    s.add_gradient_infinite_light((0.5, 0.7, 1.0), (1.0, 1.0, 1.0))
    cfg = dict(max_depth=40, rr_min_bounces=12, spp=100, adaptive_min_samples=0.12, adaptive_threshold=0.015)
    return s, cfg
