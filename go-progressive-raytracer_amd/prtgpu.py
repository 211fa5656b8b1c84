"""ctypes binding of libprt.so (the C ABI in include/prt.h) and a host-side mirror of the
reference's renderer interface for this path.

Names follow the reference: ``ProgressiveRaytracer.render_pass`` mirrors
``ProgressiveRaytracer.RenderPass`` (pkg/renderer/progressive.go:126-223),
``get_samples_for_pass`` mirrors ``getSamplesForPass`` (progressive.go:98-123) and
``Scene.hit`` mirrors the batched ``BVH.Hit`` (pkg/geometry/bvh.go:183-239).
The library must be built (``python build.py``); importing this module without it
raises immediately: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprt.so")

PRT_OK = 0
ERRORS = {1: "invalid argument", 2: "device error", 3: "out of memory", 4: "no device",
          5: "unsupported", 6: "cancelled"}

MAT_LAMBERTIAN, MAT_METAL, MAT_DIELECTRIC, MAT_EMISSIVE, MAT_LAYERED = range(5)
MAT_MIX = 6
LIGHT_QUAD, LIGHT_SPHERE, LIGHT_UNIFORM_INFINITE, LIGHT_GRADIENT_INFINITE, LIGHT_ENVMAP = range(5)
LIGHT_DISC, LIGHT_DISC_SPOT, LIGHT_POINT_SPOT = 5, 6, 7
INTEGRATOR_PATH, INTEGRATOR_BDPT = 0, 1
BUFFERS_HOST, BUFFERS_DEVICE = 0, 1


class PrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"prt error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


D3 = C.c_double * 3


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("outer", C.c_int32), ("inner", C.c_int32), ("texture", C.c_int32),
                ("albedo", D3), ("fuzz", C.c_double), ("ior", C.c_double), ("emission", D3),
                ("ratio", C.c_double)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.POINTER(C.c_float))]


class Sphere(C.Structure):
    _fields_ = [("center", D3), ("radius", C.c_double), ("material", C.c_int32), ("pad", C.c_int32)]


class Quad(C.Structure):
    _fields_ = [("corner", D3), ("u", D3), ("v", D3), ("material", C.c_int32), ("pad", C.c_int32)]


class Box(C.Structure):
    _fields_ = [("center", D3), ("size", D3), ("rotation", D3), ("material", C.c_int32), ("pad", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_double)), ("num_vertices", C.c_int64),
                ("faces", C.POINTER(C.c_int32)), ("num_faces", C.c_int64),
                ("face_materials", C.POINTER(C.c_int32)), ("material", C.c_int32), ("pad", C.c_int32),
                ("normals", C.POINTER(C.c_double)), ("rotation", C.POINTER(C.c_double)),
                ("center", C.POINTER(C.c_double)), ("uvs", C.POINTER(C.c_double))]


class Disc(C.Structure):
    _fields_ = [("center", D3), ("normal", D3), ("radius", C.c_double), ("material", C.c_int32), ("pad", C.c_int32)]


class Cylinder(C.Structure):
    _fields_ = [("base_center", D3), ("top_center", D3), ("radius", C.c_double), ("capped", C.c_int32),
                ("material", C.c_int32)]


class Cone(C.Structure):
    _fields_ = [("base_center", D3), ("base_radius", C.c_double), ("top_center", D3), ("top_radius", C.c_double),
                ("capped", C.c_int32), ("material", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("shape", C.c_int32), ("emission", D3), ("top", D3), ("bottom", D3),
                ("env_rgb", C.POINTER(C.c_float)), ("env_width", C.c_int32), ("env_height", C.c_int32),
                ("from_", D3), ("to", D3), ("cone_angle", C.c_double), ("cone_delta", C.c_double),
                ("radius", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("center", D3), ("look_at", D3), ("up", D3), ("width", C.c_int32), ("pad", C.c_int32),
                ("aspect_ratio", C.c_double), ("vfov", C.c_double), ("aperture", C.c_double),
                ("focus_distance", C.c_double)]


class SceneDescC(C.Structure):
    _fields_ = [("materials", C.POINTER(Material)), ("num_materials", C.c_int32),
                ("spheres", C.POINTER(Sphere)), ("num_spheres", C.c_int32),
                ("quads", C.POINTER(Quad)), ("num_quads", C.c_int32),
                ("boxes", C.POINTER(Box)), ("num_boxes", C.c_int32),
                ("meshes", C.POINTER(Mesh)), ("num_meshes", C.c_int32),
                ("lights", C.POINTER(Light)), ("num_lights", C.c_int32),
                ("light_weights", C.POINTER(C.c_double)),
                ("camera", Camera),
                ("discs", C.POINTER(Disc)), ("num_discs", C.c_int32),
                ("cylinders", C.POINTER(Cylinder)), ("num_cylinders", C.c_int32),
                ("cones", C.POINTER(Cone)), ("num_cones", C.c_int32),
                ("textures", C.POINTER(Texture)), ("num_textures", C.c_int32)]


class PbrtSceneC(C.Structure):
    _fields_ = [("desc", SceneDescC), ("width", C.c_int32), ("height", C.c_int32),
                ("samples_per_pixel", C.c_int32), ("max_depth", C.c_int32), ("rr_min_bounces", C.c_int32),
                ("pad", C.c_int32), ("adaptive_min_samples", C.c_double), ("adaptive_threshold", C.c_double),
                ("reference_shapes", C.c_int64), ("impl", C.c_void_p)]


class SceneInfo(C.Structure):
    _fields_ = [("num_primitives", C.c_int64), ("num_triangles", C.c_int64), ("bvh_nodes", C.c_int64),
                ("bvh_depth", C.c_int32), ("max_stack", C.c_int32), ("world_center", D3),
                ("world_radius", C.c_double), ("device_bytes", C.c_int64), ("trace_variant", C.c_int32),
                ("pt_trace_waves", C.c_int32), ("reference_primitives", C.c_int64),
                ("pt_trace_leaf_gate", C.c_int32), ("bdpt_trace_leaf_gate", C.c_int32)]


class RenderParams(C.Structure):
    _fields_ = [("integrator", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("max_depth", C.c_int32), ("rr_min_bounces", C.c_int32), ("tile_size", C.c_int32),
                ("initial_samples", C.c_int32), ("max_samples_per_pixel", C.c_int32),
                ("max_passes", C.c_int32), ("pad", C.c_int32), ("adaptive_min_samples", C.c_double),
                ("adaptive_threshold", C.c_double), ("seed", C.c_uint64)]


class PassStats(C.Structure):
    _fields_ = [("total_pixels", C.c_int64), ("total_samples", C.c_int64), ("average_samples", C.c_double),
                ("max_samples", C.c_int32), ("min_samples", C.c_int32), ("max_samples_used", C.c_int32),
                ("target_samples", C.c_int32), ("device_ms", C.c_double)]


RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<i4"), ("u", "<f4"), ("v", "<f4")])

# every symbol include/prt.h and include/prt_workload.h declare
EXPORTS = [
    "prt_last_error", "prt_version", "prt_device_count", "prt_scene_create", "prt_scene_info_get",
    "prt_scene_destroy", "prt_intersect_closest", "prt_intersect_any", "prt_traversal_counts",
    "prt_intersect_closest_static",
    "prt_samples_for_pass", "prt_film_create", "prt_render_pass", "prt_render_samples", "prt_film_read",
    "prt_film_probe",
    "prt_film_read_splats", "prt_film_read_rgba8", "prt_film_reset", "prt_film_destroy", "prt_film_size",
    "prt_workload_icosphere", "prt_workload_rays", "prt_workload_sky", "prt_ply_load", "prt_ply_free",
    "prt_server_create", "prt_server_submit", "prt_server_poll", "prt_server_cancel", "prt_server_read",
    "prt_server_read_rgba8", "prt_server_release", "prt_server_stats_get", "prt_server_destroy",
    "prt_server_submit_ex", "prt_server_read_pass", "prt_png_encode_rgba8", "prt_sse_format_pass",
    "prt_sse_format_complete", "prt_sse_format_error", "prt_sse_server_event", "prt_free",
    "prt_pbrt_load", "prt_pbrt_parse", "prt_pbrt_free",
]
SUBMIT_SNAPSHOTS = 1

EVENT_ADMITTED, EVENT_PASS_COMPLETE, EVENT_DONE, EVENT_ERROR, EVENT_CANCELLED = range(5)


class ProbeStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("extension_rays", "extension_node_visits", "extension_box_tests",
                                          "extension_prim_tests", "shadow_rays", "shadow_node_visits",
                                          "shadow_box_tests", "shadow_prim_tests", "trace_launches")] + \
               [(n, C.c_double) for n in ("camera_ms", "trace_ms", "shade_ms", "film_ms", "bdpt_paths_ms",
                                          "bdpt_walks_ms", "bdpt_connect_ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


PROBE_COUNT, PROBE_TIME = 1, 2


class PlyData(C.Structure):
    _fields_ = [("num_vertices", C.c_int64), ("num_faces", C.c_int64), ("vertices", C.POINTER(C.c_double)),
                ("faces", C.POINTER(C.c_int32)), ("normals", C.POINTER(C.c_double)),
                ("colors", C.POINTER(C.c_double)), ("texcoords", C.POINTER(C.c_double))]


class ServerEvent(C.Structure):
    _fields_ = [("request_id", C.c_int64), ("type", C.c_int32), ("pass_number", C.c_int32),
                ("target_samples", C.c_int32), ("error_code", C.c_int32), ("t_ms", C.c_double),
                ("device_ms", C.c_double)]


class SsePassInfo(C.Structure):
    _fields_ = [("pass_number", C.c_int32), ("total_passes", C.c_int32), ("target_samples", C.c_int32),
                ("max_samples_per_pixel", C.c_int32), ("tile_size", C.c_int32), ("pad", C.c_int32),
                ("primitive_count", C.c_int64), ("elapsed_ms", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int32), ("active", C.c_int32), ("completed", C.c_int64),
                ("budget_bytes", C.c_int64), ("used_bytes", C.c_int64), ("idle_bytes", C.c_int64)]

_lib = None


def lib():
    """Loads libprt.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with go-progressive-raytracer_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "prt_last_error": (C.c_char_p, []),
        "prt_version": (C.c_int, []),
        "prt_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
        "prt_scene_create": (C.c_int, [C.c_int32, C.POINTER(SceneDescC), C.POINTER(vp)]),
        "prt_scene_info_get": (C.c_int, [vp, C.POINTER(SceneInfo)]),
        "prt_scene_destroy": (C.c_int, [vp]),
        "prt_intersect_closest": (C.c_int, [vp, vp, C.c_int64, vp, C.c_uint32, vp]),
        "prt_intersect_any": (C.c_int, [vp, vp, C.c_int64, vp, C.c_uint32, vp]),
        "prt_traversal_counts": (C.c_int, [vp, vp, C.c_int64, vp, vp, vp, C.c_uint32, vp]),
        "prt_intersect_closest_static": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
        "prt_samples_for_pass": (C.c_int, [C.POINTER(RenderParams), C.c_int32, C.POINTER(C.c_int32)]),
        "prt_film_create": (C.c_int, [vp, C.POINTER(RenderParams), C.POINTER(vp)]),
        "prt_render_pass": (C.c_int, [vp, C.c_int32, C.POINTER(PassStats)]),
        "prt_render_samples": (C.c_int, [vp, C.c_int32, C.POINTER(PassStats)]),
        "prt_film_read": (C.c_int, [vp, vp, vp, vp, vp]),
        "prt_film_read_splats": (C.c_int, [vp, vp]),
        "prt_film_read_rgba8": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]),
        "prt_film_size": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "prt_film_reset": (C.c_int, [vp]),
        "prt_film_destroy": (C.c_int, [vp]),
        "prt_film_probe": (C.c_int, [vp, C.c_uint32, C.POINTER(ProbeStats)]),
        "prt_workload_icosphere": (C.c_int, [C.c_int32, C.c_double, vp, C.c_double, C.c_uint64, vp,
                                             C.POINTER(C.c_int64), vp, C.POINTER(C.c_int64)]),
        "prt_workload_rays": (C.c_int, [vp, vp, C.c_double, C.c_int64, C.c_uint64, vp]),
        "prt_workload_sky": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, C.c_double, vp]),
        "prt_ply_load": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(PlyData))]),
        "prt_ply_free": (None, [C.POINTER(PlyData)]),
        "prt_server_create": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.POINTER(vp)]),
        "prt_server_submit": (C.c_int, [vp, vp, C.POINTER(RenderParams), C.POINTER(C.c_int64)]),
        "prt_server_poll": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
        "prt_server_cancel": (C.c_int, [vp, C.c_int64]),
        "prt_server_read": (C.c_int, [vp, C.c_int64, vp, vp]),
        "prt_server_read_rgba8": (C.c_int, [vp, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp]),
        "prt_server_release": (C.c_int, [vp, C.c_int64]),
        "prt_server_stats_get": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "prt_server_destroy": (C.c_int, [vp]),
        "prt_server_submit_ex": (C.c_int, [vp, vp, C.POINTER(RenderParams), C.c_uint32, C.POINTER(C.c_int64)]),
        "prt_server_read_pass": (C.c_int, [vp, C.c_int64, C.c_int32, vp, vp]),
        "prt_png_encode_rgba8": (C.c_int, [vp, C.c_int32, C.c_int32, C.POINTER(vp), C.POINTER(C.c_int64)]),
        "prt_sse_format_pass": (C.c_int, [C.c_int32, C.c_int32, vp, vp, C.POINTER(SsePassInfo), C.POINTER(vp),
                                          C.POINTER(C.c_int64)]),
        "prt_sse_format_complete": (C.c_int, [C.POINTER(vp), C.POINTER(C.c_int64)]),
        "prt_sse_format_error": (C.c_int, [C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int64)]),
        "prt_sse_server_event": (C.c_int, [vp, vp, C.POINTER(RenderParams), C.c_int32, C.POINTER(ServerEvent),
                                           C.POINTER(vp), C.POINTER(C.c_int64)]),
        "prt_free": (None, [vp]),
        "prt_pbrt_load": (C.c_int, [C.c_char_p, C.POINTER(C.POINTER(PbrtSceneC))]),
        "prt_pbrt_parse": (C.c_int, [C.c_char_p, C.c_int64, C.POINTER(C.POINTER(PbrtSceneC))]),
        "prt_pbrt_free": (None, [C.POINTER(PbrtSceneC)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != PRT_OK:
        raise PrtError(rc, lib().prt_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class PbrtScene:
    """A scene read by the PBRT loader (prt_pbrt_load / prt_pbrt_parse; scene.NewPBRTScene). Usable
    wherever a SceneDesc is (Scene, the oracle); `sampling` holds the PBRT sampling defaults."""

    def __init__(self, handle):
        self._h = handle
        o = handle.contents
        self.camera = o.desc.camera
        self.sampling = dict(width=o.width, height=o.height, spp=o.samples_per_pixel, max_depth=o.max_depth,
                             rr_min_bounces=o.rr_min_bounces, adaptive_min_samples=o.adaptive_min_samples,
                             adaptive_threshold=o.adaptive_threshold)
        self.reference_shapes = o.reference_shapes

    @property
    def desc(self):
        return self._h.contents.desc

    @property
    def image_size(self):
        w = self.camera.width
        return w, int(w / self.camera.aspect_ratio)

    def to_c(self):
        return self._h.contents.desc, self

    def __del__(self):
        try:
            if self._h:
                lib().prt_pbrt_free(self._h)
                self._h = None
        except Exception:
            pass


def load_pbrt(path):
    """LoadPBRT + NewPBRTScene: the path must be a .pbrt file under a scenes/ directory (or the
    temporary directory), as the reference's validateFilePath requires."""
    h = C.POINTER(PbrtSceneC)()
    check(lib().prt_pbrt_load(os.fsencode(path), C.byref(h)))
    return PbrtScene(h)


def parse_pbrt(text):
    """ParsePBRT + NewPBRTScene on a string."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    h = C.POINTER(PbrtSceneC)()
    check(lib().prt_pbrt_parse(b, len(b), C.byref(h)))
    return PbrtScene(h)


def load_ply(path):
    """LoadPLY (pkg/loaders/ply.go:64-110): dict of numpy arrays (vertices (n,3), faces (m,3), and
    normals / colors (n,3), texcoords (n,2) when the file has them)."""
    h = C.POINTER(PlyData)()
    check(lib().prt_ply_load(os.fsencode(path), C.byref(h)))
    try:
        d = h.contents
        nv, nf = d.num_vertices, d.num_faces

        def arr(ptr, n, k, dt):
            if not ptr:
                return None
            return np.ctypeslib.as_array(ptr, shape=(n * k,)).astype(dt).reshape(n, k) if n else np.zeros((0, k), dt)

        out = {"vertices": arr(d.vertices, nv, 3, np.float64), "faces": arr(d.faces, nf, 3, np.int32)}
        for name, k in (("normals", 3), ("colors", 3), ("texcoords", 2)):
            a = arr(getattr(d, name), nv, k, np.float64)
            if a is not None:
                out[name] = a
        return out
    finally:
        lib().prt_ply_free(h)


def device_count():
    n = C.c_int32(0)
    check(lib().prt_device_count(C.byref(n)))
    return n.value


# ---------------------------------------------------------------------------------------
class SceneDesc:
    """Flattened scene (materials, shapes, lights, camera) in the reference's float64 terms.

    Mirrors the pieces of ``scene.Scene`` (pkg/scene/scene.go:11-19) the path needs; the
    ``add_*`` helpers follow Scene.AddSphereLight / AddQuadLight / AddGradientInfiniteLight
    (scene.go:102-141).
    """

    def __init__(self):
        self.materials = []
        self.spheres = []
        self.quads = []
        self.boxes = []
        self.discs = []
        self.cylinders = []
        self.cones = []
        self.textures = []
        self.meshes = []  # (vertices (n,3), faces (m,3), material, face_materials|None, options dict)
        self.lights = []
        self.light_weights = None
        self.camera = None

    # materials (pkg/material constructors)
    def _mat(self, **kw):
        m = Material()
        m.type = kw.get("type", MAT_LAMBERTIAN)
        m.outer, m.inner = kw.get("outer", -1), kw.get("inner", -1)
        m.albedo = D3(*kw.get("albedo", (0, 0, 0)))
        m.fuzz = kw.get("fuzz", 0.0)
        m.ior = kw.get("ior", 1.0)
        m.emission = D3(*kw.get("emission", (0, 0, 0)))
        m.texture = kw.get("texture", 0)
        m.ratio = kw.get("ratio", 0.0)
        self.materials.append(m)
        return len(self.materials) - 1

    def texture(self, rgb):
        """An ImageTexture (image_texture.go:15-22) from an (height, width, 3) array, row 0 at the top;
        returns the value to pass as `texture=` (1-based)."""
        a = np.ascontiguousarray(rgb, dtype=np.float32)
        assert a.ndim == 3 and a.shape[2] == 3
        self.textures.append(a)
        return len(self.textures)

    def lambertian(self, albedo, texture=0):
        """NewLambertian, or NewTexturedLambertian (lambertian.go:15-22) with texture > 0."""
        return self._mat(type=MAT_LAMBERTIAN, albedo=albedo, texture=texture)

    def metal(self, albedo, fuzz, texture=0):
        """NewMetal, or NewTexturedMetal (metal.go:14-35) with texture > 0."""
        return self._mat(type=MAT_METAL, albedo=albedo, fuzz=min(1.0, max(0.0, fuzz)), texture=texture)

    def mix(self, material1, material2, ratio):
        """NewMix (mix.go:17-27): material2 with probability ratio (clamped to [0, 1])."""
        return self._mat(type=MAT_MIX, outer=material1, inner=material2, ratio=min(1.0, max(0.0, ratio)))

    def dielectric(self, ior):
        return self._mat(type=MAT_DIELECTRIC, ior=ior)

    def emissive(self, emission):
        return self._mat(type=MAT_EMISSIVE, emission=emission)

    def layered(self, outer, inner):
        return self._mat(type=MAT_LAYERED, outer=outer, inner=inner)

    # shapes
    def sphere(self, center, radius, material):
        s = Sphere()
        s.center, s.radius, s.material = D3(*center), radius, material
        self.spheres.append(s)
        return len(self.spheres) - 1

    def quad(self, corner, u, v, material):
        q = Quad()
        q.corner, q.u, q.v, q.material = D3(*corner), D3(*u), D3(*v), material
        self.quads.append(q)
        return len(self.quads) - 1

    def disc(self, center, normal, radius, material):
        """NewDisc (pkg/geometry/disc.go:20-43)."""
        dsc = Disc()
        dsc.center, dsc.normal, dsc.radius, dsc.material = D3(*center), D3(*normal), radius, material
        self.discs.append(dsc)
        return len(self.discs) - 1

    def cylinder(self, base_center, top_center, radius, capped, material):
        """NewCylinder (pkg/geometry/cylinder.go:24-39)."""
        c = Cylinder()
        c.base_center, c.top_center, c.radius = D3(*base_center), D3(*top_center), radius
        c.capped, c.material = 1 if capped else 0, material
        self.cylinders.append(c)
        return len(self.cylinders) - 1

    def cone(self, base_center, base_radius, top_center, top_radius, capped, material):
        """NewCone (pkg/geometry/cone.go:27-80); its argument checks are made by prt_scene_create."""
        c = Cone()
        c.base_center, c.base_radius = D3(*base_center), base_radius
        c.top_center, c.top_radius = D3(*top_center), top_radius
        c.capped, c.material = 1 if capped else 0, material
        self.cones.append(c)
        return len(self.cones) - 1

    def box(self, center, size, rotation, material):
        b = Box()
        b.center, b.size, b.rotation, b.material = D3(*center), D3(*size), D3(*rotation), material
        self.boxes.append(b)
        return len(self.boxes) - 1

    def mesh(self, vertices, faces, material, face_materials=None, normals=None, rotation=None, center=None,
             uvs=None):
        """NewTriangleMesh (triangle_mesh.go:33-146) with TriangleMeshOptions: per-triangle normals,
        rotation (radians, X then Y then Z) about `center`, per-vertex texture coordinates."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        fm = None if face_materials is None else np.ascontiguousarray(face_materials, dtype=np.int32)
        opt = {k: (None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(-1))
               for k, x in (("normals", normals), ("rotation", rotation), ("center", center), ("uvs", uvs))}
        self.meshes.append((v, f, material, fm, opt))
        return len(self.meshes) - 1

    # lights
    def add_sphere_light(self, center, radius, emission):
        m = self.emissive(emission)
        s = self.sphere(center, radius, m)
        L = Light()
        L.type, L.shape = LIGHT_SPHERE, s
        self.lights.append(L)

    def add_quad_light(self, corner, u, v, emission):
        m = self.emissive(emission)
        q = self.quad(corner, u, v, m)
        L = Light()
        L.type, L.shape = LIGHT_QUAD, q
        self.lights.append(L)

    def add_disc_light(self, center, normal, radius, emission):
        """A disc area light (NewDiscLight, pkg/lights/disc_light.go:17-22) whose disc is a scene shape."""
        m = self.emissive(emission)
        dsc = self.disc(center, normal, radius, m)
        L = Light()
        L.type, L.shape = LIGHT_DISC, dsc
        self.lights.append(L)

    def add_spot_light(self, frm, to, emission, cone_angle, cone_delta, radius):
        """AddSpotLight (pkg/scene/scene.go:117-122): a disc spot light; its disc joins the shapes."""
        L = Light()
        L.type, L.shape, L.emission = LIGHT_DISC_SPOT, -1, D3(*emission)
        L.from_, L.to, L.cone_angle, L.cone_delta, L.radius = D3(*frm), D3(*to), cone_angle, cone_delta, radius
        self.lights.append(L)

    def add_point_spot_light(self, frm, to, emission, cone_angle, cone_delta):
        """AddPointSpotLight (pkg/scene/scene.go:125-129): a point spot light (a delta light)."""
        L = Light()
        L.type, L.shape, L.emission = LIGHT_POINT_SPOT, -1, D3(*emission)
        L.from_, L.to, L.cone_angle, L.cone_delta = D3(*frm), D3(*to), cone_angle, cone_delta
        self.lights.append(L)

    def add_uniform_infinite_light(self, emission):
        L = Light()
        L.type, L.shape, L.emission = LIGHT_UNIFORM_INFINITE, -1, D3(*emission)
        self.lights.append(L)

    def add_gradient_infinite_light(self, top, bottom):
        L = Light()
        L.type, L.shape, L.top, L.bottom = LIGHT_GRADIENT_INFINITE, -1, D3(*top), D3(*bottom)
        self.lights.append(L)

    def add_envmap_light(self, rgb):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        assert rgb.ndim == 3 and rgb.shape[2] == 3
        L = Light()
        L.type, L.shape = LIGHT_ENVMAP, -1
        L.env_width, L.env_height = rgb.shape[1], rgb.shape[0]
        L._rgb = rgb
        self.lights.append(L)

    def set_camera(self, center, look_at, up, width, aspect_ratio, vfov, aperture=0.0, focus_distance=0.0):
        c = Camera()
        c.center, c.look_at, c.up = D3(*center), D3(*look_at), D3(*up)
        c.width, c.aspect_ratio, c.vfov = width, aspect_ratio, vfov
        c.aperture, c.focus_distance = aperture, focus_distance
        self.camera = c

    @property
    def image_size(self):
        w = self.camera.width
        return w, int(w / self.camera.aspect_ratio)

    def to_c(self):
        """Returns (prt_scene_desc, keepalive) — keep the second value alive while the first is used."""
        keep = []

        def arr(ctype, items):
            if not items:
                return None
            a = (ctype * len(items))(*items)
            keep.append(a)
            return C.cast(a, C.POINTER(ctype))

        meshes = []
        for v, f, m, fm, opt in self.meshes:
            M = Mesh()
            for k, x in opt.items():
                setattr(M, k, x.ctypes.data_as(C.POINTER(C.c_double)) if x is not None else None)
                keep.append(x)
            M.vertices = v.ctypes.data_as(C.POINTER(C.c_double))
            M.num_vertices = v.shape[0]
            M.faces = f.ctypes.data_as(C.POINTER(C.c_int32))
            M.num_faces = f.shape[0]
            M.face_materials = fm.ctypes.data_as(C.POINTER(C.c_int32)) if fm is not None else None
            M.material = m
            keep += [v, f, fm]
            meshes.append(M)
        for L in self.lights:
            if L.type == LIGHT_ENVMAP:
                L.env_rgb = L._rgb.ctypes.data_as(C.POINTER(C.c_float))
                keep.append(L._rgb)
        d = SceneDescC()
        d.materials, d.num_materials = arr(Material, self.materials), len(self.materials)
        d.spheres, d.num_spheres = arr(Sphere, self.spheres), len(self.spheres)
        d.quads, d.num_quads = arr(Quad, self.quads), len(self.quads)
        d.boxes, d.num_boxes = arr(Box, self.boxes), len(self.boxes)
        d.discs, d.num_discs = arr(Disc, self.discs), len(self.discs)
        d.cylinders, d.num_cylinders = arr(Cylinder, self.cylinders), len(self.cylinders)
        d.cones, d.num_cones = arr(Cone, self.cones), len(self.cones)
        texs = []
        for a in self.textures:
            T = Texture()
            T.height, T.width = a.shape[0], a.shape[1]
            T.rgb = a.ctypes.data_as(C.POINTER(C.c_float))
            keep.append(a)
            texs.append(T)
        d.textures, d.num_textures = arr(Texture, texs), len(texs)
        d.meshes, d.num_meshes = arr(Mesh, meshes), len(meshes)
        d.lights, d.num_lights = arr(Light, self.lights), len(self.lights)
        if self.light_weights is not None:
            w = np.ascontiguousarray(self.light_weights, dtype=np.float64)
            keep.append(w)
            d.light_weights = w.ctypes.data_as(C.POINTER(C.c_double))
        d.camera = self.camera if self.camera is not None else Camera()
        return d, keep


def sampling_params(width, height, max_depth, rr_min_bounces, max_samples_per_pixel, max_passes=1,
                    initial_samples=1, seed=0, integrator=INTEGRATOR_PATH, tile_size=64,
                    adaptive_min_samples=1.0, adaptive_threshold=0.0):
    """RenderParams = SamplingConfig (scene.go:22-30) + ProgressiveConfig (progressive.go:29-46)."""
    p = RenderParams()
    p.integrator, p.width, p.height = integrator, width, height
    p.max_depth, p.rr_min_bounces, p.tile_size = max_depth, rr_min_bounces, tile_size
    p.initial_samples, p.max_samples_per_pixel, p.max_passes = initial_samples, max_samples_per_pixel, max_passes
    p.adaptive_min_samples, p.adaptive_threshold, p.seed = adaptive_min_samples, adaptive_threshold, seed
    return p


def get_samples_for_pass(params, pass_number):
    t = C.c_int32(0)
    check(lib().prt_samples_for_pass(C.byref(params), pass_number, C.byref(t)))
    return t.value


# ---------------------------------------------------------------------------------------
class Scene:
    """A scene resident on one GPU (prt_scene_create): BVH + primitives + materials + lights."""

    def __init__(self, desc, device=0):
        self.desc = desc
        d, keep = desc.to_c()
        h = C.c_void_p()
        check(lib().prt_scene_create(device, C.byref(d), C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            check(lib().prt_scene_destroy(self.handle))  # fails while a film or request still uses it
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = SceneInfo()
        check(lib().prt_scene_info_get(self.handle, C.byref(i)))
        return i

    def intersect_closest(self, rays):
        """rays: structured array of RAY_DTYPE (host). Returns HIT_DTYPE array."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.empty(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().prt_intersect_closest(self.handle, _ptr(rays), rays.shape[0], _ptr(hits), BUFFERS_HOST, None))
        return hits

    def intersect_closest_device(self, rays_ptr, n, hits_ptr, stream=None):
        check(lib().prt_intersect_closest(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                          BUFFERS_DEVICE, C.c_void_p(stream) if stream else None))

    def intersect_closest_static(self, rays_ptr, n, hits_ptr, stream=None):
        """Allocation-free closest hit on device buffers (enqueue only; the caller synchronizes)."""
        check(lib().prt_intersect_closest_static(self.handle, C.c_void_p(rays_ptr), n, C.c_void_p(hits_ptr),
                                                 C.c_void_p(stream) if stream else None))

    def intersect_any(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.empty(rays.shape[0], dtype=np.uint8)
        check(lib().prt_intersect_any(self.handle, _ptr(rays), rays.shape[0], _ptr(occ), BUFFERS_HOST, None))
        return occ

    def traversal_counts(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = rays.shape[0]
        nodes, boxes, prims = (np.empty(n, np.uint32) for _ in range(3))
        check(lib().prt_traversal_counts(self.handle, _ptr(rays), n, _ptr(nodes), _ptr(boxes), _ptr(prims),
                                         BUFFERS_HOST, None))
        return nodes, boxes, prims


class ProgressiveRaytracer:
    """Mirror of renderer.ProgressiveRaytracer (pkg/renderer/progressive.go:49-96) on one GPU."""

    def __init__(self, scene, params):
        self.scene = scene
        self.params = params
        h = C.c_void_p()
        check(lib().prt_film_create(scene.handle, C.byref(params), C.byref(h)))
        self.handle = h
        # the film's own size (the params' width/height where set, else the camera's) sizes every read
        fw, fh = C.c_int32(), C.c_int32()
        check(lib().prt_film_size(h, C.byref(fw), C.byref(fh)))
        self.width, self.height = fw.value, fh.value

    def close(self):
        if getattr(self, "handle", None):
            lib().prt_film_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_samples_for_pass(self, pass_number):
        return get_samples_for_pass(self.params, pass_number)

    def render_pass(self, pass_number):
        """RenderPass: returns (rgba uint8 image (H,W,4), PassStats)."""
        st = PassStats()
        check(lib().prt_render_pass(self.handle, pass_number, C.byref(st)))
        return self.image(), st

    def render_samples(self, spp):
        st = PassStats()
        check(lib().prt_render_samples(self.handle, spp, C.byref(st)))
        return st

    def reset(self):
        check(lib().prt_film_reset(self.handle))

    def read(self):
        """Per-pixel mean RGB (H,W,3), luminance sum, luminance^2 sum, counts."""
        n = self.width * self.height
        mean = np.empty((self.height, self.width, 3), np.float32)
        ls = np.empty((self.height, self.width), np.float32)
        lq = np.empty((self.height, self.width), np.float32)
        cnt = np.empty((self.height, self.width), np.uint32)
        check(lib().prt_film_read(self.handle, _ptr(mean), _ptr(ls), _ptr(lq), _ptr(cnt)))
        assert mean.size == 3 * n
        return mean, ls, lq, cnt

    def probe(self, flags=0):
        """prt_film_probe: statistics gathered since the last call (dict); `flags` sets what later renders
        gather (PROBE_COUNT traversal counts, PROBE_TIME per-kernel device time)."""
        st = ProbeStats()
        check(lib().prt_film_probe(self.handle, flags, C.byref(st)))
        return st.as_dict()

    def splats(self):
        """BDPT light-tracing splats accumulated so far (H,W,3); included in read()'s mean."""
        out = np.empty((self.height, self.width, 3), np.float32)
        check(lib().prt_film_read_splats(self.handle, _ptr(out)))
        return out

    def image(self, x0=0, y0=0, x1=None, y1=None):
        x1 = self.width if x1 is None else x1
        y1 = self.height if y1 is None else y1
        out = np.empty((y1 - y0, x1 - x0, 4), np.uint8)
        check(lib().prt_film_read_rgba8(self.handle, x0, y0, x1, y1, _ptr(out)))
        return out


class Server:
    """Device-resident serving loop (include/prt_server.h): continuous batching of progressive
    renders, the GPU counterpart of the web server's per-request pipeline (web/server/render.go)."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, max_active=32):
        h = C.c_void_p()
        check(lib().prt_server_create(device, hbm_budget_bytes, max_active, C.byref(h)))
        self.handle = h
        self._scenes = {}  # request id -> Scene, kept alive until the request is released

    def submit(self, scene, params, snapshots=False):
        """Queues a render; snapshots=True keeps the film of each finished pass (read_pass, sse)."""
        rid = C.c_int64()
        check(lib().prt_server_submit_ex(self.handle, scene.handle, C.byref(params),
                                         SUBMIT_SNAPSHOTS if snapshots else 0, C.byref(rid)))
        self._scenes[rid.value] = scene
        return rid.value

    def read_pass(self, rid, pass_number, width, height):
        mean = np.empty((height, width, 3), np.float32)
        cnt = np.empty((height, width), np.uint32)
        check(lib().prt_server_read_pass(self.handle, rid, pass_number, _ptr(mean), _ptr(cnt)))
        return mean, cnt

    def sse(self, event, params, tile_size=64):
        """The SSE text (bytes) of one event of a snapshots request (include/prt_sse.h)."""
        scene = self._scenes[event.request_id]
        return _take_text(lambda t, n: lib().prt_sse_server_event(self.handle, scene.handle, C.byref(params),
                                                                    tile_size, C.byref(event), t, n))

    def poll(self, max_events=256, timeout_ms=0):
        ev = (ServerEvent * max_events)()
        n = C.c_int32()
        check(lib().prt_server_poll(self.handle, ev, max_events, timeout_ms, C.byref(n)))
        return [ev[i] for i in range(n.value)]

    def cancel(self, rid):
        check(lib().prt_server_cancel(self.handle, rid))

    def read(self, rid, width, height):
        mean = np.empty((height, width, 3), np.float32)
        cnt = np.empty((height, width), np.uint32)
        check(lib().prt_server_read(self.handle, rid, _ptr(mean), _ptr(cnt)))
        return mean, cnt

    def read_rgba8(self, rid, x0, y0, x1, y1):
        out = np.empty((y1 - y0, x1 - x0, 4), np.uint8)
        check(lib().prt_server_read_rgba8(self.handle, rid, x0, y0, x1, y1, _ptr(out)))
        return out

    def release(self, rid):
        check(lib().prt_server_release(self.handle, rid))

    def stats(self):
        st = ServerStats()
        check(lib().prt_server_stats_get(self.handle, C.byref(st)))
        return st

    def close(self):
        if getattr(self, "handle", None):
            lib().prt_server_destroy(self.handle)
            self.handle = None
            self._scenes = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------
def icosphere(frequency, radius=1.0, center=(0.0, 0.0, 0.0), amplitude=0.1, seed=0):
    """Displaced geodesic icosphere: (vertices float64 (n,3), faces int32 (m,3))."""
    nv, nf = C.c_int64(0), C.c_int64(0)
    c = np.asarray(center, np.float64)
    check(lib().prt_workload_icosphere(frequency, radius, _ptr(c), amplitude, seed, None, C.byref(nv), None,
                                       C.byref(nf)))
    v = np.empty((nv.value, 3), np.float64)
    f = np.empty((nf.value, 3), np.int32)
    check(lib().prt_workload_icosphere(frequency, radius, _ptr(c), amplitude, seed, _ptr(v), C.byref(nv), _ptr(f),
                                       C.byref(nf)))
    return v, f


def random_rays(lo, hi, n, scale=2.0, seed=0):
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    rays = np.empty(n, RAY_DTYPE)
    check(lib().prt_workload_rays(_ptr(lo), _ptr(hi), scale, n, seed, _ptr(rays)))
    return rays


def hdr_sky(width=1024, height=512, seed=0, sun_peak=1.0e4):
    rgb = np.empty((height, width, 3), np.float32)
    check(lib().prt_workload_sky(width, height, seed, sun_peak, _ptr(rgb)))
    return rgb


# ---- procedural textures (pkg/material/procedural_textures.go) ----------------------------------
def checkerboard_texture(width, height, check_size, color1, color2):
    """NewCheckerboardTexture (:8-31): color1 where (x / check + y / check) is even."""
    y, x = np.mgrid[0:height, 0:width]
    even = ((x // check_size + y // check_size) % 2) == 0
    return np.where(even[..., None], np.array(color1, float), np.array(color2, float))


def uv_debug_texture(width, height):
    """NewUVDebugTexture (:34-46): (x / (width - 1), y / (height - 1), 0)."""
    y, x = np.mgrid[0:height, 0:width]
    return np.stack([x / (width - 1), y / (height - 1), np.zeros((height, width))], -1)


def gradient_texture(width, height, color1, color2):
    """NewGradientTexture (:49-64): color1 at the top row to color2 at the bottom row."""
    t = (np.arange(height) / (height - 1))[:, None, None]
    row = np.array(color1, float) * (1.0 - t) + np.array(color2, float) * t
    return np.broadcast_to(row, (height, width, 3)).copy()


# ---- SSE stream text and PNG (include/prt_sse.h) ----------------------------------------------
def _take_text(call):
    t, n = C.c_void_p(), C.c_int64()
    check(call(C.byref(t), C.byref(n)))
    if not t.value:
        return b""
    try:
        return C.string_at(t.value, n.value)
    finally:
        lib().prt_free(t)


def png_encode_rgba8(rgba):
    """PNG bytes of an (h, w, 4) uint8 image (opaque: written as 8-bit RGB)."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    return _take_text(lambda t, n: lib().prt_png_encode_rgba8(_ptr(a), a.shape[1], a.shape[0], t, n))


def sse_format_pass(mean_rgb, counts, pass_number, total_passes, target_samples, max_samples_per_pixel,
                    primitive_count, elapsed_ms, tile_size=64):
    """Tile events (twice per tile) and the passComplete event of one pass, as SSE text (bytes)."""
    m = np.ascontiguousarray(mean_rgb, dtype=np.float32)
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    h, w = c.shape
    info = SsePassInfo(pass_number, total_passes, target_samples, max_samples_per_pixel, tile_size, 0,
                       primitive_count, elapsed_ms)
    return _take_text(lambda t, n: lib().prt_sse_format_pass(w, h, _ptr(m), _ptr(c), C.byref(info), t, n))


def sse_format_complete():
    return _take_text(lambda t, n: lib().prt_sse_format_complete(t, n))


def sse_format_error(message):
    return _take_text(lambda t, n: lib().prt_sse_format_error(message.encode(), t, n))
