"""ctypes wrapper of oracle/liboracle.so — the float64 CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, d, i32, i64, u32, u64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
        sig = {
            "orc_scene_create": (vp, [vp]),
            "orc_scene_destroy": (None, [vp]),
            "orc_scene_world": (None, [vp, vp, vp, vp]),
            "orc_intersect": (None, [vp, vp, i64, vp, vp, vp, vp, i32, i32, vp]),
            "orc_render": (C.c_int, [vp, vp, i32, i32, i32, i32, i32, i32, vp, vp, vp, i32, vp, vp]),
            "orc_render_samples": (C.c_int, [vp, vp, vp, vp, i64, vp, i32, vp]),
            "orc_render_passes": (C.c_int, [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp]),
            "orc_bdpt_mis": (d, [vp, vp, vp, i32, vp, i32, vp, i32, i32]),
            "orc_camera_ray_pdfs": (None, [vp, vp, vp, vp, vp, vp, vp]),
            "orc_draw": (d, [u64, u32, u32, u32]),
            "orc_sample_cosine_hemisphere": (None, [vp, vp, vp]),
            "orc_sample_on_unit_sphere": (None, [vp, vp]),
            "orc_sample_point_in_unit_disk": (None, [vp, vp]),
            "orc_sample_point_in_unit_sphere": (None, [vp, vp]),
            "orc_power_heuristic": (d, [i32, d, i32, d]),
            "orc_reflectance": (d, [d, d]),
            "orc_shape_hit": (i32, [i32, vp, vp, vp, d, d, vp]),
            "orc_quad_bbox": (None, [vp, vp]),
            "orc_disc_frame": (None, [vp, vp]),
            "orc_axial_bbox": (None, [i32, vp, vp]),
            "orc_texture_eval": (None, [vp, i32, i32, d, d, vp]),
            "orc_light_sample_emission": (None, [vp, i32, vp, vp, vp]),
            "orc_light_pdf_le": (None, [vp, i32, vp, vp, vp]),
            "orc_aabb_hit": (i32, [vp, vp, vp, d, d]),
            "orc_camera_get_ray": (None, [vp, i32, i32, vp, vp, vp, vp]),
            "orc_camera_height": (i32, [vp]),
            "orc_camera_info": (None, [vp, vp]),
            "orc_camera_sample_from_point": (i32, [vp, vp, vp, vp]),
            "orc_camera_importance": (d, [vp, vp, vp]),
            "orc_light_sample": (None, [vp, i32, vp, vp, vp, vp]),
            "orc_light_pdf": (d, [vp, i32, vp, vp, vp]),
            "orc_material_scatter": (i32, [vp, i32, vp, vp, vp, vp, i32, u32, vp]),
            "orc_material_brdf": (None, [vp, i32, vp, vp, vp, i32, i32, vp]),
            "orc_material_pdf": (d, [vp, i32, vp, vp, vp, vp]),
            "orc_ray_color": (None, [vp, vp, vp, vp, u32, vp]),
            "orc_samples_for_pass": (i32, [i32, i32, i32, i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


ORC_VERTEX = np.dtype([("p", "<f8", 3), ("n", "<f8", 3), ("mat", "<i4"), ("light_ref", "<i4"),
                       ("light_index", "<i4"), ("flags", "<i4"), ("pdf_fwd", "<f8"), ("pdf_rev", "<f8")])
V_LIGHT, V_CAMERA, V_SPECULAR, V_INFINITE = 1, 2, 4, 8


def bdpt_mis(scene, pool, camera_path, light_path, sampled, s, t):
    """MIS weight of strategy (s, t) for hand-built vertex arrays of ORC_VERTEX records."""
    cp = np.ascontiguousarray(camera_path, dtype=ORC_VERTEX)
    lp = np.ascontiguousarray(light_path, dtype=ORC_VERTEX)
    sv = None if sampled is None else np.ascontiguousarray(sampled, dtype=ORC_VERTEX).reshape(1)
    return lib().orc_bdpt_mis(scene.h, pool.h if pool is not None else None, _p(cp) if len(cp) else None, len(cp),
                              _p(lp) if len(lp) else None, len(lp), _p(sv) if sv is not None else None, s, t)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def vec(*x):
    return np.ascontiguousarray(np.array(x, dtype=np.float64).reshape(-1))


class OracleScene:
    """The oracle's float64 scene built from the same prt_scene_desc as the device path."""

    def __init__(self, desc):
        d, keep = desc.to_c()
        self._keep = keep
        self.h = lib().orc_scene_create(C.byref(d))
        self.desc = desc

    def __del__(self):
        try:
            if self.h:
                lib().orc_scene_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def world(self):
        c = np.zeros(3)
        r = C.c_double()
        n = C.c_int64()
        lib().orc_scene_world(self.h, _p(c), C.byref(r), C.byref(n))
        return c, r.value, n.value

    def intersect(self, rays, brute=False, nthreads=0, count=False):
        n = rays.shape[0]
        t = np.empty(n)
        prim = np.empty(n, np.int64)
        u = np.empty(n)
        v = np.empty(n)
        cnt = np.zeros(2, np.uint64)
        lib().orc_intersect(self.h, _p(rays), n, _p(t), _p(prim), _p(u), _p(v), 1 if brute else 0, nthreads,
                            _p(cnt) if count else None)
        if count:
            return t, prim, u, v, cnt
        return t, prim, u, v

    def render(self, params, spp, sample_offset=0, rect=None, nthreads=0, splats=False, squares=False):
        """Per-pixel sums over samples [offset, offset+spp): (sum_rgb (H,W,3), lum_sum, lum_sq[, splat_rgb]
        [, sq_rgb]); sq_rgb (squares=True) holds the per-channel sums of squared camera-side samples."""
        W, H = self.desc.image_size
        x0, y0, x1, y1 = rect if rect else (0, 0, W, H)
        h, w = y1 - y0, x1 - x0
        s = np.zeros((h, w, 3))
        ls = np.zeros((h, w))
        lq = np.zeros((h, w))
        sp = np.zeros((H, W, 3)) if splats else None
        sq = np.zeros((h, w, 3)) if squares else None
        rc = lib().orc_render(self.h, C.byref(params), spp, sample_offset, x0, y0, x1, y1, _p(s), _p(ls), _p(lq),
                              nthreads, _p(sp) if splats else None, _p(sq) if squares else None)
        assert rc == 0, rc
        out = (s, ls, lq) + ((sp,) if splats else ()) + ((sq,) if squares else ())
        return out

    def new_film(self):
        """Per-pixel state of orc_render_passes (sums, luminance moments, counts, splats)."""
        W, H = self.desc.image_size
        return {"sum": np.zeros((H, W, 3)), "lum": np.zeros((H, W)), "lum2": np.zeros((H, W)),
                "count": np.zeros((H, W), np.int32), "splat": np.zeros((H, W, 3))}

    def render_passes(self, params, first_pass, last_pass, film=None, nthreads=0):
        """RenderPass for passes first..last with the adaptive per-pixel stop; updates and returns `film`."""
        film = film if film is not None else self.new_film()
        rc = lib().orc_render_passes(self.h, C.byref(params), first_pass, last_pass, nthreads, _p(film["sum"]),
                                     _p(film["lum"]), _p(film["lum2"]), _p(film["count"]), _p(film["splat"]))
        assert rc == 0, rc
        return film

    def render_samples(self, params, pixels, samples, nthreads=0, splats=False):
        pixels = np.ascontiguousarray(pixels, dtype=np.int32).reshape(-1, 2)
        samples = np.ascontiguousarray(samples, dtype=np.int32).reshape(-1)
        out = np.empty((pixels.shape[0], 3))
        W, H = self.desc.image_size
        sp = np.zeros((H, W, 3)) if splats else None
        lib().orc_render_samples(self.h, C.byref(params), _p(pixels), _p(samples), pixels.shape[0], _p(out),
                                 nthreads, _p(sp) if splats else None)
        if splats:
            return out, sp
        return out
