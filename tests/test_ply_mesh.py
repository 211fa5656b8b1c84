"""PLY loading (loaders.LoadPLY, pkg/loaders/ply.go) and the mesh options it feeds
(TriangleMeshOptions, pkg/geometry/triangle_mesh.go:20-199). The loader is host code in libprt; the
PLY files are written here the way createTestPLY writes them (ply_test.go:14-92). CPU only."""
import math
import struct

import numpy as np
import pytest

import oracle as O
import prtgpu as P


def write_ply(path, normals=False, colors=False, extra_face_prop=False, fmt="binary_little_endian",
              count_type="uchar", index_type="int", quad_face=False):
    head = ["ply", f"format {fmt} 1.0", "element vertex 4", "property float x", "property float y",
            "property float z"]
    if normals:
        head += ["property float nx", "property float ny", "property float nz"]
    if colors:
        head += ["property uchar red", "property uchar green", "property uchar blue"]
    head += ["element face 2", f"property list {count_type} {index_type} vertex_indices"]
    if extra_face_prop:
        head += ["property uchar flags"]
    head += ["end_header"]
    b = ("\n".join(head) + "\n").encode()
    verts = [(0, 0, 0, 255, 0, 0), (1, 0, 0, 0, 255, 0), (1, 1, 0, 0, 0, 255), (0, 1, 0, 255, 255, 0)]
    for x, y, z, r, g, bl in verts:
        b += struct.pack("<3f", x, y, z)
        if normals:
            b += struct.pack("<3f", 0, 0, 1)
        if colors:
            b += struct.pack("<3B", r, g, bl)
    faces = [(0, 1, 2), (0, 2, 3)] if not quad_face else [(0, 1, 2, 3), (0, 2, 3, 1)]
    for f in faces:
        b += struct.pack("<B" if count_type == "uchar" else "<i", len(f))
        b += struct.pack("<%d%s" % (len(f), "i" if index_type == "int" else "I"), *f)
        if extra_face_prop:
            b += struct.pack("<B", 7)
    path.write_bytes(b)
    return path


def test_basic(tmp_path):  # TestLoadPLY_Basic
    d = P.load_ply(write_ply(tmp_path / "b.ply"))
    assert np.array_equal(d["vertices"], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    assert np.array_equal(d["faces"], [[0, 1, 2], [0, 2, 3]])
    assert "normals" not in d and "colors" not in d


def test_normals_and_colors(tmp_path):  # TestLoadPLY_WithNormals / WithColors
    d = P.load_ply(write_ply(tmp_path / "n.ply", normals=True, colors=True))
    assert np.array_equal(d["normals"], [[0, 0, 1]] * 4)
    assert np.array_equal(d["colors"], [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]])


def test_int_counts_uint_indices_and_skipped_face_props(tmp_path):
    d = P.load_ply(write_ply(tmp_path / "i.ply", count_type="int", index_type="uint"))
    assert np.array_equal(d["faces"], [[0, 1, 2], [0, 2, 3]])
    d = P.load_ply(write_ply(tmp_path / "e.ply", extra_face_prop=True))
    assert np.array_equal(d["faces"], [[0, 1, 2], [0, 2, 3]])


@pytest.mark.parametrize("kw", [dict(fmt="ascii"), dict(fmt="binary_big_endian"), dict(quad_face=True)])
def test_rejected(tmp_path, kw):
    with pytest.raises(P.PrtError):
        P.load_ply(write_ply(tmp_path / "r.ply", **kw))


def test_missing_file(tmp_path):  # TestLoadPLY_NonExistentFile
    with pytest.raises(P.PrtError):
        P.load_ply(tmp_path / "nonexistent.ply")


def _mesh_scene(**opts):
    s = P.SceneDesc()
    m = s.lambertian((0.5, 0.5, 0.5))
    v = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], float)
    s.mesh(v, [[0, 1, 2]], m, **opts)
    s.set_camera(center=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=40.0)
    return s


def _hit(s, o, d):
    os_ = O.OracleScene(s)
    rays = np.zeros(1, P.RAY_DTYPE)
    rays["org"], rays["dir"], rays["tmin"], rays["tmax"] = o, d, 0.001, np.inf
    return os_.intersect(rays)


def test_rotation_about_center():
    """rotateVertex: a 90-degree Y rotation about (0,0,0) turns the z=0 triangle into the x=0 plane; a
    ray along -x now hits it at x=0 (t = 5) and a ray along -z misses."""
    s = _mesh_scene(rotation=(0.0, math.pi / 2, 0.0), center=(0.0, 0.0, 0.0))
    t, prim, _, _ = _hit(s, (5.0, -0.5, 0.0), (-1.0, 0.0, 0.0))
    assert prim[0] == 0 and abs(t[0] - 5.0) < 1e-9
    t, prim, _, _ = _hit(s, (0.3, -0.5, 5.0), (0.0, 0.0, -1.0))
    assert prim[0] == -1


def test_custom_normal_orientation_on_device_desc():
    """Custom triangle normals reach the scene description (normalized on use, triangle.go:35-49)."""
    s = _mesh_scene(normals=[[0, 0, -2.0]])
    d, keep = s.to_c()
    assert d.meshes[0].normals[2] == -2.0 and not d.meshes[0].rotation


def _emissive_triangle(normals=None):
    s = P.SceneDesc()
    e = s.emissive((2.0, 2.0, 2.0))
    v = np.array([[-3, -3, 0], [3, -3, 0], [0, 3, 0]], float)
    s.mesh(v, [[0, 1, 2]], e, normals=normals)
    s.set_camera(center=(0, 0, 5), look_at=(0, 0, 0), up=(0, 1, 0), width=8, aspect_ratio=1.0, vfov=20.0)
    return s


def test_custom_normal_decides_the_front_face():
    """Emissive emits on the front face only (emissive.go:30-37). The winding normal (+z) faces the
    camera, so the centre pixel sees the emission; a custom normal -z makes the same triangle a back
    face and the pixel black."""
    for normals, expect in ((None, 2.0), ([[0, 0, -1.0]], 0.0)):
        s = _emissive_triangle(normals)
        os_ = O.OracleScene(s)
        p = P.sampling_params(8, 8, 4, 4, 1, seed=0)
        out = os_.render_samples(p, [[4, 4]], [0])
        assert np.allclose(out[0], expect), (normals, out)
