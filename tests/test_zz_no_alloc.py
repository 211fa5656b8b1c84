"""prt_intersect_closest_static, the allocation-free closest hit: its hits equal prt_intersect_closest's
(device buffers) bit for bit, over two calls in a row that share the static scratch, the second on a
smaller batch, on a triangle-only scene (after the autotuner has picked a variant) and on a mixed
scene; the output beyond the batch is left untouched; bad arguments fail with an error."""
import numpy as np
import pytest

import devmem as D
import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu


def _hits_equal(a, b):
    return np.array_equal(a.view(np.uint32).reshape(-1, 4), b.view(np.uint32).reshape(-1, 4))


def _run(scene, rays, n_tune=0):
    n = rays.shape[0]
    d_rays = D.DeviceBuffer.from_array(rays)
    ref = D.DeviceBuffer(n * 16)
    if n_tune:  # a batch of >= 2^20 rays makes prt_intersect_closest pick the scene's kernel variant
        scene.intersect_closest_device(d_rays.ptr, n_tune, ref.ptr)
        assert scene.info().trace_variant >= 0
    scene.intersect_closest_device(d_rays.ptr, n, ref.ptr)
    want = ref.to_array(P.HIT_DTYPE, n)
    sentinel = np.full(n * 4, 0x7FC0DEAD, np.uint32)
    stream = D.Stream()
    for m in (n, n // 3 + 17):  # second call smaller: stale counter or stack contents would show
        out = D.DeviceBuffer.from_array(sentinel)
        scene.intersect_closest_static(d_rays.ptr, m, out.ptr, stream.handle)
        stream.synchronize()
        got = out.to_array(P.HIT_DTYPE, n)
        assert _hits_equal(got[:m], want[:m]), f"batch of {m}: hits differ from prt_intersect_closest"
        assert np.array_equal(got[m:].view(np.uint32).ravel(), sentinel[4 * m:]), "wrote beyond the batch"
        out.free()
    return want


def test_static_closest_hit_matches_triangle_scene():
    desc, info = scenes.mesh_scene(frequency=128)  # 327,680 triangles
    scene = P.Scene(desc)
    v = info["vertices"]
    rays = P.random_rays(v.min(0), v.max(0), (1 << 20) + 4096, seed=5)
    want = _run(scene, rays, n_tune=rays.shape[0])
    assert (want["prim_id"] >= 0).mean() > 0.05


def test_static_closest_hit_matches_mixed_scene():
    desc, cfg = scenes.default_scene(width=64)
    scene = P.Scene(desc)
    rays = P.random_rays(np.array([-3.0, 0.0, -3.0]), np.array([3.0, 2.0, 1.0]), 200_000, seed=6)
    want = _run(scene, rays)
    assert (want["prim_id"] >= 0).mean() > 0.05


def test_static_closest_hit_rejects_bad_arguments():
    desc, _ = scenes.mesh_scene(frequency=8)
    scene = P.Scene(desc)
    buf = D.DeviceBuffer(1024)
    with pytest.raises(P.PrtError):
        scene.intersect_closest_static(buf.ptr, -1, buf.ptr)
    with pytest.raises(P.PrtError):
        P.check(P.lib().prt_intersect_closest_static(scene.handle, None, 4, buf.ptr, None))
    scene.intersect_closest_static(buf.ptr, 0, buf.ptr)  # empty batch: nothing to do
