"""Serving loop (include/prt_server.h): continuous batching of progressive renders.

Isolation bar: each request's final film is bit-identical whether it ran alone or among other
requests (counter RNG per pixel/sample and per-request chunking make results independent of the mix).
"""
import numpy as np
import pytest

import prtgpu as P
import scenes

pytestmark = pytest.mark.gpu


def _params(w, h, cfg, spp, passes, seed, integrator=P.INTEGRATOR_PATH, amin=1.0):
    return P.sampling_params(w, h, cfg["max_depth"], cfg["rr_min_bounces"], spp, max_passes=passes, seed=seed,
                             integrator=integrator, adaptive_min_samples=amin, adaptive_threshold=0.3)


def _drain(server, ids, timeout_s=60, release=False):
    """Collects events until every request is done; release=True frees each finished request's
    film slot (a finished request holds its slot and its HBM budget until released)."""
    import time

    done, events = set(), {i: [] for i in ids}
    t0 = time.time()
    while len(done) < len(ids):
        for e in server.poll(timeout_ms=50):
            if e.request_id in events:
                events[e.request_id].append((e.type, e.pass_number, e.target_samples))
                if e.type in (P.EVENT_DONE, P.EVENT_ERROR, P.EVENT_CANCELLED):
                    done.add(e.request_id)
                    if release:
                        server.release(e.request_id)
        assert time.time() - t0 < timeout_s, "serving loop did not finish"
    return events


@pytest.fixture(scope="module")
def sc():
    d0, c0 = scenes.default_scene(width=96)
    d2, c2 = scenes.cornell_metre(width=64)
    d3, c3 = scenes.mesh_sky_scene(width=96)  # the benchmark's 1.8M-triangle glass icosphere + HDR sky
    return {"default": (P.Scene(d0), c0), "cornell": (P.Scene(d2), c2), "mesh_sky": (P.Scene(d3), c3)}


BD = P.INTEGRATOR_BDPT


def _requests():
    # (scene, width, height, spp, passes, seed, integrator, AdaptiveMinSamples)
    PT = P.INTEGRATOR_PATH
    return [("default", 96, 54, 24, 4, 1, PT, 1.0), ("cornell", 64, 64, 40, 5, 2, PT, 1.0),
            ("default", 128, 72, 16, 3, 3, PT, 1.0), ("cornell", 80, 48, 20, 4, 4, PT, 1.0),
            ("default", 64, 36, 12, 2, 5, PT, 1.0), ("cornell", 48, 48, 30, 6, 6, PT, 1.0),
            ("cornell", 48, 48, 12, 3, 7, BD, 1.0), ("default", 64, 36, 8, 2, 8, BD, 1.0),
            ("default", 64, 36, 32, 3, 9, PT, 0.15), ("cornell", 40, 40, 16, 2, 10, BD, 0.25),
            ("mesh_sky", 96, 96, 12, 3, 11, PT, 1.0), ("mesh_sky", 64, 64, 6, 2, 12, BD, 1.0)]


def test_isolation_alone_vs_mix(sc):
    reqs = _requests()
    srv = P.Server(max_active=8)
    ids = []
    for name, w, h, spp, passes, seed, integ, amin in reqs:
        scene, cfg = sc[name]
        ids.append(srv.submit(scene, _params(w, h, cfg, spp, passes, seed, integ, amin)))
    ev = _drain(srv, ids)
    mix = [srv.read(i, w, h) for i, (_, w, h, _, _, _, _, _) in zip(ids, reqs)]
    for rid, (name, w, h, spp, passes, seed, integ, amin) in zip(ids, reqs):
        kinds = [k for k, _, _ in ev[rid]]
        assert kinds[0] == P.EVENT_ADMITTED and kinds[-1] == P.EVENT_DONE
        passes_seen = [p for k, p, _ in ev[rid] if k == P.EVENT_PASS_COMPLETE]
        assert passes_seen == list(range(1, passes + 1))
        targets = [t for k, _, t in ev[rid] if k == P.EVENT_PASS_COMPLETE]
        assert targets[0] == 1 and targets[-1] == spp
    for rid in ids:
        srv.release(rid)
    srv.close()
    # each request alone on a fresh server
    for (name, w, h, spp, passes, seed, integ, amin), (m_mix, c_mix) in zip(reqs, mix):
        solo = P.Server(max_active=1)
        scene, cfg = sc[name]
        rid = solo.submit(scene, _params(w, h, cfg, spp, passes, seed, integ, amin))
        _drain(solo, [rid])
        m, c = solo.read(rid, w, h)
        if amin >= 1.0:
            assert np.all(c == spp) and np.all(c_mix == spp)
        else:
            assert np.array_equal(c, c_mix) and c.min() < spp
        assert np.array_equal(m, m_mix), f"{name} {w}x{h} differs between the mix and the solo render"
        solo.close()


def test_resolution_override_matches_native_camera():
    d_native, cfg = scenes.default_scene(width=128)  # 128 x 72 natively
    d_base, _ = scenes.default_scene(width=96)
    srv = P.Server(max_active=2)
    a = srv.submit(P.Scene(d_native), _params(0, 0, cfg, 4, 1, 7))
    b_scene = P.Scene(d_base)
    b = srv.submit(b_scene, _params(128, 72, cfg, 4, 1, 7))
    _drain(srv, [a, b])
    ma, _ = srv.read(a, 128, 72)
    mb, _ = srv.read(b, 128, 72)
    assert np.array_equal(ma, mb)
    srv.close()


def test_cancel_and_budget(sc):
    scene, cfg = sc["cornell"]
    # a 64 MiB budget holds one 64x64 film slot (~47 MB) at a time: requests are admitted one after another
    srv = P.Server(hbm_budget_bytes=64 << 20, max_active=8)
    ids = [srv.submit(scene, _params(64, 64, cfg, 8, 2, s)) for s in range(4)]
    big = srv.submit(scene, _params(64, 64, cfg, 100000, 1000, 9))
    srv.cancel(big)
    ev = _drain(srv, ids + [big], release=True)
    assert ev[big][-1][0] == P.EVENT_CANCELLED
    for i in ids:
        assert ev[i][-1][0] == P.EVENT_DONE
    st = srv.stats()
    assert st.completed == 4 and st.active == 0
    srv.close()


@pytest.mark.parametrize("limit", ["slots", "budget"])
def test_cancel_while_queued_behind_active_render(limit):
    """Requests cancelled (or released) while they wait behind a full slot limit or a full HBM budget
    leave the queue at once: CANCELLED arrives while the active render is still running, and the scene
    they held can be destroyed once everything is released. Cancelling the running render itself stops
    its chunk dispatch: it ends long before its pass would."""
    import time

    d, cfg = scenes.cornell_metre(width=64)
    scene = P.Scene(d)
    long_p = _params(256, 256, cfg, 1 << 20, 2, 21)  # pass 2 = 2^20 spp: minutes of device time
    if limit == "slots":
        srv = P.Server(max_active=1)
    else:  # a budget of 1.5 film slots of this size: the first request fits, the next ones wait
        probe = P.Server(max_active=1)
        r = probe.submit(scene, _params(256, 256, cfg, 1, 1, 20))
        _drain(probe, [r])
        one = probe.stats().used_bytes
        probe.release(r)
        probe.close()
        assert one > 0
        srv = P.Server(hbm_budget_bytes=int(1.5 * one), max_active=8)
    a = srv.submit(scene, long_p)
    while True:  # wait until the long render is admitted and past its first pass
        evs = [e for e in srv.poll(timeout_ms=50) if e.request_id == a]
        if any(e.type == P.EVENT_PASS_COMPLETE for e in evs):
            break
    b = srv.submit(scene, _params(256, 256, cfg, 4, 1, 22))
    c = srv.submit(scene, _params(256, 256, cfg, 4, 1, 23))
    time.sleep(0.05)
    assert srv.stats().queued == 2
    srv.cancel(b)
    srv.release(c)
    got, t0 = {}, time.time()
    while len(got) < 2:
        for e in srv.poll(timeout_ms=50):
            assert e.request_id != a or e.type == P.EVENT_ADMITTED, "the long render ended first"
            if e.request_id in (b, c):
                got[e.request_id] = e.type
        assert time.time() - t0 < 10, "queued requests were not cancelled while the long render ran"
    assert got == {b: P.EVENT_CANCELLED, c: P.EVENT_CANCELLED}
    srv.release(b)
    t1 = time.time()
    srv.cancel(a)
    ev = _drain(srv, [a], timeout_s=20)
    assert ev[a][-1][0] == P.EVENT_CANCELLED and time.time() - t1 < 5
    srv.release(a)
    deadline = time.time() + 5
    while True:  # releases are deferred to the scheduler thread
        try:
            scene.close()
            break
        except P.PrtError:
            assert time.time() < deadline, "scene still held after every request was released"
            time.sleep(0.01)
    srv.close()


def test_scene_in_use_cannot_be_destroyed():
    """A scene referenced by a film or an unreleased request refuses prt_scene_destroy (no dangling
    device pointers); the Python wrappers keep their scenes alive for as long as the C side needs them."""
    d, cfg = scenes.cornell_metre(width=32)
    scene = P.Scene(d)
    pr = P.ProgressiveRaytracer(scene, _params(0, 0, cfg, 2, 1, 1))
    with pytest.raises(P.PrtError):
        scene.close()
    pr.close()
    srv = P.Server(max_active=2)
    rid = srv.submit(scene, _params(0, 0, cfg, 2, 1, 2))
    _drain(srv, [rid])
    with pytest.raises(P.PrtError):
        scene.close()
    srv.close()
    scene.close()


@pytest.mark.parametrize("integrator,size", [(P.INTEGRATOR_PATH, (1600, 1400)), (BD, (1600, 1000))])
def test_multi_chunk_first_pass_matches_solo(sc, integrator, size):
    """Images larger than a first-pass chunk (path tracing: 2^21 paths; BDPT: the 2 GiB subpath bound)
    render their first pass in several whole-image-sized chunks and later passes in small chunks, each
    tracked by its own events. Every pixel gets the same samples as in a solo render, which adds a pass's
    samples in fewer, larger chunks; the film adds each sample with its own compensated step in sample
    order, so the grouping does not change the image: bit-identical."""
    w, h = size
    scene, cfg = sc["cornell"]
    p = _params(w, h, cfg, 3, 2, 11, integrator)
    srv = P.Server(max_active=2)
    rid = srv.submit(scene, p)
    ev = _drain(srv, [rid])[rid]
    assert [t for k, _, t in ev if k == P.EVENT_PASS_COMPLETE] == [1, 3]
    m, c = srv.read(rid, w, h)
    srv.release(rid)
    srv.close()
    pr = P.ProgressiveRaytracer(scene, p)
    pr.render_pass(1)
    pr.render_pass(2)
    m2, _, _, c2 = pr.read()
    pr.close()
    assert np.all(c == 3) and np.array_equal(c, c2)
    err = np.abs(m.astype(np.float64) - m2) / np.maximum(np.abs(m2.astype(np.float64)), 1e-3)
    print(f"{w}x{h} integrator {integrator}: max relative difference {err.max():.3g}, "
          f"identical pixels {np.all(m == m2, axis=2).mean():.5f}, mean {m2.mean():.6f}")
    assert m2.mean() > 0 and np.array_equal(m, m2)


def test_concurrent_rgba8_reads_of_a_served_film(sc):
    """prt_server_read_rgba8 from several threads at once (the device conversion's image and staging
    buffers are shared by a film's readers): every read equals a single-threaded one."""
    import threading

    scene, cfg = sc["default"]
    srv = P.Server(max_active=2)
    rid = srv.submit(scene, _params(320, 180, cfg, 4, 2, 12))
    _drain(srv, [rid])
    ref = srv.read_rgba8(rid, 0, 0, 320, 180)
    outs = [None] * 8

    def work(k):
        outs[k] = srv.read_rgba8(rid, 0, 0, 320, 180) if k % 2 == 0 else srv.read_rgba8(rid, 10, 20, 200, 150)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert ref[..., 3].min() == 255 and ref[..., :3].max() > 0
    for k, o in enumerate(outs):
        assert np.array_equal(o, ref if k % 2 == 0 else ref[20:150, 10:200]), k
    srv.release(rid)
    srv.close()


def test_reused_film_buffers_give_the_same_films(sc):
    """Released films leave their device buffers and streams to the server, and films of later requests
    take them instead of allocating: every film still matches the same request rendered on a fresh
    server bit for bit, whatever the previous owner of its buffers rendered (other scenes, PT / BDPT,
    other sizes)."""
    import time

    PT = P.INTEGRATOR_PATH
    seq = [("mesh_sky", 96, 96, 6, 2, 31, BD), ("cornell", 96, 96, 12, 2, 32, PT), ("default", 96, 96, 8, 2, 33, BD),
           ("cornell", 96, 96, 8, 2, 34, BD), ("mesh_sky", 96, 96, 8, 2, 35, PT), ("default", 80, 80, 8, 2, 36, PT)]
    srv = P.Server(max_active=1)
    for name, w, h, spp, passes, seed, integ in seq:
        scene, cfg = sc[name]
        rid = srv.submit(scene, _params(w, h, cfg, spp, passes, seed, integ))
        _drain(srv, [rid])
        m, c = srv.read(rid, w, h)
        srv.release(rid)
        t0 = time.time()
        while srv.stats().used_bytes != 0:  # the release is carried out by the scheduler thread
            assert time.time() - t0 < 10, "released film not returned"
            time.sleep(0.002)
        assert srv.stats().idle_bytes > 0
        solo = P.Server(max_active=1)
        sid = solo.submit(scene, _params(w, h, cfg, spp, passes, seed, integ))
        _drain(solo, [sid])
        ms, cs = solo.read(sid, w, h)
        solo.close()
        assert np.array_equal(c, cs) and np.all(c == spp)
        assert np.array_equal(m, ms), f"{name} {w}x{h} integrator {integ}: reused buffers changed the film"
    srv.close()
