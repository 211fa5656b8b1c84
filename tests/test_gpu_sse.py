"""Per-pass film snapshots of the serving loop and the SSE stream built from them: the snapshot of pass p
of a request equals the film of a solo render after pass p bit for bit (PT and BDPT with splats),
while the server has already moved on to later passes and other requests share the device; the
stream's tiles and passComplete statistics are those of that pass."""
import base64
import json

import numpy as np
import pytest

import prtgpu as P
import scenes
from test_sse import events, gamma8, png_decode

pytestmark = pytest.mark.gpu


def _params(desc, integrator, passes=4, spp=24, seed=5):
    w, h = desc.image_size
    return P.sampling_params(w, h, 5, 16, spp, max_passes=passes, seed=seed, integrator=integrator)


def _solo_passes(scene, params):
    pr = P.ProgressiveRaytracer(scene, params)
    out = {}
    for k in range(1, params.max_passes + 1):
        pr.render_pass(k)
        mean, _, _, cnt = pr.read()
        out[k] = (mean.copy(), cnt.copy())
    return out


@pytest.mark.parametrize("integrator", [P.INTEGRATOR_PATH, P.INTEGRATOR_BDPT])
def test_snapshots_and_stream_match_solo_passes(integrator):
    desc, _ = scenes.cornell_scene(width=96)
    scene = P.Scene(desc)
    info = scene.info()
    w, h = desc.image_size
    params = _params(desc, integrator)
    solo = _solo_passes(scene, params)
    srv = P.Server(max_active=4)
    # a second, longer request shares the device
    other = srv.submit(scene, _params(desc, P.INTEGRATOR_PATH, passes=3, spp=64, seed=9))
    rid = srv.submit(scene, params, snapshots=True)
    seen, done = [], False
    for _ in range(4000):
        for ev in srv.poll(64, timeout_ms=50):
            if ev.request_id != rid:
                continue
            if ev.type == P.EVENT_PASS_COMPLETE:
                p = ev.pass_number
                mean, cnt = srv.read_pass(rid, p, w, h)
                assert np.array_equal(mean, solo[p][0]) and np.array_equal(cnt, solo[p][1]), p
                text = srv.sse(ev, params)
                evs = events(text)
                ntiles = ((w + 63) // 64) * ((h + 63) // 64)
                assert [e for e, _ in evs] == ["tile"] * (2 * ntiles) + ["passComplete"]
                img = gamma8(solo[p][0])
                d = json.loads(evs[0][1])
                assert d["passNumber"] == p and d["totalPasses"] == params.max_passes
                assert np.array_equal(png_decode(base64.b64decode(d["imageData"])), img[:64, :64])
                pc = json.loads(evs[-1][1])
                assert pc["passNumber"] == p and pc["maxSamples"] == ev.target_samples
                assert pc["totalSamples"] == int(solo[p][1].sum()) == ev.target_samples * w * h
                assert pc["primitiveCount"] == info.reference_primitives
                seen.append(p)
            elif ev.type == P.EVENT_DONE:
                assert srv.sse(ev, params) == b"event: complete\ndata: Rendering completed\n\n"
                done = True
            elif ev.type in (P.EVENT_ERROR, P.EVENT_CANCELLED):
                pytest.fail(f"request ended with event {ev.type}")
        if done:
            break
    assert done and seen == [1, 2, 3, 4]
    # the three newest passes are kept; older ones and unknown passes are refused
    m2, _ = srv.read_pass(rid, 2, w, h)
    assert np.array_equal(m2, solo[2][0])
    with pytest.raises(P.PrtError):
        srv.read_pass(rid, 1, w, h)
    with pytest.raises(P.PrtError):
        srv.read_pass(rid, 7, w, h)
    with pytest.raises(P.PrtError):  # submitted without snapshots
        srv.read_pass(other, 1, w, h)
    srv.cancel(other)
    srv.release(other)
    srv.release(rid)
    srv.close()
