"""The SSE event text of the progressive render stream (include/prt_sse.h, the reference's
web/server/render.go): PNG tiles decode to the film's gamma-2 RGBA8 pixels, every tile is sent twice
per pass in grid order, the JSON carries the reference's fields in its order with Go's number
formatting, and passComplete holds RenderStats and the image's average luminance. Host code only."""
import base64
import json
import re
import struct
import zlib

import numpy as np
import pytest

import prtgpu as P


def png_decode(data):
    """Minimal PNG reader (8-bit RGB/RGBA, non-interlaced) -> (h, w, 3) uint8."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert depth == 8 and interlace == 0 and ctype in (2, 6)
            bpp = 3 if ctype == 2 else 4
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = w * bpp
    out = np.zeros((h, stride), np.int64)
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)], np.uint8).astype(np.int64)
        cur = np.zeros(stride, np.int64)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 255
        out[y] = cur
        prev = cur
    return out.reshape(h, w, bpp)[..., :3].astype(np.uint8)


def gamma8(mean):
    """vec3ToColor (progressive.go:482-495) in float64: uint8(255 * clamp(sqrt(c))); a negative
    channel's NaN survives Go's min/max and converts to 0 on amd64."""
    m = mean.astype(np.float64)
    v = np.where(m > 0, np.clip(np.sqrt(np.maximum(m, 0.0)), 0.0, 1.0), 0.0)
    return (255.0 * v).astype(np.uint8)


def events(text):
    out = []
    for block in text.decode().split("\n\n"):
        if not block:
            continue
        lines = block.split("\n")
        assert lines[0].startswith("event: ") and lines[1].startswith("data: ") and len(lines) == 2
        out.append((lines[0][7:], lines[1][6:]))
    return out


@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (64, 64), (37, 130)])
def test_png_round_trip(shape):
    rng = np.random.default_rng(shape[0] * 131 + shape[1])
    img = np.zeros(shape + (4,), np.uint8)
    img[..., :3] = rng.integers(0, 256, shape + (3,))
    img[: shape[0] // 2, :, :3] = np.arange(shape[1], dtype=np.uint8)[None, :, None]  # smooth rows too
    img[..., 3] = 255
    png = P.png_encode_rgba8(img)
    assert np.array_equal(png_decode(png), img[..., :3])


def _film(h, w, seed):
    rng = np.random.default_rng(seed)
    mean = rng.uniform(0, 1.3, (h, w, 3)).astype(np.float32) ** 2
    mean[0, 0] = 0
    mean[0, 1] = [4.0, -1.0, 1e-9]  # clamping both ways
    counts = rng.integers(3, 9, (h, w)).astype(np.uint32)
    return mean, counts


def test_pass_stream_tiles_and_stats():
    h, w, ts = 100, 150, 64
    mean, counts = _film(h, w, 1)
    text = P.sse_format_pass(mean, counts, pass_number=3, total_passes=7, target_samples=8,
                             max_samples_per_pixel=50, primitive_count=1234, elapsed_ms=87, tile_size=ts)
    ev = events(text)
    tx, ty = 3, 2
    assert [e for e, _ in ev] == ["tile"] * (2 * tx * ty) + ["passComplete"]
    img = gamma8(mean)
    for k, (_, data) in enumerate(ev[:-1]):
        d = json.loads(data)
        assert list(d) == ["tileX", "tileY", "imageData", "passNumber", "tileNumber", "totalTiles", "totalPasses"]
        i, j = (k % (tx * ty)) % tx, (k % (tx * ty)) // tx
        assert (d["tileX"], d["tileY"]) == (i, j)
        assert d["tileNumber"] == k + 1 and d["totalTiles"] == tx * ty
        assert d["passNumber"] == 3 and d["totalPasses"] == 7
        tile = png_decode(base64.b64decode(d["imageData"]))
        assert np.array_equal(tile, img[j * ts:(j + 1) * ts, i * ts:(i + 1) * ts])
    pc = json.loads(ev[-1][1])
    assert list(pc) == ["event", "passNumber", "totalPasses", "elapsedMs", "totalPixels", "totalSamples",
                        "averageSamples", "maxSamples", "minSamples", "maxSamplesUsed", "primitiveCount",
                        "averageLuminance"]
    assert pc["event"] == "passComplete" and pc["passNumber"] == 3 and pc["elapsedMs"] == 87
    assert pc["totalPixels"] == h * w and pc["totalSamples"] == int(counts.sum())
    assert pc["averageSamples"] == counts.sum() / (h * w)
    assert pc["maxSamples"] == 8 and pc["minSamples"] == counts.min() and pc["maxSamplesUsed"] == counts.max()
    assert pc["primitiveCount"] == 1234
    # CalculateAverageLuminance (stats.go:59-92): RGBA()/65535 of 8-bit values (= v/255 exactly),
    # summed pixel by pixel in row order; JSON's shortest digits round-trip, so the match is exact
    rgb = img.astype(np.float64) / 255.0
    per = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    total = 0.0
    for v in per.ravel().tolist():
        total += v
    assert pc["averageLuminance"] == total / (h * w)


def test_go_number_formatting():
    """encoding/json writes 4.0 as 4, 2.5 as 2.5 and a constant image's luminance without exponent."""
    mean = np.zeros((2, 2, 3), np.float32)
    counts = np.full((2, 2), 4, np.uint32)
    text = P.sse_format_pass(mean, counts, 1, 1, 4, 4, 0, 0, tile_size=64)
    raw = events(text)[-1][1]
    assert '"averageSamples":4,' in raw and '"averageLuminance":0}' in raw and '"minSamples":4,' in raw
    counts[0, 0] = 6
    raw = events(P.sse_format_pass(mean, counts, 1, 1, 4, 4, 0, 0, tile_size=64))[-1][1]
    assert '"averageSamples":4.5,' in raw and '"maxSamplesUsed":6,' in raw
    mean[:] = 1e-14  # uint8(255*sqrt(1e-14)) = 0 -> luminance 0
    mean[0, 0] = 1.0 / 255.0 ** 2 * 1.0001  # one pixel at level 1
    raw = events(P.sse_format_pass(mean, counts, 1, 1, 4, 4, 0, 0, tile_size=64))[-1][1]
    m = re.search(r'"averageLuminance":([^}]*)}', raw).group(1)
    assert "e" not in m and abs(float(m) - 1.0 / 255.0 / 4) < 1e-17


def test_complete_and_error_events():
    assert P.sse_format_complete() == b"event: complete\ndata: Rendering completed\n\n"
    assert P.sse_format_error("Rendering failed: x\ny") == b"event: error\ndata: Rendering failed: x y\n\n"
