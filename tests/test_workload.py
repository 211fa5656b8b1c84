"""Synthetic workload generators: pinned by SHA-256 of the generated buffers (they have no
reference counterpart) and by structural properties. CPU only."""
import hashlib

import numpy as np
import pytest

import prtgpu as P

# SHA-256 over vertices (float64) then faces (int32); rays as RAY_DTYPE records; sky as float32
PINNED = {
    "ico8": "880df073556a12a1514a79d283799dece724a82c0499fa1af9dbbb45f5d3cce2",
    "ico256": "c44c8bb28b383e9e6749771d850bc01565f6e80dfbf2af092e3e5a3779474f55",
    "ico300_c": "46101f342c70320b1f4834552e151df0106be2db78889b7cd3d62478412fba80",
    "rays1M": "9bf090efacb763da99e3da7eada6e9037c09cf22131b5649c487cf97f1af508a",
    "sky": "d825c92c8acb3eddd962871e0e8a229131ac40c9d35868ef424dc007e0b760b5",
}


def sha(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


@pytest.mark.parametrize("f", [1, 2, 8, 32])
def test_icosphere_counts_and_watertight(f):
    v, fc = P.icosphere(f)
    assert v.shape == (10 * f * f + 2, 3) and fc.shape == (20 * f * f, 3)
    e = np.sort(np.concatenate([fc[:, [0, 1]], fc[:, [1, 2]], fc[:, [2, 0]]]), axis=1)
    _, c = np.unique(e, axis=0, return_counts=True)
    assert np.all(c == 2)  # closed 2-manifold: every edge in exactly two faces
    assert np.all(v.astype(np.float32).astype(np.float64) == v)  # exactly representable in float32
    r = np.linalg.norm(v, axis=1)
    assert r.min() >= 0.9 - 1e-6 and r.max() <= 1.1 + 1e-6
    # consistent outward winding: signed volume positive
    a, b, c3 = v[fc[:, 0]], v[fc[:, 1]], v[fc[:, 2]]
    assert np.einsum("ij,ij->i", a, np.cross(b, c3)).sum() > 0


def test_icosphere_pinned():
    v, f = P.icosphere(8)
    assert sha(v, f) == PINNED["ico8"]
    v, f = P.icosphere(256)
    assert f.shape[0] == 1310720
    assert sha(v, f) == PINNED["ico256"]
    v, f = P.icosphere(300, center=(0, 1.15, 0))
    assert f.shape[0] == 1800000
    assert sha(v, f) == PINNED["ico300_c"]


def test_rays_pinned_and_distribution():
    v, f = P.icosphere(256)
    lo, hi = v.min(0), v.max(0)
    r = P.random_rays(lo, hi, 1 << 20)
    assert sha(r) == PINNED["rays1M"]
    d = r["dir"].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(np.abs(d.mean(0)) < 5e-3)
    c, e = (lo + hi) / 2, (hi - lo)
    o = r["org"].astype(np.float64)
    assert np.all(o >= c - e - 1e-6) and np.all(o <= c + e + 1e-6)
    assert np.all(r["tmin"] == np.float32(0.001)) and np.all(np.isinf(r["tmax"]))
    # ray i depends only on (seed, i)
    r2 = P.random_rays(lo, hi, 1000)
    assert np.array_equal(r2, r[:1000])


def test_sky_pinned():
    s = P.hdr_sky(1024, 512, 0, 1.0e4)
    assert sha(s) == PINNED["sky"]
    assert s[..., 1].max() == np.float32(1.0e4)
    assert s.min() > 0
