"""The oracle's AO ray generator pinned without retyping it: PCG32 against the
published pcg-c-basic known answer, and every ray against what the geometry
fixes -- the hit point and the viewer-facing normal recomputed in float64, a
cosine-weighted lobe whose polar angle gives back the second draw and whose
azimuth turns by 2*pi times the difference of the first draws about the normal
(kernel_InitAORay, examples/02_rtao/rtao.cpp:147-165; MapSampleToCosineDistribution
and OffsRayPos, examples/render_common.h:56-112)."""
import numpy as np

from heavyrt_amd import workloads as W
from oracle import oracle as O


def test_pcg32_known_answer():
    # pcg32-demo output for pcg32_srandom_r(&rng, 42u, 54u)
    ref = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    np.testing.assert_array_equal(O.pcg32(42, 54, 6), np.array(ref, dtype=np.uint32))


def _frame_case(w=48, h=28):
    """A scene, primary rays that see both geometry and sky, and their compact hits from the oracle."""
    v, f = W.rtao_scene(64, 32, 40)
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    eye = c + np.array([0.2, 0.35, 1.1]) * ext
    prim = W.pinhole_rays(w, h, eye=tuple(eye), target=tuple(c + np.array([0.0, 0.15 * ext, 0.0])),
                          fov_y_deg=60.0)
    t, pid, _, _ = O.OracleBVH(v, f).trace(prim)
    hits = np.zeros((prim.shape[0], 2), dtype=np.uint32)
    hits[:, 0] = t.view(np.uint32)
    hits[:, 1] = pid
    return v, f, prim, hits


def _normals64(v, f, prim, hits):
    """Viewer-facing unit geometric normal and hit point of every hit pixel, in float64."""
    live = hits[:, 1] != 0xFFFFFFFF
    tri = f.reshape(-1, 3)[hits[live, 1]]
    A, B, C = (v[tri[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(B - A, C - A)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = prim[live, 4:7].astype(np.float64)
    n[np.sum(n * d, axis=1) > 0] *= -1.0
    t = hits[live, 0].view(np.float32).astype(np.float64)
    hp = prim[live, 0:3].astype(np.float64) + (0.9999995 * t)[:, None] * d
    return live, n, hp


def test_ao_rays_geometry():
    v, f, prim, hits = _frame_case()
    spp, tmax, seed = 8, 0.75, 0x853C49E6748FEA9B
    rays = O.ao_rays(v, f, prim, hits, spp, tmax, seed).reshape(prim.shape[0], spp, 8)
    live, n, hp = _normals64(v, f, prim, hits)
    assert live.any() and not live.all()
    # a missed pixel gets the empty ray (0,0,0,0),(0,1,0,0)
    empty = np.array([0, 0, 0, 0, 0, 1, 0, 0], dtype=np.float32)
    assert (rays[~live] == empty).all()
    r = rays[live].astype(np.float64)
    np.testing.assert_array_equal(r[..., 3], 0.0)
    np.testing.assert_array_equal(rays[live][..., 7], np.float32(tmax))
    # one origin per pixel, pushed off the hit point along the normal by epsilonOfPos(hit)
    assert (rays[live][:, :, 0:3] == rays[live][:, :1, 0:3]).all()
    m = np.maximum(np.max(np.abs(hp), axis=1), 4e-5) * 2e-5
    off = r[:, 0, 0:3] - hp
    np.testing.assert_allclose(np.sum(off * n, axis=1), m, rtol=0.02, atol=1e-6)
    lat = off - np.sum(off * n, axis=1, keepdims=True) * n
    # the float32 hit point carries a few ulps of the larger of |origin| and |hit point|
    scale = np.maximum(np.max(np.abs(prim[live, 0:3]), axis=1), np.max(np.abs(hp), axis=1))
    assert np.max(np.linalg.norm(lat, axis=1) / scale) < 1e-6
    # unit directions in the normal's hemisphere
    d = r[..., 4:7]
    np.testing.assert_allclose(np.linalg.norm(d, axis=2), 1.0, atol=2e-6)
    cos = np.sum(d * n[:, None, :], axis=2)
    assert cos.min() > -1e-6


def test_ao_rays_give_back_the_draws():
    """The polar angle returns the second PCG32 draw, the azimuth the first: the
    perpendicular parts of two samples of one pixel are turned, right-handed
    about the normal, by 2*pi times the difference of their first draws."""
    v, f, prim, hits = _frame_case(24, 16)
    spp, tmax, seed = 4, 2.0, 977
    rays = O.ao_rays(v, f, prim, hits, spp, tmax, seed).reshape(prim.shape[0], spp, 8)
    live, n, _ = _normals64(v, f, prim, hits)
    pix = np.nonzero(live)[0]
    d = rays[live][..., 4:7].astype(np.float64)
    cos = np.sum(d * n[:, None, :], axis=2)
    draws = np.array([[O.pcg32(seed, p * spp + s, 2) for s in range(spp)] for p in pix], dtype=np.uint64)
    u = (draws >> np.uint64(8)).astype(np.float64) / 16777216.0
    r1, r2 = u[..., 0], u[..., 1]
    np.testing.assert_allclose(cos, np.sqrt(1.0 - r2), atol=3e-6)
    perp = d - cos[..., None] * n[:, None, :]
    big = np.linalg.norm(perp, axis=2) > 0.05  # the azimuth is well defined
    for s in range(1, spp):
        ok = big[:, 0] & big[:, s]
        a, b = perp[ok, 0], perp[ok, s]
        ang = np.arctan2(np.sum(np.cross(a, b) * n[ok], axis=1), np.sum(a * b, axis=1))
        want = 2.0 * np.pi * (r1[ok, s] - r1[ok, 0])
        diff = np.angle(np.exp(1j * (ang - want)))
        assert np.max(np.abs(diff)) < 1e-4


def test_ao_rays_cosine_lobe_moments():
    """Cosine-weighted sampling: E[cos] = 2/3, E[cos^2] = 1/2, the perpendicular part averages to zero."""
    v, f, prim, hits = _frame_case(24, 16)
    spp = 8192
    rays = O.ao_rays(v, f, prim, hits, spp, 1.0, 12345).reshape(prim.shape[0], spp, 8)
    live, n, _ = _normals64(v, f, prim, hits)
    d = rays[live][..., 4:7].astype(np.float64)
    cos = np.sum(d * n[:, None, :], axis=2).ravel()
    k = cos.size
    assert k > 100000
    assert abs(cos.mean() - 2.0 / 3.0) < 5 * np.sqrt(1.0 / 18.0 / k)
    assert abs(np.mean(cos ** 2) - 0.5) < 5 * np.sqrt(1.0 / 12.0 / k)
    perp = (d - np.sum(d * n[:, None, :], axis=2)[..., None] * n[:, None, :]).reshape(-1, 3)
    assert np.all(np.abs(perp.mean(axis=0)) < 5 * np.sqrt(0.5 / k))


def test_ao_rays_depend_on_the_seed():
    """The rays are a function of (pixel, sample, seed) only: the same call gives the same bits,
    and another seed moves every live direction."""
    v, f, prim, hits = _frame_case(12, 10)
    a = O.ao_rays(v, f, prim, hits, 6, 1.0, 99).reshape(-1, 6, 8)
    b = O.ao_rays(v, f, prim, hits, 6, 1.0, 99).reshape(-1, 6, 8)
    np.testing.assert_array_equal(a, b)
    c = O.ao_rays(v, f, prim, hits, 6, 1.0, 100).reshape(-1, 6, 8)
    live = hits[:, 1] != 0xFFFFFFFF
    assert (a[live][..., 4:7] != c[live][..., 4:7]).any(axis=2).all()
