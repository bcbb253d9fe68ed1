"""Pins the oracle's BVH4 collapse, FP16 box rounding and BVH4 traversal (CPU only).

FP16: float->half rounding is checked against numpy's IEEE binary16
conversion (round to nearest even) over many bit patterns, and the outward
step against the known answers of half.hpp's nextafter (cbvh_fat.cpp:447-448).
Collapse: every BVH4 child box equals the BVH2 node box it came from, every
triangle sits in exactly one reachable leaf, and the node count equals the
number of even-depth BVH2 internal nodes.  Traversal: FP32 and FP16 BVH4 give
the same (t, primId) as a brute force under the same (t, primId) rule.
Slab tests: both forms (the reference's (b - o) * inv and the engine's fused
fma(b, inv, -(o * inv)) of its FP16 any-hit calls) equal the correctly rounded
results of exact rational arithmetic, and any-hit answers do not depend on
the form.
"""
from fractions import Fraction
import numpy as np
import pytest

from conftest import VSGF_MESHES, vsgf_path
from heavyrt_amd import workloads as W
from oracle import oracle as O

LEAF = 0x80000000


def test_f2h_matches_numpy_binary16():
    rng = np.random.default_rng(7)
    bits = np.concatenate([
        rng.integers(0, 2**32, size=200000, dtype=np.uint64).astype(np.uint32),
        np.arange(0x33000000, 0x33100000, 97, dtype=np.uint32),   # around the smallest subnormals
        np.arange(0x387F0000, 0x38810000, 3, dtype=np.uint32),    # subnormal/normal boundary
        np.arange(0x477FE000, 0x47800100, 1, dtype=np.uint32),    # overflow boundary
    ])
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    want = vals.astype(np.float16).view(np.uint16)
    got = np.array([O.f2h(float(v)) for v in vals[:60000]], dtype=np.uint16)
    np.testing.assert_array_equal(got, want[:60000])
    sel = vals[60000:]
    got2 = np.array([O.f2h(float(v)) for v in sel], dtype=np.uint16)
    np.testing.assert_array_equal(got2, sel.astype(np.float16).view(np.uint16))


def test_outward_rounding_known_answers():
    # half.hpp nextafter toward -/+ max after RNE (cbvh_fat.cpp:447-448)
    assert (O.round_to_neg(0.0), O.round_to_pos(0.0)) == (0x8001, 0x0001)
    assert (O.round_to_neg(1.0), O.round_to_pos(1.0)) == (0x3BFF, 0x3C01)
    assert (O.round_to_neg(0.1), O.round_to_pos(0.1)) == (0x2E65, 0x2E67)
    assert (O.round_to_neg(-1.0), O.round_to_pos(-1.0)) == (0xBC01, 0xBBFF)
    # not conservative beyond the half range: both stay at +-65504
    for x in (65504.0, 65510.0, 1e5, 1e9):
        assert O.round_to_pos(x) == 0x7BFF
    for x in (-65504.0, -65510.0, -1e5, -1e9):
        assert O.round_to_neg(x) == 0xFBFF


def test_outward_step_equals_numpy_nextafter_on_every_half():
    """For every non-NaN binary16 value (both zeros and infinities included), the outward step equals
    numpy's IEEE nextafter toward -65504 and toward +65504 (a different implementation)."""
    bits = np.arange(0x10000, dtype=np.uint32).astype(np.uint16)
    h = bits.view(np.float16)
    keep = ~np.isnan(h)
    want_neg = np.nextafter(h[keep], np.float16(-65504.0)).view(np.uint16)
    want_pos = np.nextafter(h[keep], np.float16(65504.0)).view(np.uint16)
    vals = h[keep].astype(np.float32)
    got_neg = np.array([O.round_to_neg(float(x)) for x in vals], dtype=np.uint16)
    got_pos = np.array([O.round_to_pos(float(x)) for x in vals], dtype=np.uint16)
    # a float that is exactly a half rounds to that half first; -0.0 and +0.0 stay distinct
    np.testing.assert_array_equal(got_neg, want_neg)
    np.testing.assert_array_equal(got_pos, want_pos)


def test_outward_rounding_is_strictly_conservative_in_range():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-60000, 60000, 5000),
                         rng.uniform(-1e-5, 1e-5, 5000)]).astype(np.float32)
    for x in xs:
        lo = O.h2f(O.round_to_neg(float(x)))
        hi = O.h2f(O.round_to_pos(float(x)))
        assert lo < x < hi, (x, lo, hi)


def _meshes():
    for m in ("cube", "cone_low", "helix_mid"):
        yield (m,) + W.load_vsgf(vsgf_path(m))
    v, f = W.soup(20000, stream=W.STREAM_TEST)
    yield ("soup20k", v, f)
    v, f = W.icosphere(3)
    yield ("icosphere3", v, f)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    yield ("single", v, np.array([[0, 1, 2]], np.uint32))


MESHES = list(_meshes())


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_collapse_structure(name, v, f):
    b = O.OracleBVH(v, f)
    q = b.quads()
    lr = b.lr_nodes()
    L = b.leaves
    if L > 1:
        # even-depth internal nodes
        depth = {0: 0}
        stack = [0]
        while stack:
            k = stack.pop()
            for c in (int(lr[k, 3]), int(lr[k, 7])):
                if c < L - 1:
                    depth[c] = depth[k] + 1
                    stack.append(c)
        assert q.shape[0] == sum(1 for d in depth.values() if d % 2 == 0)
    # walk from quad 0: each triangle exactly once, boxes from the BVH2 nodes
    seen = np.zeros(b.ntris, np.int64)
    lr_boxes = {tuple(r[[0, 1, 2, 4, 5, 6]]) for r in lr}
    stack = [0]
    visited = 0
    while stack:
        k = stack.pop()
        visited += 1
        for c in q[k]:
            ref = int(c[3])
            empty = c[0].view(np.float32) == np.float32(1e9)
            if empty:
                assert ref == LEAF and c[4].view(np.float32) == np.float32(-1e9)
                continue
            if ref & LEAF:
                s, n = ref & 0xFFFFFF, (ref >> 24) & 0x7F
                seen[s:s + n] += 1
                if n:
                    assert tuple(c[[0, 1, 2, 4, 5, 6]]) in lr_boxes
            else:
                assert tuple(c[[0, 1, 2, 4, 5, 6]]) in lr_boxes
                stack.append(ref)
    assert visited == q.shape[0]
    assert (seen == 1).all()


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_fp16_boxes_contain_fp32_boxes(name, v, f):
    b = O.OracleBVH(v, f)
    q32 = b.quads(False)
    q16 = b.quads(True)
    lo32, hi32 = q32[..., 0:3].view(np.float32), q32[..., 4:7].view(np.float32)
    lo16, hi16 = q16[..., 0:3].view(np.float32), q16[..., 4:7].view(np.float32)
    real = lo32[..., 0] < 1e8
    assert (lo16[real] < lo32[real]).all() and (hi16[real] > hi32[real]).all()
    np.testing.assert_array_equal(q16[..., 3], q32[..., 3])


def _rays(v, n, stream):
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    r = W.uniform01(W.pcg32(n, 2, stream))
    org = (c + 1.5 * ext * W.unit_vectors(r[:, 0], r[:, 1])).astype(np.float32)
    tgt = lo + (hi - lo) * W.uniform01(W.pcg32(n, 3, stream + 1))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = W.FLT_MAX
    return rays


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_bvh4_fp32_fp16_brute_agree(name, v, f):
    rays = _rays(v, 3000, 1100)
    b = O.OracleBVH(v, f)
    t32, p32, u32, v32 = b.trace4(rays, O.MODE4_FP32)
    t16, p16, u16, v16 = b.trace4(rays, O.MODE4_FP16)
    tb, pb, ub, vb = b.trace4(rays, O.MODE4_BRUTE)
    for a, c in ((t32, tb), (t16, tb), (u32, ub), (u16, ub)):
        np.testing.assert_array_equal(a.view(np.uint32), c.view(np.uint32))
    np.testing.assert_array_equal(p32, pb)
    np.testing.assert_array_equal(p16, pb)
    # the reference's BVH2 order gives the same distance (ids may differ only on exact ties)
    t2, p2, _, _ = b.trace(rays)
    np.testing.assert_array_equal(t2.view(np.uint32), tb.view(np.uint32))


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_bvh4_any_hit_matches_bvh2(name, v, f):
    rays = _rays(v, 3000, 1200)
    rays[:, 7] = -0.4 * float(np.max(v.max(0) - v.min(0)))
    b = O.OracleBVH(v, f)
    _, p2, _, _ = b.trace(rays)
    _, p32, _, _ = b.trace4(rays, O.MODE4_FP32)
    _, p16, _, _ = b.trace4(rays, O.MODE4_FP16)
    np.testing.assert_array_equal(p32 != 0xFFFFFFFF, p2 != 0xFFFFFFFF)
    np.testing.assert_array_equal(p16 != 0xFFFFFFFF, p2 != 0xFFFFFFFF)


def _rn32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32, ties to even (exact: the candidates are compared as rationals)."""
    f = np.float32(float(x))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        e = abs(Fraction(float(c)) - x)
        if best is None or e < best[0] or (e == best[0] and int(np.float32(c).view(np.uint32)) % 2 == 0):
            best = (e, np.float32(c))
    return best[1]


def _slab_exact(o, inv, lo, hi, fused):
    ts = []
    for a in range(3):
        ia = Fraction(float(inv[a]))
        if fused:
            noi = -_rn32(Fraction(float(o[a])) * ia)
            t0 = _rn32(Fraction(float(lo[a])) * ia + Fraction(float(noi)))
            t1 = _rn32(Fraction(float(hi[a])) * ia + Fraction(float(noi)))
        else:
            t0 = _rn32(ia * Fraction(float(_rn32(Fraction(float(lo[a])) - Fraction(float(o[a]))))))
            t1 = _rn32(ia * Fraction(float(_rn32(Fraction(float(hi[a])) - Fraction(float(o[a]))))))
        ts.append((min(t0, t1), max(t0, t1)))
    tmin = max(max(ts[0][0], ts[1][0]), ts[2][0])
    tmax = min(min(ts[0][1], ts[1][1]), ts[2][1])
    return np.array([tmin, tmax], np.float32)


def test_slab_forms_equal_exact_rounding():
    rng = np.random.default_rng(31)
    for k in range(400):
        scale = 10.0 ** rng.integers(-3, 4)
        o = (rng.standard_normal(3) * scale).astype(np.float32)
        d = rng.standard_normal(3).astype(np.float32)
        inv = O.safe_inverse(d)
        lo = (rng.standard_normal(3) * scale).astype(np.float32)
        hi = (lo + np.abs(rng.standard_normal(3)) * scale).astype(np.float32)
        for fused, fn in ((False, O.slab), (True, O.slab_fused)):
            got = fn(o, inv, lo, hi)
            want = _slab_exact(o, inv, lo, hi, fused)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"{k} fused={fused}")
    # an unused slot's inverted FP16 box: the min/max form turns it into a box, the fused one
    # (near/far bound by the sign of inv) finds it empty
    for k in range(200):
        o = rng.standard_normal(3).astype(np.float32)
        inv = O.safe_inverse(rng.standard_normal(3).astype(np.float32))
        lo = np.full(3, 65504.0, np.float32)
        hi = np.full(3, -65504.0, np.float32)
        a, b = O.slab(o, inv, lo, hi)
        c, d = O.slab_fused(o, inv, lo, hi)
        assert a <= b and not (c <= d)
    # the two forms do round differently somewhere (else the fused restatement pins nothing)
    diffs = 0
    for k in range(400):
        o = rng.standard_normal(3).astype(np.float32) * np.float32(50.0)
        inv = O.safe_inverse(rng.standard_normal(3).astype(np.float32))
        lo = rng.standard_normal(3).astype(np.float32)
        hi = lo + np.float32(0.5)
        diffs += int(np.any(O.slab(o, inv, lo, hi) != O.slab_fused(o, inv, lo, hi)))
    assert diffs > 0


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_fused_box_test_keeps_any_hit_answers(name, v, f):
    rays = _rays(v, 3000, 1300)
    rays[:, 7] = -0.4 * float(np.max(v.max(0) - v.min(0)))
    b = O.OracleBVH(v, f)
    _, pb, _, _ = b.trace4(rays, O.MODE4_BRUTE)
    _, p16, _, _, n16, _ = b.trace4(rays, O.MODE4_FP16, counters=True)
    _, pf, _, _, nf, _ = b.trace4(rays, O.MODE4_FP16_FUSED, counters=True)
    np.testing.assert_array_equal(pf != 0xFFFFFFFF, pb != 0xFFFFFFFF)
    np.testing.assert_array_equal(p16 != 0xFFFFFFFF, pb != 0xFFFFFFFF)
    # the fused form prunes at most a few boxes differently
    assert abs(int(nf.sum()) - int(n16.sum())) <= 0.01 * int(n16.sum())


def far_axis_rays(n: int = 2000):
    """Rays straight down (d.x = d.z = 0) over a sphere centred at x = 500: -(o * inv)
    overflows on x and z, where safe_inverse gives +-1e36."""
    v, f = W.icosphere(3, 10.0, (500.0, 0.0, 0.0))
    r = W.uniform01(W.pcg32(n, 2, 1400))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = 488.0 + 24.0 * r[:, 0]
    rays[:, 1] = 30.0
    rays[:, 2] = -12.0 + 24.0 * r[:, 1]
    rays[:, 5] = -1.0
    rays[:, 7] = -100.0  # any-hit
    return v, f, rays


def test_fused_box_test_falls_back_where_o_inv_overflows():
    v, f, rays = far_axis_rays()
    inv = O.safe_inverse(rays[0, 4:7])
    o = rays[0, 0:3]
    lo = np.array([o[0] - 5.0, -10.0, o[2] - 5.0], np.float32)
    hi = np.array([o[0] + 5.0, 10.0, o[2] + 5.0], np.float32)
    a = O.slab(o, inv, lo, hi)
    assert a[0] <= a[1]  # the ray passes through the box
    b = O.slab_fused(o, inv, lo, hi)
    assert not (b[0] <= b[1])  # the raw fused form loses it: -(o * inv) is -inf
    bvh = O.OracleBVH(v, f)
    _, pb, _, _ = bvh.trace4(rays, O.MODE4_BRUTE)
    _, pf, _, _ = bvh.trace4(rays, O.MODE4_FP16_FUSED)
    hit = pb != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 1.0
    np.testing.assert_array_equal(pf != 0xFFFFFFFF, hit)
