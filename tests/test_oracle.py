"""Pins the CPU oracle (CPU only, no GPU).

Each check is independent of the oracle's own code path: known answers,
closed forms, a second construction of the same object (numpy lexsort for
the sort, np.unique for the runs, a top-down radix-tree split for the Karras
hierarchy, box unions over leaf ranges for the refit, a numpy float32 brute
force for the traversal) or the reference's own test ray
(core/builders_tests/tests_traversal.cpp:106-107, tolerance :120).
"""
import numpy as np
import pytest

from conftest import VSGF_MESHES, vsgf_path
from heavyrt_amd import workloads as W
from oracle import oracle as O

LEAF = 0x80000000


def interleave3(x: int, y: int, z: int) -> int:
    """Bit-by-bit Morton interleave (independent of the magic-mask spread)."""
    c = 0
    for b in range(10):
        c |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return c


def deinterleave3(c: int):
    x = y = z = 0
    for b in range(10):
        x |= ((c >> (3 * b)) & 1) << b
        y |= ((c >> (3 * b + 1)) & 1) << b
        z |= ((c >> (3 * b + 2)) & 1) << b
    return x, y, z


# ---------------------------------------------------------------------------
# Morton codes (lbvh.cpp:25-51, 259-287)
# ---------------------------------------------------------------------------
def test_zindex_known_answers():
    assert O.zindex3d(1023, 1023, 1023) == 0x3FFFFFFF
    assert O.zindex3d(1, 0, 0) == 1
    assert O.zindex3d(0, 1, 0) == 2
    assert O.zindex3d(0, 0, 1) == 4
    assert O.zindex3d(0, 0, 0) == 0


def test_zindex_matches_bitwise_interleave():
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, size=(500, 3)):
        assert O.zindex3d(int(x), int(y), int(z)) == interleave3(int(x), int(y), int(z))


def test_morton_triangle_known_answer():
    # triangle.vsgf: (-0.5,0,0.5) (0.5,0,0.5) (0,0,-0.5).  Scene box x,z in
    # [-0.5,0.5], y = 0.  One triangle (< 256): centroid rule.  Extended
    # extent 1.0002 -> scale 1023.7952; centroid (0, 0, 0.1666667) lands in
    # cell x = trunc(511.8976) = 511, y = 0, z = trunc(682.53015) = 682.
    v, f = W.load_vsgf(vsgf_path("triangle"))
    b = O.OracleBVH(v, f)
    codes, _, _ = b.codes()
    assert int(codes[0]) == interleave3(511, 0, 682)


@pytest.mark.parametrize("mesh", VSGF_MESHES)
def test_morton_cells_close_to_fp64(mesh):
    """Every code is < 2^30 and decodes to the float64 cell of the triangle
    centre within one cell (the float32 rounding can move a boundary case)."""
    v, f = W.load_vsgf(vsgf_path(mesh))
    b = O.OracleBVH(v, f)
    codes, _, _ = b.codes()
    assert int(codes.max()) < (1 << 30)
    v64 = v.astype(np.float64)
    lo, hi = v64.min(0), v64.max(0)
    ext = (hi + 1e-4) - (lo - 1e-4)
    tri = v64[f]
    if f.shape[0] < 256:
        c = tri.sum(axis=1) / 3.0
    else:
        c = 0.5 * (tri.min(axis=1) + tri.max(axis=1))
    cells = np.floor(1024.0 * (c - lo) / ext).astype(np.int64)
    for k in range(f.shape[0]):
        got = np.array(deinterleave3(int(codes[k])))
        assert np.all(np.abs(got - cells[k]) <= 1), (mesh, k, got, cells[k])


# ---------------------------------------------------------------------------
# sort / runs / Karras / refit
# ---------------------------------------------------------------------------
def _meshes():
    for m in VSGF_MESHES:
        yield (m,) + W.load_vsgf(vsgf_path(m))
    v, f = W.soup(20000, stream=W.STREAM_TEST)
    yield ("soup20k", v, f)
    v, f = W.icosphere(3)
    yield ("icosphere3", v, f)


MESHES = list(_meshes())


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_sort_is_stable_code_then_id(name, v, f):
    b = O.OracleBVH(v, f)
    codes, sc, sid = b.codes()
    order = np.lexsort((np.arange(codes.size), codes))
    np.testing.assert_array_equal(sid, order)
    np.testing.assert_array_equal(sc, codes[order])


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_runs_are_unique_codes(name, v, f):
    b = O.OracleBVH(v, f)
    _, sc, _ = b.codes()
    rs, uc = b.runs()
    u, first = np.unique(sc, return_index=True)
    np.testing.assert_array_equal(uc, u)
    np.testing.assert_array_equal(rs, first)


def radix_tree(ucodes: np.ndarray):
    """Top-down radix tree over sorted unique codes: each node splits its range
    where its first differing bit flips.  Returns {internal index: (left, right)}
    in the reference's numbering (internal k, leaf m -> L-1+m; a left child is
    named by its last key, a right child by its first, lbvh.cpp:435-439)."""
    L = ucodes.size
    out = {}
    stack = [(0, L - 1, 0)]
    while stack:
        a, bnd, idx = stack.pop()
        x = int(ucodes[a]) ^ int(ucodes[bnd])
        bit = 31 - (32 - x.bit_length())  # highest differing bit position
        # split: last key in [a, bnd] whose `bit` equals that of ucodes[a]
        s = a
        while s + 1 <= bnd and ((int(ucodes[s + 1]) >> bit) & 1) == ((int(ucodes[a]) >> bit) & 1):
            s += 1
        left = s if s > a else L - 1 + s
        right = s + 1 if s + 1 < bnd else L - 1 + s + 1
        out[idx] = (left, right)
        if s > a:
            stack.append((a, s, s))
        if s + 1 < bnd:
            stack.append((s + 1, bnd, s + 1))
    return out


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_karras_matches_radix_tree(name, v, f):
    b = O.OracleBVH(v, f)
    if b.leaves < 2:
        pytest.skip("single leaf")
    _, uc = b.runs()
    lr = b.lr_nodes()
    tree = radix_tree(uc)
    assert len(tree) == b.leaves - 1
    for k, (l, r) in tree.items():
        assert (int(lr[k, 3]), int(lr[k, 7])) == (l, r), (name, k)


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_boxes_are_unions_of_triangle_boxes(name, v, f):
    """Leaf boxes and refit boxes equal the union of the triangle boxes below
    them (exact), and leaves pack (start, count) as raytrace_common.h:48-51."""
    b = O.OracleBVH(v, f)
    L = b.leaves
    lr = b.lr_nodes()
    rs, _ = b.runs()
    perm = b.perm()
    tri = v.astype(np.float32)[f.astype(np.int64)]
    tmin, tmax = tri.min(axis=1), tri.max(axis=1)
    ends = np.append(rs[1:], b.ntris)
    lo_leaf = np.zeros((L, 3), np.float32)
    hi_leaf = np.zeros((L, 3), np.float32)
    for m in range(L):
        ids = perm[rs[m]:ends[m]]
        lo_leaf[m] = tmin[ids].min(axis=0)
        hi_leaf[m] = tmax[ids].max(axis=0)
        node = lr[L - 1 + m]
        assert int(node[3]) == (LEAF | ((int(ends[m] - rs[m]) << 24) & 0x7F000000) | int(rs[m]))
        np.testing.assert_array_equal(node[0:3].view(np.float32), lo_leaf[m])
        np.testing.assert_array_equal(node[4:7].view(np.float32), hi_leaf[m])
    if L < 2:
        return
    # internal node k covers a contiguous leaf range; walk the tree
    def rng(k):
        if k >= L - 1:
            m = k - (L - 1)
            return m, m
        a, _ = rng(int(lr[k, 3]))
        _, c = rng(int(lr[k, 7]))
        return a, c
    import sys
    sys.setrecursionlimit(10000)
    for k in range(L - 1):
        a, c = rng(k)
        np.testing.assert_array_equal(lr[k, 0:3].view(np.float32), lo_leaf[a:c + 1].min(axis=0))
        np.testing.assert_array_equal(lr[k, 4:7].view(np.float32), hi_leaf[a:c + 1].max(axis=0))


# ---------------------------------------------------------------------------
# traversal
# ---------------------------------------------------------------------------
def brute_force_np(v, f, rays):
    """numpy float32 Moller-Trumbore over every triangle, operation order of
    BruteForceRT.cpp:241-252, first strictly closer hit wins."""
    v = v.astype(np.float32)
    A, B, Cc = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1, e2 = B - A, Cc - A
    n = rays.shape[0]
    best_t = rays[:, 7].copy()
    best_p = np.full(n, 0xFFFFFFFF, dtype=np.uint64)
    o, d, tn = rays[:, 0:3], rays[:, 4:7], rays[:, 3]
    for k in range(f.shape[0]):
        a, x1, x2 = A[k], e1[k], e2[k]
        px = d[:, 1] * x2[2] - d[:, 2] * x2[1]
        py = d[:, 2] * x2[0] - d[:, 0] * x2[2]
        pz = d[:, 0] * x2[1] - d[:, 1] * x2[0]
        tx, ty, tz = o[:, 0] - a[0], o[:, 1] - a[1], o[:, 2] - a[2]
        qx = ty * x1[2] - tz * x1[1]
        qy = tz * x1[0] - tx * x1[2]
        qz = tx * x1[1] - ty * x1[0]
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.float32(1.0) / (x1[0] * px + x1[1] * py + x1[2] * pz)
            vv = (tx * px + ty * py + tz * pz) * inv
            uu = (qx * d[:, 0] + qy * d[:, 1] + qz * d[:, 2]) * inv
            t = (x2[0] * qx + x2[1] * qy + x2[2] * qz) * inv
            ok = (vv >= np.float32(-1e-6)) & (uu >= np.float32(-1e-6)) & (uu + vv <= np.float32(1.0) + np.float32(1e-6)) \
                & (t > tn) & (t < best_t)
        best_t = np.where(ok, t, best_t)
        best_p = np.where(ok, k, best_p)
    return best_t, best_p.astype(np.uint32)


def rays_around(v, n, stream):
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    r = W.uniform01(W.pcg32(n, 4, stream))
    org = (c + 1.5 * ext * W.unit_vectors(r[:, 0], r[:, 1])).astype(np.float32)
    tgt = lo + (hi - lo) * W.uniform01(W.pcg32(n, 3, stream + 1))
    d = (tgt - org)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = W.FLT_MAX
    return rays


@pytest.mark.parametrize("name,v,f", MESHES[:6], ids=[m[0] for m in MESHES[:6]])
def test_traversal_equals_brute_force(name, v, f):
    rays = rays_around(v, 4000, 100)
    b = O.OracleBVH(v, f)
    t, p, _, _ = b.trace(rays, O.MODE_BVH2COMMON)
    bt, bp = brute_force_np(v, f, rays)
    tb, pb, _, _ = b.trace(rays, O.MODE_BRUTE)
    np.testing.assert_array_equal(tb.view(np.uint32), bt.view(np.uint32))
    np.testing.assert_array_equal(pb, bp)
    # same nearest distance; same triangle unless two triangles tie at that t
    np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    diff = p != bp
    assert diff.sum() <= max(2, rays.shape[0] // 500), f"{diff.sum()} differing ids"


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_peach_and_lr_layouts_agree(name, v, f):
    """The reference traverses its peach layout (BVH2CommonRT) or the raw
    left/right layout (BVH2_LRFT); same tree, same child order -> same hits."""
    rays = rays_around(v, 3000, 200)
    b = O.OracleBVH(v, f)
    t0, p0, u0, v0 = b.trace(rays, O.MODE_BVH2COMMON)
    t1, p1, u1, v1 = b.trace(rays, O.MODE_LR)
    np.testing.assert_array_equal(t0.view(np.uint32), t1.view(np.uint32))
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(u0.view(np.uint32), u1.view(np.uint32))


@pytest.mark.parametrize("name,v,f", MESHES, ids=[m[0] for m in MESHES])
def test_any_hit_is_existence(name, v, f):
    rays = rays_around(v, 3000, 300)
    rays[:, 7] = 0.5 * float(np.max(v.max(0) - v.min(0)))
    b = O.OracleBVH(v, f)
    _, pn, _, _ = b.trace(rays, O.MODE_BRUTE)
    anyr = rays.copy()
    anyr[:, 7] = -anyr[:, 7]
    _, pa, _, _ = b.trace(anyr, O.MODE_BVH2COMMON)
    np.testing.assert_array_equal(pa != 0xFFFFFFFF, pn != 0xFFFFFFFF)


@pytest.mark.parametrize("mesh", VSGF_MESHES)
def test_reference_test_ray(mesh):
    """tests_traversal.cpp:106-107: pos (-0.00264156517, -0.00264156517,
    0.00543617224), dir (-0.203098252, 0.733451784, -0.648690641), tFar=+inf;
    BVH hit t must equal brute force within 1e-5 (:120)."""
    v, f = W.load_vsgf(vsgf_path(mesh))
    b = O.OracleBVH(v, f)
    ray = np.array([[-0.00264156517, -0.00264156517, 0.00543617224, 0.0,
                     -0.203098252, 0.733451784, -0.648690641, np.inf]], dtype=np.float32)
    t, p, _, _ = b.trace(ray, O.MODE_BVH2COMMON)
    tb, pb, _, _ = b.trace(ray, O.MODE_BRUTE)
    if np.isinf(tb[0]):
        assert np.isinf(t[0])
    else:
        assert abs(float(t[0]) - float(tb[0])) <= 1e-5


def test_t_error_against_fp64():
    v, f = W.soup(5000, stream=W.STREAM_TEST + 1)
    rays = W.incoherent_rays(20000, stream=W.STREAM_TEST + 2)
    b = O.OracleBVH(v, f)
    t, p, _, _ = b.trace(rays)
    hit = p != 0xFFFFFFFF
    assert hit.sum() > 100
    t64 = b.t_fp64(rays[hit], p[hit])
    rel = np.abs(t[hit].astype(np.float64) - t64) / np.abs(t64)
    assert rel.max() < 1e-4
    assert np.median(rel) < 1e-6


def test_traversal_counters_are_plausible():
    v, f = W.eye_scene()
    rays = W.eye_rays(64, 64)
    b = O.OracleBVH(v, f)
    t, p, u, vv, pairs, tris = b.trace(rays, counters=True)
    assert (pairs >= 1).all()
    assert ((tris >= 1) | (p == 0xFFFFFFFF)).all()
    assert pairs.max() < 2 * b.leaves


def test_degenerate_single_leaf():
    """Four copies of one triangle share one Morton code: one leaf holding all
    four, root is a leaf (peach layout: leaf + dummy, lbvh.cpp:497-504)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]] * 4, np.uint32)
    b = O.OracleBVH(v, f)
    assert b.leaves == 1 and b.max_run == 4
    pn = b.peach_nodes()
    assert int(pn[0, 3]) == (LEAF | (4 << 24))
    assert int(pn[1, 3]) == LEAF  # dummy
    ray = np.array([[0.2, 0.2, 1.0, 0.0, 0.0, 0.0, -1.0, W.FLT_MAX]], np.float32)
    t, p, _, _ = b.trace(ray)
    assert p[0] == 0 and abs(t[0] - 1.0) < 1e-6


def test_leaf_count_wraps_past_127_like_the_reference():
    """130 equal codes: the 7-bit count field keeps 130 & 0x7F = 2 (raytrace_common.h:48-51)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]] * 130, np.uint32)
    b = O.OracleBVH(v, f)
    assert b.max_run == 130
    assert O.pack_leaf(0, 130) == (LEAF | (2 << 24))


def test_as_shipped_metrics_mode_matches_the_counters():
    """The reference's metrics bookkeeping (the CPU baseline's "as shipped" run) changes no hit, and its
    NC / TC equal twice the pair visits / the triangle tests of the counting traversal."""
    v, f = W.eye_scene()
    rays = W.eye_rays()[::7]
    ob = O.OracleBVH(v, f)
    t, p, m = ob.trace_metrics(rays)
    t2, p2, _, _, pairs, tris = ob.trace(rays, counters=True)
    np.testing.assert_array_equal(t.view(np.uint32), t2.view(np.uint32))
    np.testing.assert_array_equal(p, p2)
    assert m["NC"] == 2 * int(pairs.sum()) and m["TC"] == int(tris.sum())
    assert m["BLB"] == 32 * m["NC"] + 60 * m["TC"]
    assert 0 < m["WSS128"] <= m["NC"]
