"""Pins the oracle's instance path (CPU only).

Checks, each independent of the code they check:
- the instance inverse against numpy (M @ M^-1 = I) and exact closed forms;
- instance boxes against exact transforms and vertex containment;
- the top-level tree over instance boxes against the same structural
  invariants as the triangle LBVH (every instance in exactly one leaf, boxes
  are unions);
- the two-level traversal against the per-instance closed form (each
  instance traced alone, smallest t, then smallest instance id), against
  brute force, and against a mesh pre-transformed to world space in float64
  (which pins the direction of the matrix and its transposition).
"""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O

LEAF = 0x80000000


def col_major(M: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).T.reshape(16)).astype(np.float32)


def rot(axis: int, deg: float) -> np.ndarray:
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    R = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


def affine(seed: int) -> np.ndarray:
    r = W.uniform01(W.pcg32(1, 12, 900 + seed))[0]
    A = np.eye(4)
    q = r[0:4] - 0.5
    q /= np.linalg.norm(q)
    x, y, z, w = q
    A[:3, :3] = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                          [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                          [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    A[:3, :3] = A[:3, :3] @ np.diag(0.5 + 2.0 * r[4:7])
    A[:3, 3] = 4.0 * (r[7:10] - 0.5)
    return A


# ---------------------------------------------------------------------------
# inverse and instance boxes
# ---------------------------------------------------------------------------
def test_inverse_closed_forms():
    v, f = W.unit_cube()
    s = O.OracleScene()
    g = s.add_geom(v, f)
    cases = [np.eye(4)]
    T = np.eye(4)
    T[:3, 3] = [3.0, -2.5, 0.25]
    cases.append(T)
    S = np.diag([2.0, 0.5, 4.0, 1.0])
    cases.append(S)
    P = np.eye(4)[[1, 2, 0, 3]]  # axis permutation
    cases.append(P)
    for M in cases:
        s.add_instance(g, col_major(M))
    inv, _ = s.inst_info()
    for k, M in enumerate(cases):
        np.testing.assert_array_equal(inv[k], np.linalg.inv(M)[:3].astype(np.float32), err_msg=str(k))


def test_inverse_matches_numpy_for_general_affine():
    v, f = W.unit_cube()
    s = O.OracleScene()
    g = s.add_geom(v, f)
    Ms = [affine(k) for k in range(32)]
    for M in Ms:
        s.add_instance(g, col_major(M))
    inv, _ = s.inst_info()
    for k, M in enumerate(Ms):
        Mf = col_major(M).reshape(4, 4).T.astype(np.float64)  # the float32 matrix the scene received
        ref = np.linalg.inv(Mf)[:3]
        np.testing.assert_allclose(inv[k], ref, rtol=1e-6, atol=1e-6)


def test_singular_matrix_is_rejected():
    v, f = W.unit_cube()
    s = O.OracleScene()
    g = s.add_geom(v, f)
    with pytest.raises(ValueError):
        s.add_instance(g, col_major(np.diag([1.0, 0.0, 1.0, 1.0])))


def test_instance_boxes():
    v, f = W.unit_cube()
    s = O.OracleScene()
    g = s.add_geom(v, f)
    T = np.eye(4)
    T[:3, 3] = [1.0, 2.0, -3.0]
    R = rot(2, 90.0)
    R[np.abs(R) < 1e-12] = 0.0
    s.add_instance(g, col_major(np.eye(4)))
    s.add_instance(g, col_major(T))
    s.add_instance(g, col_major(R @ np.diag([1.0, 2.0, 4.0, 1.0])))
    Ms = [affine(k) for k in range(16)]
    for M in Ms:
        s.add_instance(g, col_major(M))
    _, box = s.inst_info()
    np.testing.assert_array_equal(box[0], [[-0.5, -0.5, -0.5, 1.0], [0.5, 0.5, 0.5, 1.0]])
    np.testing.assert_array_equal(box[1], [[0.5, 1.5, -3.5, 1.0], [1.5, 2.5, -2.5, 1.0]])
    # 90 degrees about z maps (x, y) -> (-y, x): extents 2 (y) -> x, 1 (x) -> y
    np.testing.assert_array_equal(box[2], [[-1.0, -0.5, -2.0, 1.0], [1.0, 0.5, 2.0, 1.0]])
    for k, M in enumerate(Ms):
        Mf = col_major(M).reshape(4, 4).T.astype(np.float64)
        w = (Mf[:3, :3] @ v.T.astype(np.float64)).T + Mf[:3, 3]
        lo, hi = box[3 + k, 0, :3], box[3 + k, 1, :3]
        assert np.all(w >= lo - 1e-6) and np.all(w <= hi + 1e-6)
        np.testing.assert_allclose(lo, w.min(0), atol=1e-6)
        np.testing.assert_allclose(hi, w.max(0), atol=1e-6)


# ---------------------------------------------------------------------------
# top-level tree
# ---------------------------------------------------------------------------
def instanced(nboxes=300, nx=32, nz=16, extra=True):
    geoms, inst = W.rtao_instanced_scene(nx, nz, nboxes)
    s = O.OracleScene()
    for v, f in geoms:
        s.add_geom(v, f)
    if extra:
        iv, if_ = W.icosphere(2, 0.1, (0.0, 0.0, 0.0))
        gi = s.add_geom(iv, if_)
        geoms = geoms + [(iv, if_)]
        for k in range(8):
            M = affine(40 + k)
            M[:3, 3] = [0.25 * k, 0.3, 0.1 * k]
            inst.append((gi, col_major(M)))
    for g, m in inst:
        s.add_instance(g, m)
    s.commit()
    return s, geoms, inst


def test_tlas_structure():
    s, _, inst = instanced()
    t = s.tlas
    n = len(inst)
    assert t.ntris == n
    perm = t.perm()
    np.testing.assert_array_equal(np.sort(perm), np.arange(n))
    sc = t.codes()[1]
    assert np.all(np.diff(sc.astype(np.int64)) >= 0)
    _, box = s.inst_info()
    lr = t.lr_nodes()
    L = t.leaves
    rs, _ = t.runs()
    ends = np.append(rs[1:], n)
    for m in range(L):
        ids = perm[rs[m]:ends[m]]
        node = lr[L - 1 + m]
        assert int(node[3]) == (LEAF | ((int(ends[m] - rs[m]) << 24) & 0x7F000000) | int(rs[m]))
        np.testing.assert_array_equal(node[0:3].view(np.float32), box[ids, 0, :3].min(0))
        np.testing.assert_array_equal(node[4:7].view(np.float32), box[ids, 1, :3].max(0))
    for k in range(L - 1):
        a, b = lr[int(lr[k, 3])], lr[int(lr[k, 7])]
        np.testing.assert_array_equal(lr[k, 0:3].view(np.float32),
                                      np.minimum(a[0:3].view(np.float32), b[0:3].view(np.float32)))
        np.testing.assert_array_equal(lr[k, 4:7].view(np.float32),
                                      np.maximum(a[4:7].view(np.float32), b[4:7].view(np.float32)))


def test_tlas_keeps_instances_with_equal_codes():
    """Instances sharing a Morton cell stay in one leaf (the reference's LBVH
    top level keeps only the first of them, lbvh.cpp:344-346)."""
    v, f = W.unit_cube()
    s = O.OracleScene()
    g = s.add_geom(v, f)
    for k in range(5):
        s.add_instance(g, col_major(np.eye(4)))
    T = np.eye(4)
    T[:3, 3] = [10.0, 0.0, 0.0]
    s.add_instance(g, col_major(T))
    s.commit()
    assert s.tlas.leaves == 2 and s.tlas.max_run == 5
    rays = np.zeros((6, 8), np.float32)
    rays[:, 0:3] = [[-5, 0, 0], [5, 0, 0], [15, 0, 0], [0, 5, 0], [10, 5, 0], [3, 3, 3]]
    rays[:, 4:7] = [[1, 0, 0], [-1, 0, 0], [-1, 0, 0], [0, -1, 0], [0, -1, 0], [1, 0, 0]]
    rays[:, 7] = W.FLT_MAX
    t, p, i, g_, _, _ = s.trace(rays)
    # the five coincident instances tie at equal t: the smallest id wins
    assert list(i[:4]) == [0, 0, 5, 0] and i[4] == 5 and i[5] == 0xFFFFFFFF
    np.testing.assert_array_equal(t[:5], np.float32([4.5, 4.5, 4.5, 4.5, 4.5]))


# ---------------------------------------------------------------------------
# two-level traversal
# ---------------------------------------------------------------------------
def scene_rays(n, stream, lo=(-0.2, -0.2, -0.2), hi=(2.2, 0.6, 1.2), tfar=W.FLT_MAX):
    r = W.uniform01(W.pcg32(n, 7, stream))
    lo, hi = np.array(lo), np.array(hi)
    org = lo + (hi - lo) * r[:, 0:3]
    tgt = lo + (hi - lo) * r[:, 3:6]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = tfar
    return rays


def test_two_level_equals_per_instance_closed_form():
    s, _, _ = instanced()
    for stream, tfar in ((300, W.FLT_MAX), (310, 0.3)):
        rays = scene_rays(6000, stream, tfar=tfar)
        a = s.trace(rays, O.SCENE_TWO_LEVEL)
        b = s.trace(rays, O.SCENE_PER_INSTANCE)
        hit = a[1] != 0xFFFFFFFF
        assert hit.mean() > 0.15
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))


def test_two_level_any_hit_is_existence():
    s, _, _ = instanced()
    rays = scene_rays(6000, 320, tfar=0.2)
    near = s.trace(rays, O.SCENE_TWO_LEVEL)
    ra = rays.copy()
    ra[:, 7] = -ra[:, 7]
    anyh = s.trace(ra, O.SCENE_TWO_LEVEL)
    np.testing.assert_array_equal(anyh[1] != 0xFFFFFFFF, near[1] != 0xFFFFFFFF)


def test_two_level_equals_brute_force():
    s, _, _ = instanced(nboxes=60, nx=8, nz=4)
    rays = scene_rays(3000, 330)
    a = s.trace(rays, O.SCENE_TWO_LEVEL)
    b = s.trace(rays, O.SCENE_BRUTE)
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    np.testing.assert_array_equal(a[2], b[2])
    same = a[1] == b[1]
    assert same.mean() > 0.999, same.mean()


def test_instance_equals_pretransformed_mesh():
    """A mesh placed by an instance matrix and the same mesh transformed to
    world space in float64 give the same triangles at nearly the same t."""
    iv, if_ = W.icosphere(3, 1.0, (0.0, 0.0, 0.0))
    M = affine(7)
    s = O.OracleScene()
    g = s.add_geom(iv, if_)
    s.add_instance(g, col_major(M))
    s.commit()
    Mf = col_major(M).reshape(4, 4).T.astype(np.float64)
    wv = ((Mf[:3, :3] @ iv.T.astype(np.float64)).T + Mf[:3, 3]).astype(np.float32)
    w = O.OracleBVH(wv, if_)
    c = wv.mean(0)
    rays = scene_rays(4000, 340, lo=c - 4.0, hi=c + 4.0)
    a = s.trace(rays)
    t, p, _, _ = w.trace(rays)
    hit = p != 0xFFFFFFFF
    assert hit.mean() > 0.1
    assert np.mean(a[1] == p) > 0.995
    both = hit & (a[1] == p)
    np.testing.assert_allclose(a[0][both], t[both], rtol=1e-4, atol=1e-6)


def test_identity_instance_matches_single_geometry():
    v, f = W.icosphere(3, 1.0, (0.0, 1.0, 0.0))
    s = O.OracleScene()
    g = s.add_geom(v, f)
    s.add_instance(g, col_major(np.eye(4)))
    s.commit()
    b = O.OracleBVH(v, f)
    rays = scene_rays(4000, 350, lo=(-3, -2, -3), hi=(3, 4, 3))
    a = s.trace(rays, O.SCENE_TWO_LEVEL, counters=True)
    t, p, u, vv, pairs, tris = b.trace(rays, O.MODE_BVH2COMMON, counters=True)
    np.testing.assert_array_equal(a[0].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(a[1], p)
    # one extra pair visit for the top level (its root is a leaf); a ray that
    # misses the instance box never enters the object's tree, which the
    # single-geometry walk rejects at its root pair
    entered = a[6] == pairs + 1
    assert np.all(entered | ((a[6] == 1) & (pairs == 1)))
    assert entered.mean() > 0.3
    np.testing.assert_array_equal(a[7], tris)


def test_scene_without_instances_misses():
    v, f = W.unit_cube()
    s = O.OracleScene()
    s.add_geom(v, f)
    s.commit()
    rays = scene_rays(10, 360)
    t, p, i, g, _, _ = s.trace(rays)
    assert np.all(p == 0xFFFFFFFF) and np.all(i == 0xFFFFFFFF)
    np.testing.assert_array_equal(t, rays[:, 7])
