"""The oracle's agglomerative hierarchy (PLOC) for the any-hit tree: the rule restated independently in Python
on small inputs, the tree invariants, and query answers that do not depend on the tree."""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_oracle import brute_force_np, rays_around

INVALID = 0xFFFFFFFF


def ploc_python(lo, hi, radius):
    """The rule, written again: clusters in order, nearest by union half area (ties to the smaller j),
    mutual pairs merge at the left position, ids count down from L - 2. Returns {id: (left, right)} in
    left/right numbering (leaves L-1+m)."""
    L = lo.shape[0]
    cl = [(L - 1 + m, lo[m].astype(np.float32), hi[m].astype(np.float32)) for m in range(L)]
    nodes = {}
    next_id = L - 2
    while len(cl) > 1:
        n = len(cl)
        nn = []
        for i in range(n):
            best, bj = None, None
            for j in range(max(0, i - radius), min(n - 1, i + radius) + 1):
                if j == i:
                    continue
                ulo = np.minimum(cl[i][1], cl[j][1])
                uhi = np.maximum(cl[i][2], cl[j][2])
                e = (uhi - ulo).astype(np.float32)
                a = np.float32(np.float32(e[0] * e[1]) + np.float32(e[1] * e[2])) + np.float32(e[2] * e[0])
                if bj is None or a < best:
                    best, bj = a, j
            nn.append(bj)
        out, merged = [], 0
        for i in range(n):
            j = nn[i]
            mutual = nn[j] == i
            if mutual and i > j:
                continue
            if not mutual:
                out.append(cl[i])
                continue
            nid = next_id - merged
            merged += 1
            nodes[nid] = (cl[i][0], cl[j][0])
            out.append((nid, np.minimum(cl[i][1], cl[j][1]), np.maximum(cl[i][2], cl[j][2])))
        next_id -= merged
        cl = out
    return nodes


def depth_of(lr, L):
    depth = np.zeros(2 * L - 1, dtype=np.int64)
    for i in range(L - 1):  # parents have smaller ids than their children
        for c in (int(lr[i, 3]), int(lr[i, 7])):
            depth[c] = depth[i] + 1
    return depth


SMALL = [("soup300", lambda: W.soup(300, stream=W.STREAM_TEST)), ("icosphere2", lambda: W.icosphere(2)),
         ("heightfield", lambda: W.heightfield(12, 9)[:2])]


@pytest.mark.parametrize("name,gen", SMALL, ids=[s[0] for s in SMALL])
@pytest.mark.parametrize("radius", [1, 3, 8])
def test_ploc_matches_the_rule(name, gen, radius):
    v, f = gen()
    b = O.OracleBVH(v, f, cubic=True)
    L = b.leaves
    lr0 = b.lr_nodes()
    lo = lr0[L - 1:, 0:3].view(np.float32)
    hi = lr0[L - 1:, 4:7].view(np.float32)
    want = ploc_python(lo, hi, radius)
    b.ploc(radius)
    lr = b.lr_nodes()
    assert len(want) == L - 1
    for k in range(L - 1):
        assert (int(lr[k, 3]), int(lr[k, 7])) == want[k]
    # leaves untouched
    np.testing.assert_array_equal(lr[L - 1:], lr0[L - 1:])


@pytest.mark.parametrize("name,gen", SMALL + [("soup20k", lambda: W.soup(20_000, stream=W.STREAM_TEST))],
                         ids=[s[0] for s in SMALL] + ["soup20k"])
def test_ploc_tree_invariants(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f, cubic=True)
    b.ploc(8)
    L = b.leaves
    lr = b.lr_nodes()
    kids = np.concatenate([lr[:L - 1, 3], lr[:L - 1, 7]]).astype(np.int64)
    # every node but the root is exactly one node's child; parents come first
    assert np.array_equal(np.sort(kids), np.arange(1, 2 * L - 1))
    for k in range(L - 1):
        assert lr[k, 3] > k and lr[k, 7] > k
        a, c = lr[int(lr[k, 3])], lr[int(lr[k, 7])]
        np.testing.assert_array_equal(lr[k, 0:3].view(np.float32),
                                      np.minimum(a[0:3].view(np.float32), c[0:3].view(np.float32)))
        np.testing.assert_array_equal(lr[k, 4:7].view(np.float32),
                                      np.maximum(a[4:7].view(np.float32), c[4:7].view(np.float32)))
    assert depth_of(lr, L).max() <= 40


def test_ploc_answers_do_not_depend_on_the_tree():
    v, f = W.heightfield(64, 32)[:2]
    rays = rays_around(v, 4000, 4700)
    ref = O.OracleBVH(v, f)
    t, p, _, _ = ref.trace(rays)
    b = O.OracleBVH(v, f, cubic=True)
    b.ploc(8)
    # BVH2 on the reference order: same t (ids may differ only on equal-t ties)
    t2, p2, _, _ = b.trace(rays)
    np.testing.assert_array_equal(t2.view(np.uint32), t.view(np.uint32))
    # BVH4 (t, primId): identical to the reference tree's BVH4
    t4, p4, _, _ = b.trace4(rays, O.MODE4_FP16)
    r4, q4, _, _ = ref.trace4(rays, O.MODE4_FP32)
    np.testing.assert_array_equal(t4.view(np.uint32), r4.view(np.uint32))
    np.testing.assert_array_equal(p4, q4)
    bt, _ = brute_force_np(v, f, rays)
    np.testing.assert_array_equal(t4.view(np.uint32), bt.view(np.uint32))
    # any-hit booleans
    anyr = rays.copy()
    anyr[:, 7] = -0.3 * float(np.max(v.max(0) - v.min(0)))
    _, pa, _, _ = ref.trace(anyr)
    _, pb, _, _ = b.trace4(anyr, O.MODE4_FP16)
    np.testing.assert_array_equal(pa != INVALID, pb != INVALID)


def surface_rays(v, f, n, stream, tfar):
    """Short any-hit rays leaving the surface upwards (tFar <= 0 selects any-hit)."""
    r = W.uniform01(W.pcg32(n, 4, stream))
    tri = f[(r[:, 0] * f.shape[0]).astype(np.int64) % f.shape[0]]
    c = (v[tri[:, 0]] + v[tri[:, 1]] + v[tri[:, 2]]) / 3
    d = W.unit_vectors(r[:, 1], r[:, 2])
    d[:, 1] = np.abs(d[:, 1])
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = c + np.array([0.0, 1e-3, 0.0])
    rays[:, 4:7] = d
    rays[:, 7] = -tfar
    return rays


def test_ploc_visits_fewer_nodes_on_ao_rays():
    """On a heightfield with boxes (the RTAO scene's make-up) the agglomerative tree needs fewer node visits
    per short any-hit ray than the cubic Morton tree it starts from, with the same answers."""
    v, f = W.rtao_scene(256, 128, 400)
    rays = surface_rays(v, f, 4000, 4900, 0.1 * float(np.max(v.max(0) - v.min(0))))
    b = O.OracleBVH(v, f, cubic=True)
    _, p0, _, _, n0, _ = b.trace4(rays, O.MODE4_FP16, counters=True)
    b.ploc(8)
    _, p1, _, _, n1, _ = b.trace4(rays, O.MODE4_FP16, counters=True)
    np.testing.assert_array_equal(p0 != INVALID, p1 != INVALID)
    assert n1.mean() < 0.92 * n0.mean()
