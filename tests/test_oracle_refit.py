"""Pins the oracle's refit (BUILD_REFIT: new vertex positions, same
hierarchy; CPU only): unchanged vertices give back the original nodes, node
boxes are the unions of the new triangle boxes over each node's unchanged
leaf range, and traversal of the refitted tree equals brute force on the new
geometry."""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_oracle import brute_force_np, rays_around

LEAF = 0x80000000


def moved(v, seed):
    r = W.uniform01(W.pcg32(v.shape[0], 3, seed)) - 0.5
    ext = float(np.max(v.max(0) - v.min(0)))
    return (v + 0.02 * ext * r).astype(np.float32)


CASES = [("icosphere3", lambda: W.icosphere(3)), ("soup", lambda: W.soup(3000, stream=W.STREAM_TEST)),
         ("heightfield", lambda: W.heightfield(32, 16)[:2])]


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_refit_with_same_vertices_is_identity(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f)
    lr0, pe0 = b.lr_nodes().copy(), b.peach_nodes().copy()
    b.refit(v)
    np.testing.assert_array_equal(b.lr_nodes(), lr0)
    np.testing.assert_array_equal(b.peach_nodes(), pe0)


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_refit_boxes_are_unions_over_kept_leaves(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f)
    perm0, rs0 = b.perm().copy(), b.runs()[0].copy()
    lr0 = b.lr_nodes().copy()
    v2 = moved(v, 77)
    b.refit(v2)
    np.testing.assert_array_equal(b.perm(), perm0)
    lr = b.lr_nodes()
    np.testing.assert_array_equal(lr[:, 3], lr0[:, 3])  # hierarchy kept
    np.testing.assert_array_equal(lr[:, 7], lr0[:, 7])
    L = b.leaves
    tri = v2[f.astype(np.int64)]
    tmin, tmax = tri.min(axis=1), tri.max(axis=1)
    ends = np.append(rs0[1:], b.ntris)
    lo_leaf = np.stack([tmin[perm0[rs0[m]:ends[m]]].min(0) for m in range(L)])
    hi_leaf = np.stack([tmax[perm0[rs0[m]:ends[m]]].max(0) for m in range(L)])
    np.testing.assert_array_equal(lr[L - 1:, 0:3].view(np.float32), lo_leaf)
    np.testing.assert_array_equal(lr[L - 1:, 4:7].view(np.float32), hi_leaf)
    for k in range(L - 1):
        a, c = lr[int(lr[k, 3])], lr[int(lr[k, 7])]
        np.testing.assert_array_equal(lr[k, 0:3].view(np.float32),
                                      np.minimum(a[0:3].view(np.float32), c[0:3].view(np.float32)))
    bmin, bmax, _ = b.scene_box()
    np.testing.assert_array_equal(bmin[:3], v2.min(0))
    np.testing.assert_array_equal(bmax[:3], v2.max(0))


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_refit_traversal_equals_brute_force(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f)
    v2 = moved(v, 78)
    b.refit(v2)
    rays = rays_around(v2, 3000, 4300)
    t, p, _, _ = b.trace(rays)
    bt, bp = brute_force_np(v2, f, rays)
    np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert np.mean(p == bp) > 0.999
    anyr = rays.copy()
    anyr[:, 7] = -0.5 * float(np.max(v2.max(0) - v2.min(0)))
    _, pa, _, _ = b.trace(anyr)
    pos = anyr.copy()
    pos[:, 7] = -pos[:, 7]
    bt2, _ = brute_force_np(v2, f, pos)
    np.testing.assert_array_equal(pa != 0xFFFFFFFF, bt2 < pos[:, 7])


def test_refit_refuses_a_different_triangle_count():
    v, f = W.icosphere(2)
    b = O.OracleBVH(v, f)
    with pytest.raises(ValueError):
        b.refit(v, f[:-1])
