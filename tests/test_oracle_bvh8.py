"""The oracle's BVH8 nodes (the engine's wide any-hit nodes; no reference counterpart): structure, outward
FP16 boxes that contain their children's exact boxes, and hits equal to brute force under the (t, primId)
rule on every tree (reference tree, cubic grid, clustered hierarchy)."""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_oracle import brute_force_np, rays_around
from test_oracle_ploc import surface_rays

INVALID = 0xFFFFFFFF
LEAF = 0x80000000
EMPTY = 0x80000000


def trees(v, f):
    yield "reference", O.OracleBVH(v, f)
    yield "cubic", O.OracleBVH(v, f, cubic=True)
    b = O.OracleBVH(v, f, cubic=True)
    b.ploc(8)
    yield "clustered", b


SCENES = [("icosphere3", lambda: W.icosphere(3)), ("soup5k", lambda: W.soup(5000, stream=W.STREAM_TEST)),
          ("rtao_small", lambda: W.rtao_scene(64, 32, 40))]


@pytest.mark.parametrize("name,gen", SCENES, ids=[s[0] for s in SCENES])
def test_bvh8_structure(name, gen):
    v, f = gen()
    for tname, b in trees(v, f):
        o = b.octs()
        L = b.leaves
        refs = o[:, :, 3]
        internal = refs[(refs & LEAF) == 0]
        # every node but the root is referenced exactly once
        assert np.array_equal(np.sort(internal), np.arange(1, o.shape[0])), tname
        # every leaf appears exactly once, with the leaf runs' starts
        leaves = refs[((refs & LEAF) != 0) & (refs != EMPTY)]
        assert leaves.size == L, tname
        assert np.array_equal(np.sort(leaves & 0xFFFFFF), b.runs()[0]), tname
        # FP16 boxes contain the exact boxes of the binary nodes they stand for
        lr = b.lr_nodes()
        box_of_leaf = {int(lr[L - 1 + m, 3]): lr[L - 1 + m] for m in range(L)}
        for i in range(0, o.shape[0], max(1, o.shape[0] // 200)):
            for k in range(8):
                r = int(refs[i, k])
                if r & LEAF and r != EMPTY:
                    ex = box_of_leaf[r]
                    assert np.all(o[i, k, 0:3].view(np.float32) <= ex[0:3].view(np.float32))
                    assert np.all(o[i, k, 4:7].view(np.float32) >= ex[4:7].view(np.float32))


@pytest.mark.parametrize("name,gen", SCENES, ids=[s[0] for s in SCENES])
def test_bvh8_hits_equal_brute_force(name, gen):
    v, f = gen()
    rays = rays_around(v, 3000, 5100)
    bt, bp = brute_force_np(v, f, rays)
    anyr = rays.copy()
    anyr[:, 7] = -0.4 * float(np.max(v.max(0) - v.min(0)))
    ref = O.OracleBVH(v, f)
    _, pa, _, _ = ref.trace(anyr)
    for tname, b in trees(v, f):
        t8, p8, _, _ = b.trace8(rays)
        t4, p4, _, _ = b.trace4(rays, O.MODE4_BRUTE)
        np.testing.assert_array_equal(t8.view(np.uint32), bt.view(np.uint32), err_msg=tname)
        np.testing.assert_array_equal(p8, p4, err_msg=tname)
        _, q8, _, _ = b.trace8(anyr)
        np.testing.assert_array_equal(q8 != INVALID, pa != INVALID, err_msg=tname)


def test_bvh8_visits_fewer_nodes_than_bvh4():
    v, f = W.rtao_scene(256, 128, 400)
    rays = surface_rays(v, f, 4000, 5200, 0.1 * float(np.max(v.max(0) - v.min(0))))
    b = O.OracleBVH(v, f, cubic=True)
    b.ploc(8)
    *_, n4, _ = b.trace4(rays, O.MODE4_FP16, counters=True)
    *_, n8, _ = b.trace8(rays, counters=True)
    assert n8.mean() < 0.8 * n4.mean()
