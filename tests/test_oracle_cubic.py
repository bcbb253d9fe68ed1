"""Pins the oracle's cubic-grid tree (the engine's any-hit tree; CPU only).

The cubic grid scales every axis by the largest extent.  Checks:
- codes decode to the float64 cubic cell of each triangle centre (+-1 cell);
- the tree satisfies the same invariants as the reference tree (stable
  (code, id) order, Karras = radix split, boxes = unions);
- nearest hits under the (t, primId) rule and any-hit answers do not depend
  on the tree: cubic FP32 / FP16 BVH4 = reference-tree BVH4 = brute force.
"""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_oracle import deinterleave3, radix_tree, rays_around

SCENES = [
    ("heightfield", lambda: W.heightfield(64, 32)[:2]),
    ("rtao_small", lambda: W.rtao_scene(64, 32, 60)),
    ("soup", lambda: W.soup(20_000, stream=W.STREAM_TEST)),
]


@pytest.mark.parametrize("name,gen", SCENES, ids=[s[0] for s in SCENES])
def test_cubic_codes_are_cubic_cells(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f, cubic=True)
    codes, sc, ids = b.codes()
    assert int(codes.max()) < (1 << 30)
    v64 = v.astype(np.float64)
    lo, hi = v64.min(0), v64.max(0)
    ext = ((hi + 1e-4) - (lo - 1e-4)).max()
    tri = v64[f.astype(np.int64)]
    c = 0.5 * (tri.min(axis=1) + tri.max(axis=1)) if f.shape[0] >= 256 else tri.mean(axis=1)
    cells = np.floor(1024.0 * (c - lo) / ext).astype(np.int64)
    got = np.array([deinterleave3(int(x)) for x in codes])
    assert np.all(np.abs(got - cells) <= 1)
    # the thin axis of a flat scene uses only part of its 1024 cells
    if name == "heightfield":
        assert got[:, 1].max() < 1024 * (hi[1] - lo[1]) / ext + 2
    order = np.lexsort((np.arange(codes.size), codes))
    np.testing.assert_array_equal(ids, order)
    np.testing.assert_array_equal(sc, codes[order])


@pytest.mark.parametrize("name,gen", SCENES, ids=[s[0] for s in SCENES])
def test_cubic_tree_invariants(name, gen):
    v, f = gen()
    b = O.OracleBVH(v, f, cubic=True)
    _, uc = b.runs()
    lr = b.lr_nodes()
    tree = radix_tree(uc)
    for k, (l, r) in tree.items():
        assert (int(lr[k, 3]), int(lr[k, 7])) == (l, r)
    L = b.leaves
    for k in range(L - 1):
        a, c = lr[int(lr[k, 3])], lr[int(lr[k, 7])]
        np.testing.assert_array_equal(lr[k, 0:3].view(np.float32),
                                      np.minimum(a[0:3].view(np.float32), c[0:3].view(np.float32)))
        np.testing.assert_array_equal(lr[k, 4:7].view(np.float32),
                                      np.maximum(a[4:7].view(np.float32), c[4:7].view(np.float32)))


@pytest.mark.parametrize("name,gen", SCENES, ids=[s[0] for s in SCENES])
def test_cubic_tree_answers_equal_reference_tree(name, gen):
    v, f = gen()
    ref = O.OracleBVH(v, f)
    cub = O.OracleBVH(v, f, cubic=True)
    rays = rays_around(v, 4000, 4100)
    tb, pb, _, _ = ref.trace4(rays, O.MODE4_BRUTE)
    for mode in (O.MODE4_FP32, O.MODE4_FP16):
        t, p, _, _ = cub.trace4(rays, mode)
        np.testing.assert_array_equal(t.view(np.uint32), tb.view(np.uint32))
        np.testing.assert_array_equal(p, pb)
    anyr = rays.copy()
    anyr[:, 7] = -0.3 * float(np.max(v.max(0) - v.min(0)))
    _, pr, _, _ = ref.trace(anyr)
    _, pc, _, _ = cub.trace4(anyr, O.MODE4_FP16)
    np.testing.assert_array_equal(pc != 0xFFFFFFFF, pr != 0xFFFFFFFF)


def test_cubic_tree_visits_fewer_nodes_on_flat_scenes():
    v, f = W.rtao_scene(128, 64, 200)
    ref = O.OracleBVH(v, f)
    cub = O.OracleBVH(v, f, cubic=True)
    rays = W.incoherent_rays(20_000, stream=4200)
    rays[:, 0:3] = rays[:, 0:3] * np.float32([2.0, 0.3, 1.0])
    rays[:, 7] = -0.1
    *_, q_ref, _ = ref.trace4(rays, O.MODE4_FP16, counters=True)
    *_, q_cub, _ = cub.trace4(rays, O.MODE4_FP16, counters=True)
    assert q_cub.mean() < 0.9 * q_ref.mean()
