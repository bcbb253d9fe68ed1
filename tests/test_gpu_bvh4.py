"""GPU parity for the BVH4 path (FP32 and FP16 boxes) against the CPU oracle.

Bar: BVH4 nodes bit-exact (FP32 boxes and references; FP16 boxes equal the
oracle's outward-rounded halves); nearest hits (t, primId, u, v) bit-exact
with the oracle's BVH4 traversal under the (t, primId) rule, identical
between FP32 and FP16 nodes; traversal counters exact; any-hit booleans equal
the BVH2 path's.
"""
import numpy as np
import pytest

import heavyrt_amd as H
from conftest import vsgf_path
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_gpu_parity import check_build, rays_for

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def _cases():
    for m in ("triangle", "cube", "cone_low", "helix_mid"):
        yield m, lambda m=m: W.load_vsgf(vsgf_path(m))
    yield "eye_scene", W.eye_scene
    yield "soup_200k", lambda: W.soup(200_000, stream=W.STREAM_TEST + 3)


CASES = list(_cases())


def scene_with(ctx, v, f, layout):
    s = H.Scene(ctx)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit(layout=layout)
    return s


def gpu_quads32_as_oracle(q: np.ndarray) -> np.ndarray:
    """(nq, 32) SoA words -> (nq, 4, 8) per-child words like the oracle's."""
    nq = q.shape[0]
    out = np.zeros((nq, 4, 8), np.uint32)
    lox, hix, loy, hiy, loz, hiz, ref = (q[:, 4 * k:4 * k + 4] for k in range(7))
    out[:, :, 0], out[:, :, 1], out[:, :, 2], out[:, :, 3] = lox, loy, loz, ref
    out[:, :, 4], out[:, :, 5], out[:, :, 6] = hix, hiy, hiz
    return out


def gpu_quads16_as_floats(q: np.ndarray):
    """(nq, 16) words -> boxes (nq, 4, 6) float32 and refs (nq, 4)."""
    w = q.reshape(-1, 4, 4)
    halves = w[:, :, 0:3].copy().view(np.uint16).reshape(-1, 4, 6)
    f = halves.view(np.float16).astype(np.float32)
    # per child: minx, miny, minz, maxx, maxy, maxz
    return f, w[:, :, 3]


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_bvh4_nodes_bit_exact(gpu_ctx, name, gen):
    v, f = gen()
    say(f"bvh4 nodes {name}: {f.shape[0]} tris")
    ob = O.OracleBVH(v, f)
    s = scene_with(gpu_ctx, v, f, H.LAYOUT_BVH4)
    q32 = s.dump(0, H.DUMP_QUADS32)
    oq = ob.quads(False)
    assert q32.shape[0] == oq.shape[0]
    np.testing.assert_array_equal(gpu_quads32_as_oracle(q32), oq)
    boxes16, refs16 = gpu_quads16_as_floats(s.dump(0, H.DUMP_QUADS16))
    oq16 = ob.quads(True)
    np.testing.assert_array_equal(refs16, oq16[:, :, 3])
    want = np.concatenate([oq16[:, :, 0:3], oq16[:, :, 4:7]], axis=2).view(np.float32)
    np.testing.assert_array_equal(boxes16.view(np.uint32), want.view(np.uint32))
    s.close()


FP16_LAYOUTS = (H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC_FP16)
NEAREST4 = [H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16]
NEAREST4_IDS = ["fp32", "fp16", "ploc", "ploc_fp16"]


def walked_tree(scene, layout, v, f, ob):
    """The oracle tree the scene's BVH4 traversals walk: the second (clustered) tree for the
    AUTO_FP16 and PLOC layouts when the commit built one, else the reference tree ob."""
    if layout in (H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16) and has_any_tree(scene):
        return oracle_any_tree(v, f)[0]
    return ob


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("layout", NEAREST4, ids=NEAREST4_IDS)
def test_bvh4_nearest_bit_exact(gpu_ctx, name, gen, layout):
    """Nearest hits bit-exact with the oracle's BVH4 traversal of the reference tree under the
    (t, primId) rule, on every BVH4 layout: the PLOC layouts walk the clustered second tree, whose
    hits are the same; their counters equal the oracle's walk of that tree."""
    v, f = gen()
    rays = rays_for(v, 20000, 1500)
    say(f"bvh4 nearest {name} layout {layout}")
    ob = O.OracleBVH(v, f)
    s = scene_with(gpu_ctx, v, f, layout)
    mode = O.MODE4_FP16 if layout in FP16_LAYOUTS else O.MODE4_FP32
    t, p, u, vv, nq, nt = ob.trace4(rays, mode, counters=True)
    walked = walked_tree(s, layout, v, f, ob)
    if walked is not ob:
        t2, p2, u2, v2, nq, nt = walked.trace4(rays, mode, counters=True)
        np.testing.assert_array_equal(t2.view(np.uint32), t.view(np.uint32))
        np.testing.assert_array_equal(p2, p)
        np.testing.assert_array_equal(u2.view(np.uint32), u.view(np.uint32))
        np.testing.assert_array_equal(v2.view(np.uint32), vv.view(np.uint32))
    g = trace(s, rays, "full", stats=True)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), vv.view(np.uint32))
    np.testing.assert_array_equal(g["pairs"], nq)
    np.testing.assert_array_equal(g["tris"], nt)
    g2 = trace(s, rays, "compact")
    np.testing.assert_array_equal(g2["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g2["prim"], p)
    # FP16 and FP32 nodes and the brute force agree under the (t, primId) rule
    tb, pb, _, _ = ob.trace4(rays, O.MODE4_BRUTE)
    np.testing.assert_array_equal(tb.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(pb, p)
    s.close()


@pytest.mark.parametrize("layout", NEAREST4, ids=NEAREST4_IDS)
def test_bvh4_nearest_sorted_waves(gpu_ctx, layout):
    """A batch of 2^20 rays or more is sorted into coherent waves before the trace (each thread
    traces the ray the order gives it): full hits equal the oracle's BVH4 traversal of the tree
    the layout walks under the (t, primId) rule on every ray, at every ray's own index, including
    rays that are any-hit by their own tFar <= 0 (their hit is the first one the walk finds, so
    it depends on the tree; their hit/miss answer and every nearest hit do not), empty and
    reversed intervals, a NaN bound and a NaN direction."""
    v, f = W.soup(150_000, stream=W.STREAM_TEST + 5)
    n = 1_100_000
    rays = W.incoherent_rays(n, stream=W.STREAM_TEST + 6)
    kind = np.arange(n) % 10
    rays[kind == 1, 7] = -rays[kind == 1, 7]      # any-hit by the ray's own convention
    rays[kind == 2, 7] = 0.0                      # tFar = tNear = 0: empty
    rays[kind == 3, 3] = 0.5                      # tNear > tFar
    rays[kind == 3, 7] = 0.25
    rays[kind == 4, 7] = np.float32(np.nan)
    rays[kind == 5, 7] = 0.05                     # short rays
    rays[kind == 6, 4] = np.float32(np.nan)       # a NaN direction sorts into cell 0
    say(f"bvh4 nearest sorted waves, layout {layout}")
    s = scene_with(gpu_ctx, v, f, layout)
    g = trace(s, rays, "full")
    ref = O.OracleBVH(v, f)
    walked = walked_tree(s, layout, v, f, ref)
    s.close()
    mode = O.MODE4_FP16 if layout in FP16_LAYOUTS else O.MODE4_FP32
    t, p, u, vv = walked.trace4(rays, mode)
    if walked is not ref:
        tr, pr, _, _ = ref.trace4(rays, mode)
        near = kind != 1
        np.testing.assert_array_equal(p[near], pr[near])
        np.testing.assert_array_equal(t[near].view(np.uint32), tr[near].view(np.uint32))
        np.testing.assert_array_equal(p[~near] != INVALID, pr[~near] != INVALID)
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), vv.view(np.uint32))
    hit = p != INVALID
    assert 0.05 < hit[kind == 0].mean() < 0.999
    assert not hit[(kind == 2) | (kind == 3) | (kind == 4) | (kind == 6)].any()


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("layout", [H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO, H.LAYOUT_AUTO_FP16,
                                    H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16],
                         ids=["fp32", "fp16", "auto", "auto_fp16", "ploc", "ploc_fp16"])
def test_bvh4_any_hit(gpu_ctx, name, gen, layout):
    v, f = gen()
    rays = rays_for(v, 20000, 1700)
    rays[:, 7] = 0.5 * float(np.max(v.max(0) - v.min(0)))
    say(f"bvh4 any {name} layout {layout}")
    ob = O.OracleBVH(v, f)
    s = scene_with(gpu_ctx, v, f, layout)
    anyr = rays.copy()
    anyr[:, 7] = -anyr[:, 7]
    _, p2, _, _ = ob.trace(anyr)
    g = trace(s, rays, "any")
    np.testing.assert_array_equal(g["hit"], p2 != INVALID)
    if layout != H.LAYOUT_AUTO:
        # hrt_trace_any on FP16 nodes tests boxes in the fused form
        mode = O.MODE4_FP16_FUSED if layout in FP16_LAYOUTS else O.MODE4_FP32
        ob4 = walked_tree(s, layout, v, f, ob)  # AUTO_FP16 and PLOC answer any-hit on the second tree
        _, p4, _, _, nq, nt = ob4.trace4(anyr, mode, counters=True)
        gs = trace(s, rays, "any", stats=True)
        np.testing.assert_array_equal(gs["hit"], p4 != INVALID)
        np.testing.assert_array_equal(gs["pairs"], nq)
        np.testing.assert_array_equal(gs["tris"], nt)
    s.close()


def has_any_tree(scene) -> bool:
    try:
        scene.dump(0, H.DUMP_SIZES | H.DUMP_ANY_TREE)
        return True
    except H.HrtError:
        return False


def oracle_any_tree(v, f):
    """The oracle's any-hit tree: leaves of the cubic-grid LBVH, hierarchy by agglomerative clustering
    (radius 8) unless that tree has more than 42 internal levels (the BVH4 stack bound)."""
    ob = O.OracleBVH(v, f, cubic=True)
    if ob.max_run > 127 or ob.leaves < 3:
        return ob, False
    pl = O.OracleBVH(v, f, cubic=True)
    pl.ploc(8)
    lr = pl.lr_nodes()
    L = pl.leaves
    depth = np.zeros(2 * L - 1, dtype=np.int64)
    for i in range(L - 1):  # parents have smaller ids than their children
        depth[int(lr[i, 3])] = depth[i] + 1
        depth[int(lr[i, 7])] = depth[i] + 1
    return (pl, True) if depth.max() <= 42 else (ob, False)


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_any_tree_bit_exact(gpu_ctx, name, gen):
    """HRT_LAYOUT_AUTO_FP16 builds a second tree for any-hit queries: leaves on
    a cubic Morton grid, the hierarchy above them by agglomerative clustering.
    Its build arrays, node layout and FP16 BVH4 nodes equal the oracle's bit
    for bit."""
    v, f = gen()
    s = H.Scene(gpu_ctx, keep_build_arrays=True)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=H.LAYOUT_AUTO_FP16)
    ob, clustered = oracle_any_tree(v, f)
    if ob.max_run > 127:
        assert not has_any_tree(s)
        s.close()
        return
    assert int(s.dump(0, H.DUMP_SIZES | H.DUMP_ANY_TREE)[5]) == int(clustered)
    check_build(s, ob, flag=H.DUMP_ANY_TREE)
    boxes16, refs16 = gpu_quads16_as_floats(s.dump(0, H.DUMP_QUADS16 | H.DUMP_ANY_TREE))
    oq16 = ob.quads(True)
    np.testing.assert_array_equal(refs16, oq16[:, :, 3])
    want = np.concatenate([oq16[:, :, 0:3], oq16[:, :, 4:7]], axis=2).view(np.float32)
    np.testing.assert_array_equal(boxes16.view(np.uint32), want.view(np.uint32))
    s.close()


@pytest.mark.parametrize("layout", [H.LAYOUT_AUTO, H.LAYOUT_AUTO_FP16], ids=["auto", "auto_fp16"])
def test_auto_layout_nearest_is_reference_order(gpu_ctx, layout):
    v, f = W.eye_scene()
    rays = W.eye_rays(256, 256)
    ob = O.OracleBVH(v, f)
    s = scene_with(gpu_ctx, v, f, layout)
    g = trace(s, rays, "full")
    t, p, u, vv = ob.trace(rays)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["prim"], p)
    s.close()


def test_fp16_layout_rejects_out_of_range_scene(gpu_ctx):
    v, f = W.icosphere(2)
    v = v * np.float32(1e5)
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    with pytest.raises(H.HrtError) as ei:
        s.commit(layout=H.LAYOUT_BVH4_FP16)
    assert ei.value.status == 5
    s.commit(layout=H.LAYOUT_BVH4)  # FP32 nodes are fine
    s.close()


def test_precision_soup_fp16_vs_fp32_identical(gpu_ctx):
    """Precision workload shape at reduced size: random soup, incoherent rays;
    FP16 and FP32 BVH4 nodes return identical hits, on the reference tree and
    on the clustered second tree (PLOC layouts); t error vs FP64 reported."""
    v, f = W.soup(500_000, stream=W.STREAM_SOUP)
    rays = W.incoherent_rays(500_000, stream=W.STREAM_RAYS)
    say("precision soup 500k")
    res = {}
    for lay in (H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16):
        sc = scene_with(gpu_ctx, v, f, lay)
        try:
            res[lay] = trace(sc, rays, "full")
        finally:
            sc.close()
    a = res[H.LAYOUT_BVH4]
    for lay in (H.LAYOUT_BVH4_FP16, H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16):
        b = res[lay]
        for k in ("t", "u", "v"):
            np.testing.assert_array_equal(a[k].view(np.uint32), b[k].view(np.uint32))
        np.testing.assert_array_equal(a["prim"], b["prim"])
    ob = O.OracleBVH(v, f)
    hit = a["prim"] != INVALID
    t64 = ob.t_fp64(rays[hit], a["prim"][hit])
    rel = np.abs(a["t"][hit].astype(np.float64) - t64) / np.abs(t64)
    say(f"precision: {hit.sum()} hits, t rel err vs fp64 max {rel.max():.3e} mean {rel.mean():.3e} "
        f"p50 {np.median(rel):.3e} p99.9 {np.percentile(rel, 99.9):.3e}")
    # FP32 Moller-Trumbore error grows without bound at grazing angles; the
    # bulk of the distribution stays at float32 rounding level
    assert np.median(rel) < 1e-6
    assert rel.mean() < 1e-5


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC_FP16],
                         ids=["bvh4_fp16", "auto_fp16", "ploc_fp16"])
def test_any_hit_axis_parallel_far_from_origin(gpu_ctx, layout):
    """Rays with d.x = d.z = 0 over geometry at x = 500: -(o * inv) overflows there, so these lanes
    keep the unfused box test; answers equal brute force and counts equal the oracle's."""
    from test_oracle_bvh4 import far_axis_rays

    v, f, anyr = far_axis_rays()
    rays = anyr.copy()
    rays[:, 7] = -rays[:, 7]
    say(f"axis-parallel any-hit layout {layout}")
    s = scene_with(gpu_ctx, v, f, layout)
    _, pb, _, _ = O.OracleBVH(v, f).trace4(anyr, O.MODE4_BRUTE)
    g = trace(s, rays, "any")
    np.testing.assert_array_equal(g["hit"], pb != INVALID)
    assert 0.2 < g["hit"].mean() < 1.0
    gs = trace(s, rays, "any", stats=True)
    np.testing.assert_array_equal(gs["hit"], pb != INVALID)
    if layout == H.LAYOUT_BVH4_FP16:
        ob = O.OracleBVH(v, f)
        *_, nq, nt = ob.trace4(anyr, O.MODE4_FP16_FUSED, counters=True)
    elif has_any_tree(s):
        ob = oracle_any_tree(v, f)[0]
        *_, nq, nt = ob.trace4(anyr, O.MODE4_FP16_FUSED, counters=True)
    else:
        s.close()
        return
    np.testing.assert_array_equal(gs["pairs"], nq)
    np.testing.assert_array_equal(gs["tris"], nt)
    s.close()


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO_FP16,
                                    H.LAYOUT_PLOC, H.LAYOUT_PLOC_FP16],
                         ids=["bvh2", "bvh4", "bvh4_fp16", "auto_fp16", "ploc", "ploc_fp16"])
def test_any_hit_interval_edge_cases(gpu_ctx, layout):
    """Empty and reversed (tNear, tFar) intervals, a negative tNear and a NaN bound: the any-hit
    kernels finish such rays at once; answers equal the oracle's brute force."""
    v, f = W.icosphere(3, 1.0, (0.0, 0.0, 0.0))
    rays = rays_for(v, 6000, 2100)
    n = rays.shape[0]
    span = float(np.max(v.max(0) - v.min(0)))
    kind = np.arange(n) % 6
    rays[:, 3] = 0.0
    rays[:, 7] = 3.0 * span                       # origins sit 1.5 extents from the centre
    rays[kind == 1, 7] = 0.0                      # tFar = tNear = 0
    rays[kind == 2, 3] = 0.5 * span               # tNear > tFar
    rays[kind == 2, 7] = 0.25 * span
    rays[kind == 3, 3] = -span                    # negative tNear
    rays[kind == 4, 7] = np.float32(np.nan)       # NaN tFar
    rays[kind == 5, 3] = 1.2 * span               # a proper interval that starts late
    say(f"any-hit interval edge cases, layout {layout}")
    s = scene_with(gpu_ctx, v, f, layout)
    anyr = rays.copy()
    anyr[:, 7] = -anyr[:, 7]
    _, pb, _, _ = O.OracleBVH(v, f).trace4(anyr, O.MODE4_BRUTE)
    g = trace(s, rays, "any")
    np.testing.assert_array_equal(g["hit"], pb != INVALID)
    assert not g["hit"][(kind == 1) | (kind == 2) | (kind == 4)].any()
    assert g["hit"][kind == 3].any() and g["hit"][kind == 0].any()
    s.close()
