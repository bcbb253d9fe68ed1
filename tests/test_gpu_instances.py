"""GPU parity for instanced scenes: the top-level LBVH over instance boxes and
the two-level traversal, against the oracle's restatement (bit-exact: tree
arrays, t, primId, instId, geomId, u, v, any-hit booleans and traversal
counters)."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_gpu_parity import check_build
from test_oracle_scene import affine, col_major, scene_rays

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def build_both(ctx, geoms, inst, layout=H.LAYOUT_BVH2):
    s = H.Scene(ctx, keep_build_arrays=True)
    o = O.OracleScene()
    for v, f in geoms:
        s.add_triangles(v, f)
        o.add_geom(v, f)
    for g, m in inst:
        s.add_instance(g, m)
        o.add_instance(g, m)
    s.commit(layout=layout)
    o.commit()
    return s, o


def mixed_scene(nboxes=300, nx=32, nz=16):
    geoms, inst = W.rtao_instanced_scene(nx, nz, nboxes)
    iv, if_ = W.icosphere(2, 0.1, (0.0, 0.0, 0.0))
    geoms = geoms + [(iv, if_)]
    for k in range(8):
        M = affine(40 + k)
        M[:3, 3] = [0.25 * k, 0.3, 0.1 * k]
        inst.append((2, col_major(M)))
    return geoms, inst


def check_hits(res, ref, any_hit=False):
    if any_hit:
        np.testing.assert_array_equal(res["hit"], ref[1] != INVALID)
        return
    np.testing.assert_array_equal(res["t"].view(np.uint32), ref[0].view(np.uint32))
    np.testing.assert_array_equal(res["prim"], ref[1])
    np.testing.assert_array_equal(res["inst"], ref[2])
    np.testing.assert_array_equal(res["geom"], ref[3])
    np.testing.assert_array_equal(res["u"].view(np.uint32), ref[4].view(np.uint32))
    np.testing.assert_array_equal(res["v"].view(np.uint32), ref[5].view(np.uint32))


def test_tlas_build_bit_exact(gpu_ctx):
    geoms, inst = mixed_scene()
    say(f"tlas build: {len(inst)} instances")
    s, o = build_both(gpu_ctx, geoms, inst)
    check_build(s, o.tlas, geom=H.TLAS_ID)
    s.close()


@pytest.mark.parametrize("tfar", [W.FLT_MAX, 0.3], ids=["inf", "short"])
def test_two_level_nearest_and_counters(gpu_ctx, tfar):
    geoms, inst = mixed_scene()
    s, o = build_both(gpu_ctx, geoms, inst)
    rays = scene_rays(100_000, 400, tfar=tfar)
    say(f"two-level nearest, tfar={tfar}")
    res = trace(s, rays, "full", stats=True)
    ref = o.trace(rays, O.SCENE_TWO_LEVEL, counters=True)
    check_hits(res, ref)
    assert (ref[1] != INVALID).mean() > 0.15
    np.testing.assert_array_equal(res["pairs"], ref[6])
    np.testing.assert_array_equal(res["tris"], ref[7])
    # without counters the same kernel gives the same hits
    check_hits(trace(s, rays, "full"), ref)
    c = trace(s, rays, "compact")
    np.testing.assert_array_equal(c["t"].view(np.uint32), ref[0].view(np.uint32))
    np.testing.assert_array_equal(c["prim"], ref[1])
    s.close()


def test_two_level_any_hit(gpu_ctx):
    geoms, inst = mixed_scene()
    s, o = build_both(gpu_ctx, geoms, inst)
    rays = scene_rays(100_000, 410, tfar=0.2)
    ref = o.trace(rays, O.SCENE_TWO_LEVEL)
    res = trace(s, rays, "any")
    check_hits(res, ref, any_hit=True)
    # negative tFar in a nearest-hit call stops at the first hit
    ra = rays.copy()
    ra[:, 7] = -ra[:, 7]
    near_any = trace(s, ra, "full")
    np.testing.assert_array_equal(near_any["prim"] != INVALID, ref[1] != INVALID)
    s.close()


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO], ids=["bvh4", "fp16", "auto"])
def test_instanced_scenes_use_two_level_for_every_layout(gpu_ctx, layout):
    geoms, inst = mixed_scene(nboxes=100)
    s, o = build_both(gpu_ctx, geoms, inst, layout=layout)
    rays = scene_rays(20_000, 420)
    check_hits(trace(s, rays, "full"), o.trace(rays, O.SCENE_TWO_LEVEL))
    s.close()


def test_rtao_boxes_as_instances(gpu_ctx):
    """The RTAO workload with its 2,000 boxes as instances of one cube: primary
    nearest hits bit-exact with the oracle; against the merged-mesh scene the
    same object class is hit (heightfield or box) and heightfield hits have
    the same t bit for bit (identical triangles under the identity instance)."""
    geoms, inst = W.rtao_instanced_scene()
    s, o = build_both(gpu_ctx, geoms, inst)
    rays = W.rtao_primary_rays(512, 512)
    say(f"rtao instanced: {len(inst)} instances, {rays.shape[0]} rays")
    res = trace(s, rays, "full")
    ref = o.trace(rays, O.SCENE_TWO_LEVEL, threads=16)
    check_hits(res, ref)
    v, f = W.rtao_scene()
    merged = H.Scene(gpu_ctx)
    merged.add_instance(merged.add_triangles(v, f))
    merged.commit()
    m = trace(merged, rays, "compact")
    nfield = 2 * 1024 * 512
    field_i = res["inst"] == 0
    field_m = m["prim"] < nfield
    assert field_i.mean() > 0.2 and (res["inst"] > 0).mean() > 0.5
    assert np.mean(field_i == field_m) > 0.999
    both = field_i & field_m
    np.testing.assert_array_equal(res["t"][both].view(np.uint32), m["t"][both].view(np.uint32))
    np.testing.assert_array_equal(res["prim"][both], m["prim"][both])
    merged.close()
    s.close()


def test_identity_pair_matches_single_instance(gpu_ctx):
    """Identity instance plus a far-away copy: the rays near the first copy get
    the single-instance scene's answers (t, primId, u, v bit-exact)."""
    v, f = W.icosphere(4, 1.0, (0.0, 1.0, 0.0))
    single = H.Scene(gpu_ctx)
    g = single.add_triangles(v, f)
    single.add_instance(g)
    single.commit()
    two = H.Scene(gpu_ctx)
    g2 = two.add_triangles(v, f)
    two.add_instance(g2)
    T = np.eye(4)
    T[:3, 3] = [100.0, 0.0, 0.0]
    two.add_instance(g2, col_major(T))
    two.commit()
    rays = scene_rays(50_000, 430, lo=(-3, -2, -3), hi=(3, 4, 3))
    a = trace(single, rays, "full")
    b = trace(two, rays, "full")
    # rays that miss the first copy may go on to hit the far one
    far = b["inst"] == 1
    assert np.all(a["prim"][far] == INVALID)
    near = ~far
    assert near.mean() > 0.99
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(a[k][near].view(np.uint32), b[k][near].view(np.uint32))
    np.testing.assert_array_equal(a["prim"][near], b["prim"][near])
    np.testing.assert_array_equal(b["inst"][near & (b["prim"] != INVALID)], 0)
    single.close()
    two.close()


def test_instance_edge_cases(gpu_ctx):
    v, f = W.unit_cube()
    # no instances: every ray misses, t = tFar
    s = H.Scene(gpu_ctx)
    s.add_triangles(v, f)
    s.commit()
    rays = scene_rays(1000, 440, lo=(-1, -1, -1), hi=(1, 1, 1))
    r = trace(s, rays, "full")
    assert np.all(r["prim"] == INVALID) and np.all(r["inst"] == INVALID)
    np.testing.assert_array_equal(r["t"], rays[:, 7])
    assert not trace(s, rays, "any")["hit"].any()
    # singular matrix
    with pytest.raises(H.HrtError) as e:
        s.add_instance(0, col_major(np.diag([1.0, 1.0, 0.0, 1.0])))
    assert e.value.status == 1
    # more than 127 instances in one Morton cell of the top level
    for _ in range(130):
        s.add_instance(0)
    T = np.eye(4)
    T[:3, 3] = [5.0, 0.0, 0.0]
    s.add_instance(0, col_major(T))
    with pytest.raises(H.HrtError) as e:
        s.commit()
    assert e.value.status == 6
    # 127 coincident instances are fine; ties go to the smallest id
    s.clear_instances()
    for _ in range(127):
        s.add_instance(0)
    s.add_instance(0, col_major(T))
    s.commit()
    ray = np.array([[-5, 0, 0, 0, 1, 0, 0, W.FLT_MAX], [10, 0, 0, 0, -1, 0, 0, W.FLT_MAX]], np.float32)
    r = trace(s, ray, "full")
    assert list(r["inst"]) == [0, 127]
    np.testing.assert_array_equal(r["t"], np.float32([4.5, 4.5]))
    s.close()
