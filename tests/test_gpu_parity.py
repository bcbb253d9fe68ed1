"""GPU parity: the HIP engine against the CPU oracle, through the C-ABI.

Bar: builder arrays (Morton codes, sort order, runs, Karras children, boxes,
leaf permutation) bit-exact; nearest-hit (t, primId, u, v) bit-exact; any-hit
booleans exact; traversal counters exact.
"""
import numpy as np
import pytest

import heavyrt_amd as H
from conftest import VSGF_MESHES, vsgf_path
from gpu_util import make_scene, say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def _cases():
    for m in VSGF_MESHES:
        yield m, lambda m=m: W.load_vsgf(vsgf_path(m))
    yield "eye_scene", W.eye_scene
    yield "soup_100k", lambda: W.soup(100_000, stream=W.STREAM_TEST)
    yield "icosphere4", lambda: W.icosphere(4)


CASES = list(_cases())


def rays_for(v, n, stream):
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


def check_build(scene, ob: O.OracleBVH, geom: int = 0, flag: int = 0):
    sizes = scene.dump(geom, flag | H.DUMP_SIZES)
    assert int(sizes[0]) == ob.ntris
    assert int(sizes[1]) == ob.leaves
    assert int(sizes[2]) == ob.max_run
    bmin, bmax, zs = ob.scene_box()
    box = scene.dump(geom, flag | H.DUMP_SCENE_BOX)
    np.testing.assert_array_equal(box[0:4].view(np.uint32), bmin.view(np.uint32))
    np.testing.assert_array_equal(box[4:8].view(np.uint32), bmax.view(np.uint32))
    np.testing.assert_array_equal(box[8:12].view(np.uint32), zs.view(np.uint32))
    c_in, c_sorted, ids = ob.codes()
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_CODES_INPUT), c_in)
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_CODES_SORTED), c_sorted)
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_IDS_SORTED), ids)
    rs, uc = ob.runs()
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_RUN_START), rs)
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_UNIQUE_CODES), uc)
    np.testing.assert_array_equal(scene.dump(geom, flag | H.DUMP_PERM), ob.perm())
    lr = scene.dump(geom, flag | H.DUMP_LR_NODES)
    np.testing.assert_array_equal(lr, ob.lr_nodes())


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_build_bit_exact(gpu_ctx, name, gen):
    v, f = gen()
    say(f"build {name}: {f.shape[0]} tris")
    ob = O.OracleBVH(v, f)
    scene = make_scene(gpu_ctx, v, f)
    check_build(scene, ob)
    scene.close()


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_nearest_bit_exact(gpu_ctx, name, gen):
    v, f = gen()
    rays = rays_for(v, 20000, 500)
    say(f"nearest {name}: {f.shape[0]} tris, {rays.shape[0]} rays")
    ob = O.OracleBVH(v, f)
    scene = make_scene(gpu_ctx, v, f, keep=False)
    g = trace(scene, rays, "full", stats=True)
    t, p, u, vv, pairs, tris = ob.trace(rays, counters=True)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), vv.view(np.uint32))
    hit = p != INVALID
    assert (g["inst"][hit] == 0).all() and (g["inst"][~hit] == INVALID).all()
    np.testing.assert_array_equal(g["pairs"], pairs)
    np.testing.assert_array_equal(g["tris"], tris)
    # the persistent kernel (no counters) gives the same records
    g2 = trace(scene, rays, "full")
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(g2[k].view(np.uint32), g[k].view(np.uint32))
    for k in ("prim", "inst", "geom"):
        np.testing.assert_array_equal(g2[k], g[k])
    # compact format carries the same (t, primId)
    c = trace(scene, rays, "compact")
    np.testing.assert_array_equal(c["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(c["prim"], p)
    scene.close()


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_any_hit_exact(gpu_ctx, name, gen):
    v, f = gen()
    rays = rays_for(v, 20000, 700)
    rays[:, 7] = 0.6 * float(np.max(v.max(0) - v.min(0)))
    say(f"any-hit {name}")
    ob = O.OracleBVH(v, f)
    scene = make_scene(gpu_ctx, v, f, keep=False)
    g = trace(scene, rays, "any", stats=True)
    anyr = rays.copy()
    anyr[:, 7] = -anyr[:, 7]
    _, p, _, _, pairs, tris = ob.trace(anyr, counters=True)
    np.testing.assert_array_equal(g["hit"], p != INVALID)
    np.testing.assert_array_equal(g["pairs"], pairs)
    np.testing.assert_array_equal(g["tris"], tris)
    # persistent speculative any-hit kernel (no counters): same booleans
    g2 = trace(scene, rays, "any")
    np.testing.assert_array_equal(g2["hit"], p != INVALID)
    # negative tFar through the nearest-hit entry follows the same stop-at-first-hit rule
    n = trace(scene, anyr, "compact")
    np.testing.assert_array_equal(n["prim"] != INVALID, p != INVALID)
    scene.close()


def test_eye_rays_full_config(gpu_ctx):
    """512x512 eye rays on the 81,922-triangle icosphere + quad."""
    v, f = W.eye_scene()
    rays = W.eye_rays()
    say("eye config 512x512")
    ob = O.OracleBVH(v, f)
    scene = make_scene(gpu_ctx, v, f)
    check_build(scene, ob)
    g = trace(scene, rays, "full")
    t, p, u, vv = ob.trace(rays)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), vv.view(np.uint32))
    scene.close()


@pytest.mark.parametrize("layout", ["bvh2", "auto_fp16"])
def test_rtao_config_parity(gpu_ctx, layout):
    """1M-triangle heightfield + 2,000 boxes; 1024x1024 primaries (nearest),
    16 AO rays per hit made on the device, any-hit, all compared exactly; on
    the reference tree's BVH2 and on the timed layout (nearest on BVH2, any-hit
    on the FP16 BVH4 nodes of the any-hit tree)."""
    import torch

    v, f = W.rtao_scene()
    say(f"rtao build {f.shape[0]} tris, layout {layout}")
    ob = O.OracleBVH(v, f)
    scene = H.Scene(gpu_ctx, keep_build_arrays=True)
    scene.add_instance(scene.add_triangles(v, f))
    scene.commit(layout={"bvh2": H.LAYOUT_BVH2, "auto_fp16": H.LAYOUT_AUTO_FP16}[layout])
    check_build(scene, ob)
    prim = W.rtao_primary_rays()
    npix = prim.shape[0]
    spp = 16
    tmax = 0.1 * W.scene_diagonal(v)
    stream = torch.cuda.current_stream().cuda_stream
    dprim = torch.from_numpy(prim).cuda()
    dhits = torch.zeros((npix, 2), dtype=torch.int32, device="cuda")
    scene.trace_nearest(dprim, npix, dhits, H.HIT_COMPACT, stream=stream)
    dao = torch.empty((npix * spp, 8), dtype=torch.float32, device="cuda")
    scene.rtao_make_rays(dprim, dhits, npix, spp, tmax, 1234, dao, stream=stream)
    dvis = torch.zeros(npix * spp, dtype=torch.uint8, device="cuda")
    scene.trace_any(dao, npix * spp, dvis, stream=stream)
    torch.cuda.synchronize()
    say("rtao oracle compare")
    hits = dhits.cpu().numpy().view(np.uint32)
    t, p, _, _ = ob.trace(prim)
    np.testing.assert_array_equal(hits[:, 0].view(np.float32).view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(hits[:, 1], p)
    ao = dao.cpu().numpy()
    vis = dvis.cpu().numpy().astype(bool)
    # AO rays: finite, unit length, in the hemisphere of the hit, tFar = tmax
    live = np.repeat(p != INVALID, spp)
    assert np.isfinite(ao).all()
    assert np.allclose(np.linalg.norm(ao[live, 4:7], axis=1), 1.0, atol=1e-5)
    assert np.all(ao[live, 7] == np.float32(tmax))
    anyr = ao.copy()
    anyr[:, 7] = -np.abs(anyr[:, 7])
    _, pa, _, _ = ob.trace(anyr)
    np.testing.assert_array_equal(vis, pa != INVALID)
    occl = vis[live].mean()
    say(f"rtao occlusion fraction {occl:.3f}")
    assert 0.01 < occl < 0.99
    # the fused AO pass (rays made inside the traversal) gives the same bytes
    dvis2 = torch.full((npix * spp,), 7, dtype=torch.uint8, device="cuda")
    scene.rtao_trace(dprim, dhits, npix, spp, tmax, 1234, dvis2, stream=stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dvis2.cpu().numpy(), dvis.cpu().numpy())
    scene.close()


def test_single_triangle_and_duplicates(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    for copies in (1, 4, 127):
        f = np.array([[0, 1, 2]] * copies, np.uint32)
        ob = O.OracleBVH(v, f)
        scene = make_scene(gpu_ctx, v, f)
        check_build(scene, ob)
        rays = np.array([[0.2, 0.2, 1.0, 0.0, 0.0, 0.0, -1.0, W.FLT_MAX],
                         [2.0, 2.0, 1.0, 0.0, 0.0, 0.0, -1.0, W.FLT_MAX]], np.float32)
        g = trace(scene, rays, "full")
        t, p, u, vv = ob.trace(rays)
        np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
        np.testing.assert_array_equal(g["prim"], p)
        assert p[0] == 0 and p[1] == INVALID
        scene.close()


def test_leaf_overflow_is_an_error(gpu_ctx):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]] * 130, np.uint32)
    s = H.Scene(gpu_ctx)
    with pytest.raises(H.HrtError) as ei:
        s.add_triangles(v, f)
    assert ei.value.status == 6
    assert s.geom_count == 0
    s.close()


def test_api_errors_and_empty_batches(gpu_ctx):
    import torch

    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(v, f)
    dummy = torch.zeros(8, dtype=torch.float32, device="cuda")
    out = torch.zeros(8, dtype=torch.int32, device="cuda")
    with pytest.raises(H.HrtError) as ei:
        s.trace_nearest(dummy, 1, out)
    assert ei.value.status == 4  # not committed
    s.add_instance(g)
    s.add_instance(g)
    s.commit()
    s.trace_nearest(dummy, 1, out)  # two instances: traced two-level
    with pytest.raises(H.HrtError) as ei:  # AO ray generation needs one identity instance
        s.rtao_make_rays(dummy, out, 1, 1, 1.0, 0, dummy)
    assert ei.value.status == 5
    s.clear_instances()
    s.add_instance(g)
    s.commit()
    s.trace_nearest(dummy, 0, out)  # empty batch is a no-op
    with pytest.raises(H.HrtError):
        s.add_instance(99)
    with pytest.raises(H.HrtError):
        s.add_triangles(v, np.array([[0, 1, 10_000_000]], np.uint32))
    s.close()


def test_query_one_matches_batch(gpu_ctx):
    v, f = W.icosphere(3)
    scene = make_scene(gpu_ctx, v, f, keep=False)
    rays = rays_for(v, 16, 900)
    g = trace(scene, rays, "full")
    for i in range(rays.shape[0]):
        t, prim, inst, geom, u, vv = scene.query_nearest_one(rays[i, 0:4], rays[i, 4:8])
        assert np.float32(t).view(np.uint32) == g["t"][i].view(np.uint32)
        assert prim == g["prim"][i]
        if prim != INVALID:
            assert inst == 0 and geom == 0
    scene.close()


def test_issue_probe_runs(gpu_ctx):
    """The issue-rate calibration kernel launches and completes on the context's device."""
    import torch

    gpu_ctx.issue_probe(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
