"""GPU parity of the watertight triangle test (HRT_TRI_WATERTIGHT) against the
oracle's restatement of Woop, Benthin and Wald (JCGT 2(1), 2013), on every
single-scene layout: nearest hits (t, primId, u, v) bit-exact with the
oracle's traversal of the same tree under the same test (the BVH2 order and
its first strictly closer hit; the BVH4 (t, primId) rule), any-hit answers
equal to brute force, no ray escaping a closed mesh, the fused AO pass equal
to the two-step one, and snapshots keeping the test."""
import numpy as np
import pytest

import heavyrt_amd as H
from conftest import vsgf_path
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_gpu_parity import rays_for
from test_oracle_watertight import SPHERES, closed_sphere_rays

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF
LAYOUTS = [H.LAYOUT_BVH2, H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO_FP16]
LAYOUT_IDS = ["bvh2", "bvh4", "bvh4_fp16", "auto_fp16"]


def _cases():
    for m in ("cube", "torus18", "helix_mid"):
        yield m, lambda m=m: W.load_vsgf(vsgf_path(m))
    yield "eye_scene", W.eye_scene
    yield "soup_200k", lambda: W.soup(200_000, stream=W.STREAM_TEST + 7)


CASES = list(_cases())


def wt_scene(ctx, v, f, layout):
    s = H.Scene(ctx)
    s.add_instance(s.add_triangles(v, f))
    s.set_triangle_test(H.TRI_WATERTIGHT)
    s.commit(layout=layout)
    return s


def oracle_nearest(ob, rays, layout):
    with O.triangle_test(O.TRI_WATERTIGHT):
        if layout in (H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16):
            return ob.trace4(rays, O.MODE4_FP16 if layout == H.LAYOUT_BVH4_FP16 else O.MODE4_FP32)
        return ob.trace(rays)


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_watertight_nearest_and_any_equal_oracle(gpu_ctx, name, gen, layout):
    v, f = gen()
    rays = rays_for(v, 20_000, 1700)
    say(f"watertight {name} layout {layout}")
    s = wt_scene(gpu_ctx, v, f, layout)
    ob = O.OracleBVH(v, f)
    g = trace(s, rays, "full")
    t, p, u, vv = oracle_nearest(ob, rays, layout)
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), vv.view(np.uint32))
    anyr = rays.copy()
    anyr[:, 7] = 0.3 * float(np.max(v.max(0) - v.min(0)))
    ga = trace(s, anyr, "any")["hit"]
    counts = O.tri_hit_counts(O.TRI_WATERTIGHT, v, f, anyr)
    np.testing.assert_array_equal(ga, counts > 0)
    s.close()


@pytest.mark.parametrize("level,radius,center", SPHERES, ids=["unit", "tiny", "far", "offset"])
@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_no_ray_escapes_a_closed_mesh(gpu_ctx, layout, level, radius, center):
    """Rays aimed exactly at every vertex and edge midpoint of a closed icosphere from inside it:
    every one hits, nearest hits equal the oracle's traversal bit for bit (the conservative box test
    included), and every any-hit query answers yes."""
    v, f, rays = closed_sphere_rays(level, radius, center)
    s = wt_scene(gpu_ctx, v, f, layout)
    g = trace(s, rays, "full")
    assert np.all(g["prim"] != INVALID)
    t, p, _, _ = oracle_nearest(O.OracleBVH(v, f), rays, layout)
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    anyr = rays.copy()
    anyr[:, 7] = -anyr[:, 7]
    assert np.all(trace(s, anyr, "any")["hit"])
    s.close()


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16], ids=["fp32", "fp16"])
def test_watertight_sorted_waves(gpu_ctx, layout):
    """A batch large enough to be sorted into coherent waves, under the watertight test."""
    v, f = W.soup(150_000, stream=W.STREAM_TEST + 8)
    rays = W.incoherent_rays(1_100_000, stream=W.STREAM_TEST + 9)
    s = wt_scene(gpu_ctx, v, f, layout)
    g = trace(s, rays, "full")
    s.close()
    t, p, u, vv = oracle_nearest(O.OracleBVH(v, f), rays, layout)
    np.testing.assert_array_equal(g["prim"], p)
    np.testing.assert_array_equal(g["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g["u"].view(np.uint32), u.view(np.uint32))


@pytest.mark.parametrize("layout", LAYOUTS, ids=LAYOUT_IDS)
def test_watertight_ao_pass(gpu_ctx, layout):
    """hrt_rtao_trace under the watertight test equals hrt_rtao_make_rays + hrt_trace_any byte for byte,
    and the answers equal brute force with the same test."""
    import torch

    v, f = W.rtao_scene(128, 64, 60)
    s = wt_scene(gpu_ctx, v, f, layout)
    prim = W.rtao_primary_rays(64, 64)
    npix, spp = prim.shape[0], 4
    tmax = 0.1 * W.scene_diagonal(v)
    stream = torch.cuda.current_stream().cuda_stream
    dprim = torch.from_numpy(prim).cuda()
    dhits = torch.zeros((npix, 2), dtype=torch.int32, device="cuda")
    s.trace_nearest(dprim, npix, dhits, H.HIT_COMPACT, stream=stream)
    dao = torch.empty((npix * spp, 8), dtype=torch.float32, device="cuda")
    s.rtao_make_rays(dprim, dhits, npix, spp, tmax, 77, dao, stream=stream)
    a = torch.zeros(npix * spp, dtype=torch.uint8, device="cuda")
    s.trace_any(dao, npix * spp, a, stream=stream)
    b = torch.full((npix * spp,), 9, dtype=torch.uint8, device="cuda")
    s.rtao_trace(dprim, dhits, npix, spp, tmax, 77, b, stream=stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    ao = dao.cpu().numpy()
    counts = O.tri_hit_counts(O.TRI_WATERTIGHT, v, f, ao)
    np.testing.assert_array_equal(a.cpu().numpy().astype(bool), counts > 0)
    s.close()


def test_watertight_snapshot_and_limits(gpu_ctx):
    v, f = W.icosphere(4)
    rays = rays_for(v, 5000, 1800)
    s = wt_scene(gpu_ctx, v, f, H.LAYOUT_BVH4_FP16)
    r = H.Scene.deserialize(gpu_ctx, s.serialize())
    a, b = trace(s, rays, "full"), trace(r, rays, "full")
    for k in ("t", "prim", "u", "v"):
        np.testing.assert_array_equal(a[k], b[k])
    r.close()
    # back to the reference's test on the next commit
    s.set_triangle_test(H.TRI_MT)
    s.commit(layout=H.LAYOUT_BVH4_FP16)
    t, p, _, _ = O.OracleBVH(v, f).trace4(rays, O.MODE4_FP16)
    np.testing.assert_array_equal(trace(s, rays, "compact")["t"].view(np.uint32), t.view(np.uint32))
    with pytest.raises(H.HrtError):
        s.set_triangle_test(7)
    s.set_triangle_test(H.TRI_WATERTIGHT)
    s.commit()
    srv = H.Server(gpu_ctx, 16 << 20)
    with pytest.raises(H.HrtError) as e:  # the server keeps the reference's test
        srv.add_scene(s)
    assert e.value.status == 5
    srv.close()
    s.close()
    inst = H.Scene(gpu_ctx)
    g = inst.add_triangles(v, f)
    inst.add_instance(g)
    inst.add_instance(g, np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 3, 0, 0, 1], np.float32))
    inst.set_triangle_test(H.TRI_WATERTIGHT)
    with pytest.raises(H.HrtError) as e:
        inst.commit()
    assert e.value.status == 5
    inst.close()
