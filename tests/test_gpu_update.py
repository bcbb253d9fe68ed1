"""GPU parity for UpdateGeom_Triangles3f / UpdateInstance (hrt_update_*):
a refit keeps the hierarchy and matches the oracle's refit bit for bit; a
rebuild equals a fresh build of the new geometry; a moved instance traces
like the oracle scene with the new matrix."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O
from test_gpu_parity import rays_for
from test_oracle_refit import moved
from test_oracle_scene import affine, col_major, scene_rays

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF
H_BUILD_HIGH = 2   # BuildQuality (CrossRT.h:9-15)
H_BUILD_REFIT = 3
CASES = [("icosphere4", lambda: W.icosphere(4)), ("soup50k", lambda: W.soup(50_000, stream=W.STREAM_TEST)),
         ("rtao_small", lambda: W.rtao_scene(128, 64, 100))]


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
def test_refit_bit_exact(gpu_ctx, name, gen):
    v, f = gen()
    s = H.Scene(gpu_ctx, keep_build_arrays=True)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit()
    perm0 = s.dump(0, H.DUMP_PERM).copy()
    ob = O.OracleBVH(v, f)
    v2 = moved(v, 91)
    say(f"refit {name}: {f.shape[0]} tris")
    s.update_triangles(g, v2, quality=H_BUILD_REFIT)
    with pytest.raises(H.HrtError) as e:  # traced again only after a commit
        trace(s, rays_for(v2, 10, 1), "compact")
    assert e.value.status == 4
    s.commit()
    ob.refit(v2)
    np.testing.assert_array_equal(s.dump(0, H.DUMP_PERM), perm0)
    np.testing.assert_array_equal(s.dump(0, H.DUMP_LR_NODES), ob.lr_nodes())
    box = s.dump(0, H.DUMP_SCENE_BOX)
    bmin, bmax, _ = ob.scene_box()
    np.testing.assert_array_equal(box[0:4].view(np.uint32), bmin.view(np.uint32))
    np.testing.assert_array_equal(box[4:8].view(np.uint32), bmax.view(np.uint32))
    rays = rays_for(v2, 30_000, 4400)
    g_ = trace(s, rays, "full")
    t, p, u, vv = ob.trace(rays)
    np.testing.assert_array_equal(g_["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(g_["prim"], p)
    np.testing.assert_array_equal(g_["u"].view(np.uint32), u.view(np.uint32))
    # BVH4 nodes of the refitted tree
    s.commit(layout=H.LAYOUT_BVH4)
    q = s.dump(0, H.DUMP_QUADS32)
    oq = ob.quads(False)
    assert q.shape[0] == oq.shape[0]
    r4 = trace(s, rays, "compact")
    t4, p4, _, _ = ob.trace4(rays, O.MODE4_FP32)
    np.testing.assert_array_equal(r4["t"].view(np.uint32), t4.view(np.uint32))
    np.testing.assert_array_equal(r4["prim"], p4)
    # any-hit on the (rebuilt) cubic-grid tree of the new vertices
    s.commit(layout=H.LAYOUT_AUTO_FP16)
    anyr = rays.copy()
    anyr[:, 7] = -0.4 * float(np.max(v2.max(0) - v2.min(0)))
    _, pa, _, _ = ob.trace(anyr)
    pos = anyr.copy()
    pos[:, 7] = -pos[:, 7]
    np.testing.assert_array_equal(trace(s, pos, "any")["hit"], pa != INVALID)
    s.close()


def test_rebuild_equals_fresh_build(gpu_ctx):
    v, f = W.soup(40_000, stream=W.STREAM_TEST)
    s = H.Scene(gpu_ctx, keep_build_arrays=True)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit()
    v2 = moved(v, 92)
    f2 = f[:30_000]
    s.update_triangles(g, v2, f2, quality=H_BUILD_HIGH)  # smaller: allowed
    s.commit()
    fresh = H.Scene(gpu_ctx, keep_build_arrays=True)
    fresh.add_instance(fresh.add_triangles(v2, f2))
    fresh.commit()
    for what in (H.DUMP_SIZES, H.DUMP_PAIRS, H.DUMP_PERM, H.DUMP_LR_NODES):
        np.testing.assert_array_equal(s.dump(g, what), fresh.dump(0, what))
    rays = rays_for(v2, 20_000, 4500)
    a, b = trace(s, rays, "compact"), trace(fresh, rays, "compact")
    np.testing.assert_array_equal(a["prim"], b["prim"])
    s.close()
    fresh.close()


def test_update_errors(gpu_ctx):
    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit()
    bigger = np.concatenate([v, v[:3]])
    with pytest.raises(H.HrtError):  # may not grow
        s.update_triangles(g, bigger, quality=H_BUILD_REFIT)
    with pytest.raises(H.HrtError):  # a refit keeps the triangle count
        s.update_triangles(g, v, f[:-1], quality=H_BUILD_REFIT)
    with pytest.raises(H.HrtError):
        s.update_triangles(g + 1, v, quality=H_BUILD_REFIT)
    with pytest.raises(H.HrtError):
        s.update_instance(5, col_major(np.eye(4)))
    s.close()


def test_update_instance_moves_it(gpu_ctx):
    iv, if_ = W.icosphere(3, 0.3, (0.0, 0.0, 0.0))
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(iv, if_)
    mats = []
    for k in range(6):
        M = affine(60 + k)
        M[:3, 3] = [0.6 * k, 0.2, 0.3 * (k % 2)]
        mats.append(col_major(M))
        s.add_instance(g, mats[-1])
    s.commit()
    M = affine(99)
    M[:3, 3] = [1.0, 0.8, 0.5]
    s.update_instance(3, col_major(M))
    s.commit()
    o = O.OracleScene()
    o.add_geom(iv, if_)
    for k, m in enumerate(mats):
        o.add_instance(0, col_major(M) if k == 3 else m)
    o.commit()
    rays = scene_rays(30_000, 4600, lo=(-0.5, -0.5, -0.5), hi=(3.5, 1.5, 1.0))
    r = trace(s, rays, "full")
    ref = o.trace(rays)
    np.testing.assert_array_equal(r["t"].view(np.uint32), ref[0].view(np.uint32))
    np.testing.assert_array_equal(r["prim"], ref[1])
    np.testing.assert_array_equal(r["inst"], ref[2])
    assert (r["inst"] == 3).sum() > 50
    s.close()
