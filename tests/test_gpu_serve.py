"""GPU tests of the serving loop: results are the same alone and in any mix
(bit-exact with hrt_trace_nearest / hrt_trace_any), the HBM budget bounds
the device memory in flight, requests that arrive together share steps, and
bad requests are refused before anything runs."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def scenes_for(ctx):
    out = []
    for k, (v, f) in enumerate([W.soup(20_000, stream=7000), W.icosphere(5), W.heightfield(64, 32)[:2],
                                W.soup(3_000, stream=7001)]):
        s = H.Scene(ctx)
        s.add_instance(s.add_triangles(v, f))
        s.commit()
        out.append((s, v))
    return out


def request_rays(v, n, stream):
    lo, hi = v.min(0), v.max(0)
    r = W.uniform01(W.pcg32(n, 6, stream))
    org = lo - 0.5 * (hi - lo) + 2.0 * (hi - lo) * r[:, 0:3]
    tgt = lo + (hi - lo) * r[:, 3:6]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = H.pinned_array((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = 0.5 * float(np.max(hi - lo)) if stream % 3 == 0 else W.FLT_MAX
    return rays


def make_requests(scenes, count, seed):
    sizes = W.log_uniform_int(count, 1, 40_000, seed)
    reqs = []
    for i in range(count):
        sid = int(i * 7 % len(scenes))
        anyh = (i % 2) == 1
        rays = request_rays(scenes[sid][1], int(sizes[i]), seed + 1 + i)
        out = H.pinned_array((rays.shape[0],), np.uint8) if anyh else H.pinned_array((rays.shape[0], 2), np.uint32)
        reqs.append((sid, anyh, rays, out))
    return reqs


def solo_result(scene, rays, anyh):
    if anyh:
        return trace(scene, np.asarray(rays), "any")["hit"].astype(np.uint8)
    r = trace(scene, np.asarray(rays), "compact")
    return np.stack([r["t"].view(np.uint32), r["prim"]], axis=1)


def test_mix_equals_solo(gpu_ctx):
    scenes = scenes_for(gpu_ctx)
    srv = H.Server(gpu_ctx, 256 << 20, steps_in_flight=3)
    ids = [srv.add_scene(s) for s, _ in scenes]
    reqs = make_requests(scenes, 60, 8000)
    say(f"{len(reqs)} requests, {sum(r[2].shape[0] for r in reqs)} rays")
    rid = [srv.submit(ids[sid], rays, out, any_hit=anyh) for sid, anyh, rays, out in reqs]
    lat = [srv.wait(r) for r in rid]
    assert all(l >= q >= 0 for l, q in lat)
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["mean_requests_per_step"] > 1.0  # arrivals shared steps
    for sid, anyh, rays, out in reqs:
        np.testing.assert_array_equal(np.asarray(out), solo_result(scenes[sid][0], rays, anyh))
    # each request alone through the server gives the same result again
    for sid, anyh, rays, out in reqs[:12]:
        again = H.pinned_array(out.shape, out.dtype)
        srv.wait(srv.submit(ids[sid], rays, again, any_hit=anyh))
        np.testing.assert_array_equal(np.asarray(again), np.asarray(out))
    srv.close()
    for s, _ in scenes:
        s.close()


def test_budget_bounds_memory_in_flight(gpu_ctx):
    scenes = scenes_for(gpu_ctx)
    budget = 2 << 20  # 2 MiB: a few requests at a time
    srv = H.Server(gpu_ctx, budget, steps_in_flight=2)
    ids = [srv.add_scene(s) for s, _ in scenes]
    reqs = make_requests(scenes, 40, 9000)
    rid = [srv.submit(ids[sid], rays, out, any_hit=anyh) for sid, anyh, rays, out in reqs]
    for r in rid:
        srv.wait(r)
    st = srv.stats()
    assert 0 < st["peak_bytes_in_flight"] <= budget
    for sid, anyh, rays, out in reqs:
        np.testing.assert_array_equal(np.asarray(out), solo_result(scenes[sid][0], rays, anyh))
    # a request larger than the whole budget is refused
    big = request_rays(scenes[0][1], 70_000, 9999)
    big_out = H.pinned_array((70_000, 2), np.uint32)
    with pytest.raises(H.HrtError) as e:
        srv.submit(ids[0], big, big_out)
    assert e.value.status == 3
    srv.close()
    for s, _ in scenes:
        s.close()


def test_refuses_bad_requests(gpu_ctx):
    scenes = scenes_for(gpu_ctx)
    srv = H.Server(gpu_ctx, 16 << 20)
    sid = srv.add_scene(scenes[0][0])
    rays = request_rays(scenes[0][1], 100, 12)
    out = H.pinned_array((100, 2), np.uint32)
    with pytest.raises(H.HrtError):  # pageable host memory
        srv.submit(sid, np.asarray(rays).copy(), out)
    with pytest.raises(H.HrtError):  # unknown scene
        srv.submit(sid + 5, rays, out)
    inst = H.Scene(gpu_ctx)
    g = inst.add_triangles(*W.icosphere(1))
    inst.add_instance(g)
    inst.add_instance(g)
    inst.commit()
    with pytest.raises(H.HrtError):  # the server traces single-instance scenes
        srv.add_scene(inst)
    srv.wait(srv.submit(sid, rays, out))
    np.testing.assert_array_equal(np.asarray(out), solo_result(scenes[0][0], rays, False))
    srv.close()
    inst.close()
    for s, _ in scenes:
        s.close()


def test_device_requests_any_hit_tree_and_frozen_scenes(gpu_ctx):
    """Requests with device buffers (no staging) mix with pinned-host ones; any-hit requests on an
    AUTO_FP16 scene walk its any-hit tree and equal hrt_trace_any; a served scene refuses changes
    until the server is gone."""
    import torch

    v, f = W.soup(30_000, stream=7100)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=H.LAYOUT_AUTO_FP16)
    srv = H.Server(gpu_ctx, 8 << 20, steps_in_flight=2)
    sid = srv.add_scene(s)
    with pytest.raises(H.HrtError) as e:
        s.commit(layout=H.LAYOUT_BVH2)
    assert e.value.status == 1
    with pytest.raises(H.HrtError):
        s.add_triangles(v, f)
    reqs = []
    for i in range(24):
        n = 1000 + 997 * i
        anyh = i % 2 == 1
        rays = request_rays(v, n, 7200 + i)
        if i % 3 == 0:  # device buffers
            d_rays = torch.from_numpy(np.asarray(rays)).cuda()
            d_out = torch.zeros(n, dtype=torch.uint8, device="cuda") if anyh else \
                torch.zeros((n, 2), dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            reqs.append((anyh, rays, d_out, srv.submit(sid, d_rays, d_out, any_hit=anyh), d_rays))
        else:
            out = H.pinned_array((n,), np.uint8) if anyh else H.pinned_array((n, 2), np.uint32)
            reqs.append((anyh, rays, out, srv.submit(sid, rays, out, any_hit=anyh), None))
    for anyh, rays, out, rid, _ in reqs:
        srv.wait(rid)
        got = out.cpu().numpy().view(np.uint32) if hasattr(out, "cpu") and not anyh else \
            (out.cpu().numpy() if hasattr(out, "cpu") else np.asarray(out))
        np.testing.assert_array_equal(got, solo_result(s, rays, anyh))
    st = srv.stats()
    assert st["completed"] == len(reqs)
    # nearest-hit rays walked the node pairs, every any-hit ray the FP16 any-hit tree
    tr = srv.tree_rays()
    assert tr["bvh2"] == sum(r[1].shape[0] for r in reqs if not r[0])
    assert tr["bvh4_fp16"] == sum(r[1].shape[0] for r in reqs if r[0])
    assert tr["bvh4_fp32"] == 0
    srv.close()
    s.commit(layout=H.LAYOUT_BVH2)  # free again once the server is gone
    s.close()
