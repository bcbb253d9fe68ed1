"""Parity at the full size of the batched, precision and serving workloads.

  precision  4,194,304-triangle soup x 4,194,304 incoherent rays, BVH4 with
             FP32 and with FP16 boxes: hits bit-identical to each other and to
             the oracle's BVH4 traversal (BVH4HalfRT.cpp:73-256, nearest by
             (t, primId)) on every ray; t against FP64 Moller-Trumbore
             reported.
  batched    256 soups of 1k-200k triangles with 4k-256k rays each, built and
             traced in one graph launch: every scene's hits equal the oracle's
             BVH2CommonRT restatement (BVH2CommonRT.cpp:228-279) on every ray.
  serving    16 resident scenes (50k-2M triangles, soups and heightfields) and
             a Poisson stream of 200 requests of 1k-1M rays (one of 2^20):
             every request's result in the mix equals its solo trace, and the
             solo trace equals the oracle on a sample of every request.
"""
import time

import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def host_threads() -> int:
    import os

    return len(os.sched_getaffinity(0))


def scene_with(ctx, v, f, layout):
    s = H.Scene(ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    return s


@pytest.mark.timeout(900)
def test_precision_config_full_size(gpu_ctx):
    n = 4_194_304
    v, f = W.soup(n, stream=W.STREAM_SOUP)
    rays = W.incoherent_rays(n, stream=W.STREAM_RAYS)
    say(f"precision: {n} tris x {n} rays")
    # the timed layout (FP16 nodes of the clustered tree), its FP32 twin, and
    # the reference tree's FP16 and FP32 BVH4
    res = {}
    for lay in (H.LAYOUT_PLOC_FP16, H.LAYOUT_PLOC, H.LAYOUT_BVH4_FP16, H.LAYOUT_BVH4):
        sc = scene_with(gpu_ctx, v, f, lay)
        try:
            res[lay] = trace(sc, rays, "full")
        finally:
            sc.close()
    a = res[H.LAYOUT_PLOC_FP16]
    for lay in (H.LAYOUT_PLOC, H.LAYOUT_BVH4_FP16, H.LAYOUT_BVH4):
        b = res[lay]
        for k in ("t", "u", "v"):
            np.testing.assert_array_equal(a[k].view(np.uint32), b[k].view(np.uint32), err_msg=f"{k} layout {lay}")
        np.testing.assert_array_equal(a["prim"], b["prim"], err_msg=f"layout {lay}")
    del res
    say(f"precision: FP16 = FP32 on both trees, all rays; oracle BVH4 on {host_threads()} threads")
    t0 = time.perf_counter()
    ob = O.OracleBVH(v, f)
    t, p, u, vv = ob.trace4(rays, O.MODE4_FP32, threads=host_threads())
    say(f"precision: oracle {time.perf_counter() - t0:.1f} s")
    np.testing.assert_array_equal(a["prim"], p)
    np.testing.assert_array_equal(a["t"].view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(a["u"].view(np.uint32), u.view(np.uint32))
    np.testing.assert_array_equal(a["v"].view(np.uint32), vv.view(np.uint32))
    hit = p != INVALID
    t64 = ob.t_fp64(rays[hit], p[hit])
    rel = np.abs(t[hit].astype(np.float64) - t64) / np.abs(t64)
    say(f"precision: {int(hit.sum())} hits; t vs FP64 rel. error max {rel.max():.3e} mean {rel.mean():.3e} "
        f"median {np.median(rel):.3e}")
    assert np.median(rel) < 1e-6


@pytest.mark.timeout(900)
def test_batched_config_full_size(gpu_ctx):
    import torch

    scenes = W.batched_workload()
    ntris = sum(x[1].shape[0] for x in scenes)
    nrays = sum(x[2].shape[0] for x in scenes)
    say(f"batched: {len(scenes)} scenes, {ntris} tris, {nrays} rays")
    b = H.Batch(gpu_ctx, [x[0].shape[0] for x in scenes], [x[1].shape[0] for x in scenes],
                [x[2].shape[0] for x in scenes])
    b.upload(scenes)
    b.autotune(reps=1)
    b.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    hits = b.hits()
    b.close()
    threads = host_threads()
    bad = []
    for s, (v, f, rays) in enumerate(scenes):
        t, p, _, _ = O.OracleBVH(v, f).trace(rays, threads=threads)
        hb = hits[b.ray_off[s]:b.ray_off[s + 1]]
        if not (np.array_equal(hb[:, 0], t.view(np.uint32)) and np.array_equal(hb[:, 1], p)):
            bad.append(s)
    assert not bad, f"scenes whose hits differ from the oracle: {bad[:20]}"


def _serving_pools(ctx, scenes):
    """Per scene: 1024x1024 eye rays and one AO ray per eye hit (made on the
    device, as the serving workload's request pools are)."""
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    pools = []
    for k, (sc, (v, f)) in enumerate(scenes):
        eye, target, radius = W.serving_camera(v)
        eye_rays = W.pinhole_rays(1024, 1024, eye=eye, target=target, fov_y_deg=50.0)
        n = eye_rays.shape[0]
        d_eye = torch.from_numpy(eye_rays).cuda()
        d_hit = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        d_ao = torch.empty((n, 8), dtype=torch.float32, device="cuda")
        sc.trace_nearest(d_eye, n, d_hit, H.HIT_COMPACT, stream=stream)
        sc.rtao_make_rays(d_eye, d_hit, n, 1, radius, 0x5EED + k, d_ao, stream=stream)
        torch.cuda.synchronize()
        pool = {"eye": H.pinned_array((n, 8), np.float32), "ao": H.pinned_array((n, 8), np.float32)}
        pool["eye"][:] = eye_rays
        pool["ao"][:] = d_ao.cpu().numpy()
        pools.append(pool)
    return pools


@pytest.mark.timeout(900)
def test_serving_config_full_size(gpu_ctx):
    geo = W.serving_scenes()
    say(f"serving: {len(geo)} scenes, {sum(f.shape[0] for _, f in geo)} tris")
    scenes = []
    for v, f in geo:
        s = H.Scene(gpu_ctx)
        s.add_instance(s.add_triangles(v, f))
        s.commit(layout=H.LAYOUT_AUTO_FP16)
        scenes.append((s, (v, f)))
    pools = _serving_pools(gpu_ctx, scenes)
    req = W.serving_requests(200, 2000.0, stream=W.STREAM_SERVE + 150)
    big = int(np.argmax(req["size"]))
    req["size"][big] = 1 << 20  # one request of a whole pool
    req["start"][big] = 0
    say(f"serving: {len(req['size'])} requests, {int(req['size'].sum())} rays")
    srv = H.Server(gpu_ctx, 4 << 30, steps_in_flight=3)
    ids = [srv.add_scene(s) for s, _ in scenes]
    outs, rids = [], []
    t0 = time.perf_counter()
    for i in range(len(req["size"])):
        while time.perf_counter() - t0 < req["arrival_s"][i]:
            pass
        k, a, n, anyh = int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]), bool(req["any_hit"][i])
        rays = pools[k]["ao" if anyh else "eye"][a:a + n]
        out = H.pinned_array((n,), np.uint8) if anyh else H.pinned_array((n, 2), np.uint32)
        outs.append(out)
        rids.append(srv.submit(ids[k], rays, out, any_hit=anyh))
    lat = np.array([srv.wait(r)[0] for r in rids])
    st = srv.stats()
    srv.close()
    say(f"serving: p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms, "
        f"{st['mean_requests_per_step']:.2f} requests per step")
    assert st["completed"] == len(rids)
    # mix = solo on every ray of every request
    for i in range(len(rids)):
        k, a, n, anyh = int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]), bool(req["any_hit"][i])
        rays = np.asarray(pools[k]["ao" if anyh else "eye"][a:a + n])
        solo = trace(scenes[k][0], rays, "any" if anyh else "compact")
        if anyh:
            np.testing.assert_array_equal(np.asarray(outs[i]).astype(bool), solo["hit"], err_msg=f"request {i}")
        else:
            got = np.asarray(outs[i])
            np.testing.assert_array_equal(got[:, 0], solo["t"].view(np.uint32), err_msg=f"request {i}")
            np.testing.assert_array_equal(got[:, 1], solo["prim"], err_msg=f"request {i}")
    # solo = oracle on 512 rays of every request
    threads = host_threads()
    oracles = {}
    for i in range(len(rids)):
        k, a, n, anyh = int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]), bool(req["any_hit"][i])
        if k not in oracles:
            oracles[k] = O.OracleBVH(*scenes[k][1])
        pick = np.unique(np.linspace(0, n - 1, min(n, 512)).astype(np.int64))
        rays = np.asarray(pools[k]["ao" if anyh else "eye"][a:a + n])[pick].copy()
        got = np.asarray(outs[i])[pick]
        if anyh:
            rays[:, 7] = -np.abs(rays[:, 7])
            _, p, _, _ = oracles[k].trace(rays, threads=threads)
            np.testing.assert_array_equal(got.astype(bool), p != INVALID, err_msg=f"request {i}")
        else:
            t, p, _, _ = oracles[k].trace(rays, threads=threads)
            np.testing.assert_array_equal(got[:, 0], t.view(np.uint32), err_msg=f"request {i}")
            np.testing.assert_array_equal(got[:, 1], p, err_msg=f"request {i}")
    for s, _ in scenes:
        s.close()
