"""Ray-query throughput on one MI355X: the five workloads of the ray-query path.

`python bench.py` runs every workload and prints ONE JSON line.  Its `value`
is the headline RTAO workload; the other four are sub-objects of `configs`:

  rtao       (value) a 1024x512-quad heightfield (value noise, amp 0.2) plus
             2,000 rotated boxes = 1,072,576 triangles; 1024x1024 primary
             rays traced nearest-hit on the LBVH2 in the reference's order
             (bit-exact with BVH2CommonRT); 16 cosine-weighted AO rays per hit
             made on the device inside the any-hit traversal (tMax = 0.1 x the
             scene diagonal), traced on the FP16 BVH4 nodes of the any-hit tree
  eye        512x512 pinhole rays made on the device, nearest hit on the
             icosphere (81,920 tris) + ground quad; also the per-ray drop-in
             path: one RayQuery_NearestHit per ray from an OpenMP loop
  precision  4,194,304-triangle random soup, 4,194,304 incoherent nearest-hit
             rays, BVH4 with FP16 boxes rounded outward vs FP32 boxes; t error
             against FP64; the watertight triangle test beside the reference's
  batched    256 random soups (1k-200k tris) with 4k-256k rays each, every
             build and every trace in one hipGraph launch, trace kernel
             picked by the autotuner
  serving    16 resident scenes (50k-2M tris), Poisson requests of 1k-1M rays,
             half nearest-hit eye rays, half any-hit AO rays; p50/p99 latency
             per offered rate, the plain per-request path on the same stream
             and device-resident requests

Every scene and ray batch is resident in HBM before its timed region; times
come from HIP events on the stream the work runs on.

`roofline` (per workload, for its dominant kernel) is an HBM fraction:
  achieved  HBM bytes per launch / mean launch time.  The bytes are measured
            in this run by rocprofv3 (FETCH_SIZE and WRITE_SIZE of the same
            kernel in a child process, each scaled by what the counter reads
            for a 1 GiB device copy); where the counters cannot be collected,
            the algorithmic bytes below stand in and `traffic` is null.
  algorithmic_bytes  what one launch must move at least: its rays (or AO
            frames) in, its answers out and the tree's nodes and triangles
            once.
  bus_load  node and triangle bytes the traversal requests per launch (the
            reference's BLB metric, core/CrossRT.h:52, BVH2CommonRT.cpp:
            69-70,125: 64 B per BVH2 pair or FP16 BVH4 node, 128 B per FP32
            BVH4 node, 60 B per triangle tested, counted by the engine's
            counting variant), against the L2-to-CU fill bandwidth (64 B per
            clock per CU, 256 CUs, 2.4 GHz); these are mostly L2 hits.
`cpu_baseline` is the CPU oracle (a C++ restatement of the reference's
BVH2CommonRT / BVH4HalfRT traversal) on this host, on every CPU this process
may use, plus one thread, plus one thread with the reference's metrics on
(CrossRT.h:31 defines ENABLE_METRICS as shipped).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Mrays/s closest/any-hit on 1 MI355X, % HBM roofline; p50/p99 query latency"
HBM_PEAK_GBPS = 8000.0
L2_FILL_PEAK_GBPS = 64 * 256 * 2.4  # 64 B/clk/CU x 256 CUs x 2.4 GHz
SPP = 16
SEED = 0x2545F4914F6CDD1D
INVALID = 0xFFFFFFFF
WORKLOADS = ("rtao", "eye", "precision", "batched", "serving")
# kernels whose HBM bytes the counter child measures, by workload (substrings
# of the demangled names rocprofv3 reports)
COUNTER_KERNELS = {"rtao": ("AoRays",), "eye": ("k_trace_simple",), "precision": ("k_trace4_simple",),
                   "batched": ("k_b_trace",), "serving": ("k_serve", "ServeRays")}
# workloads whose counters are summed over every matching dispatch of one
# replayed request stream (the others: mean of the last three launches)
COUNTER_SUM = ("serving",)
# the request stream the serving counter pass replays: the default run's top
# rate (four times the highest host rate), device-resident requests
SERVE_REPLAY_QPS, SERVE_REPLAY_REQUESTS, SERVE_REPLAY_LEVEL = 48000.0, 1500, 3


def counter_match(w: str, name: str) -> bool:
    return any(p in name for p in COUNTER_KERNELS[w])


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# host facts and helpers
# ---------------------------------------------------------------------------
def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, p = fh.read().split()
            if q != "max":
                quota = int(q) / int(p)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(math.ceil(quota))))
    return {"model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "cgroup_cpu_quota": quota, "threads": usable}


def timed(stream, fn, reps: int) -> np.ndarray:
    """Milliseconds of each of `reps` calls of fn, from HIP events on `stream`."""
    import torch

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    torch.cuda.synchronize()
    for k in range(reps):
        ev[k].record(stream)
        fn()
    ev[reps].record(stream)
    torch.cuda.synchronize()
    return np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(reps)])


def lat(ms: np.ndarray) -> dict:
    return {"p50": round(float(np.percentile(ms, 50)), 4), "p99": round(float(np.percentile(ms, 99)), 4)}


def roofline(kernel: str, ms: float, algorithmic: float, bus_load: float | None) -> dict:
    """The roofline object of one kernel; `traffic` is filled from counters later when available."""
    ach = algorithmic / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBPS, 5), "traffic": None,
           "traffic_source": "not measured: achieved uses the algorithmic bytes",
           "mean_launch_ms": round(ms, 4), "algorithmic_bytes": float(algorithmic),
           "algorithmic_frac": round(ach / HBM_PEAK_GBPS, 5)}
    if bus_load is not None:
        bl = bus_load / (ms * 1e-3) / 1e9
        out["bus_load"] = {"bytes": float(bus_load), "achieved": round(bl, 1), "peak": round(L2_FILL_PEAK_GBPS, 1),
                           "unit": "GB/s", "frac": round(bl / L2_FILL_PEAK_GBPS, 4)}
    return out


def line(value, ms, steps, warmup, config: dict, extra: dict) -> dict:
    out = {"metric": METRIC, "value": value, "unit": "Mrays/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


def cpu_baseline(label: str, full, single, shipped, cpu: dict) -> dict:
    """full(threads) / single() / shipped() -> (rays, seconds) of the oracle on this host."""
    rays, sec = full(cpu["threads"])
    out = {"value": round(rays / sec / 1e6, 3), "unit": "Mrays/s", "cores": cpu["threads"], "kind": "port",
           "sample": f"{label}: {rays} rays in {sec:.2f} s", "model": cpu["model"], "nproc": cpu["nproc"],
           "affinity_cpus": cpu["affinity_cpus"], "cgroup_cpu_quota": cpu["cgroup_cpu_quota"]}
    if single is not None:
        r1, s1 = single()
        out["single_core"] = {"value": round(r1 / s1 / 1e6, 4), "rays": r1, "seconds": round(s1, 3)}
    if shipped is not None:
        r2, s2, what = shipped()
        out["as_shipped"] = {"value": round(r2 / s2 / 1e6, 4), "rays": r2, "seconds": round(s2, 3),
                             "note": what}
    return out


def clock(fn):
    t0 = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t0


# ---------------------------------------------------------------------------
# rtao (the headline workload)
# ---------------------------------------------------------------------------
def bench_rtao(args, cpu) -> dict:
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    layouts = {"bvh2": H.LAYOUT_BVH2, "bvh4": H.LAYOUT_BVH4, "bvh4_fp16": H.LAYOUT_BVH4_FP16,
               "bvh4_fp16_anytree": H.LAYOUT_AUTO_FP16}
    node_bytes = {"bvh2": 64, "bvh4": 128, "bvh4_fp16": 64, "bvh4_fp16_anytree": 64}
    log("rtao: generating the scene")
    verts, faces = W.rtao_scene()
    prim_rays = W.rtao_primary_rays(order="superblock")
    npix = prim_rays.shape[0]
    tmax = 0.1 * W.scene_diagonal(verts)
    n_ao = npix * SPP
    ctx = H.Context(0)
    scene = H.Scene(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene.add_instance(scene.add_triangles(verts, faces))
    build_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    scene.commit(layout=H.LAYOUT_BVH4)
    collapse_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    scene.commit(layout=H.LAYOUT_AUTO_FP16)
    any_tree_ms = (time.perf_counter() - t0) * 1e3
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    d_prim = torch.from_numpy(prim_rays).cuda()
    d_hits = torch.zeros((npix, 2), dtype=torch.int32, device="cuda")
    d_vis = torch.zeros(n_ao, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        scene.trace_nearest(d_prim, npix, d_hits, H.HIT_COMPACT, stream=sptr)
        if ev is not None:
            ev[1].record(stream)
        scene.rtao_trace(d_prim, d_hits, npix, SPP, tmax, SEED, d_vis, stream=sptr)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    step_ms = np.array([e[0].elapsed_time(e[2]) for e in events])
    near_ms = np.array([e[0].elapsed_time(e[1]) for e in events])
    ao_ms = np.array([e[1].elapsed_time(e[2]) for e in events])
    rays_per_step = npix + n_ao
    value = rays_per_step / (ms_per_step * 1e-3) / 1e6
    log(f"rtao: step {ms_per_step:.3f} ms, nearest {near_ms.mean():.3f} ms, fused AO pass {ao_ms.mean():.3f} ms")

    # the same AO rays written to HBM, traced by hrt_trace_any: the counters, the
    # other layouts and the CPU baseline read them
    d_vis_timed = d_vis.clone()
    d_ao = torch.empty((n_ao, 8), dtype=torch.float32, device="cuda")
    scene.rtao_make_rays(d_prim, d_hits, npix, SPP, tmax, SEED, d_ao, stream=sptr)
    d_vis2 = torch.zeros_like(d_vis)
    two_ms = timed(stream, lambda: scene.trace_any(d_ao, n_ao, d_vis2, stream=sptr), 5)
    ao_identical = bool(torch.equal(d_vis2, d_vis_timed))
    # the whole 02_rtao frame on the device
    cam_pi, cam_wi = W.eye_camera(1024, 1024, eye=W.RTAO_EYE, target=W.RTAO_TARGET, fov_y_deg=W.RTAO_FOV)
    d_img = torch.zeros(npix, dtype=torch.int32, device="cuda")
    frame_ms = timed(stream, lambda: scene.rtao_render(cam_pi, cam_wi, 1024, 1024, SPP, tmax, SEED, d_img,
                                                       stream=sptr), 5)

    def counts(layout, any_hit):
        scene.commit(layout=layout)
        st = torch.zeros(((n_ao if any_hit else npix), 2), dtype=torch.int32, device="cuda")
        if any_hit:
            chk = torch.zeros_like(d_vis)
            scene.trace_any(d_ao, n_ao, chk, stream=sptr, stats=st)
        else:
            chk = torch.zeros_like(d_hits)
            scene.trace_nearest(d_prim, npix, chk, H.HIT_COMPACT, stream=sptr, stats=st)
        torch.cuda.synchronize()
        return st.cpu().numpy().view(np.uint32).astype(np.int64), bool(torch.equal(chk, d_vis_timed if any_hit
                                                                                   else d_hits))

    sp, same_p = counts(H.LAYOUT_AUTO_FP16, False)
    sa, same_a = counts(H.LAYOUT_AUTO_FP16, True)
    variants = {}
    for name, lay in layouts.items():
        scene.commit(layout=lay)
        out = torch.zeros_like(d_vis)
        ms = timed(stream, lambda: scene.trace_any(d_ao, n_ao, out, stream=sptr), 5)
        variants[name] = {"any_hit_ms": round(float(ms.mean()), 4),
                          "any_hit_mrays_s": round(n_ao / (ms.mean() * 1e-3) / 1e6, 1),
                          "identical_to_timed": bool(torch.equal(out, d_vis_timed))}
    scene.commit(layout=H.LAYOUT_AUTO_FP16)
    sizes = scene.dump(0, H.DUMP_SIZES)
    any_sizes = scene.dump(0, H.DUMP_SIZES | H.DUMP_ANY_TREE)
    ntris = int(faces.shape[0])
    # algorithmic bytes: the fused pass reads one 32-byte frame per pixel and
    # writes one byte per AO ray; the any-hit tree's FP16 nodes and triangles once
    alg_any = 32.0 * npix + n_ao + 64.0 * int(any_sizes[4]) + 48.0 * ntris
    blb_any = float(np.sum(32.0 / SPP + 1 + 64 * sa[:, 0] + 60 * sa[:, 1]))
    alg_near = 40.0 * npix + 64.0 * int(sizes[3]) + 48.0 * ntris
    blb_near = float(np.sum(40 + 64 * sp[:, 0] + 60 * sp[:, 1]))
    res = line(round(value, 2), round(ms_per_step, 4), args.steps, args.warmup,
               {"workload": "rtao: 1024x1024 primary nearest-hit (LBVH2, reference order; 8x8-superblock pixel "
                            "order) + 16 AO any-hit rays per hit made inside the any-hit traversal (FP16 BVH4 of "
                            "the any-hit tree), 1,072,576-tri heightfield+boxes",
                "triangles": ntris, "primary_rays": int(npix), "ao_rays": int(n_ao), "ao_tmax": round(tmax, 6)},
               {"roofline": roofline("any-hit traversal, fused AO pass (k_trace4_any_persistent, FP16 BVH4 of the "
                                     "any-hit tree)", float(ao_ms.mean()), alg_any, blb_any),
                "nearest_roofline": roofline("nearest-hit traversal (k_trace_simple, BVH2, reference order)",
                                             float(near_ms.mean()), alg_near, blb_near),
                "nearest_mrays_s": round(npix / (near_ms.mean() * 1e-3) / 1e6, 2),
                "anyhit_mrays_s": round(n_ao / (ao_ms.mean() * 1e-3) / 1e6, 2),
                "latency_ms": lat(step_ms),
                "visits_per_ray": {"primary_pairs": round(float(sp[:, 0].mean()), 3),
                                   "ao_anytree_nodes": round(float(sa[:, 0].mean()), 3)},
                "tris_per_ray": {"primary": round(float(sp[:, 1].mean()), 3), "ao": round(float(sa[:, 1].mean()), 3)},
                "build_ms": round(build_ms, 3), "bvh4_collapse_ms": round(collapse_ms, 3),
                "any_hit_tree_build_ms": round(any_tree_ms, 3),
                "counting_variant_identical": same_p and same_a,
                "ao_pass": {"fused_ms": round(float(ao_ms.mean()), 4), "two_step_any_hit_ms": round(float(
                    two_ms.mean()), 4), "identical": ao_identical},
                "device_frame": {"frame_ms": round(float(frame_ms.mean()), 4),
                                 "mrays_s": round(rays_per_step / (frame_ms.mean() * 1e-3) / 1e6, 1)},
                "any_hit_layouts": variants, "cpu_baseline": None})
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        log(f"rtao: cpu baseline on {cpu['threads']} threads")
        ob = O.OracleBVH(verts, faces)
        ao_host = d_ao.cpu().numpy()
        ao_host[:, 7] = -np.abs(ao_host[:, 7])  # any-hit: negative tFar
        hits_host = d_hits.cpu().numpy().view(np.uint32)
        vis_host = d_vis_timed.cpu().numpy().astype(bool)
        mism = {}

        def full(threads):
            (tp, pp, _, _), s1 = clock(lambda: ob.trace(prim_rays, threads=threads))
            (_, pa, _, _), s2 = clock(lambda: ob.trace(ao_host, threads=threads))
            mism["n"] = int(np.sum(pp != hits_host[:, 1]) + np.sum(tp.view(np.uint32) != hits_host[:, 0]) +
                            np.sum((pa != INVALID) != vis_host))
            return rays_per_step, s1 + s2

        # one thread and "as shipped": the same mix on a sample (1/16 of the pixels with all their AO rays)
        sel = np.arange(0, npix, 16)
        ao_sel = (sel[:, None] * SPP + np.arange(SPP)[None, :]).reshape(-1)

        def single():
            _, s1 = clock(lambda: ob.trace(prim_rays[sel], threads=1))
            _, s2 = clock(lambda: ob.trace(ao_host[ao_sel], threads=1))
            return int(sel.size + ao_sel.size), s1 + s2

        def shipped():
            _, s1 = clock(lambda: ob.trace_metrics(prim_rays[sel]))
            _, s2 = clock(lambda: ob.trace_metrics(ao_host[ao_sel]))
            return int(sel.size + ao_sel.size), s1 + s2, "one thread, metrics on (CrossRT.h:31)"

        res["cpu_baseline"] = cpu_baseline("one full step (1 M primary nearest + 16.8 M AO any-hit, BVH2CommonRT "
                                           "restatement, OpenMP dynamic,256)", full, single, shipped, cpu)
        res["parity_mismatches_vs_oracle"] = mism["n"]
    scene.close()
    ctx.close()
    return res


# ---------------------------------------------------------------------------
# eye rays (and the per-ray drop-in path)
# ---------------------------------------------------------------------------
def per_ray_loop(v, f, rays, threads) -> dict | None:
    """The C++ harness: one RayQuery_NearestHit per ray through CreateSceneRT, from an OpenMP loop."""
    from heavyrt_amd import build as B

    if not os.path.exists(B.HARNESS_BIN):
        return None
    tmp = tempfile.mkdtemp(prefix="hrt_perray_")
    try:
        with open(os.path.join(tmp, "mesh.bin"), "wb") as fh:
            v4 = np.ones((v.shape[0], 4), np.float32)
            v4[:, :3] = v
            fh.write(np.array([v.shape[0], f.size], np.uint32).tobytes())
            fh.write(v4.tobytes())
            fh.write(np.ascontiguousarray(f, np.uint32).tobytes())
        with open(os.path.join(tmp, "rays.bin"), "wb") as fh:
            fh.write(np.array([rays.shape[0]], np.uint32).tobytes())
            fh.write(np.ascontiguousarray(rays, np.float32).tobytes())
        r = subprocess.run([B.HARNESS_BIN, "HipBVH2", os.path.join(tmp, "mesh.bin"), os.path.join(tmp, "rays.bin"),
                            os.path.join(tmp, "out.bin"), str(threads)], capture_output=True, text=True, timeout=300)
        for ln in r.stdout.splitlines():
            if ln.startswith("{"):
                return json.loads(ln)["per_ray_loop"]
        log(f"per-ray harness printed no result: {r.stderr[-300:]}")
        return None
    except (OSError, subprocess.SubprocessError, ValueError) as e:
        log(f"per-ray harness failed: {e}")
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def bench_eye(args, cpu) -> dict:
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    v, f = W.eye_scene()
    width = height = 512
    cam = W.eye_camera(width, height)
    n = width * height
    ctx = H.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    sc = H.Scene(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sc.add_instance(sc.add_triangles(v, f))
    sc.commit()
    build_ms = (time.perf_counter() - t0) * 1e3
    d_rays = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    ctx.eye_make_rays(cam[0], cam[1], width, height, d_rays, stream=sptr)
    torch.cuda.synchronize()
    rays = d_rays.cpu().numpy()
    # the rays in the reference's superblock order, checked against the
    # oracle below, are traced in the XCD-stripe order hrt_eye_render uses
    # (each XCD's workgroups get 4-row stripes of the screen), which traces
    # fastest of the orders tried
    sb = W.superblock_order(width, height)
    xs = W.xcd_stripe_order(width, height)
    by_pixel = torch.empty_like(d_rays)
    by_pixel[torch.from_numpy(sb).cuda()] = d_rays
    d_ord = by_pixel[torch.from_numpy(xs).cuda()].contiguous()
    del by_pixel
    d_hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    for _ in range(args.warmup):
        sc.trace_nearest(d_ord, n, d_hits, H.HIT_COMPACT, stream=sptr)
    ms = timed(stream, lambda: sc.trace_nearest(d_ord, n, d_hits, H.HIT_COMPACT, stream=sptr), args.steps)
    hits_by_pixel = np.empty((n, 2), dtype=np.uint32)
    hits_by_pixel[xs] = d_hits.cpu().numpy().view(np.uint32)
    hits = hits_by_pixel[sb]  # in the superblock order of `rays`
    d_st = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    sc.trace_nearest(d_ord, n, torch.empty_like(d_hits), H.HIT_COMPACT, stream=sptr, stats=d_st)
    torch.cuda.synchronize()
    st = d_st.cpu().numpy().view(np.uint32).astype(np.int64)
    mean = float(ms.mean())
    d_color = torch.empty(n, dtype=torch.int32, device="cuda")
    frame_ms = timed(stream, lambda: sc.eye_render(cam[0], cam[1], width, height, W.EYE_PALETTE, d_color,
                                                   stream=sptr), args.steps)
    sizes = sc.dump(0, H.DUMP_SIZES)
    alg = 40.0 * n + 64.0 * int(sizes[3]) + 48.0 * int(f.shape[0])
    blb = float(np.sum(40 + 64 * st[:, 0] + 60 * st[:, 1]))
    res = line(round(n / (mean * 1e-3) / 1e6, 2), round(mean, 4), args.steps, args.warmup,
               {"workload": "eye: 512x512 pinhole rays (made on the device in the reference's 8x8-superblock "
                            "order, traced in XCD stripes), nearest hit, "
                            "LBVH2 of an icosphere (81,920 tris) + ground quad", "triangles": int(f.shape[0]),
                "rays": n},
               {"roofline": roofline("nearest-hit traversal (k_trace_simple, BVH2, reference order)", mean, alg, blb),
                "latency_ms": lat(ms), "build_ms": round(build_ms, 3),
                "visits_per_ray": round(float(st[:, 0].mean()), 3), "frame_ms": round(float(frame_ms.mean()), 4),
                "cpu_baseline": None})
    sc.close()
    ctx.close()
    # the per-ray drop-in path on a quarter of the frame (65,536 rays)
    pr = per_ray_loop(v, f, rays[:65536], cpu["threads"])
    if pr is not None:
        res["per_ray_drop_in"] = pr
    from oracle import oracle as O

    ob = O.OracleBVH(v, f)
    ref_rays, _ = O.eye_rays(cam[0], cam[1], width, height)
    res["rays_bit_exact_vs_oracle"] = bool(np.array_equal(ref_rays.view(np.uint32), rays.view(np.uint32)))
    mism = {}

    def full(threads):
        (t, p, _, _), s = clock(lambda: ob.trace(ref_rays, threads=threads))
        mism["n"] = int(np.sum(t.view(np.uint32) != hits[:, 0]) + np.sum(p != hits[:, 1]))
        return n, s

    res["cpu_baseline"] = cpu_baseline(
        "the full 512x512 frame", full,
        lambda: (n, clock(lambda: ob.trace(ref_rays, threads=1))[1]),
        lambda: (n, clock(lambda: ob.trace_metrics(ref_rays))[1], "one thread, metrics on (CrossRT.h:31)"), cpu)
    res["parity_mismatches_vs_oracle"] = mism["n"]
    if pr is not None:
        res["per_ray_drop_in"]["cpu_baseline_mrays_s"] = res["cpu_baseline"]["value"]
    return res


# ---------------------------------------------------------------------------
# precision: FP16 vs FP32 BVH4 nodes, 4M x 4M
# ---------------------------------------------------------------------------
def t_fp64(v, f, rays, prim):
    """Float64 Moller-Trumbore t of each ray against its hit triangle (NaN on a miss)."""
    out = np.full(rays.shape[0], np.nan)
    hit = prim != INVALID
    tri = f[prim[hit].astype(np.int64)].astype(np.int64)
    A, B, C = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[hit, 0:3].astype(np.float64)
    d = rays[hit, 4:7].astype(np.float64)
    e1, e2 = B - A, C - A
    out[hit] = np.sum(e2 * np.cross(o - A, e1), axis=1) / np.sum(e1 * np.cross(d, e2), axis=1)
    return out


def rel_stats(t, t64, hit):
    rel = np.abs(t[hit].astype(np.float64) - t64[hit]) / np.abs(t64[hit])
    return {"max": float(rel.max()), "mean": float(rel.mean()), "median": float(np.median(rel)),
            "p99": float(np.percentile(rel, 99))}


def bench_precision(args, cpu) -> dict:
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    n = 4 * 1024 * 1024
    log("precision: generating the 4M-triangle soup and 4M rays")
    v, f = W.soup(n, stream=W.STREAM_SOUP)
    rays = W.incoherent_rays(n, stream=W.STREAM_RAYS)
    ctx = H.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    sc = H.Scene(ctx)
    sc.add_instance(sc.add_triangles(v, f))
    d_rays = torch.from_numpy(rays).cuda()
    res, timing = {}, {}
    out = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    # the reference tree's BVH4 (FP32, FP16 boxes), then the clustered second
    # tree's (the timed layout last: the counters below walk it)
    for name, lay in (("lbvh_fp32", H.LAYOUT_BVH4), ("lbvh_fp16", H.LAYOUT_BVH4_FP16), ("fp32", H.LAYOUT_PLOC),
                      ("fp16", H.LAYOUT_PLOC_FP16)):
        sc.commit(layout=lay)
        for _ in range(args.warmup):
            sc.trace_nearest(d_rays, n, out, H.HIT_COMPACT, stream=sptr)
        timing[name] = timed(stream, lambda: sc.trace_nearest(d_rays, n, out, H.HIT_COMPACT, stream=sptr),
                             max(3, args.steps // 2))
        res[name] = out.cpu().numpy().view(np.uint32).copy()
        log(f"precision: {name} nodes {timing[name].mean():.3f} ms")
    # counting variant on a sample of the FP16 walk: node and triangle bytes per ray
    ns = 262144
    st = torch.empty((ns, 2), dtype=torch.int32, device="cuda")
    sc.trace_nearest(d_rays[:ns], ns, torch.empty((ns, 2), dtype=torch.int32, device="cuda"), H.HIT_COMPACT,
                     stream=sptr, stats=st)
    torch.cuda.synchronize()
    stc = st.cpu().numpy().view(np.uint32).astype(np.int64)
    clustered = True
    try:
        sizes = sc.dump(0, H.DUMP_SIZES | H.DUMP_ANY_TREE)
    except H.HrtError:  # no second tree: the PLOC layouts walk the reference tree
        sizes, clustered = sc.dump(0, H.DUMP_SIZES), False
    same = all(bool(np.array_equal(res["fp16"], res[k])) for k in ("fp32", "lbvh_fp32", "lbvh_fp16"))
    mism = int(np.sum(np.any(res["fp32"] != res["fp16"], axis=1) | np.any(res["lbvh_fp32"] != res["fp16"], axis=1) |
                      np.any(res["lbvh_fp16"] != res["fp16"], axis=1)))
    prim = res["fp16"][:, 1]
    hit = prim != INVALID
    t = res["fp16"][:, 0].view(np.float32)
    t64 = t_fp64(v, f, rays, prim)
    # the watertight test on the same tree and rays
    sc.set_triangle_test(H.TRI_WATERTIGHT)
    sc.commit(layout=H.LAYOUT_PLOC_FP16)
    wt_ms = timed(stream, lambda: sc.trace_nearest(d_rays, n, out, H.HIT_COMPACT, stream=sptr), 3)
    wt = out.cpu().numpy().view(np.uint32).copy()
    wprim = wt[:, 1]
    whit = wprim != INVALID
    wt64 = t_fp64(v, f, rays, wprim)
    differ = (wprim != prim)
    watertight = {"ms": round(float(wt_ms.mean()), 4), "mrays_s": round(n / (wt_ms.mean() * 1e-3) / 1e6, 2),
                  "rays_whose_hit_differs_from_mt": int(differ.sum()),
                  "disagreement_rate": float(differ.mean()),
                  "hit_only_mt": int(np.sum(hit & ~whit)), "hit_only_watertight": int(np.sum(whit & ~hit)),
                  "t_rel_error_vs_fp64": rel_stats(wt[:, 0].view(np.float32), wt64, whit)}
    sc.set_triangle_test(H.TRI_MT)
    ms16, ms32 = float(timing["fp16"].mean()), float(timing["fp32"].mean())
    nq = int(sizes[4])
    alg = 40.0 * n + 64.0 * nq + 48.0 * n
    blb = float(np.mean(40 + 64 * stc[:, 0] + 60 * stc[:, 1])) * n

    def rate(k):
        ms = float(timing[k].mean())
        return {"ms": round(ms, 4), "mrays_s": round(n / (ms * 1e-3) / 1e6, 2)}

    out_res = line(round(n / (ms16 * 1e-3) / 1e6, 2), round(ms16, 4), len(timing["fp16"]), args.warmup,
                   {"workload": "precision: 4,194,304-tri random soup, 4,194,304 incoherent nearest-hit rays, BVH4 "
                                "with FP16 boxes rounded outward (value) vs FP32 boxes, on the clustered (PLOC) "
                                "tree; the reference LBVH's BVH4 beside", "triangles": n, "rays": n},
                   {"roofline": roofline("nearest-hit traversal (BVH4, FP16 nodes of the clustered tree)", ms16, alg,
                                         blb),
                    "fp32_nodes": rate("fp32"), "fp16_nodes": rate("fp16"),
                    "reference_tree": {"fp32_nodes": rate("lbvh_fp32"), "fp16_nodes": rate("lbvh_fp16")},
                    "clustered_tree": clustered,
                    "visits_per_ray": round(float(stc[:, 0].mean()), 2),
                    "tris_per_ray": round(float(stc[:, 1].mean()), 2),
                    "hits_bit_identical_fp16_fp32_both_trees": same, "mismatching_rays": mism,
                    "hit_fraction": round(float(hit.mean()), 4),
                    "t_rel_error_vs_fp64": rel_stats(t, t64, hit),
                    "watertight": watertight,
                    "latency_ms": lat(timing["fp16"]), "cpu_baseline": None})
    sc.close()
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        log("precision: cpu baseline (oracle build of the 4M soup, FP16 BVH4 traversal of a sample)")
        ob = O.OracleBVH(v, f)
        sample = 262144
        par = {}

        def full(threads):
            (tc, pc, _, _), s = clock(lambda: ob.trace4(rays[:sample], O.MODE4_FP16, threads=threads))
            par["n"] = int(np.sum(tc.view(np.uint32) != res["fp16"][:sample, 0]) + np.sum(pc != res["fp16"][:sample, 1]))
            return sample, s

        out_res["cpu_baseline"] = cpu_baseline(
            f"the first {sample} rays, FP16 BVH4 traversal (BVH4HalfRT restatement)", full,
            lambda: (16384, clock(lambda: ob.trace4(rays[:16384], O.MODE4_FP16, threads=1))[1]), None, cpu)
        out_res["parity_mismatches_vs_oracle"] = par["n"]
        out_res["parity_sample_rays"] = sample
    return out_res


# ---------------------------------------------------------------------------
# batched: 256 scenes built and traced in one graph launch
# ---------------------------------------------------------------------------
def bench_batched(args, cpu) -> dict:
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    log("batched: generating 256 scenes")
    scenes = W.batched_workload()
    nv = [x[0].shape[0] for x in scenes]
    nt = [x[1].shape[0] for x in scenes]
    nr = [x[2].shape[0] for x in scenes]
    T, R = int(sum(nt)), int(sum(nr))
    ctx = H.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    b = H.Batch(ctx, nv, nt, nr)
    b.upload(scenes)
    chosen, cand_ms = b.autotune(reps=3)
    cands = H.batch_candidates()
    for _ in range(args.warmup):
        b.run(sptr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ms = timed(stream, lambda: b.run(sptr), args.steps)
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    build_ms = timed(stream, lambda: b.build(sptr), 5)
    trace_ms = timed(stream, lambda: b.trace(sptr), 5)
    hits = b.hits()
    ms = float(np.median(step_ms))
    log(f"batched: graph step {ms:.3f} ms (build {build_ms.mean():.3f} + trace {trace_ms.mean():.3f})")
    # the plain path: one scene at a time through the single-scene API
    d_rays = [torch.from_numpy(x[2]).cuda() for x in scenes]
    d_out = [torch.empty((k, 2), dtype=torch.int32, device="cuda") for k in nr]
    d_st = [torch.empty((k, 2), dtype=torch.int32, device="cuda") for k in nr]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solo = []
    for s, (v, f, _) in enumerate(scenes):
        sc = H.Scene(ctx)
        sc.add_instance(sc.add_triangles(v, f))
        sc.commit()
        sc.trace_nearest(d_rays[s], nr[s], d_out[s], H.HIT_COMPACT, stream=sptr)
        solo.append(sc)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3
    mism = 0
    blb = 0.0
    pairs = 0
    for s, sc in enumerate(solo):
        mism += int(np.sum(d_out[s].cpu().numpy().view(np.uint32) != hits[b.ray_off[s]:b.ray_off[s + 1]]))
        sc.trace_nearest(d_rays[s], nr[s], d_out[s], H.HIT_COMPACT, stream=sptr, stats=d_st[s])
        pairs += int(sc.dump(0, H.DUMP_SIZES)[3])
    torch.cuda.synchronize()
    for s in range(len(scenes)):
        st = d_st[s].cpu().numpy().view(np.uint32).astype(np.int64)
        blb += float(np.sum(40 + 64 * st[:, 0] + 60 * st[:, 1]))
    for sc in solo:
        sc.close()
    alg = 40.0 * R + 64.0 * pairs + 48.0 * T
    res = line(round(R / (ms * 1e-3) / 1e6, 2), round(ms, 4), args.steps, args.warmup,
               {"workload": "batched: 256 random soups (1k-200k tris, log-uniform) with 4k-256k incoherent rays "
                            "each; all LBVH builds + nearest-hit traces in one hipGraph launch",
                "scenes": len(scenes), "triangles": T, "rays": R},
               {"roofline": roofline("batched nearest-hit traversal (k_b_trace)", float(trace_ms.mean()), alg, blb),
                "latency_ms": lat(step_ms), "wall_ms_per_step": round(wall_ms, 4),
                "build_ms": round(float(build_ms.mean()), 4),
                "build_mtris_s": round(T / (build_ms.mean() * 1e-3) / 1e6, 1),
                "trace_ms": round(float(trace_ms.mean()), 4),
                "trace_mrays_s": round(R / (trace_ms.mean() * 1e-3) / 1e6, 1),
                "autotune": {"chosen": {"waves": cands[chosen][0], "lds_stack": cands[chosen][1],
                                        "refill": cands[chosen][2], "claim": cands[chosen][3]},
                             "candidates_ms": {f"w{c[0]}_lds{c[1]}_r{c[2]}_c{c[3]}": round(float(m), 4)
                                               for c, m in zip(cands, cand_ms)}},
                "plain_per_scene_path": {"ms": round(plain_ms, 2), "mrays_s": round(R / (plain_ms * 1e-3) / 1e6, 2),
                                         "note": "256 x (hrt_add_triangles from host memory + commit + "
                                                 "hrt_trace_nearest), rays resident"},
                "parity_mismatches_vs_plain": mism, "cpu_baseline": None})
    b.close()
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        sample = scenes[:24]
        par = {}

        def full(threads):
            (_, s) = clock(lambda: [O.OracleBVH(v, f).trace(r, threads=threads) for v, f, r in sample])
            return sum(x[2].shape[0] for x in sample), s

        def full_checked(threads):
            rays_n, s = full(threads)
            bad = 0
            for k, (v, f, r) in enumerate(sample):
                t, p, _, _ = O.OracleBVH(v, f).trace(r, threads=threads)
                hb = hits[b.ray_off[k]:b.ray_off[k + 1]]
                bad += int(np.sum(hb[:, 0] != t.view(np.uint32)) + np.sum(hb[:, 1] != p))
            par["n"] = bad
            return rays_n, s

        res["cpu_baseline"] = cpu_baseline("the first 24 of the 256 scenes: oracle LBVH build + nearest-hit trace",
                                           full_checked, lambda: full(1), None, cpu)
        res["parity_mismatches_vs_oracle"] = par["n"]
    return res


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
def bench_serving(args, cpu) -> dict:
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    ctx = H.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    log("serving: building 16 resident scenes")
    geo = W.serving_scenes()
    scenes, pools = [], []
    t_build = 0.0
    for k, (v, f) in enumerate(geo):
        torch.cuda.synchronize()
        t = time.perf_counter()
        sc = H.Scene(ctx)
        sc.add_instance(sc.add_triangles(v, f))
        sc.commit(layout=H.LAYOUT_AUTO_FP16)
        torch.cuda.synchronize()
        t_build += time.perf_counter() - t
        eye, target, radius = W.serving_camera(v)
        eye_rays = W.pinhole_rays(1024, 1024, eye=eye, target=target, fov_y_deg=50.0)
        n = eye_rays.shape[0]
        d_eye = torch.from_numpy(eye_rays).cuda()
        d_hit = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        d_ao = torch.empty((n, 8), dtype=torch.float32, device="cuda")
        sc.trace_nearest(d_eye, n, d_hit, H.HIT_COMPACT, stream=sptr)
        sc.rtao_make_rays(d_eye, d_hit, n, 1, radius, SEED + k, d_ao, stream=sptr)
        torch.cuda.synchronize()
        pool = {"eye": H.pinned_array((n, 8), np.float32), "ao": H.pinned_array((n, 8), np.float32),
                "d_eye": d_eye, "d_ao": d_ao}
        pool["eye"][:] = eye_rays
        pool["ao"][:] = d_ao.cpu().numpy()
        scenes.append(sc)
        pools.append(pool)
    tris = int(sum(f.shape[0] for _, f in geo))
    npool = 1024 * 1024
    srv = H.Server(ctx, 4 << 30, steps_in_flight=3)
    ids = [srv.add_scene(sc) for sc in scenes]

    def run_stream(req, device: bool):
        nreq = len(req["size"])
        outs, rids, t_sub = [], [], []
        t0 = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t0 < req["arrival_s"][i]:
                pass
            k, a, n, anyh = int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]), bool(req["any_hit"][i])
            p = pools[k]
            if device:
                rays = (p["d_ao"] if anyh else p["d_eye"])[a:a + n]
                out = torch.empty(n, dtype=torch.uint8, device="cuda") if anyh else \
                    torch.empty((n, 2), dtype=torch.int32, device="cuda")
            else:
                rays = (p["ao"] if anyh else p["eye"])[a:a + n]
                out = H.pinned_array((n,), np.uint8) if anyh else H.pinned_array((n, 2), np.uint32)
            outs.append(out)
            rids.append(srv.submit(ids[k], rays, out, any_hit=anyh))
            t_sub.append(time.perf_counter() - t0)
        lat_ms = np.array([srv.wait(r)[0] for r in rids])
        t_end = max(ts + l * 1e-3 for ts, l in zip(t_sub, lat_ms))
        return outs, lat_ms, t_end

    def run_levels(device: bool, rates, stream0):
        out, best = [], None
        for li, qps in enumerate(rates):
            req = W.serving_requests(args.requests, qps, stream=stream0 + li)
            log(f"serving: {'device' if device else 'pinned host'} requests at {qps:.0f}/s: {args.requests} "
                f"requests, {int(req['size'].sum())} rays")
            s0 = srv.stats()
            outs, lat_ms, t_end = run_stream(req, device=device)
            s1 = srv.stats()
            steps = max(1, s1["steps"] - s0["steps"])
            lv = {"qps_offered": qps, "qps_achieved": round(args.requests / t_end, 1),
                  "mrays_s": round(float(req["size"].sum()) / t_end / 1e6, 1), "p50_ms": round(float(np.percentile(
                      lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                  "mean_requests_per_step": round(args.requests / steps, 3)}
            out.append(lv)
            log(f"  -> {lv}")
            if best is None or lv["mrays_s"] >= best[2]["mrays_s"]:
                best = (req, outs, lv)
        return out, best

    rates = [float(q) for q in args.qps.split(",")]
    # device-resident requests (the value), up to four times the highest host rate to find the ceiling
    dev_levels, (best_req, best_outs, best_level) = run_levels(True, rates + [4 * max(rates)], W.STREAM_SERVE + 100)
    # pinned host requests: the same server with every request's rays and answers crossing PCIe
    levels, (host_req, host_outs, host_best) = run_levels(False, rates, W.STREAM_SERVE + 200)
    # mix = solo on every request of the best device rate and of the best host rate
    mism = 0
    for i in range(len(best_req["size"])):
        k, a, n, anyh = (int(best_req["scene"][i]), int(best_req["start"][i]), int(best_req["size"][i]),
                         bool(best_req["any_hit"][i]))
        p = pools[k]
        d = (p["d_ao"] if anyh else p["d_eye"])[a:a + n]
        if anyh:
            o = torch.empty(n, dtype=torch.uint8, device="cuda")
            scenes[k].trace_any(d, n, o, stream=sptr)
        else:
            o = torch.empty((n, 2), dtype=torch.int32, device="cuda")
            scenes[k].trace_nearest(d, n, o, H.HIT_COMPACT, stream=sptr)
        torch.cuda.synchronize()
        ref = o.cpu().numpy().view(np.uint32) if not anyh else o.cpu().numpy()
        got = best_outs[i].cpu().numpy() if hasattr(best_outs[i], "cpu") else np.asarray(best_outs[i])
        mism += int(np.sum(got.view(np.uint32).reshape(ref.shape) != ref)) if not anyh else \
            int(np.sum(got.reshape(ref.shape) != ref))
    for i in range(len(host_req["size"])):  # the host-request answers too
        k, a, n, anyh = (int(host_req["scene"][i]), int(host_req["start"][i]), int(host_req["size"][i]),
                         bool(host_req["any_hit"][i]))
        d = (pools[k]["d_ao"] if anyh else pools[k]["d_eye"])[a:a + n]
        o = torch.empty(n, dtype=torch.uint8, device="cuda") if anyh else \
            torch.empty((n, 2), dtype=torch.int32, device="cuda")
        if anyh:
            scenes[k].trace_any(d, n, o, stream=sptr)
        else:
            scenes[k].trace_nearest(d, n, o, H.HIT_COMPACT, stream=sptr)
        torch.cuda.synchronize()
        ref = o.cpu().numpy().view(np.uint32) if not anyh else o.cpu().numpy()
        mism += int(np.sum(np.asarray(host_outs[i]).reshape(ref.shape) != ref))

    def plain_run(req, device: bool):
        """One hrt_trace_* call per request in arrival order, back to back (no queueing)."""
        calls = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(len(req["size"])):
            k, a, n, anyh = (int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]),
                             bool(req["any_hit"][i]))
            r0 = time.perf_counter()
            if device:
                d = (pools[k]["d_ao"] if anyh else pools[k]["d_eye"])[a:a + n]
                o = torch.empty(n, dtype=torch.uint8, device="cuda") if anyh else \
                    torch.empty((n, 2), dtype=torch.int32, device="cuda")
                if anyh:
                    scenes[k].trace_any(d, n, o, stream=sptr)
                else:
                    scenes[k].trace_nearest(d, n, o, H.HIT_COMPACT, stream=sptr)
                stream.synchronize()
            else:
                rays = (pools[k]["ao"] if anyh else pools[k]["eye"])[a:a + n]
                if anyh:
                    scenes[k].trace_any_host(rays)
                else:
                    scenes[k].trace_nearest_host(rays)
            calls.append((time.perf_counter() - r0) * 1e3)
        s = time.perf_counter() - t0
        return {"mrays_s": round(float(req["size"].sum()) / s / 1e6, 1), "seconds": round(s, 3),
                "call_ms": lat(np.array(calls))}

    plain = dict(plain_run(best_req, True), note="the best device rate's requests, one hrt_trace_nearest / "
                 "hrt_trace_any call each on device buffers, each waited for, back to back")
    plain_host = dict(plain_run(host_req, False), note="the best host rate's requests, one hrt_trace_nearest_host "
                      "/ hrt_trace_any_host call each, back to back")
    srv_tree_rays = srv.tree_rays()
    # warm start: snapshots
    t_ser = t_de = 0.0
    nbytes, same = 0, True
    for sc in scenes:
        t = time.perf_counter()
        data = sc.serialize()
        t_ser += time.perf_counter() - t
        t = time.perf_counter()
        r = H.Scene.deserialize(ctx, data)
        torch.cuda.synchronize()
        t_de += time.perf_counter() - t
        nbytes += len(data)
        same = same and r.serialize() == data
        r.close()
    srv.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    # algorithmic bytes of the best device stream: rays in, answers out
    alg = float(best_req["size"].sum()) * 32 + float(np.sum(np.where(best_req["any_hit"], 1, 8) * best_req["size"]))
    dev_ms = float(best_req["size"].sum()) / (best_level["mrays_s"] * 1e6) * 1e3
    tr = srv_tree_rays
    res = line(best_level["mrays_s"], None, args.requests, 0,
               {"workload": "serving: 16 resident scenes (50k-2M tris, soups and heightfields, any-hit tree "
                            "committed), Poisson requests of 1k-1M rays (log-uniform), 50/50 nearest-hit eye rays "
                            "/ any-hit AO rays, 4 GiB HBM budget, 3 steps in flight, device-resident requests "
                            "(value) and pinned host requests",
                "triangles": tris, "requests_per_rate": args.requests},
               {"roofline": dict(roofline("served traces, device-resident requests (whole stream)", dev_ms, alg, None),
                                 replayed_stream=bool(best_level["qps_offered"] == SERVE_REPLAY_QPS and
                                                      args.requests == SERVE_REPLAY_REQUESTS and
                                                      dev_levels.index(best_level) == SERVE_REPLAY_LEVEL)),
                "latency_ms": {"p50": best_level["p50_ms"], "p99": best_level["p99_ms"]},
                "device_rates": dev_levels, "parity_mismatches_vs_solo": mism, "plain_path_same_requests": plain,
                "served_vs_plain": round(best_level["mrays_s"] / max(plain["mrays_s"], 1e-9), 2),
                "host_requests": {"rates": levels, "plain_path_same_requests": plain_host,
                                  "served_vs_plain": round(host_best["mrays_s"] / max(plain_host["mrays_s"], 1e-9),
                                                           2),
                                  "note": "rays and answers cross PCIe inside the timed region"},
                "rays_per_tree": tr,
                "warm_start": {"build_from_triangles_ms": round(1e3 * t_build, 1),
                               "serialize_ms": round(1e3 * t_ser, 1), "deserialize_ms": round(1e3 * t_de, 1),
                               "snapshot_mb": round(nbytes / 1e6, 1), "restored_bytes_identical": bool(same)},
                "cpu_baseline": None})
    res["roofline"]["traffic_source"] = "not measured (a host-driven stream); achieved uses the algorithmic bytes"
    return res


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 counters (a child process per counter pass)
# ---------------------------------------------------------------------------
def counter_child(workloads) -> int:
    """Runs under rocprofv3: a 1 GiB device copy (the byte calibration), two
    launches of the issue-rate calibration kernel (hrt_issue_probe), then
    three launches of each workload's dominant kernel after its setup."""
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    torch.cuda.init()
    src = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    ctx = H.Context(0)
    for _ in range(2):
        ctx.issue_probe()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    # the serving replay first: its setup traces and served any-hit kernel
    # share kernels with other workloads, whose last three launches are read
    for w in sorted(workloads, key=lambda x: x != "serving"):
        if w == "rtao":
            v, f = W.rtao_scene()
            prim = torch.from_numpy(W.rtao_primary_rays(order="superblock")).cuda()
            npix = prim.shape[0]
            sc = H.Scene(ctx)
            sc.add_instance(sc.add_triangles(v, f))
            sc.commit(layout=H.LAYOUT_AUTO_FP16)
            hits = torch.zeros((npix, 2), dtype=torch.int32, device="cuda")
            vis = torch.zeros(npix * SPP, dtype=torch.uint8, device="cuda")
            sc.trace_nearest(prim, npix, hits, H.HIT_COMPACT, stream=sp)
            for _ in range(3):
                sc.rtao_trace(prim, hits, npix, SPP, 0.1 * W.scene_diagonal(v), SEED, vis, stream=sp)
        elif w == "eye":
            v, f = W.eye_scene()
            cam = W.eye_camera(512, 512)
            sc = H.Scene(ctx)
            sc.add_instance(sc.add_triangles(v, f))
            sc.commit()
            made = torch.empty((512 * 512, 8), dtype=torch.float32, device="cuda")
            ctx.eye_make_rays(cam[0], cam[1], 512, 512, made, stream=sp)
            by_pixel = torch.empty_like(made)  # in XCD stripes, as bench_eye traces them
            by_pixel[torch.from_numpy(W.superblock_order(512, 512)).cuda()] = made
            rays = by_pixel[torch.from_numpy(W.xcd_stripe_order(512, 512)).cuda()].contiguous()
            hits = torch.empty((512 * 512, 2), dtype=torch.int32, device="cuda")
            for _ in range(3):
                sc.trace_nearest(rays, 512 * 512, hits, H.HIT_COMPACT, stream=sp)
        elif w == "precision":
            n = 4 << 20
            v, f = W.soup(n, stream=W.STREAM_SOUP)
            rays = torch.from_numpy(W.incoherent_rays(n, stream=W.STREAM_RAYS)).cuda()
            sc = H.Scene(ctx)
            sc.add_instance(sc.add_triangles(v, f))
            sc.commit(layout=H.LAYOUT_PLOC_FP16)
            hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
            for _ in range(3):
                sc.trace_nearest(rays, n, hits, H.HIT_COMPACT, stream=sp)
        elif w == "serving":
            # the 16 resident scenes and their ray pools as bench_serving builds
            # them, then the top rate's request stream, submitted at once
            geo = W.serving_scenes()
            srv = H.Server(ctx, 4 << 30, steps_in_flight=3)
            held, ids, pools = [], [], []
            for k, (v, f) in enumerate(geo):
                s_k = H.Scene(ctx)
                s_k.add_instance(s_k.add_triangles(v, f))
                s_k.commit(layout=H.LAYOUT_AUTO_FP16)
                eye, target, radius = W.serving_camera(v)
                d_eye = torch.from_numpy(W.pinhole_rays(1024, 1024, eye=eye, target=target, fov_y_deg=50.0)).cuda()
                n_k = d_eye.shape[0]
                d_hit = torch.empty((n_k, 2), dtype=torch.int32, device="cuda")
                d_ao = torch.empty((n_k, 8), dtype=torch.float32, device="cuda")
                s_k.trace_nearest(d_eye, n_k, d_hit, H.HIT_COMPACT, stream=sp)
                s_k.rtao_make_rays(d_eye, d_hit, n_k, 1, radius, SEED + k, d_ao, stream=sp)
                torch.cuda.synchronize()
                held.append(s_k)
                pools.append((d_eye, d_ao))
                ids.append(srv.add_scene(s_k))
            req = W.serving_requests(SERVE_REPLAY_REQUESTS, SERVE_REPLAY_QPS,
                                     stream=W.STREAM_SERVE + 100 + SERVE_REPLAY_LEVEL)
            outs, rids = [], []
            for i in range(len(req["size"])):
                k, a, n_r, anyh = (int(req["scene"][i]), int(req["start"][i]), int(req["size"][i]),
                                   bool(req["any_hit"][i]))
                rays_r = (pools[k][1] if anyh else pools[k][0])[a:a + n_r]
                o = torch.empty(n_r, dtype=torch.uint8, device="cuda") if anyh else \
                    torch.empty((n_r, 2), dtype=torch.int32, device="cuda")
                outs.append(o)
                rids.append(srv.submit(ids[k], rays_r, o, any_hit=anyh))
            for r_id in rids:
                srv.wait(r_id)
            srv.close()
            for s_k in held:
                s_k.close()
            continue
        elif w == "batched":
            scenes = W.batched_workload()
            b = H.Batch(ctx, [x[0].shape[0] for x in scenes], [x[1].shape[0] for x in scenes],
                        [x[2].shape[0] for x in scenes])
            b.upload(scenes)
            b.autotune(reps=1)
            b.build(sp)
            for _ in range(3):
                b.trace(sp)
            torch.cuda.synchronize()
            b.close()
            continue
        torch.cuda.synchronize()
        sc.close()
    ctx.close()
    print("counter child done", flush=True)
    return 0


def read_counter_csv(directory: str):
    """(kernel name, counter name, value) rows in dispatch order from rocprofv3's CSV output."""
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                try:
                    rows.append((int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0), r["Kernel_Name"],
                                 r["Counter_Name"], float(r["Counter_Value"])))
                except (KeyError, ValueError):
                    continue
    rows.sort(key=lambda x: x[0])
    return rows


def measure_traffic(workloads, timeout_s: float = 240.0):
    """HBM bytes per launch of each workload's dominant kernel, from two
    rocprofv3 passes (FETCH_SIZE; WRITE_SIZE with the L2 hit/miss counts),
    each scaled by what the counter reads for the 1 GiB copy in the same
    child; and its VALU issue rate from a third pass (SQ_INSTS_VALU per
    GRBM_GUI_ACTIVE cycle) against the calibration kernel hrt_issue_probe's
    in the same child.  Returns ({workload: {...}}, reason when nothing was
    measured)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return {}, "rocprofv3 not found"
    out = {w: {} for w in workloads}
    tmp = tempfile.mkdtemp(prefix="hrt_counters_")
    env = dict(os.environ, TMPDIR="/tmp")
    passes = [("fetch", ["FETCH_SIZE"]), ("write", ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"]),
              ("issue", ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "GRBM_GUI_ACTIVE"])]
    issue = {w: [] for w in workloads}  # per launch: (VALU instructions, SALU instructions, GRBM cycles)
    probe = []
    try:
        for tag, counters in passes:
            d = os.path.join(tmp, tag)
            cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", tag, "--", sys.executable,
                   os.path.abspath(__file__), "--counter-child", ",".join(workloads)]
            log(f"counters: rocprofv3 --pmc {' '.join(counters)}")
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, env=env, cwd=ROOT)
            except subprocess.TimeoutExpired:
                return {}, f"rocprofv3 pass {tag} timed out after {timeout_s:.0f} s"
            rows = read_counter_csv(d)
            if not rows:
                return {}, f"rocprofv3 pass {tag} wrote no counters (rc {r.returncode}): {r.stderr[-200:]}"
            if tag == "issue":
                per = {}
                for did, name, cname, val in rows:
                    per.setdefault(did, [name, {}])[1][cname] = val
                for did in sorted(per):
                    name, c = per[did]
                    if not all(k in c for k in counters) or c["GRBM_GUI_ACTIVE"] <= 0:
                        continue
                    rec = (c["SQ_INSTS_VALU"], c["SQ_INSTS_SALU"], c["GRBM_GUI_ACTIVE"])
                    if "k_issue_probe" in name:
                        probe.append(rec)
                    for w in workloads:
                        if counter_match(w, name):
                            issue[w].append(rec)
                continue
            by_counter = {}
            for _, name, cname, val in rows:
                by_counter.setdefault(cname, []).append((name, val))
            for cname in counters:
                vals = by_counter.get(cname, [])
                calib = [v for n_, v in vals if "copyBuffer" in n_ or "copy" in n_.lower()]
                if not calib:
                    return {}, f"no calibration copy in the {cname} pass"
                kib_per_copy = max(calib)
                for w in workloads:
                    ks = [v for n_, v in vals if counter_match(w, n_)]
                    if not ks:
                        continue
                    agg = float(np.sum(ks)) if w in COUNTER_SUM else float(np.mean(ks[-3:]))
                    out[w][cname] = {"kernel_value": agg, "copy_value": kib_per_copy}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    res = {}
    for w, c in out.items():
        if "FETCH_SIZE" not in c or "WRITE_SIZE" not in c:
            continue
        # the copy moves 1 GiB in and 1 GiB out
        fetch = c["FETCH_SIZE"]["kernel_value"] / c["FETCH_SIZE"]["copy_value"] * float(1 << 30)
        write = c["WRITE_SIZE"]["kernel_value"] / c["WRITE_SIZE"]["copy_value"] * float(1 << 30)
        hit = c.get("TCC_HIT_sum", {}).get("kernel_value")
        miss = c.get("TCC_MISS_sum", {}).get("kernel_value")
        res[w] = {"hbm_read_bytes": fetch, "hbm_write_bytes": write,
                  "l2_hit_rate": (hit / (hit + miss)) if hit is not None and miss and hit + miss > 0 else None}
        if probe and issue[w]:
            peak = max(v / g for v, _, g in probe)
            if w in COUNTER_SUM:  # the stream's instructions over its cycles
                recs = issue[w]
                rate = float(sum(v for v, _, _ in recs) / sum(g for _, _, g in recs))
                valu, salu = sum(v for v, _, _ in recs), sum(s for _, s, _ in recs)
            else:
                recs = issue[w][-3:]
                rate = float(np.mean([v / g for v, _, g in recs]))
                valu, salu = np.mean([v for v, _, _ in recs]), np.mean([s for _, s, _ in recs])
            res[w]["issue"] = {"valu_instructions": round(float(valu)), "salu_instructions": round(float(salu)),
                               "valu_per_cycle": round(rate, 2), "probe_valu_per_cycle": round(peak, 2),
                               "frac": round(rate / peak, 4)}
    return res, ("" if res else "no workload kernel found in the counter output")


def attach_traffic(res: dict, meas: dict | None, reason: str) -> None:
    if not meas:
        res["roofline"]["traffic_source"] = f"not measured ({reason}): achieved uses the algorithmic bytes"
        return
    roof = res["roofline"]
    traffic = meas["hbm_read_bytes"] + meas["hbm_write_bytes"]
    ach = traffic / (roof["mean_launch_ms"] * 1e-3) / 1e9
    source = ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel in this run (mean of 3 launches), each scaled by "
              "the counter's reading for a 1 GiB device copy")
    replayed = roof.pop("replayed_stream", False)
    if replayed:
        source = ("rocprofv3 FETCH_SIZE + WRITE_SIZE summed over every served-trace dispatch (k_serve and the "
                  "served any-hit kernel) of the same request stream replayed in a counter child with all "
                  "requests submitted at once, each scaled by the counter's reading for a 1 GiB device copy; "
                  "achieved = those bytes over the stream's measured time")
    roof.update({"achieved": round(ach, 2), "frac": round(ach / HBM_PEAK_GBPS, 5), "traffic": round(traffic),
                 "traffic_source": source,
                 "hbm_read_bytes": round(meas["hbm_read_bytes"]), "hbm_write_bytes": round(meas["hbm_write_bytes"]),
                 "l2_hit_rate": None if meas["l2_hit_rate"] is None else round(meas["l2_hit_rate"], 4)})
    if "issue" in meas:
        # VALU instructions issued per GPU cycle (summed over every SIMD) against the calibration kernel's
        # rate in the same run: how close the kernel is to the instruction-issue bound
        roof["issue"] = dict(meas["issue"], source="rocprofv3 SQ_INSTS_VALU / GRBM_GUI_ACTIVE of this kernel "
                                                   + ("(summed over the replayed stream's dispatches)" if replayed
                                                      else "(mean of 3 launches)") +
                                                   " over the same ratio for hrt_issue_probe")
        top = max(meas["issue"]["frac"], roof["frac"])
        roof["limit"] = ("latency" if top < 0.5 else
                         "valu_issue" if meas["issue"]["frac"] > roof["frac"] else "hbm")


# ---------------------------------------------------------------------------
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("all",) + WORKLOADS, default="all",
                    help="all (default): rtao as the value, the other four as sub-objects of `configs`")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--counters", choices=("auto", "off"), default="auto",
                    help="auto: measure each dominant kernel's HBM bytes with rocprofv3 in a child process")
    ap.add_argument("--qps", type=str, default="1000,4000,12000", help="serving: request rates to run")
    ap.add_argument("--requests", type=int, default=1500, help="serving: requests per rate")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (scales rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--counter-child", type=str, default="", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.counter_child:
        return counter_child([w for w in args.counter_child.split(",") if w])
    if args.copy_calibration_mb:
        import torch

        n = args.copy_calibration_mb << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box in this setup has one GPU, and the traversal "
              "shards by rays with no exchange. Run with --gpus 1.", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        log("no GPU visible")
        return 1
    torch.cuda.init()
    cpu = host_cpu()
    fns = {"rtao": bench_rtao, "eye": bench_eye, "precision": bench_precision, "batched": bench_batched,
           "serving": bench_serving}
    todo = WORKLOADS if args.workload == "all" else (args.workload,)
    results = {}
    t_all = time.perf_counter()
    for w in todo:
        t0 = time.perf_counter()
        results[w] = fns[w](args, cpu)
        log(f"{w}: {results[w]['value']} Mrays/s ({time.perf_counter() - t0:.1f} s)")
    if args.counters == "auto":
        ws = [w for w in todo if w in COUNTER_KERNELS]
        meas, reason = measure_traffic(ws) if ws else ({}, "")
        for w in todo:
            m = meas.get(w)
            if w == "serving" and not results[w]["roofline"].get("replayed_stream"):
                m = None  # the best stream was not the one the counter pass replays
            attach_traffic(results[w], m, reason or "no counter pass for this workload or stream")
    for w in todo:
        results[w]["roofline"].pop("replayed_stream", None)
    head = results[todo[0]]
    if len(todo) > 1:
        head["configs"] = {w: results[w] for w in todo[1:]}
    head["host_cpu"] = cpu
    head["bench_seconds"] = round(time.perf_counter() - t_all, 1)
    print(json.dumps(head), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
