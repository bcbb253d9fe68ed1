"""Ray-query throughput on one MI355X: the RTAO workload.

Workload (one step): a 1024x512-quad heightfield (value noise, amp 0.2) plus
2,000 randomly rotated boxes = 1,072,576 triangles in one LBVH; 1024x1024
primary rays traced nearest-hit (BVH2 node pairs, the reference's visiting
order, bit-exact with BVH2CommonRT); 16 cosine-weighted AO rays per primary
hit made on the device (tMax = 0.1 x scene diagonal); all AO rays traced
any-hit on FP16 BVH4 nodes (BVH4HalfRT's node format, boxes rounded outward)
of a second tree over the same triangles: leaves from an LBVH on a cubic
Morton grid, the hierarchy above them by agglomerative clustering (any-hit
answers do not depend on the tree; the tests check them against the
oracle and every run checks them against the reference tree's BVH2, FP32 BVH4
and FP16 BVH4).  Scene and rays
are resident in HBM before the timed region starts.

Prints ONE JSON line.  `value` is total rays (primary + AO) per second over
the timed steps.  `roofline` is for the dominant kernel (the any-hit
traversal): algorithmic bytes per launch = sum over rays of 32 (ray) + 1
(result) + node bytes per visit x nodes visited + 60 x triangles tested --
the per-ray figure of the reference's bus-load metric (core/CrossRT.h:52,
core/BVH2CommonRT.cpp:69-70,125), with 64 B per BVH2 pair, 128 B per FP32 BVH4
node, 64 B per FP16 BVH4 node; visit and test counts come from the engine's
counting variant of the same traversal -- divided by the kernel's mean
duration, measured with HIP events on the stream the kernel runs on.
`cpu_baseline` is the CPU oracle (a C++ restatement of the reference's
BVH2CommonRT traversal) timed on this host on the same rays.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SPP = 16
SEED = 0x2545F4914F6CDD1D
NODE_BYTES = {"bvh2": 64, "bvh4": 128, "bvh4_fp16": 64, "bvh4_fp16_anytree": 64, "bvh8_fp16_anytree": 128}
# HBM bytes per any-hit launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
# this workload, scaled by the same counters on a device copy of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rtao_anyhit_traffic.json")


def measured_traffic(layout: str, ao_pass: str):
    """Profiled HBM bytes per launch for `layout` and AO pass kind ("fused" or
    "two_step"), or None if not profiled."""
    try:
        with open(TRAFFIC_FILE) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("any_hit_layout") != layout or rec.get("ao_pass", "two_step") != ao_pass:
        return None
    return float(rec["hbm_bytes_per_launch"])


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--any-layout", choices=["bvh2", "bvh4", "bvh4_fp16_anytree", "bvh8_fp16_anytree"],
                    default="bvh4_fp16_anytree",
                    help="tree for the AO any-hit rays: bvh2 (the reference tree's pairs), bvh4 (its FP32 "
                         "BVH4), bvh4_fp16_anytree (FP16 BVH4 of the any-hit tree: cubic-grid leaves, clustered hierarchy), "
                         "bvh8_fp16_anytree (the same tree in 8-wide FP16 nodes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip timing the other any-hit layouts")
    ap.add_argument("--two-step", action="store_true",
                    help="rtao: write the AO rays to HBM and trace them (hrt_rtao_make_rays + hrt_trace_any) "
                         "instead of the fused AO pass (hrt_rtao_trace)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, os.cpu_count())")
    ap.add_argument("--workload", choices=["rtao", "eye", "batched", "precision", "serving"], default="rtao",
                    help="rtao (default): the headline RTAO step; eye: 512x512 eye rays on the icosphere "
                         "scene; batched: 256 scenes built and traced in one graph launch; precision: 4M-tri "
                         "soup, FP16 vs FP32 BVH4 nodes; serving: Poisson request stream over 16 resident scenes")
    ap.add_argument("--qps", type=str, default="1000,4000,12000",
                    help="serving: comma-separated request rates to run")
    ap.add_argument("--requests", type=int, default=1500, help="serving: requests per rate")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (scales rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
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
              "shards by rays with no exchange (see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    if not torch.cuda.is_available():
        log("no GPU visible")
        return 1
    torch.cuda.init()
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "eye":
        return bench_eye(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "serving":
        return bench_serving(args)
    dev = torch.device("cuda:0")
    layouts = {"bvh2": H.LAYOUT_BVH2, "bvh4": H.LAYOUT_BVH4, "bvh4_fp16": H.LAYOUT_BVH4_FP16,
               "bvh4_fp16_anytree": H.LAYOUT_AUTO_FP16, "bvh8_fp16_anytree": H.LAYOUT_AUTO_FP16_WIDE}
    # the committed layout of a step: nearest hits on BVH2 in the reference
    # order, any-hit on the chosen tree
    step_layout = {"bvh2": H.LAYOUT_BVH2, "bvh4": H.LAYOUT_AUTO, "bvh4_fp16_anytree": H.LAYOUT_AUTO_FP16,
                   "bvh8_fp16_anytree": H.LAYOUT_AUTO_FP16_WIDE}

    log("generating scene")
    verts, faces = W.rtao_scene()
    # one eye ray per pixel, laid out in the renderers' 8x8-superblock order
    prim_rays = W.rtao_primary_rays(order="superblock")
    npix = prim_rays.shape[0]
    tmax = 0.1 * W.scene_diagonal(verts)
    n_ao = npix * SPP

    ctx = H.Context(0)
    scene = H.Scene(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gid = scene.add_triangles(verts, faces)
    build_ms = (time.perf_counter() - t0) * 1e3
    scene.add_instance(gid)
    t0 = time.perf_counter()
    scene.commit(layout=H.LAYOUT_BVH4)  # BVH4 collapse (FP32 + FP16 nodes)
    collapse_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    scene.commit(layout=H.LAYOUT_AUTO_FP16)  # the any-hit tree + its BVH4
    any_tree_ms = (time.perf_counter() - t0) * 1e3
    log(f"build {faces.shape[0]} tris: {build_ms:.2f} ms (host copy + LBVH), BVH4 collapse {collapse_ms:.2f} ms, "
        f"any-hit tree {any_tree_ms:.2f} ms")

    # one explicit stream (default priority) for every copy, kernel and event
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    d_prim = torch.from_numpy(prim_rays).to(dev)
    d_hits = torch.zeros((npix, 2), dtype=torch.int32, device=dev)
    d_ao = torch.empty((n_ao, 8), dtype=torch.float32, device=dev)
    d_vis = torch.zeros(n_ao, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def set_any_layout(name: str) -> None:
        scene.commit(layout=step_layout[name])

    def nearest():
        scene.trace_nearest(d_prim, npix, d_hits, H.HIT_COMPACT, stream=sptr)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        nearest()
        if ev is not None:
            ev[1].record(stream)
        if args.two_step:
            scene.rtao_make_rays(d_prim, d_hits, npix, SPP, tmax, SEED, d_ao, stream=sptr)
            if ev is not None:
                ev[2].record(stream)
            scene.trace_any(d_ao, n_ao, d_vis, stream=sptr)
        else:
            if ev is not None:
                ev[2].record(stream)
            scene.rtao_trace(d_prim, d_hits, npix, SPP, tmax, SEED, d_vis, stream=sptr)
        if ev is not None:
            ev[3].record(stream)

    set_any_layout(args.any_layout)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    step_ms = np.array([e[0].elapsed_time(e[3]) for e in events])
    near_ms = np.array([e[0].elapsed_time(e[1]) for e in events])
    gen_ms = np.array([e[1].elapsed_time(e[2]) for e in events])
    any_ms = np.array([e[2].elapsed_time(e[3]) for e in events])
    rays_per_step = npix + n_ao
    value = rays_per_step / (ms_per_step * 1e-3) / 1e6
    ao_kind = "raygen + any-hit" if args.two_step else "fused AO pass"
    log(f"step {ms_per_step:.3f} ms wall; nearest {near_ms.mean():.3f} ms, raygen {gen_ms.mean():.3f} ms, "
        f"any-hit[{args.any_layout}, {ao_kind}] {any_ms.mean():.3f} ms")

    # the same AO rays through HBM (hrt_rtao_make_rays + hrt_trace_any): the
    # counters, the layout comparison and the CPU baseline read them
    d_vis_timed = d_vis.clone()
    scene.rtao_make_rays(d_prim, d_hits, npix, SPP, tmax, SEED, d_ao, stream=sptr)
    d_vis2 = torch.zeros_like(d_vis)
    scene.trace_any(d_ao, n_ao, d_vis2, stream=sptr)
    torch.cuda.synchronize()
    ao_identical = bool(torch.equal(d_vis2, d_vis_timed))
    ev2 = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(5)]
    for e in ev2:
        e[0].record(stream)
        scene.rtao_make_rays(d_prim, d_hits, npix, SPP, tmax, SEED, d_ao, stream=sptr)
        e[1].record(stream)
        scene.trace_any(d_ao, n_ao, d_vis2, stream=sptr)
        e[2].record(stream)
    ev3 = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(5)]
    for e in ev3:
        e[0].record(stream)
        scene.rtao_trace(d_prim, d_hits, npix, SPP, tmax, SEED, d_vis2, stream=sptr)
        e[1].record(stream)
    torch.cuda.synchronize()
    ao_pass = {
        "fused_ms": round(float(np.mean([e[0].elapsed_time(e[1]) for e in ev3])), 4),
        "two_step_raygen_ms": round(float(np.mean([e[0].elapsed_time(e[1]) for e in ev2])), 4),
        "two_step_any_hit_ms": round(float(np.mean([e[1].elapsed_time(e[2]) for e in ev2])), 4),
        "identical": ao_identical,
    }
    log(f"AO pass: fused {ao_pass['fused_ms']:.3f} ms; two-step raygen {ao_pass['two_step_raygen_ms']:.3f} + "
        f"any-hit {ao_pass['two_step_any_hit_ms']:.3f} ms; identical {ao_identical}")

    # the whole 02_rtao frame on the device (hrt_rtao_render): eye rays of the
    # same camera in the renderer's 8x8-superblock order, nearest hits, the
    # fused AO pass and the AO image
    cam_pi, cam_wi = W.eye_camera(1024, 1024, eye=W.RTAO_EYE, target=W.RTAO_TARGET, fov_y_deg=W.RTAO_FOV)
    d_img = torch.zeros(npix, dtype=torch.int32, device=dev)
    for _ in range(2):
        scene.rtao_render(cam_pi, cam_wi, 1024, 1024, SPP, tmax, SEED, d_img, stream=sptr)
    evf = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(5)]
    for e in evf:
        e[0].record(stream)
        scene.rtao_render(cam_pi, cam_wi, 1024, 1024, SPP, tmax, SEED, d_img, stream=sptr)
        e[1].record(stream)
    torch.cuda.synchronize()
    frame_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evf]))
    frame = {"frame_ms": round(frame_ms, 4), "mrays_s": round(rays_per_step / (frame_ms * 1e-3) / 1e6, 1),
             "note": "hrt_rtao_render, 1024x1024, 16 AO rays per hit, eye rays made on the device"}
    log(f"RTAO frame on the device: {frame_ms:.3f} ms")

    # traversal counters (counting variant, untimed)
    def counters(layout_name: str, any_hit: bool):
        scene.commit(layout=step_layout[layout_name])
        st = torch.zeros(((n_ao if any_hit else npix), 2), dtype=torch.int32, device=dev)
        if any_hit:
            chk = torch.zeros_like(d_vis)
            scene.trace_any(d_ao, n_ao, chk, stream=sptr, stats=st)
        else:
            chk = torch.zeros_like(d_hits)
            scene.trace_nearest(d_prim, npix, chk, H.HIT_COMPACT, stream=sptr, stats=st)
        torch.cuda.synchronize()
        same = bool(torch.equal(chk, d_vis_timed if any_hit else d_hits))
        return st.cpu().numpy().view(np.uint32).astype(np.int64), same

    sp, same_p = counters("bvh2", False)
    sa, same_a = counters(args.any_layout, True)
    # per AO ray: its 32-byte record (the fused pass reads a 32-byte frame per
    # pixel instead), the 1-byte answer, node and triangle bytes
    ray_bytes = 32.0 if args.two_step else 32.0 / SPP
    bytes_any = float(np.sum(ray_bytes + 1 + NODE_BYTES[args.any_layout] * sa[:, 0] + 60 * sa[:, 1]))
    bytes_near = float(np.sum(40 + 64 * sp[:, 0] + 60 * sp[:, 1]))
    achieved_any = bytes_any / (any_ms.mean() * 1e-3) / 1e9
    achieved_near = bytes_near / (near_ms.mean() * 1e-3) / 1e9

    # the other any-hit layouts, same rays, same answers
    variants = {}
    if not args.no_variants:
        ref_vis = d_vis_timed
        for name in ("bvh2", "bvh4", "bvh4_fp16", "bvh4_fp16_anytree", "bvh8_fp16_anytree"):
            scene.commit(layout=layouts[name])
            out = torch.zeros_like(d_vis)
            for _ in range(2):
                scene.trace_any(d_ao, n_ao, out, stream=sptr)
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
            for k in range(5):
                evs[k].record(stream)
                scene.trace_any(d_ao, n_ao, out, stream=sptr)
            evs[5].record(stream)
            torch.cuda.synchronize()
            ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(5)])
            variants[name] = {"any_hit_ms": round(float(ms.mean()), 4),
                              "any_hit_mrays_s": round(n_ao / (ms.mean() * 1e-3) / 1e6, 1),
                              "identical_to_timed": bool(torch.equal(out, ref_vis))}
            log(f"any-hit variant {name}: {ms.mean():.3f} ms")

    result = {
        "metric": "Mrays/s closest/any-hit on 1 MI355X, % HBM roofline; p50/p99 query latency",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "rtao: 1024x1024 primary nearest-hit (LBVH2, reference order; rays in 8x8-superblock "
                        "pixel order) + 16 AO any-hit rays per "
                        f"hit ({args.any_layout}; "
                        + ("rays written to HBM, then traced" if args.two_step else
                           "rays made inside the any-hit traversal") + "), 1,072,576-tri heightfield+boxes",
            "triangles": int(faces.shape[0]),
            "primary_rays": int(npix),
            "ao_rays": int(n_ao),
            "ao_tmax": round(tmax, 6),
            "any_hit_layout": args.any_layout,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"any-hit traversal ({args.any_layout}, {ao_kind})",
            "achieved": round(achieved_any, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved_any / HBM_PEAK_GBPS, 4),
            "traffic": measured_traffic(args.any_layout, "two_step" if args.two_step else "fused"),
            "traffic_source": "profiles/rtao_anyhit_traffic.json (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
            "bytes_per_launch": bytes_any,
            "mean_launch_ms": round(float(any_ms.mean()), 4),
        },
        "nearest_mrays_s": round(npix / (near_ms.mean() * 1e-3) / 1e6, 2),
        "anyhit_mrays_s": round(n_ao / (any_ms.mean() * 1e-3) / 1e6, 2),
        "nearest_roofline_frac": round(achieved_near / HBM_PEAK_GBPS, 4),
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "visits_per_ray": {"primary_pairs": round(float(sp[:, 0].mean()), 3),
                           f"ao_{args.any_layout}_nodes": round(float(sa[:, 0].mean()), 3)},
        "tris_per_ray": {"primary": round(float(sp[:, 1].mean()), 3), "ao": round(float(sa[:, 1].mean()), 3)},
        "build_ms": round(build_ms, 3),
        "bvh4_collapse_ms": round(collapse_ms, 3),
        "any_hit_tree_build_ms": round(any_tree_ms, 3),
        "counting_variant_identical": same_p and same_a,
        "ao_pass": ao_pass,
        "device_frame": frame,
        "any_hit_layouts": variants,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle build + trace on {threads} threads")
        ob = O.OracleBVH(verts, faces)
        ao_host = d_ao.cpu().numpy()
        ao_host[:, 7] = -np.abs(ao_host[:, 7])  # any-hit convention: negative tFar
        hits_host = d_hits.cpu().numpy().view(np.uint32)
        vis_host = d_vis_timed.cpu().numpy().astype(bool)
        t0 = time.perf_counter()
        tp, pp, _, _ = ob.trace(prim_rays, threads=threads)
        _, pa, _, _ = ob.trace(ao_host, threads=threads)
        cpu_s = time.perf_counter() - t0
        mism = int(np.sum(pp != hits_host[:, 1]) + np.sum(tp.view(np.uint32) != hits_host[:, 0])
                   + np.sum((pa != 0xFFFFFFFF) != vis_host))
        result["cpu_baseline"] = {
            "value": round(rays_per_step / cpu_s / 1e6, 3),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"one full step on the host: {npix} primary nearest-hit + {n_ao} AO any-hit rays "
                      f"(BVH2CommonRT restatement, OpenMP dynamic,256), {cpu_s:.2f} s",
        }
        result["parity_mismatches_vs_oracle"] = mism
        log(f"cpu {cpu_s:.2f} s, mismatches {mism}")

    scene.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


def bench_batched(args) -> int:
    """256 random-soup scenes (1k-200k triangles, log-uniform) with 4k-256k
    incoherent rays each: every LBVH build and every nearest-hit trace in one
    hipGraph launch per step, trace kernel picked by the autotuner."""
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    log("generating 256 scenes")
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
    log(f"{len(scenes)} scenes, {T} triangles, {R} rays; autotuning the trace kernel")
    chosen, cand_ms = b.autotune(reps=3)
    cands = H.batch_candidates()
    log(f"chosen candidate {chosen} {cands[chosen]}: {cand_ms[chosen]:.3f} ms")

    def timed(fn, reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        torch.cuda.synchronize()
        for k in range(reps):
            ev[k].record(stream)
            fn()
        ev[reps].record(stream)
        torch.cuda.synchronize()
        return np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(reps)])

    for _ in range(args.warmup):
        b.run(sptr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ms = timed(lambda: b.run(sptr), args.steps)
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    build_ms = timed(lambda: b.build(sptr), 5)
    trace_ms = timed(lambda: b.trace(sptr), 5)
    hits = b.hits()
    ms = float(np.median(step_ms))
    log(f"graph step {ms:.3f} ms (build {build_ms.mean():.3f} + trace {trace_ms.mean():.3f} as plain launches)")

    # the plain path: one scene at a time through the single-scene API, rays
    # already on the device; also the hit-parity reference and the counters
    log("plain per-scene path")
    d_rays = [torch.from_numpy(x[2]).cuda() for x in scenes]
    d_out = [torch.empty((n, 2), dtype=torch.int32, device="cuda") for n in nr]
    d_st = [torch.empty((n, 2), dtype=torch.int32, device="cuda") for n in nr]
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
    bytes_trace = 0.0
    for s, sc in enumerate(solo):
        ref = d_out[s].cpu().numpy().view(np.uint32)
        mism += int(np.sum(ref != hits[b.ray_off[s]:b.ray_off[s + 1]]))
        sc.trace_nearest(d_rays[s], nr[s], d_out[s], H.HIT_COMPACT, stream=sptr, stats=d_st[s])
    torch.cuda.synchronize()
    for s in range(len(scenes)):
        st = d_st[s].cpu().numpy().view(np.uint32).astype(np.int64)
        bytes_trace += float(np.sum(40 + 64 * st[:, 0] + 60 * st[:, 1]))
    for sc in solo:
        sc.close()
    achieved = bytes_trace / (trace_ms.mean() * 1e-3) / 1e9
    log(f"plain path {plain_ms:.1f} ms, hit mismatches {mism}")

    result = {
        "metric": "Mrays/s closest/any-hit on 1 MI355X, % HBM roofline; p50/p99 query latency",
        "value": round(R / (ms * 1e-3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: 256 random soups (1k-200k tris, log-uniform) with 4k-256k incoherent "
                               "rays each; all LBVH builds + nearest-hit traces in one hipGraph launch",
                   "scenes": len(scenes), "triangles": T, "rays": R},
        "roofline": {"bound": "hbm", "kernel": "batched nearest-hit traversal",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "bytes_per_launch": bytes_trace, "mean_launch_ms": round(float(trace_ms.mean()), 4)},
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "wall_ms_per_step": round(wall_ms, 4),
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
        "parity_mismatches_vs_plain": mism,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = scenes[:24]
        rs = sum(x[2].shape[0] for x in sample)
        t0 = time.perf_counter()
        for v, f, rays in sample:
            O.OracleBVH(v, f).trace(rays, threads=threads)
        cpu_s = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": round(rs / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"first 24 of the 256 scenes: oracle LBVH build + nearest-hit trace of "
                                            f"{rs} rays, {cpu_s:.2f} s"}
    b.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


def _timed(stream, fn, reps):
    import torch

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    torch.cuda.synchronize()
    for k in range(reps):
        ev[k].record(stream)
        fn()
    ev[reps].record(stream)
    torch.cuda.synchronize()
    return np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(reps)])


def _line(value, unit, ms, steps, warmup, workload, extra):
    out = {"metric": "Mrays/s closest/any-hit on 1 MI355X, % HBM roofline; p50/p99 query latency",
           "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic", "config": {"workload": workload}}
    out.update(extra)
    return out


def bench_eye(args) -> int:
    """The CPU-runnable eye-ray configuration: 512x512 pinhole rays, nearest
    hit on the LBVH2 of an icosphere (6 subdivisions, 81,920 tris) on a
    2-triangle ground quad; primId and t per pixel, bit-exact with the oracle.
    The rays are made on the device in the renderer's 8x8-superblock order
    (hrt_eye_make_rays); the whole frame (rays, trace, colours) is timed too."""
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
    d_hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    d_st = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    for _ in range(args.warmup):
        sc.trace_nearest(d_rays, n, d_hits, H.HIT_COMPACT, stream=sptr)
    ms = _timed(stream, lambda: sc.trace_nearest(d_rays, n, d_hits, H.HIT_COMPACT, stream=sptr), args.steps)
    hits = d_hits.cpu().numpy().view(np.uint32)
    sc.trace_nearest(d_rays, n, torch.empty_like(d_hits), H.HIT_COMPACT, stream=sptr, stats=d_st)
    torch.cuda.synchronize()
    st = d_st.cpu().numpy().view(np.uint32).astype(np.int64)
    nbytes = float(np.sum(40 + 64 * st[:, 0] + 60 * st[:, 1]))
    mean = float(ms.mean())
    # the whole frame: rays, trace, colours (hrt_eye_render)
    d_color = torch.empty(n, dtype=torch.int32, device="cuda")
    frame = lambda: sc.eye_render(cam[0], cam[1], width, height, W.EYE_PALETTE, d_color, stream=sptr)
    frame()
    frame_ms = _timed(stream, frame, args.steps)
    # the same rays in row-major pixel order, for the effect of the tile order
    xy_rows = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.eye_make_rays(cam[0], cam[1], width, height, d_rays, xy_rows, stream=sptr)
    torch.cuda.synchronize()
    xy = xy_rows.cpu().numpy().view(np.uint32)
    order = np.argsort((xy >> 16).astype(np.int64) * width + (xy & 0xFFFF))
    d_rows = torch.from_numpy(rays[order]).cuda()
    rows_ms = _timed(stream, lambda: sc.trace_nearest(d_rows, n, d_hits, H.HIT_COMPACT, stream=sptr), args.steps)
    result = _line(round(n / (mean * 1e-3) / 1e6, 2), "Mrays/s", round(mean, 4), args.steps, args.warmup,
                   "eye: 512x512 pinhole rays (made on the device in 8x8 superblock order), nearest hit, "
                   "LBVH2 of an icosphere (81,920 tris) + ground quad",
                   {"roofline": {"bound": "hbm", "kernel": "nearest-hit traversal (BVH2, reference order)",
                                 "achieved": round(nbytes / (mean * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBPS,
                                 "unit": "GB/s", "frac": round(nbytes / (mean * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                                 "traffic": None, "bytes_per_launch": nbytes},
                    "latency_ms": {"p50": round(float(np.percentile(ms, 50)), 4),
                                   "p99": round(float(np.percentile(ms, 99)), 4)},
                    "build_ms": round(build_ms, 3), "triangles": int(f.shape[0]), "rays": n,
                    "visits_per_ray": round(float(st[:, 0].mean()), 3),
                    "frame_ms": round(float(frame_ms.mean()), 4),
                    "row_major_order_ms": round(float(rows_ms.mean()), 4), "cpu_baseline": None})
    from oracle import oracle as O

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    ob = O.OracleBVH(v, f)
    ref_rays, _ = O.eye_rays(cam[0], cam[1], width, height)
    result["rays_bit_exact_vs_oracle"] = bool(np.array_equal(ref_rays.view(np.uint32), rays.view(np.uint32)))
    t0 = time.perf_counter()
    t, p, _, _ = ob.trace(ref_rays, threads=threads)
    cpu_s = time.perf_counter() - t0
    result["cpu_baseline"] = {"value": round(n / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                              "kind": "port", "sample": f"the full 512x512 frame, {cpu_s * 1e3:.1f} ms"}
    result["parity_mismatches_vs_oracle"] = int(np.sum(t.view(np.uint32) != hits[:, 0]) + np.sum(p != hits[:, 1]))
    sc.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


def _t_fp64(v, f, rays, prim):
    """Float64 Moller-Trumbore t of each ray against its hit triangle (NaN on a miss)."""
    out = np.full(rays.shape[0], np.nan)
    hit = prim != 0xFFFFFFFF
    tri = f[prim[hit].astype(np.int64)].astype(np.int64)
    A, B, C = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[hit, 0:3].astype(np.float64)
    d = rays[hit, 4:7].astype(np.float64)
    e1, e2 = B - A, C - A
    pv = np.cross(d, e2)
    tv = o - A
    qv = np.cross(tv, e1)
    out[hit] = np.sum(e2 * qv, axis=1) / np.sum(e1 * pv, axis=1)
    return out


def bench_precision(args) -> int:
    """4M-triangle random soup, 4M incoherent rays, nearest hit on BVH4 with
    FP32 boxes and with FP16 boxes rounded outward: hits must be identical;
    t is compared with a float64 Moller-Trumbore on the hit triangle."""
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    n = 4 * 1024 * 1024
    log("generating the 4M-triangle soup and 4M rays")
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
    for name, lay in (("fp32", H.LAYOUT_BVH4), ("fp16", H.LAYOUT_BVH4_FP16)):
        sc.commit(layout=lay)
        out = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        for _ in range(args.warmup):
            sc.trace_nearest(d_rays, n, out, H.HIT_COMPACT, stream=sptr)
        timing[name] = _timed(stream, lambda: sc.trace_nearest(d_rays, n, out, H.HIT_COMPACT, stream=sptr), args.steps)
        res[name] = out.cpu().numpy().view(np.uint32).copy()
        log(f"{name} nodes: {timing[name].mean():.3f} ms")
    # counting variant on the FP16 nodes: algorithmic bytes of the timed kernel
    st = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    sc.trace_nearest(d_rays, n, torch.empty((n, 2), dtype=torch.int32, device="cuda"), H.HIT_COMPACT, stream=sptr,
                     stats=st)
    torch.cuda.synchronize()
    stc = st.cpu().numpy().view(np.uint32).astype(np.int64)
    nbytes = float(np.sum(40 + 64 * stc[:, 0] + 60 * stc[:, 1]))
    del st
    same = bool(np.array_equal(res["fp32"], res["fp16"]))
    mism = int(np.sum(np.any(res["fp32"] != res["fp16"], axis=1)))
    prim = res["fp16"][:, 1]
    t = res["fp16"][:, 0].view(np.float32).astype(np.float64)
    t64 = _t_fp64(v, f, rays, prim)
    hit = prim != 0xFFFFFFFF
    rel = np.abs(t[hit] - t64[hit]) / np.abs(t64[hit])
    ms16 = float(timing["fp16"].mean())
    result = _line(round(n / (ms16 * 1e-3) / 1e6, 2), "Mrays/s", round(ms16, 4), args.steps, args.warmup,
                   "precision: 4,194,304-tri random soup, 4,194,304 incoherent nearest-hit rays, BVH4 with FP16 "
                   "boxes rounded outward (value) vs FP32 boxes",
                   {"fp32_nodes": {"ms": round(float(timing["fp32"].mean()), 4),
                                   "mrays_s": round(n / (timing["fp32"].mean() * 1e-3) / 1e6, 2)},
                    "fp16_nodes": {"ms": round(ms16, 4), "mrays_s": round(n / (ms16 * 1e-3) / 1e6, 2)},
                    "roofline": {"bound": "hbm", "kernel": "nearest-hit traversal (BVH4, FP16 nodes)",
                                 "achieved": round(nbytes / (ms16 * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBPS,
                                 "unit": "GB/s", "frac": round(nbytes / (ms16 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                                 "traffic": None, "bytes_per_launch": nbytes},
                    "visits_per_ray": round(float(stc[:, 0].mean()), 2),
                    "tris_per_ray": round(float(stc[:, 1].mean()), 2),
                    "hits_bit_identical": same, "mismatching_rays": mism, "hit_fraction": round(float(hit.mean()), 4),
                    "t_rel_error_vs_fp64": {"max": float(rel.max()), "mean": float(rel.mean()),
                                            "median": float(np.median(rel)), "p99": float(np.percentile(rel, 99))},
                    "latency_ms": {"p50": round(float(np.percentile(timing["fp16"], 50)), 4),
                                   "p99": round(float(np.percentile(timing["fp16"], 99)), 4)},
                    "cpu_baseline": None})
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline: oracle build of the 4M-tri soup, FP16 BVH4 trace of a 262,144-ray sample")
        ob = O.OracleBVH(v, f)
        ns = 262144
        t0 = time.perf_counter()
        tc, pc, _, _ = ob.trace4(rays[:ns], O.MODE4_FP16, threads=threads)
        cpu_s = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": round(ns / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                                  "kind": "port", "sample": f"first {ns} rays, FP16 BVH4 traversal (BVH4HalfRT "
                                                            f"restatement), {cpu_s:.2f} s"}
        result["parity_mismatches_vs_oracle"] = int(np.sum(tc.view(np.uint32) != res["fp16"][:ns, 0]) +
                                                    np.sum(pc != res["fp16"][:ns, 1]))
    sc.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


def bench_serving(args) -> int:
    """16 resident scenes (50k-2M triangles, log-uniform; even ids random
    soups, odd ids heightfields); requests of 1k-1M rays (log-uniform) arrive
    as a Poisson stream, half nearest-hit eye rays and half any-hit AO rays,
    each for a uniformly chosen scene.  Requests point into per-scene pinned
    ray pools (1024x1024 eye rays, and one AO ray per eye hit made on the
    device), so nothing is copied on the host inside the measurement."""
    import torch

    import heavyrt_amd as H
    from heavyrt_amd import workloads as W

    ctx = H.Context(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    nt = W.log_uniform_int(16, 5e4, 2e6, W.STREAM_SERVE)
    scenes, pools = [], []
    t_build = 0.0
    log("building 16 resident scenes")
    for k in range(16):
        if k % 2 == 0:
            v, f = W.soup(int(nt[k]), stream=W.STREAM_SERVE + 10 + k)
        else:
            nz = max(8, int(round(math.sqrt(nt[k] / 4.0))))
            v, f, _ = W.heightfield(2 * nz, nz)
        torch.cuda.synchronize()
        t = time.perf_counter()
        sc = H.Scene(ctx)
        sc.add_instance(sc.add_triangles(v, f))
        sc.commit()
        torch.cuda.synchronize()
        t_build += time.perf_counter() - t
        lo, hi = v.min(0), v.max(0)
        c = 0.5 * (lo + hi)
        ext = float(np.max(hi - lo))
        eye = c + np.array([0.35, 0.8, 1.6]) * ext
        eye_rays = W.pinhole_rays(1024, 1024, eye=tuple(eye), target=tuple(c), fov_y_deg=50.0)
        d_eye = torch.from_numpy(eye_rays).cuda()
        d_hit = torch.empty((eye_rays.shape[0], 2), dtype=torch.int32, device="cuda")
        d_ao = torch.empty((eye_rays.shape[0], 8), dtype=torch.float32, device="cuda")
        sc.trace_nearest(d_eye, eye_rays.shape[0], d_hit, H.HIT_COMPACT, stream=sptr)
        sc.rtao_make_rays(d_eye, d_hit, eye_rays.shape[0], 1, 0.1 * ext, SEED + k, d_ao, stream=sptr)
        torch.cuda.synchronize()
        n = eye_rays.shape[0]
        pool = {"eye": H.pinned_array((n, 8), np.float32), "ao": H.pinned_array((n, 8), np.float32),
                "eye_out": H.pinned_array((n, 2), np.uint32), "ao_out": H.pinned_array((n,), np.uint8)}
        pool["eye"][:] = eye_rays
        pool["ao"][:] = d_ao.cpu().numpy()
        scenes.append(sc)
        pools.append(pool)
    tris = int(sum(int(x) for x in nt))
    srv = H.Server(ctx, 4 << 30, steps_in_flight=3)
    ids = [srv.add_scene(sc) for sc in scenes]
    npool = 1024 * 1024
    levels = []
    for li, qps in enumerate(float(q) for q in args.qps.split(",")):
        nreq = args.requests
        sizes = W.log_uniform_int(nreq, 1e3, 1e6, W.STREAM_SERVE + 100 + li)
        u = W.uniform01(W.pcg32(nreq, 4, W.STREAM_SERVE + 200 + li))
        gaps = -np.log(np.maximum(u[:, 0], 1e-12)) / qps
        arrivals = np.cumsum(gaps)
        sc_id = np.minimum((u[:, 1] * 16).astype(np.int64), 15)
        kind_any = u[:, 2] < 0.5
        starts = (u[:, 3] * (npool - sizes)).astype(np.int64)
        log(f"qps {qps:.0f}: {nreq} requests, {int(sizes.sum())} rays")
        rids, t_sub = [], []
        t0 = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            p = pools[sc_id[i]]
            a, n = int(starts[i]), int(sizes[i])
            if kind_any[i]:
                rid = srv.submit(ids[sc_id[i]], p["ao"][a:a + n], p["ao_out"][a:a + n], any_hit=True)
            else:
                rid = srv.submit(ids[sc_id[i]], p["eye"][a:a + n], p["eye_out"][a:a + n], any_hit=False)
            t_sub.append(time.perf_counter() - t0)
            rids.append(rid)
        lat = np.array([srv.wait(r)[0] for r in rids])
        t_end = max(ts + l * 1e-3 for ts, l in zip(t_sub, lat))
        st = srv.stats()
        levels.append({"qps_offered": qps, "qps_achieved": round(nreq / t_end, 1),
                       "mrays_s": round(float(sizes.sum()) / t_end / 1e6, 1),
                       "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3),
                       "mean_requests_per_step": st["mean_requests_per_step"],
                       "peak_bytes_in_flight": st["peak_bytes_in_flight"]})
        log(f"  -> {levels[-1]}")
    # parity: every pool entry written by some request equals a solo trace
    mism = 0
    for k, (sc, p) in enumerate(zip(scenes, pools)):
        d = torch.from_numpy(np.asarray(p["eye"])).cuda()
        o = torch.empty((npool, 2), dtype=torch.int32, device="cuda")
        sc.trace_nearest(d, npool, o, H.HIT_COMPACT, stream=sptr)
        da = torch.from_numpy(np.asarray(p["ao"])).cuda()
        oa = torch.empty(npool, dtype=torch.uint8, device="cuda")
        sc.trace_any(da, npool, oa, stream=sptr)
        torch.cuda.synchronize()
        ref = o.cpu().numpy().view(np.uint32)
        got = np.asarray(p["eye_out"])
        written = np.any(got != 0, axis=1)
        mism += int(np.sum(np.any(got[written] != ref[written], axis=1)))
        refa = oa.cpu().numpy()
        gota = np.asarray(p["ao_out"])
        mism += int(np.sum(gota[gota != 0] != refa[gota != 0]))
    # warm start: every resident scene saved (hrt_scene_serialize) and restored
    # (hrt_scene_deserialize) from host memory, against building it from its triangles
    log("warm start: snapshots of the 16 scenes")
    t_ser = t_de = 0.0
    nbytes, restored_same = 0, True
    for sc in scenes:
        torch.cuda.synchronize()
        t = time.perf_counter()
        data = sc.serialize()
        t_ser += time.perf_counter() - t
        t = time.perf_counter()
        r = H.Scene.deserialize(ctx, data)
        torch.cuda.synchronize()
        t_de += time.perf_counter() - t
        nbytes += len(data)
        restored_same = restored_same and r.serialize() == data
        r.close()
        del data
    warm = {"build_from_triangles_ms": round(1e3 * t_build, 1), "serialize_ms": round(1e3 * t_ser, 1),
            "deserialize_ms": round(1e3 * t_de, 1), "snapshot_mb": round(nbytes / 1e6, 1),
            "restored_bytes_identical": bool(restored_same)}
    log(f"  -> {warm}")
    best = max(levels, key=lambda x: x["mrays_s"])
    result = {
        "metric": "Mrays/s closest/any-hit on 1 MI355X, % HBM roofline; p50/p99 query latency",
        "value": best["mrays_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serving: 16 resident scenes (50k-2M tris, soups and heightfields), Poisson "
                               "requests of 1k-1M rays (log-uniform), 50/50 nearest-hit eye rays / any-hit AO "
                               "rays, 4 GiB HBM budget, 3 steps in flight",
                   "triangles": tris, "requests_per_rate": args.requests},
        "latency_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"]},
        "rates": levels,
        "parity_mismatches_vs_solo": mism,
        "warm_start": warm,
        "cpu_baseline": None,
    }
    srv.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
