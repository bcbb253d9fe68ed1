"""Ray-query throughput on one MI355X: the RTAO workload.

Workload (one step): a 1024x512-quad heightfield (value noise, amp 0.2) plus
2,000 randomly rotated boxes = 1,072,576 triangles in one LBVH2; 1024x1024
primary rays traced nearest-hit; 16 cosine-weighted AO rays per primary hit
made on the device (tMax = 0.1 x scene diagonal); all AO rays traced any-hit.
Rays and scene are resident in HBM before the timed region starts.

Prints ONE JSON line.  `value` is total rays (primary + AO) per second over
the timed steps.  `roofline` is for the dominant kernel (the any-hit
traversal): algorithmic bytes per launch = sum over rays of
32 (ray) + 1 (result) + 64 x node pairs visited + 60 x triangles tested,
the per-ray figure the reference's bus-load metric defines
(core/CrossRT.h:52, core/BVH2CommonRT.cpp:69-70,125), with the visit and test
counts taken from the engine's counting variant of the same kernel; divided
by the kernel's mean duration, measured with HIP events on the stream the
kernel runs on.  `cpu_baseline` is the CPU oracle (a C++ restatement of the
reference's BVH2CommonRT traversal) timed on this host on the same rays.
"""
from __future__ import annotations

import argparse
import json
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, os.cpu_count())")
    args = ap.parse_args()
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
    dev = torch.device("cuda:0")

    log("generating scene")
    verts, faces = W.rtao_scene()
    prim_rays = W.rtao_primary_rays()
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
    scene.commit()
    log(f"build {faces.shape[0]} tris: {build_ms:.2f} ms (host->device copy + LBVH)")

    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    d_prim = torch.from_numpy(prim_rays).to(dev)
    d_hits = torch.zeros((npix, 2), dtype=torch.int32, device=dev)
    d_ao = torch.empty((n_ao, 8), dtype=torch.float32, device=dev)
    d_vis = torch.zeros(n_ao, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        scene.trace_nearest(d_prim, npix, d_hits, H.HIT_COMPACT, stream=sptr)
        if ev is not None:
            ev[1].record(stream)
        scene.rtao_make_rays(d_prim, d_hits, npix, SPP, tmax, SEED, d_ao, stream=sptr)
        if ev is not None:
            ev[2].record(stream)
        scene.trace_any(d_ao, n_ao, d_vis, stream=sptr)
        if ev is not None:
            ev[3].record(stream)

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
    log(f"step {ms_per_step:.3f} ms wall; nearest {near_ms.mean():.3f} ms, raygen {gen_ms.mean():.3f} ms, "
        f"any-hit {any_ms.mean():.3f} ms")

    # traversal counters (counting variant of the same kernels, untimed)
    st_prim = torch.zeros((npix, 2), dtype=torch.int32, device=dev)
    st_ao = torch.zeros((n_ao, 2), dtype=torch.int32, device=dev)
    hits_chk = torch.zeros_like(d_hits)
    vis_chk = torch.zeros_like(d_vis)
    scene.trace_nearest(d_prim, npix, hits_chk, H.HIT_COMPACT, stream=sptr, stats=st_prim)
    scene.trace_any(d_ao, n_ao, vis_chk, stream=sptr, stats=st_ao)
    torch.cuda.synchronize()
    sp = st_prim.cpu().numpy().view(np.uint32).astype(np.int64)
    sa = st_ao.cpu().numpy().view(np.uint32).astype(np.int64)
    same = bool(torch.equal(hits_chk, d_hits) and torch.equal(vis_chk, d_vis))
    bytes_any = float(np.sum(32 + 1 + 64 * sa[:, 0] + 60 * sa[:, 1]))
    bytes_near = float(np.sum(32 + 8 + 64 * sp[:, 0] + 60 * sp[:, 1]))
    achieved_any = bytes_any / (any_ms.mean() * 1e-3) / 1e9
    achieved_near = bytes_near / (near_ms.mean() * 1e-3) / 1e9

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
            "workload": "rtao: 1024x1024 primary nearest-hit + 16 AO any-hit rays per hit, "
                        "1,072,576-tri heightfield+boxes, LBVH2",
            "triangles": int(faces.shape[0]),
            "primary_rays": int(npix),
            "ao_rays": int(n_ao),
            "ao_tmax": round(tmax, 6),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace_bvh2<any-hit>",
            "achieved": round(achieved_any, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved_any / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "bytes_per_launch": bytes_any,
            "mean_launch_ms": round(float(any_ms.mean()), 4),
        },
        "nearest_mrays_s": round(npix / (near_ms.mean() * 1e-3) / 1e6, 2),
        "anyhit_mrays_s": round(n_ao / (any_ms.mean() * 1e-3) / 1e6, 2),
        "nearest_roofline_frac": round(achieved_near / HBM_PEAK_GBPS, 4),
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "avg_pairs_per_ray": {"primary": round(float(sp[:, 0].mean()), 3), "ao": round(float(sa[:, 0].mean()), 3)},
        "avg_tris_per_ray": {"primary": round(float(sp[:, 1].mean()), 3), "ao": round(float(sa[:, 1].mean()), 3)},
        "build_ms": round(build_ms, 3),
        "repeat_identical": same,
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
        vis_host = d_vis.cpu().numpy().astype(bool)
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


if __name__ == "__main__":
    sys.exit(main())
