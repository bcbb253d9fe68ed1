"""Helpers for the GPU parity tests: upload rays, run the engine, fetch hits."""
import sys

import numpy as np

import heavyrt_amd as H


def say(msg: str) -> None:
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def make_scene(ctx, v, f, keep=True):
    s = H.Scene(ctx, keep_build_arrays=keep)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit()
    return s


def trace(scene, rays: np.ndarray, mode: str = "compact", stats: bool = False):
    """mode: 'compact' -> (t, prim); 'full' -> (t, prim, inst, geom, u, v); 'any' -> bool."""
    import torch

    rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    n = rays.shape[0]
    dev = torch.from_numpy(rays).cuda()
    st = torch.zeros((max(n, 1), 2), dtype=torch.int32, device="cuda") if stats else None
    stream = torch.cuda.current_stream().cuda_stream
    assert stream, "tests run on an explicit torch stream (see conftest.gpu_ctx)"
    if mode == "any":
        out = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")
        scene.trace_any(dev, n, out, stream=stream, stats=st)
    elif mode == "full":
        out = torch.zeros((max(n, 1), 8), dtype=torch.int32, device="cuda")
        scene.trace_nearest(dev, n, out, H.HIT_FULL, stream=stream, stats=st)
    else:
        out = torch.zeros((max(n, 1), 2), dtype=torch.int32, device="cuda")
        scene.trace_nearest(dev, n, out, H.HIT_COMPACT, stream=stream, stats=st)
    torch.cuda.synchronize()
    o = out.cpu().numpy()[:n]
    res = {}
    if mode == "any":
        res["hit"] = o.astype(bool)
    else:
        o = o.view(np.uint32)
        res["t"] = o[:, 0].view(np.float32)
        res["prim"] = o[:, 1].copy()
        if mode == "full":
            res["inst"] = o[:, 2].copy()
            res["geom"] = o[:, 3].copy()
            res["u"] = o[:, 4].view(np.float32)
            res["v"] = o[:, 5].view(np.float32)
    if stats:
        s = st.cpu().numpy()[:n].view(np.uint32)
        res["pairs"] = s[:, 0].copy()
        res["tris"] = s[:, 1].copy()
    return res
