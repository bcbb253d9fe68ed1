"""ISceneObject drop-in: a C++ caller written against CrossRT.h drives the GPU
engine the way the reference's renderers do, and gets the oracle's hits.

The harness (tests/cpp/scene_harness.cpp) builds through CreateSceneRT /
ClearGeom / AddGeom_Triangles3f (float4 vertices, stride 16) / AddInstance /
CommitScene, asks one RayQuery_NearestHit and one RayQuery_AnyHit per ray, and
repeats both through the batch extensions.
"""
import os
import subprocess

import numpy as np
import pytest

from heavyrt_amd import build as B
from heavyrt_amd import workloads as W
from gpu_util import say
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HIT = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("geom", "<u4"), ("c", "<f4", (4,))])


def run_harness(tmp_path, impl, v, f, rays):
    mesh = tmp_path / "mesh.bin"
    with open(mesh, "wb") as fh:
        v4 = np.ones((v.shape[0], 4), np.float32)
        v4[:, :3] = v
        fh.write(np.array([v.shape[0], f.size], np.uint32).tobytes())
        fh.write(v4.tobytes())
        fh.write(np.ascontiguousarray(f, np.uint32).tobytes())
    rf = tmp_path / "rays.bin"
    with open(rf, "wb") as fh:
        fh.write(np.array([rays.shape[0]], np.uint32).tobytes())
        fh.write(np.ascontiguousarray(rays, np.float32).tobytes())
    out = tmp_path / f"out_{impl}.bin"
    r = subprocess.run([B.HARNESS_BIN, impl, str(mesh), str(rf), str(out)], capture_output=True, text=True,
                       timeout=120)
    say(r.stdout.strip())
    assert r.returncode == 0, r.stderr + r.stdout
    n = rays.shape[0]
    data = open(out, "rb").read()
    off = 0
    per_ray = np.frombuffer(data, HIT, n, off)
    off += n * HIT.itemsize
    batch = np.frombuffer(data, HIT, n, off)
    off += n * HIT.itemsize
    any_per_ray = np.frombuffer(data, np.uint8, n, off).astype(bool)
    off += n
    any_batch = np.frombuffer(data, np.uint8, n, off).astype(bool)
    off += n
    counts = np.frombuffer(data, np.uint32, 2, off)
    return per_ray, batch, any_per_ray, any_batch, counts, r.stdout


@pytest.mark.parametrize("impl", ["HipBVH2", "HipBVH4", "HipBVH4Half", "HipAuto", "HipAutoHalf"])
def test_renderer_call_pattern(tmp_path, gpu_ctx, impl):
    assert os.path.exists(B.HARNESS_BIN), "build the harness first (python -m heavyrt_amd.build)"
    v, f = W.merge(W.icosphere(3), W.ground_quad())
    rays = W.pinhole_rays(40, 40, eye=(0.0, 1.5, 4.0), target=(0.0, 1.0, 0.0))
    say(f"harness {impl}: {f.shape[0]} tris, {rays.shape[0]} rays")
    per_ray, batch, any_pr, any_b, counts, _ = run_harness(tmp_path, impl, v, f, rays)
    ob = O.OracleBVH(v, f)
    if impl in ("HipBVH2", "HipAuto", "HipAutoHalf"):
        t, p, u, vv = ob.trace(rays)
    else:
        t, p, u, vv = ob.trace4(rays, O.MODE4_FP16 if impl == "HipBVH4Half" else O.MODE4_FP32)
    for rec in (per_ray, batch):
        np.testing.assert_array_equal(rec["t"].view(np.uint32), t.view(np.uint32))
        np.testing.assert_array_equal(rec["prim"], p)
        np.testing.assert_array_equal(rec["c"][:, 0].view(np.uint32), u.view(np.uint32))
        np.testing.assert_array_equal(rec["c"][:, 1].view(np.uint32), vv.view(np.uint32))
        hit = p != 0xFFFFFFFF
        assert (rec["geom"][hit] == 0).all() and (rec["inst"][hit] == 0).all()
        assert (rec["geom"][~hit] == 0xFFFFFFFF).all()
    np.testing.assert_array_equal(any_pr, p != 0xFFFFFFFF)
    np.testing.assert_array_equal(any_b, p != 0xFFFFFFFF)
    assert list(counts) == [1, 1]


def test_unknown_name_falls_back_to_default_engine(tmp_path, gpu_ctx):
    v, f = W.icosphere(2)
    rays = W.pinhole_rays(16, 16, eye=(0.0, 1.0, 4.0), target=(0.0, 1.0, 0.0))
    per_ray, _, _, _, _, out = run_harness(tmp_path, "Embree", v, f, rays)
    assert "unknown implementation" in out
    t, p, _, _ = O.OracleBVH(v, f).trace(rays)
    np.testing.assert_array_equal(per_ray["prim"], p)
