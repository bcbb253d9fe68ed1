"""hrt_rtao_trace: the AO pass in one call equals hrt_rtao_make_rays followed by
hrt_trace_any, byte for byte, on every layout (BVH4 layouts make the rays
inside the traversal; the others go through a ray buffer)."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say
from heavyrt_amd import workloads as W

pytestmark = pytest.mark.gpu

LAYOUTS = [H.LAYOUT_BVH2, H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO, H.LAYOUT_AUTO_FP16,
           H.LAYOUT_AUTO_FP16_WIDE]
IDS = ["bvh2", "bvh4", "bvh4_fp16", "auto", "auto_fp16", "auto_fp16_wide"]


def two_step_and_fused(scene, prim, spp, tmax, seed):
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    npix = prim.shape[0]
    dprim = torch.from_numpy(prim).cuda()
    dhits = torch.zeros((npix, 2), dtype=torch.int32, device="cuda")
    scene.trace_nearest(dprim, npix, dhits, H.HIT_COMPACT, stream=stream)
    n = npix * spp
    dao = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    scene.rtao_make_rays(dprim, dhits, npix, spp, tmax, seed, dao, stream=stream)
    a = torch.zeros(n, dtype=torch.uint8, device="cuda")
    scene.trace_any(dao, n, a, stream=stream)
    b = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    scene.rtao_trace(dprim, dhits, npix, spp, tmax, seed, b, stream=stream)
    torch.cuda.synchronize()
    miss = dhits.cpu().numpy().view(np.uint32)[:, 1] == 0xFFFFFFFF
    return a.cpu().numpy(), b.cpu().numpy(), miss


@pytest.mark.parametrize("layout", LAYOUTS, ids=IDS)
def test_rtao_trace_equals_make_rays_then_any(gpu_ctx, layout):
    v, f = W.rtao_scene(128, 64, 150)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    # a camera that sees the scene and the sky, so some pixels miss
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    eye = c + np.array([0.2, 0.35, 1.1]) * ext
    prim = W.pinhole_rays(75, 40, eye=tuple(eye), target=tuple(c + np.array([0.0, 0.15 * ext, 0.0])),
                          fov_y_deg=60.0)
    tmax = 0.1 * W.scene_diagonal(v)
    for spp, seed in ((1, 5), (3, 77), (16, 0x2545F4914F6CDD1D)):
        say(f"rtao_trace layout {layout} spp {spp}")
        a, b, miss = two_step_and_fused(s, prim, spp, tmax, seed)
        np.testing.assert_array_equal(b, a)
        assert miss.any() and not miss.all()
        assert not b[np.repeat(miss, spp)].any()
        live = b[np.repeat(~miss, spp)]
        assert 0 < live.mean() < 1
    s.close()


def test_rtao_trace_errors(gpu_ctx):
    import torch

    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.commit(layout=H.LAYOUT_AUTO_FP16)
    d = torch.zeros(64, dtype=torch.float32, device="cuda")
    o = torch.zeros(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(H.HrtError) as e:  # 2^32 rays
        s.rtao_trace(d, d, 1 << 28, 16, 1.0, 0, o)
    assert e.value.status == 1
    s.rtao_trace(d, d, 0, 16, 1.0, 0, o)  # nothing to do
    s.add_instance(g, np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32).reshape(16))
    s.commit()
    with pytest.raises(H.HrtError) as e:  # needs one identity instance
        s.rtao_trace(d, d, 1, 1, 1.0, 0, o)
    assert e.value.status == 5
    s.close()
