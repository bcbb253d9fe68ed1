"""hrt_trace_any_noalloc (the any-hit trace with its scratch in a static device
array, no allocation) through the C-ABI: its answers equal hrt_trace_any's on
two calls in a row, the second with fewer rays, on the BVH2 kernel and on the
FP16 BVH4 kernel of the any-hit tree; an instanced scene is refused."""
import ctypes as C

import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say
from heavyrt_amd import workloads as W

pytestmark = pytest.mark.gpu


def any_rays(v, n, stream):
    lo, hi = v.min(0), v.max(0)
    r = W.uniform01(W.pcg32(n, 6, stream))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = lo + (hi - lo) * r[:, 0:3]
    d = W.unit_vectors(r[:, 3], r[:, 4])
    rays[:, 4:7] = d
    rays[:, 7] = 0.25 * float(np.max(hi - lo))
    return rays


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC],
                         ids=["bvh2", "auto_fp16", "ploc"])
def test_noalloc_any_hit_equals_trace_any(gpu_ctx, layout):
    import torch

    v, f = W.rtao_scene(256, 128, 200)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    lib = H.lib()
    stream = torch.cuda.current_stream().cuda_stream
    for call, n in enumerate((700_000, 90_001)):
        rays = any_rays(v, n, 9100 + call)
        d_rays = torch.from_numpy(rays).cuda()
        want = torch.zeros(n, dtype=torch.uint8, device="cuda")
        got = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
        H._check(lib.hrt_trace_any(s._h, C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(want.data_ptr()),
                                   C.c_void_p(stream)))
        H._check(lib.hrt_trace_any_noalloc(s._h, C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(got.data_ptr()),
                                           C.c_void_p(stream)))
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        say(f"noalloc layout {layout} call {call}: {n} rays, {int(w.sum())} occluded")
        assert 0.05 < w.mean() < 0.95
        np.testing.assert_array_equal(g, w)
    s.close()


def test_noalloc_refuses_instanced_scenes(gpu_ctx):
    import torch

    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(v, f)
    s.add_instance(g)
    s.add_instance(g, np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 3, 0, 0, 1], np.float32))
    s.commit()
    d = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    o = torch.zeros(4, dtype=torch.uint8, device="cuda")
    st = H.lib().hrt_trace_any_noalloc(s._h, C.c_void_p(d.data_ptr()), 4, C.c_void_p(o.data_ptr()), None)
    assert st == 5
    s.close()
