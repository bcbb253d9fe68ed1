"""hrt_rtao_trace: the AO pass in one call equals hrt_rtao_make_rays followed by
hrt_trace_any, byte for byte, on every layout (BVH4 layouts make the rays
inside the traversal; the others go through a ray buffer)."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say
from heavyrt_amd import workloads as W

pytestmark = pytest.mark.gpu

LAYOUTS = [H.LAYOUT_BVH2, H.LAYOUT_BVH4, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO, H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC,
           H.LAYOUT_PLOC_FP16]
IDS = ["bvh2", "bvh4", "bvh4_fp16", "auto", "auto_fp16", "ploc", "ploc_fp16"]


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


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_AUTO_FP16], ids=["bvh2", "auto_fp16"])
def test_rtao_render_matches_oracle(gpu_ctx, layout):
    """hrt_rtao_render's image equals kernel_AO2Color applied to the oracle's answers: primary hits
    of the device eye rays and any-hit answers for the AO rays hrt_rtao_make_rays makes from them."""
    import torch

    from oracle import oracle as O

    v, f = W.rtao_scene(128, 64, 150)
    width, height, spp, seed = 64, 48, 8, 99
    n = width * height
    tmax = 0.1 * W.scene_diagonal(v)
    # a camera over the heightfield toward the horizon: about half the pixels see sky
    pi, wi = W.eye_camera(width, height, eye=(1.0, 0.45, -0.3), target=(1.0, 0.25, 0.8), fov_y_deg=60.0)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    stream = torch.cuda.current_stream().cuda_stream
    rays = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    xy = torch.empty(n, dtype=torch.int32, device="cuda")
    gpu_ctx.eye_make_rays(pi, wi, width, height, rays, xy, stream=stream)
    img = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    s.rtao_render(pi, wi, width, height, spp, tmax, seed, img, stream=stream)
    img2 = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    s.rtao_render(pi, wi, width, height, spp, tmax, seed, img2, power=2.0, stream=stream)
    hits = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    s.trace_nearest(rays, n, hits, H.HIT_COMPACT, stream=stream)
    ao = torch.empty((n * spp, 8), dtype=torch.float32, device="cuda")
    s.rtao_make_rays(rays, hits, n, spp, tmax, seed, ao, stream=stream)
    torch.cuda.synchronize()
    say(f"rtao render layout {layout}")
    ob = O.OracleBVH(v, f)
    t, p, _, _ = ob.trace(rays.cpu().numpy())
    h = hits.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(h[:, 1], p)
    np.testing.assert_array_equal(h[:, 0], t.view(np.uint32))
    anyr = ao.cpu().numpy()
    anyr[:, 7] = -np.abs(anyr[:, 7])
    _, pa, _, _ = ob.trace(anyr)
    count = (pa != 0xFFFFFFFF).reshape(n, spp).sum(axis=1)
    frac = count.astype(np.float32) * (np.float32(1.0) / np.float32(spp))
    packed = xy.cpu().numpy().view(np.uint32)
    px, py = packed & 0xFFFF, packed >> 16

    def image(val):
        g = (np.clip(val, np.float32(0.0), np.float32(1.0)) * np.float32(255.0)).astype(np.uint32)
        c = np.where(p == 0xFFFFFFFF, 0, g | (g << 8) | (g << 16) | (255 << 24)).astype(np.uint32)
        out = np.zeros(n, np.uint32)
        out[py * width + px] = c
        return out

    np.testing.assert_array_equal(img.cpu().numpy().view(np.uint32), image(frac))
    assert (p == 0xFFFFFFFF).any() and (p != 0xFFFFFFFF).any() and 0 < (count > 0).mean() < 1
    # another power goes through powf: within one gray level of numpy's float32 power
    got2 = img2.cpu().numpy().view(np.uint32) & 0xFF
    want2 = image(np.power(frac, np.float32(2.0))) & 0xFF
    assert np.max(np.abs(got2.astype(np.int64) - want2.astype(np.int64))) <= 1
    s.close()


def test_rtao_render_errors(gpu_ctx):
    import torch

    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit()
    pi, wi = W.eye_camera(64, 64)
    img = torch.zeros(64 * 64, dtype=torch.int32, device="cuda")
    for w, hgt, spp in ((60, 64, 4), (64, 64, 0)):  # width not a multiple of 8, no AO rays
        with pytest.raises(H.HrtError) as e:
            s.rtao_render(pi, wi, w, hgt, spp, 1.0, 0, img)
        assert e.value.status == 1
    s.close()


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_AUTO_FP16], ids=["bvh2", "auto_fp16"])
def test_rtao_make_rays_match_the_oracle(gpu_ctx, layout):
    """hrt_rtao_make_rays against the oracle's generator on the oracle's own primary hits: origins,
    tNear and tFar bit for bit, directions within 4e-7 (sincosf is the device math library's on one
    side and the C library's on the other, both within about an ulp)."""
    import torch

    from oracle import oracle as O

    v, f = W.rtao_scene(128, 64, 150)
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    eye = c + np.array([0.2, 0.35, 1.1]) * ext
    prim = W.pinhole_rays(64, 40, eye=tuple(eye), target=tuple(c + np.array([0.0, 0.15 * ext, 0.0])),
                          fov_y_deg=60.0)
    t, pid, _, _ = O.OracleBVH(v, f).trace(prim)
    hits = np.zeros((prim.shape[0], 2), dtype=np.uint32)
    hits[:, 0] = t.view(np.uint32)
    hits[:, 1] = pid
    live = pid != 0xFFFFFFFF
    assert live.any() and not live.all()
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    stream = torch.cuda.current_stream().cuda_stream
    npix = prim.shape[0]
    for spp, seed in ((1, 3), (5, 0xDA3E39CB94B95BDB)):
        say(f"rtao_make_rays vs oracle layout {layout} spp {spp}")
        tmax = 0.1 * W.scene_diagonal(v)
        ref = O.ao_rays(v, f, prim, hits, spp, tmax, seed)
        dprim = torch.from_numpy(prim).cuda()
        dhits = torch.from_numpy(hits.view(np.int32)).cuda()
        out = torch.full((npix * spp, 8), 7.0, dtype=torch.float32, device="cuda")
        s.rtao_make_rays(dprim, dhits, npix, spp, tmax, seed, out, stream=stream)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        np.testing.assert_array_equal(got[:, 0:4].view(np.uint32), ref[:, 0:4].view(np.uint32))
        np.testing.assert_array_equal(got[:, 7].view(np.uint32), ref[:, 7].view(np.uint32))
        np.testing.assert_allclose(got[:, 4:7], ref[:, 4:7], rtol=0, atol=4e-7)
        miss = np.repeat(~live, spp)
        np.testing.assert_array_equal(got[miss], ref[miss])
    s.close()
