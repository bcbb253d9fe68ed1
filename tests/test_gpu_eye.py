"""GPU parity for the 01_eye_rays path on the device: hrt_eye_make_rays gives the oracle's rays and pixel
order bit for bit (kernel_PackXY / kernel_InitEyeRay, examples/01_eye_rays/eye_ray.cpp:26-47, 114-138), and
hrt_eye_render gives the reference's colours (eye_ray.cpp:50-75) for the oracle's hits."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def device_rays(ctx, cam, w, h):
    import torch

    rays = torch.zeros((w * h, 8), dtype=torch.float32, device="cuda")
    xy = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    ctx.eye_make_rays(cam[0], cam[1], w, h, rays, xy, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rays.cpu().numpy(), xy.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("w,h,eye,target", [(512, 512, W.EYE_POS, W.EYE_TARGET), (64, 40, (0.3, 2.0, 3.0), (0.0, 0.8, 0.0)),
                                            (8, 8, (0.0, 0.0, 5.0), (0.0, 0.0, 0.0))])
def test_eye_rays_bit_exact(gpu_ctx, w, h, eye, target):
    cam = W.eye_camera(w, h, eye=eye, target=target)
    say(f"eye rays {w}x{h}")
    rays, xy = device_rays(gpu_ctx, cam, w, h)
    ref, ref_xy = O.eye_rays(cam[0], cam[1], w, h)
    np.testing.assert_array_equal(xy, ref_xy)
    np.testing.assert_array_equal(rays.view(np.uint32), ref.view(np.uint32))


def test_eye_render_full_config(gpu_ctx):
    import torch

    v, f = W.eye_scene()
    w = h = 512
    cam = W.eye_camera(w, h)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit()
    say("eye render 512x512")
    out = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    s.eye_render(cam[0], cam[1], w, h, W.EYE_PALETTE, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    img = out.cpu().numpy().view(np.uint32)
    rays, xy = O.eye_rays(cam[0], cam[1], w, h)
    _, prim, _, _ = O.OracleBVH(v, f).trace(rays)
    want = np.zeros(w * h, dtype=np.uint32)
    pix = (xy >> 16).astype(np.int64) * w + (xy & 0xFFFF)
    hit = prim != 0xFFFFFFFF
    want[pix[hit]] = W.EYE_PALETTE[prim[hit] % W.EYE_PALETTE.size]
    np.testing.assert_array_equal(img, want)
    assert 0.3 < hit.mean() < 1.0  # sphere and ground fill part of the frame
    # the same rays through the plain trace give the same hits
    r = trace(s, rays, "compact")
    np.testing.assert_array_equal(r["prim"], prim)
    s.close()


@pytest.mark.parametrize("w,h", [(640, 480), (72, 40), (192, 64)])
def test_eye_render_any_frame_size(gpu_ctx, w, h):
    """hrt_eye_render traces in XCD stripes where the frame splits into whole workgroups per XCD
    (640x480, 192x64) and in row order otherwise (72x40); every pixel gets its own ray's colour."""
    import torch

    v, f = W.eye_scene()
    cam = W.eye_camera(w, h)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit()
    say(f"eye render {w}x{h}")
    out = torch.full((w * h,), 7, dtype=torch.int32, device="cuda")
    s.eye_render(cam[0], cam[1], w, h, W.EYE_PALETTE, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    img = out.cpu().numpy().view(np.uint32)
    rays, xy = O.eye_rays(cam[0], cam[1], w, h)
    _, prim, _, _ = O.OracleBVH(v, f).trace(rays)
    want = np.zeros(w * h, dtype=np.uint32)
    pix = (xy >> 16).astype(np.int64) * w + (xy & 0xFFFF)
    hit = prim != 0xFFFFFFFF
    want[pix[hit]] = W.EYE_PALETTE[prim[hit] % W.EYE_PALETTE.size]
    np.testing.assert_array_equal(img, want)
    s.close()


def test_eye_render_instanced(gpu_ctx):
    """A two-level scene (two spheres as instances of one geometry) renders like its merged mesh."""
    import torch

    iv, if_ = W.icosphere(3, 0.5, (0.0, 0.0, 0.0))
    s = H.Scene(gpu_ctx)
    g = s.add_triangles(iv, if_)
    mats = []
    for dx in (-0.6, 0.6):
        M = np.eye(4, dtype=np.float32)
        M[:3, 3] = [dx, 1.0, 0.0]
        mats.append(M.flatten(order="F"))
        s.add_instance(g, mats[-1])
    s.commit()
    w, h = 128, 96
    cam = W.eye_camera(w, h)
    out = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    s.eye_render(cam[0], cam[1], w, h, W.EYE_PALETTE, out, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    img = out.cpu().numpy().view(np.uint32)
    o = O.OracleScene()
    o.add_geom(iv, if_)
    for m in mats:
        o.add_instance(0, m)
    o.commit()
    rays, xy = O.eye_rays(cam[0], cam[1], w, h)
    _, prim, _ = o.trace(rays)[:3]
    want = np.zeros(w * h, dtype=np.uint32)
    pix = (xy >> 16).astype(np.int64) * w + (xy & 0xFFFF)
    hit = prim != 0xFFFFFFFF
    want[pix[hit]] = W.EYE_PALETTE[prim[hit] % W.EYE_PALETTE.size]
    np.testing.assert_array_equal(img, want)
    assert hit.mean() > 0.1
    s.close()


def test_eye_errors(gpu_ctx):
    import torch

    cam = W.eye_camera(64, 64)
    rays = torch.zeros((64 * 64, 8), dtype=torch.float32, device="cuda")
    with pytest.raises(H.HrtError) as e:
        gpu_ctx.eye_make_rays(cam[0], cam[1], 60, 64, rays)
    assert e.value.status == 1
    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    out = torch.zeros(64 * 64, dtype=torch.int32, device="cuda")
    with pytest.raises(H.HrtError) as e:  # not committed
        s.eye_render(cam[0], cam[1], 64, 64, W.EYE_PALETTE, out)
    assert e.value.status == 4
    s.commit()
    with pytest.raises(H.HrtError) as e:
        s.eye_render(cam[0], cam[1], 64, 64, np.zeros(65, dtype=np.uint32), out)
    assert e.value.status == 1
    s.close()
