"""Oracle eye rays (EyeRayCaster::kernel_PackXY / kernel_InitEyeRay, examples/01_eye_rays/eye_ray.cpp:26-47,
114-138): the superblock order is a bijection onto 8x8 tiles, and the rays equal a float64 pinhole camera."""
import numpy as np
import pytest

from heavyrt_amd import workloads as W
from oracle import oracle as O


def test_superblock_offset_known_answers():
    # 4x4 blocks, 2x2 blocks per superblock, superblocks row by row
    assert O.superblock_offset(0, 0, 16) == 0
    assert O.superblock_offset(3, 0, 16) == 3
    assert O.superblock_offset(0, 1, 16) == 4
    assert O.superblock_offset(4, 0, 16) == 16
    assert O.superblock_offset(0, 4, 16) == 32
    assert O.superblock_offset(4, 4, 16) == 48
    assert O.superblock_offset(8, 0, 16) == 64
    assert O.superblock_offset(0, 8, 16) == 128
    assert O.superblock_offset(15, 15, 16) == 255


@pytest.mark.parametrize("w,h", [(16, 8), (64, 40), (512, 512)])
def test_superblock_order_is_tiles_of_64(w, h):
    _, xy = O.eye_rays(*W.eye_camera(w, h), w, h)
    x, y = xy & 0xFFFF, xy >> 16
    assert np.unique(xy).size == w * h
    assert x.max() == w - 1 and y.max() == h - 1
    # each run of 64 thread indices is one 8x8 tile
    tx, ty = (x // 8).reshape(-1, 64), (y // 8).reshape(-1, 64)
    assert np.all(tx == tx[:, :1]) and np.all(ty == ty[:, :1])
    for t in range(0, w * h, 997):
        assert O.superblock_offset(int(x[t]), int(y[t]), w) == t


def test_eye_rays_equal_float64_pinhole():
    w = h = 64
    rays, xy = O.eye_rays(*W.eye_camera(w, h), w, h)
    ref = W.pinhole_rays(w, h, eye=W.EYE_POS, target=W.EYE_TARGET, fov_y_deg=45.0)
    pix = (xy >> 16).astype(np.int64) * w + (xy & 0xFFFF)
    ref = ref[pix]
    np.testing.assert_allclose(rays[:, 0:3], ref[:, 0:3], atol=2e-6)
    np.testing.assert_allclose(rays[:, 4:7], ref[:, 4:7], atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(rays[:, 4:7].astype(np.float64), axis=1), 1.0, atol=1e-6)
    assert np.all(rays[:, 3] == 0.0) and np.all(rays[:, 7] == np.float32(3.402823466e38))


def test_eye_rays_mirror_about_the_centre():
    # y is flipped (row 0 at the top): the top row looks up, the left column looks left
    w = h = 32
    rays, xy = O.eye_rays(*W.eye_camera(w, h, eye=(0, 0, 5), target=(0, 0, 0)), w, h)
    x, y = xy & 0xFFFF, xy >> 16
    d = rays[:, 4:7]
    assert np.all(d[y == 0, 1] > 0) and np.all(d[y == h - 1, 1] < 0)
    assert np.all(d[x == 0, 0] < 0) and np.all(d[x == w - 1, 0] > 0)
    # pixels mirrored about the centre get mirrored directions
    k = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(x, y))}
    for (a, b), i in list(k.items())[::37]:
        j = k[(w - 1 - a, h - 1 - b)]
        np.testing.assert_allclose(d[i, 0:2], -d[j, 0:2], atol=1e-6)
        np.testing.assert_allclose(d[i, 2], d[j, 2], atol=1e-6)


def test_superblock_order_inverts_the_oracle_packing():
    """workloads.superblock_order (the host-side ray layout of the RTAO bench) is the inverse of the
    oracle's SuperBlockIndex2DOpt restatement on every pixel."""
    for w, h in ((64, 48), (128, 128)):
        order = W.superblock_order(w, h)
        assert sorted(order.tolist()) == list(range(w * h))
        tid = np.array([O.superblock_offset(int(i % w), int(i // w), w) for i in order])
        np.testing.assert_array_equal(tid, np.arange(w * h))


@pytest.mark.parametrize("w,h", [(512, 512), (640, 480), (64, 32), (100, 96), (512, 40)])
def test_xcd_stripe_order_gives_each_xcd_its_stripes(w, h):
    """workloads.xcd_stripe_order (the order hrt_eye_render traces its rays in): a permutation of
    the pixels; where the frame splits into whole workgroups per XCD, workgroup b (256 rays, run
    on XCD b % 8) holds only rows r with (r // 4) % 8 == b % 8, and each XCD takes its rows in
    order; otherwise plain row order."""
    order = W.xcd_stripe_order(w, h)
    assert np.array_equal(np.sort(order), np.arange(w * h))
    if h % 32 or w % 64:
        assert np.array_equal(order, np.arange(w * h))
        return
    rows = (order // w).reshape(-1, 256)
    xcd = np.arange(rows.shape[0]) % 8
    assert np.all((rows // 4) % 8 == xcd[:, None])
    for x in range(8):
        share = order.reshape(-1, 256)[xcd == x].reshape(-1)
        assert np.all(np.diff(share) > 0)  # row order within the XCD's share
