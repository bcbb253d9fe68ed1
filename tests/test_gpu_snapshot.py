"""Scene snapshots (hrt_scene_serialize / hrt_scene_deserialize): a restored scene holds the saved trees word
for word and traces the saved answers bit for bit, without rebuilding them."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from test_gpu_parity import rays_for
from test_oracle_refit import moved
from test_oracle_scene import affine, col_major, scene_rays

pytestmark = pytest.mark.gpu


def same_traces(a, b, rays):
    ra, rb = trace(a, rays, "full"), trace(b, rays, "full")
    for k in ("t", "prim", "inst", "geom", "u", "v"):
        np.testing.assert_array_equal(ra[k].view(np.uint32), rb[k].view(np.uint32))
    anyr = rays.copy()
    anyr[:, 7] = 0.3 * float(np.max(np.abs(rays[:, 0:3])) + 1.0)
    np.testing.assert_array_equal(trace(a, anyr, "any")["hit"], trace(b, anyr, "any")["hit"])


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_BVH4_FP16, H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC],
                         ids=["bvh2", "bvh4_fp16", "auto_fp16", "ploc"])
def test_snapshot_single_geometry(gpu_ctx, layout):
    v, f = W.rtao_scene(256, 128, 300)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit(layout=layout)
    data = s.serialize()
    say(f"snapshot layout {layout}: {len(data)} bytes")
    r = H.Scene.deserialize(gpu_ctx, data)
    np.testing.assert_array_equal(r.dump(0, H.DUMP_SIZES), s.dump(0, H.DUMP_SIZES))
    for what in (H.DUMP_PAIRS, H.DUMP_PERM, H.DUMP_SCENE_BOX):
        np.testing.assert_array_equal(r.dump(0, what).view(np.uint32), s.dump(0, what).view(np.uint32))
    if layout == H.LAYOUT_BVH4_FP16:
        np.testing.assert_array_equal(r.dump(0, H.DUMP_QUADS16), s.dump(0, H.DUMP_QUADS16))
    if layout in (H.LAYOUT_AUTO_FP16, H.LAYOUT_PLOC):
        for what in (H.DUMP_SIZES, H.DUMP_PAIRS, H.DUMP_QUADS16, H.DUMP_QUADS32):
            np.testing.assert_array_equal(r.dump(0, what | H.DUMP_ANY_TREE), s.dump(0, what | H.DUMP_ANY_TREE))
    rays = rays_for(v, 20_000, 5300)
    same_traces(s, r, rays)
    # a second snapshot of the restored scene is the same bytes
    assert r.serialize() == data
    # the kept refit arrays came along: a refit of both gives the same trees
    v2 = moved(v, 93)
    s.update_triangles(0, v2, quality=3)
    r.update_triangles(0, v2, quality=3)
    s.commit(layout=layout)
    r.commit(layout=layout)
    np.testing.assert_array_equal(r.dump(0, H.DUMP_PAIRS), s.dump(0, H.DUMP_PAIRS))
    same_traces(s, r, rays_for(v2, 10_000, 5400))
    s.close()
    r.close()


def test_snapshot_instanced(gpu_ctx):
    geoms = [W.icosphere(2, 0.3, (0.0, 0.0, 0.0)), W.unit_cube(), W.soup(2000, stream=W.STREAM_TEST)]
    s = H.Scene(gpu_ctx)
    ids = [s.add_triangles(v, f) for v, f in geoms]
    for k in range(12):
        M = affine(70 + k)
        M[:3, 3] = [0.2 * k, 0.1 * (k % 3), 0.15 * (k % 4)]
        s.add_instance(ids[k % 3], col_major(M))
    s.commit()
    r = H.Scene.deserialize(gpu_ctx, s.serialize())
    assert H.lib().hrt_instance_count(r._h) == 12 and H.lib().hrt_geom_count(r._h) == 3
    # the top level is rebuilt from the saved instances: the same tree
    np.testing.assert_array_equal(r.dump(H.TLAS_ID, H.DUMP_PAIRS), s.dump(H.TLAS_ID, H.DUMP_PAIRS))
    for g in range(3):
        np.testing.assert_array_equal(r.dump(g, H.DUMP_PAIRS), s.dump(g, H.DUMP_PAIRS))
    rays = scene_rays(20_000, 5500, lo=(-0.5, -0.5, -0.5), hi=(2.8, 0.8, 1.0))
    same_traces(s, r, rays)
    s.close()
    r.close()


def snap_hash(body: bytes) -> int:
    """FNV-1a over little-endian 64-bit words, then the tail bytes."""
    h, m = 0xCBF29CE484222325, 0xFFFFFFFFFFFFFFFF
    nw = len(body) // 8
    for w in np.frombuffer(body[:8 * nw], dtype="<u8").tolist():
        h = ((h ^ w) * 0x100000001B3) & m
    for b in body[8 * nw:]:
        h = ((h ^ b) * 0x100000001B3) & m
    return h


def rehashed(body: bytes) -> bytes:
    """A snapshot body with a matching hash trailer, so the record checks behind the hash are reached."""
    return bytes(body) + snap_hash(body).to_bytes(8, "little")


def status_of(ctx, data: bytes) -> int:
    with pytest.raises(H.HrtError) as e:
        H.Scene.deserialize(ctx, data)
    return e.value.status


def test_snapshot_errors(gpu_ctx):
    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    with pytest.raises(H.HrtError) as e:
        s.serialize()
    assert e.value.status == 4  # not committed
    s.commit()
    data = bytearray(s.serialize())
    body = bytes(data[:-8])
    assert snap_hash(body).to_bytes(8, "little") == bytes(data[-8:])
    # a bad hash, a bad magic, a flipped bit
    assert status_of(gpu_ctx, bytes(data[:-9]) + bytes(data[-8:])) == 1
    assert status_of(gpu_ctx, b"XXXX" + bytes(data[4:])) == 1
    flipped = bytearray(data)
    flipped[len(flipped) // 2] ^= 0x40
    assert status_of(gpu_ctx, bytes(flipped)) == 1
    # behind a matching hash: truncated records, trailing bytes, an unknown version
    for cut in (12, 30, 60, len(body) // 3, len(body) - 70, len(body) - 1):
        assert status_of(gpu_ctx, rehashed(body[:cut])) == 1, cut
    assert status_of(gpu_ctx, rehashed(body + b"\0\0\0\0")) == 1
    assert status_of(gpu_ctx, rehashed(body[:8] + (2).to_bytes(4, "little") + body[12:])) == 5
    # the untouched bytes still load
    H.Scene.deserialize(gpu_ctx, bytes(data)).close()
    s.close()


def test_snapshot_crafted_records_are_refused(gpu_ctx):
    """Records behind a matching hash whose references would lead a kernel out of its arrays (a child
    past the node pairs, a leaf past the triangles, the root as a child, a duplicated leaf-order entry,
    an index past the vertices, a triangle id past the triangles, a leaf of 200 triangles) are refused
    on the host before anything is uploaded."""
    v, f = W.icosphere(2)
    s = H.Scene(gpu_ctx)
    s.add_instance(s.add_triangles(v, f))
    s.commit()
    body = bytes(s.serialize()[:-8])
    s.close()
    head = np.frombuffer(body[24:56], dtype="<u4")
    nverts, ntris, leaves = int(head[0]), int(head[1]), int(head[2])
    np_alloc = max(ntris - 1, 1)
    at = 24 + 80
    off = {}
    for name, size in (("verts", 16 * nverts), ("idx", 12 * ntris), ("pairs", 64 * np_alloc),
                       ("tris", 48 * ntris), ("perm", 4 * ntris)):
        off[name] = at
        at += size

    def with_word(byte_off: int, value: int) -> bytes:
        b = bytearray(body)
        b[byte_off:byte_off + 4] = int(value).to_bytes(4, "little")
        return bytes(b)

    pair = lambda k, side: off["pairs"] + 64 * k + 32 * side + 12  # child reference word
    cases = {
        "child past the pairs": with_word(pair(0, 0), leaves + 3),
        "root as a child": with_word(pair(0, 1), 0),
        "leaf past the triangles": with_word(pair(0, 0), 0x80000000 | (5 << 24) | (ntris - 2)),
        "longest leaf of 200": with_word(24 + 12, 200),
        "duplicated leaf-order entry": with_word(off["perm"] + 4, int(np.frombuffer(body[off["perm"]:off["perm"] + 4],
                                                                                    dtype="<u4")[0])),
        "index past the vertices": with_word(off["idx"], nverts),
        "triangle id past the triangles": with_word(off["tris"] + 12, ntris),
    }
    for name, bad in cases.items():
        assert bad != body, name
        assert status_of(gpu_ctx, rehashed(bad)) == 1, name
    H.Scene.deserialize(gpu_ctx, rehashed(body)).close()
