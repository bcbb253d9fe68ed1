"""GPU parity for batches (many scenes built and traced in one graph launch):
per scene the node pairs, the leaf permutation and the hits must be
bit-identical to a solo build + trace of the same scene, for every trace-kernel
candidate, and the solo path is itself checked against the oracle."""
import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fetch_hits(batch):
    return batch.hits()


def small_batch():
    """Mixed sizes: a 1-triangle scene, a 2-triangle one, a < 256-triangle one
    (centroid Morton rule), soups up to 30k triangles, a scene without rays."""
    scenes = []
    for s, n in enumerate([1, 2, 100, 255, 256, 3000, 30000, 12345, 700]):
        v, f = W.soup(n, stream=2000 + s)
        nr = 0 if s == 8 else 5000 + 997 * s
        rays = W.incoherent_rays(nr, stream=3000 + s) if nr else np.zeros((0, 8), np.float32)
        scenes.append((v, f, rays))
    scenes += W.batched_workload(24, tris=(1e3, 2e4), rays=(4096, 16384), stream=4000)
    return scenes


def make_batch(ctx, scenes):
    b = H.Batch(ctx, [x[0].shape[0] for x in scenes], [x[1].shape[0] for x in scenes],
                [x[2].shape[0] for x in scenes])
    b.upload(scenes)
    return b


def test_batch_build_and_hits_equal_solo(gpu_ctx):
    import torch

    scenes = small_batch()
    say(f"batch of {len(scenes)} scenes")
    b = make_batch(gpu_ctx, scenes)
    stream = torch.cuda.current_stream().cuda_stream
    b.run(stream)
    torch.cuda.synchronize()
    hits = fetch_hits(b)
    for s, (v, f, rays) in enumerate(scenes):
        solo = H.Scene(gpu_ctx)
        g = solo.add_triangles(v, f)
        solo.add_instance(g)
        solo.commit()
        bs = b.dump(s, H.DUMP_SIZES)
        ss = solo.dump(0, H.DUMP_SIZES)
        assert (bs[0], bs[1], bs[2], bs[3]) == (ss[0], ss[1], ss[2], ss[3]), s
        np.testing.assert_array_equal(b.dump(s, H.DUMP_PAIRS), solo.dump(0, H.DUMP_PAIRS), err_msg=str(s))
        np.testing.assert_array_equal(b.dump(s, H.DUMP_PERM), solo.dump(0, H.DUMP_PERM), err_msg=str(s))
        if rays.shape[0]:
            r = trace(solo, rays, "compact")
            hb = hits[b.ray_off[s]:b.ray_off[s + 1]]
            np.testing.assert_array_equal(hb[:, 0], r["t"].view(np.uint32), err_msg=str(s))
            np.testing.assert_array_equal(hb[:, 1], r["prim"], err_msg=str(s))
        solo.close()
    b.close()


def test_every_candidate_gives_the_same_hits(gpu_ctx):
    import torch

    scenes = small_batch()
    b = make_batch(gpu_ctx, scenes)
    stream = torch.cuda.current_stream().cuda_stream
    b.run(stream)
    torch.cuda.synchronize()
    ref = fetch_hits(b)
    cands = H.batch_candidates()
    say(f"{len(cands)} candidates")
    for i in range(len(cands)):
        b.set_candidate(i)
        b.trace(stream)  # plain launch
        torch.cuda.synchronize()
        np.testing.assert_array_equal(fetch_hits(b), ref, err_msg=f"candidate {cands[i]}")
        b.run(stream)  # graph (re-recorded for the new candidate)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(fetch_hits(b), ref, err_msg=f"graph, candidate {cands[i]}")
    chosen, ms = b.autotune(reps=2)
    assert 0 <= chosen < len(cands) and b.candidate == chosen
    assert np.all(ms > 0) and ms[chosen] == ms.min()
    b.run(stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(fetch_hits(b), ref)
    b.close()


def test_batch_hits_match_oracle(gpu_ctx):
    import torch

    scenes = W.batched_workload(6, tris=(1e3, 3e4), rays=(4096, 8192), stream=5000)
    b = make_batch(gpu_ctx, scenes)
    b.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    hits = fetch_hits(b)
    for s, (v, f, rays) in enumerate(scenes):
        t, p, _, _ = O.OracleBVH(v, f).trace(rays)
        hb = hits[b.ray_off[s]:b.ray_off[s + 1]]
        np.testing.assert_array_equal(hb[:, 0], t.view(np.uint32))
        np.testing.assert_array_equal(hb[:, 1], p)
    b.close()


def test_batch_rerun_with_new_inputs(gpu_ctx):
    """Refilling the input buffers and replaying the same graph rebuilds and
    retraces from the new data."""
    import torch

    a = W.batched_workload(5, tris=(500, 5000), rays=(1000, 3000), stream=6000)
    sizes = [(x[0].shape[0], x[1].shape[0], x[2].shape[0]) for x in a]
    b = make_batch(gpu_ctx, a)
    stream = torch.cuda.current_stream().cuda_stream
    b.run(stream)
    torch.cuda.synchronize()
    first = fetch_hits(b).copy()
    # same sizes, different triangles: shift every soup
    moved = [(v + np.float32(0.25), f, r) for v, f, r in a]
    b.upload(moved)
    b.run(stream)
    torch.cuda.synchronize()
    second = fetch_hits(b)
    assert not np.array_equal(first, second)
    for s, (v, f, rays) in enumerate(moved):
        t, p, _, _ = O.OracleBVH(v, f).trace(rays)
        hb = second[b.ray_off[s]:b.ray_off[s + 1]]
        np.testing.assert_array_equal(hb[:, 0], t.view(np.uint32))
        np.testing.assert_array_equal(hb[:, 1], p)
    assert sizes == [(x[0].shape[0], x[1].shape[0], x[2].shape[0]) for x in moved]
    b.close()


def test_batch_reports_a_leaf_overflow(gpu_ctx):
    """A scene with 130 triangles on one Morton code cannot be stored in 7-bit leaf counts: the batch's
    status names it (hrt_add_triangles refuses the same scene); the other scenes' hits stay right."""
    import torch

    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    bad = (tri, np.array([[0, 1, 2]] * 130, np.uint32), W.incoherent_rays(500, stream=6100))
    good = W.batched_workload(2, tris=(500, 2000), rays=(1000, 2000), stream=6200)
    scenes = [good[0], bad, good[1]]
    b = make_batch(gpu_ctx, scenes)
    b.run(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(H.HrtError) as e:
        b.status()
    assert e.value.status == 6 and e.value.scene == 1
    hits = b.hits()
    for s in (0, 2):
        v, f, rays = scenes[s]
        t, p, _, _ = O.OracleBVH(v, f).trace(rays)
        hb = hits[b.ray_off[s]:b.ray_off[s + 1]]
        np.testing.assert_array_equal(hb[:, 0], t.view(np.uint32))
        np.testing.assert_array_equal(hb[:, 1], p)
    b.close()
    ok = make_batch(gpu_ctx, good)
    ok.run(torch.cuda.current_stream().cuda_stream)
    ok.status()
    ok.close()
