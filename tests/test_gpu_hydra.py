"""A Hydra XML scene file on the device: the scene built from the file (geometries in
geometry_lib order, instances with the file's row-major matrices) traced with the file's
first camera equals the oracle's two-level traversal of the same file, bit for bit
(RTAO::LoadSceneHydra, examples/02_rtao/rtao_host.cpp:66-171)."""
import os

import numpy as np
import pytest

import heavyrt_amd as H
from gpu_util import say, trace
from heavyrt_amd import hydra as Y
from oracle import oracle as O
from test_gpu_instances import check_hits

pytestmark = pytest.mark.gpu

SCENE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hydra_scene.xml")


@pytest.mark.parametrize("layout", [H.LAYOUT_BVH2, H.LAYOUT_AUTO_FP16], ids=["bvh2", "auto_fp16"])
def test_hydra_file_matches_the_oracle(gpu_ctx, layout):
    hs = Y.load_hydra(SCENE)
    meshes = Y.load_meshes(hs)
    s = Y.build_scene(gpu_ctx, hs, layout=layout, meshes=meshes)
    o = O.OracleScene()
    for v, f in meshes:
        o.add_geom(v, f)
    for g, m in Y.instance_list(hs):
        o.add_instance(g, m)
    o.commit()
    assert s.instance_count == len(hs.instances) and s.geom_count == len(meshes)
    pi, wi = hs.cameras[0].matrices(160, 120)
    rays, _ = O.eye_rays(pi, wi, 160, 120)
    say(f"hydra scene, layout {layout}: nearest")
    ref = o.trace(rays, O.SCENE_TWO_LEVEL)
    assert 0.3 < (ref[1] != 0xFFFFFFFF).mean() < 1.0
    check_hits(trace(s, rays, "full"), ref)
    say(f"hydra scene, layout {layout}: any-hit")
    check_hits(trace(s, rays, "any"), ref, any_hit=True)
    s.close()
