"""Hydra XML scene files (hydra_xml::HydraScene, examples/loader/hydraxml.cpp:90-198;
RTAO::LoadSceneHydra, examples/02_rtao/rtao_host.cpp:66-171): the mesh list, the
instance list with row-major matrices, the first camera, the AO box size, and the
refusals of LoadState -- checked against closed forms, and the scene the file
describes traced through the oracle against a float64 brute force over the
instances' transformed triangles."""
import os

import numpy as np
import pytest

from heavyrt_amd import hydra as Y
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SCENE = os.path.join(HERE, "golden", "hydra_scene.xml")


def test_mesh_and_instance_lists():
    hs = Y.load_hydra(SCENE)
    names = [os.path.basename(p) for p in hs.mesh_files]
    assert names == ["cube.vsgf", "helix_mid.vsgf", "cone_low.vsgf", "triangle.vsgf"]
    assert all(os.path.isfile(p) for p in hs.mesh_files)
    # instance_light and remap_lists are not geometry instances
    assert [g for g, _ in hs.instances] == [3, 0, 1, 2, 0, 2]
    assert len(hs.cameras) == 2
    c = hs.cameras[0]
    assert (c.fov, c.near, c.far) == (45.0, np.float32(0.01), 100.0)
    np.testing.assert_allclose(c.pos, (0.3, 2.2, 5.5), rtol=1e-7)
    np.testing.assert_allclose(c.look_at, (0.0, 0.4, 0.0), rtol=1e-7)


def test_matrices_are_rows():
    """The 4th number of each row is the translation: a unit cube box [-0.5,0.5]x[0,1]x[-0.5,0.5]
    scaled by 0.5 and moved by (1.6, 0, 0.2)."""
    hs = Y.load_hydra(SCENE)
    meshes = Y.load_meshes(hs)
    b = Y.instance_boxes(hs, meshes)[1]
    np.testing.assert_allclose(b, [[1.35, 0.0, -0.05], [1.85, 0.5, 0.45]], atol=1e-6)
    _, cm = Y.instance_list(hs)[1]
    np.testing.assert_array_equal(cm[12:15], np.float32([1.6, 0.0, 0.2]))
    # the floor triangle spans [-0.5, 0.5] in x and z, scaled by 12
    assert Y.ao_box_size(hs, meshes) == pytest.approx(12.0)


def test_camera_matrices():
    cam = Y.load_hydra(SCENE).cameras[0]
    pi, wi = cam.matrices(160, 120)
    P = pi.reshape(4, 4, order="F").astype(np.float64)
    V = wi.reshape(4, 4, order="F").astype(np.float64)
    np.testing.assert_allclose(V @ [0, 0, 0, 1], [*cam.pos, 1], atol=1e-6)
    fwd = (V @ [0, 0, -1, 0])[:3]
    want = np.subtract(cam.look_at, cam.pos)
    np.testing.assert_allclose(fwd, want / np.linalg.norm(want), atol=1e-6)
    # the top edge of the image (ndc y = 1) is fov/2 above the axis; the aspect ratio widens x
    top = P @ [0, 1, 1, 1]
    right = P @ [1, 0, 1, 1]
    assert np.degrees(np.arctan2(top[1], -top[2])) == pytest.approx(cam.fov / 2, abs=1e-4)
    assert right[0] / -right[2] == pytest.approx(np.tan(np.radians(cam.fov / 2)) * 160 / 120, rel=1e-5)


def test_load_refusals(tmp_path):
    with open(SCENE) as fh:
        good = fh.read()
    p = tmp_path / "no_cams.xml"
    p.write_text(good.replace("<cam_lib>", "<cam_lib_x>").replace("</cam_lib>", "</cam_lib_x>"))
    with pytest.raises(ValueError, match="cam_lib"):
        Y.load_hydra(str(p))
    with pytest.raises(ValueError, match="Error loading"):
        Y.load_hydra(str(tmp_path / "absent.xml"))
    p = tmp_path / "broken.xml"
    p.write_text(good.replace("</scenes>", ""))
    with pytest.raises(ValueError, match="Error loading"):
        Y.load_hydra(str(p))
    p = tmp_path / "short_matrix.xml"
    p.write_text(good.replace('matrix="12 0 0 0 0 1 0 0 0 0 12 0 0 0 0 1 "', 'matrix="12 0 0"'))
    with pytest.raises(ValueError, match="matrix"):
        Y.load_hydra(str(p))
    hs = Y.load_hydra(SCENE)
    hs.instances.append((7, np.eye(4, dtype=np.float32)))
    with pytest.raises(ValueError, match="mesh_id 7"):
        Y.build_scene(None, hs)


def camera_rays(hs, w=96, h=72):
    pi, wi = hs.cameras[0].matrices(w, h)
    rays, _ = O.eye_rays(pi, wi, w, h)
    return rays


def test_oracle_scene_from_file_equals_brute_force():
    hs = Y.load_hydra(SCENE)
    meshes = Y.load_meshes(hs)
    o = O.OracleScene()
    for v, f in meshes:
        o.add_geom(v, f)
    for g, m in Y.instance_list(hs):
        o.add_instance(g, m)
    o.commit()
    rays = camera_rays(hs)
    t, prim, inst = o.trace(rays, O.SCENE_TWO_LEVEL)[:3]
    assert 0.3 < (prim != 0xFFFFFFFF).mean() < 1.0
    # float64 brute force over every instance's transformed triangles
    org = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    best = np.full(rays.shape[0], np.inf)
    best_inst = np.full(rays.shape[0], -1)
    for i, (g, m) in enumerate(hs.instances):
        v, f = meshes[g]
        w = (np.c_[v.astype(np.float64), np.ones(len(v))] @ m.astype(np.float64).T)[:, :3]
        A, B, C = w[f[:, 0]], w[f[:, 1]], w[f[:, 2]]
        e1, e2 = B - A, C - A
        for r in range(0, rays.shape[0], 512):
            o_, d_ = org[r:r + 512, None, :], d[r:r + 512, None, :]
            p = np.cross(d_, e2[None])
            det = np.sum(e1[None] * p, axis=2)
            ok = np.abs(det) > 1e-12
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            s = o_ - A[None]
            u = np.sum(s * p, axis=2) * inv
            q = np.cross(s, e1[None])
            vv = np.sum(d_ * q, axis=2) * inv
            tt = np.sum(e2[None] * q, axis=2) * inv
            hit = ok & (u >= 0) & (vv >= 0) & (u + vv <= 1) & (tt > 0)
            tt = np.where(hit, tt, np.inf)
            j = tt.min(axis=1)
            better = j < best[r:r + 512]
            best[r:r + 512][better] = j[better]
            best_inst[r:r + 512][better] = i
    live = np.isfinite(best)
    got = prim != 0xFFFFFFFF
    # edge-grazing rays may flip between float32 and float64; everything else agrees
    assert (live != got).mean() < 0.003
    both = live & got
    np.testing.assert_allclose(t[both], best[both], rtol=1e-4)
    assert (inst[both] == best_inst[both]).mean() > 0.995
