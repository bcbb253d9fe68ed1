"""Hydra XML scene files on the engine: the mesh list, the instance list and the
first camera of a scene file, and a committed Scene built from them the way the
renderers load one (RTAO::LoadSceneHydra, examples/02_rtao/rtao_host.cpp:66-171;
EyeRayCaster::LoadSceneHydra, examples/01_eye_rays/eye_ray_host.cpp:66-130).

File format, as hydra_xml::HydraScene reads it (examples/loader/hydraxml.cpp:90-198,
hydraxml.h:72-185, 226-294):
  - the root holds textures_lib, materials_lib, lights_lib, cam_lib,
    geometry_lib, render_lib and scenes; a file without one of them is refused;
  - geometry_lib's children are the meshes, in geometry-id order, each with a
    ``loc`` attribute: a VSGF file relative to the scene file's directory;
  - the first ``scene`` under ``scenes`` lists ``instance`` nodes with
    ``mesh_id`` (the geometry id) and ``matrix``: 16 numbers, row by row;
  - cam_lib's first camera gives ``fov`` (degrees, vertical),
    ``nearClipPlane``, ``farClipPlane``, ``position``, ``look_at`` and ``up``.

"""
from __future__ import annotations

import os
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import workloads as W

REQUIRED_LIBS = ("textures_lib", "materials_lib", "lights_lib", "cam_lib", "geometry_lib", "render_lib", "scenes")


@dataclass
class Camera:
    fov: float
    near: float
    far: float
    pos: tuple
    look_at: tuple
    up: tuple

    def matrices(self, width: int, height: int):
        """(inverse projection, inverse world-view), 16 column-major float32 each, from
        perspectiveMatrix(fov, width/height, near, far) and lookAt(pos, look_at, up)
        (rtao_host.cpp:83-91)."""
        proj = W.perspective(self.fov, width / height, self.near, self.far)
        view = W.look_at(self.pos, self.look_at, self.up)
        pi = np.linalg.inv(proj).astype(np.float32)
        wi = np.linalg.inv(view).astype(np.float32)
        return pi.flatten(order="F"), wi.flatten(order="F")


@dataclass
class HydraScene:
    path: str
    mesh_files: list = field(default_factory=list)   # absolute paths, geometry-id order
    instances: list = field(default_factory=list)    # (geometry id, 4x4 float32 row-major matrix)
    cameras: list = field(default_factory=list)


def _floats(text: str | None, n: int, what: str) -> np.ndarray:
    vals = (text or "").split()
    if len(vals) < n:
        raise ValueError(f"{what}: expected {n} numbers, found {len(vals)}")
    return np.array([float(x) for x in vals[:n]], dtype=np.float32)


def _child_float(node, tag: str) -> float:
    c = node.find(tag)
    if c is None or c.text is None:
        raise ValueError(f"camera: missing <{tag}>")
    return float(c.text.split()[0])


def _child3(node, tag: str) -> tuple:
    c = node.find(tag)
    v = _floats(None if c is None else c.text, 3, f"camera <{tag}>")
    return tuple(float(x) for x in v)


def load_hydra(path: str) -> HydraScene:
    """Parse a Hydra XML scene file (HydraScene::LoadState, hydraxml.cpp:90-126). Raises
    ValueError where LoadState returns -1: an unreadable file or a missing library node."""
    # The libraries are top-level elements side by side, which pugixml accepts as one document
    # and a strict XML parser does not: the body is parsed inside a wrapping element.
    try:
        with open(path, "rb") as fh:
            raw = fh.read()
        text = raw.decode("utf-16") if raw[:2] in (b"\xff\xfe", b"\xfe\xff") else raw.decode("utf-8-sig")
        text = re.sub(r"^\s*<\?xml[^>]*\?>", "", text)
        root = ET.fromstring("<hydra_file>" + text + "</hydra_file>")
    except (OSError, UnicodeDecodeError, ET.ParseError) as e:
        raise ValueError(f"Error loading scene from: {path}: {e}") from e
    libs = {}
    for c in root:
        libs.setdefault(c.tag, c)  # pugixml's child(name) is the first of that name
    missing = [t for t in REQUIRED_LIBS if t not in libs]
    if missing:
        raise ValueError(f"Loaded state ({path}) doesn't have one of {', '.join(missing)}")
    base = os.path.dirname(os.path.abspath(path))
    hs = HydraScene(path=path)
    for mesh in libs["geometry_lib"]:
        hs.mesh_files.append(os.path.join(base, mesh.get("loc", "")))
    scene = libs["scenes"].find("scene")
    if scene is not None:
        for inst in scene.findall("instance"):
            m = _floats(inst.get("matrix"), 16, "instance matrix").reshape(4, 4)
            hs.instances.append((int(inst.get("mesh_id", "0")), m))
    for cam in libs["cam_lib"]:
        hs.cameras.append(Camera(fov=_child_float(cam, "fov"), near=_child_float(cam, "nearClipPlane"),
                                 far=_child_float(cam, "farClipPlane"), pos=_child3(cam, "position"),
                                 look_at=_child3(cam, "look_at"), up=_child3(cam, "up")))
    return hs


def instance_boxes(hs: HydraScene, meshes) -> np.ndarray:
    """World box of every instance: the 8 corners of its mesh's box through the instance matrix
    (rtao_host.cpp:148-164). meshes: [(verts, faces)] in geometry-id order. Returns (n, 2, 3)."""
    out = np.zeros((len(hs.instances), 2, 3), dtype=np.float64)
    for i, (gid, m) in enumerate(hs.instances):
        v = np.asarray(meshes[gid][0], dtype=np.float64)
        lo, hi = v.min(axis=0), v.max(axis=0)
        corners = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        w = corners @ m.astype(np.float64).T
        out[i, 0], out[i, 1] = w[:, :3].min(axis=0), w[:, :3].max(axis=0)
    return out


def ao_box_size(hs: HydraScene, meshes) -> float:
    """The largest side of the scene box over all instances: the renderer's m_aoBoxSize, which
    scales the AO radius (rtao_host.cpp:30, 168-170)."""
    b = instance_boxes(hs, meshes)
    if b.shape[0] == 0:
        return 0.0
    size = b[:, 1].max(axis=0) - b[:, 0].min(axis=0)
    return float(size.max())


def load_meshes(hs: HydraScene):
    """Every mesh of the file as (verts (n,3) float32, faces (m,3) uint32), geometry-id order."""
    return [W.load_vsgf(p) for p in hs.mesh_files]


def instance_list(hs: HydraScene):
    """[(geometry id, 16 column-major float32)]: the file's instances in the form AddInstance takes."""
    return [(gid, np.ascontiguousarray(m.T).reshape(16)) for gid, m in hs.instances]


def build_scene(ctx, hs: HydraScene, layout: int | None = None, meshes=None):
    """A committed Scene holding the file's meshes as geometries (ids in geometry_lib order) and
    its instances in file order, as LoadSceneHydra fills the accelerator (AddGeom_Triangles3f per
    mesh, AddInstance(mesh_id, matrix) per instance, CommitScene)."""
    from . import Scene

    meshes = load_meshes(hs) if meshes is None else meshes
    for gid, _ in hs.instances:
        if not 0 <= gid < len(meshes):
            raise ValueError(f"instance refers to mesh_id {gid}; the file has {len(meshes)} meshes")
    s = Scene(ctx)
    try:
        for v, f in meshes:
            s.add_triangles(v, f)
        for gid, m in instance_list(hs):
            s.add_instance(gid, m)
        if layout is None:
            s.commit()
        else:
            s.commit(layout=layout)
    except BaseException:
        s.close()
        raise
    return s
