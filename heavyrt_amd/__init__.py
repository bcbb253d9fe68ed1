"""heavyrt_amd -- MI355X ray-query engine (LBVH build + BVH2 nearest/any-hit).

Python view of the C-ABI in include/hrt.h (libhrt.so, built in-tree by
``python -m heavyrt_amd.build``).  The classes mirror the reference's
ISceneObject (core/CrossRT.h:64-193): ``Scene.add_triangles`` is
AddGeom_Triangles3f, ``add_instance`` is AddInstance, ``commit`` is
CommitScene, ``trace_nearest`` / ``trace_any`` are batched
RayQuery_NearestHit / RayQuery_AnyHit over device buffers.

There is no CPU fallback: if libhrt.so is missing or no GPU is visible the
calls raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhrt.so")

HIT_FULL = 0
HIT_COMPACT = 1
LAYOUT_BVH2 = 0
LAYOUT_BVH4 = 1
LAYOUT_BVH4_FP16 = 2
LAYOUT_AUTO = 3
LAYOUT_AUTO_FP16 = 4
LAYOUT_PLOC = 5  # nearest and any-hit on the FP32 BVH4 of the clustered second tree
LAYOUT_PLOC_FP16 = 6  # the same with FP16 boxes
TRI_MT = 0
TRI_WATERTIGHT = 1

DUMP_SIZES = 0
DUMP_CODES_INPUT = 1
DUMP_CODES_SORTED = 2
DUMP_IDS_SORTED = 3
DUMP_RUN_START = 4
DUMP_UNIQUE_CODES = 5
DUMP_LR_NODES = 6
DUMP_PERM = 7
DUMP_SCENE_BOX = 8
DUMP_QUADS32 = 9
DUMP_QUADS16 = 10
DUMP_PAIRS = 11
DUMP_ANY_TREE = 0x100  # OR'ed into a dump kind: the second tree of HRT_LAYOUT_AUTO_FP16 / the PLOC layouts
# hrt_debug_dump geometry id that selects the top-level tree of an instanced scene
TLAS_ID = 0xFFFFFFFF

STATUS_NAMES = {
    0: "HRT_OK",
    1: "HRT_ERR_INVALID_ARG",
    2: "HRT_ERR_HIP",
    3: "HRT_ERR_OUT_OF_MEMORY",
    4: "HRT_ERR_NOT_COMMITTED",
    5: "HRT_ERR_UNSUPPORTED",
    6: "HRT_ERR_LEAF_OVERFLOW",
}

# Every symbol include/hrt.h declares (checked by the CPU test suite).
EXPORTS = [
    "hrt_init", "hrt_shutdown", "hrt_ctx_stream", "hrt_scene_create", "hrt_scene_destroy",
    "hrt_scene_set_keep_build_arrays", "hrt_scene_clear_geom", "hrt_scene_clear_instances",
    "hrt_add_triangles", "hrt_add_instance", "hrt_scene_set_triangle_test", "hrt_commit", "hrt_trace_nearest",
    "hrt_trace_any", "hrt_trace_any_noalloc",
    "hrt_trace_nearest_stats", "hrt_trace_any_stats", "hrt_query_nearest_one", "hrt_geom_count",
    "hrt_instance_count", "hrt_debug_dump", "hrt_last_error", "hrt_rtao_make_rays", "hrt_rtao_trace", "hrt_rtao_render", "hrt_eye_make_rays", "hrt_eye_render", "hrt_scene_serialize", "hrt_scene_deserialize",
    "hrt_trace_nearest_host", "hrt_trace_any_host", "hrt_update_triangles", "hrt_update_instance",
    "hrt_batch_create", "hrt_batch_destroy", "hrt_batch_buffers", "hrt_batch_run", "hrt_batch_build",
    "hrt_batch_trace", "hrt_batch_candidate_count", "hrt_batch_candidate_info", "hrt_batch_set_candidate",
    "hrt_batch_get_candidate", "hrt_batch_autotune", "hrt_batch_dump", "hrt_batch_status",
    "hrt_issue_probe", "hrt_host_alloc", "hrt_host_free", "hrt_server_create", "hrt_server_destroy", "hrt_server_add_scene",
    "hrt_server_submit", "hrt_server_poll", "hrt_server_wait", "hrt_server_stats", "hrt_server_tree_rays",
]


class HrtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None
# Native objects still alive; closed (scenes first) before the interpreter
# exits, while the HIP runtime is still loaded.
_live_scenes: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


def _close_order(objs):
    # servers first: they trace scenes they do not own
    return sorted(objs, key=lambda o: 0 if type(o).__name__ == "Server" else 1)


@atexit.register
def _close_all() -> None:
    for s in _close_order(list(_live_scenes)):
        s.close()
    for c in list(_live_contexts):
        c.close()


def lib():
    """Loads libhrt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m heavyrt_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, sz, u32, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
        L.hrt_last_error.restype = C.c_char_p
        L.hrt_init.argtypes = [i32, C.POINTER(vp)]
        L.hrt_shutdown.argtypes = [vp]
        L.hrt_shutdown.restype = None
        L.hrt_ctx_stream.argtypes = [vp]
        L.hrt_ctx_stream.restype = vp
        L.hrt_scene_create.argtypes = [vp, C.POINTER(vp)]
        L.hrt_scene_destroy.argtypes = [vp]
        L.hrt_scene_destroy.restype = None
        L.hrt_scene_set_keep_build_arrays.argtypes = [vp, i32]
        L.hrt_scene_clear_geom.argtypes = [vp]
        L.hrt_scene_clear_instances.argtypes = [vp]
        L.hrt_add_triangles.argtypes = [vp, vp, sz, sz, vp, sz, i32, C.POINTER(u32)]
        L.hrt_add_instance.argtypes = [vp, u32, vp, C.POINTER(u32)]
        L.hrt_commit.argtypes = [vp, i32, i32]
        L.hrt_scene_set_triangle_test.argtypes = [vp, i32]
        L.hrt_trace_nearest.argtypes = [vp, vp, sz, vp, i32, vp]
        L.hrt_trace_any.argtypes = [vp, vp, sz, vp, vp]
        L.hrt_trace_any_noalloc.argtypes = [vp, vp, sz, vp, vp]
        L.hrt_trace_nearest_stats.argtypes = [vp, vp, sz, vp, i32, vp, vp]
        L.hrt_trace_any_stats.argtypes = [vp, vp, sz, vp, vp, vp]
        L.hrt_query_nearest_one.argtypes = [vp, vp, vp, vp]
        L.hrt_geom_count.argtypes = [vp]
        L.hrt_geom_count.restype = u32
        L.hrt_instance_count.argtypes = [vp]
        L.hrt_instance_count.restype = u32
        L.hrt_debug_dump.argtypes = [vp, u32, i32, vp, sz]
        L.hrt_rtao_make_rays.argtypes = [vp, vp, vp, sz, u32, C.c_float, C.c_uint64, vp, vp]
        L.hrt_rtao_trace.argtypes = [vp, vp, vp, sz, u32, C.c_float, C.c_uint64, vp, vp]
        L.hrt_rtao_render.argtypes = [vp, vp, vp, u32, u32, u32, C.c_float, C.c_uint64, C.c_float, vp, vp]
        L.hrt_eye_make_rays.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp]
        L.hrt_eye_render.argtypes = [vp, vp, vp, u32, u32, vp, u32, vp, vp]
        L.hrt_scene_serialize.argtypes = [vp, vp, sz, C.POINTER(C.c_size_t)]
        L.hrt_scene_deserialize.argtypes = [vp, vp, sz, C.POINTER(C.c_void_p)]
        L.hrt_trace_nearest_host.argtypes = [vp, vp, sz, vp, vp]
        L.hrt_trace_any_host.argtypes = [vp, vp, sz, vp]
        L.hrt_update_triangles.argtypes = [vp, u32, vp, sz, sz, vp, sz, i32]
        L.hrt_update_instance.argtypes = [vp, u32, vp]
        L.hrt_batch_create.argtypes = [vp, u32, vp, vp, vp, C.POINTER(vp)]
        L.hrt_batch_destroy.argtypes = [vp]
        L.hrt_batch_destroy.restype = None
        L.hrt_batch_buffers.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
        L.hrt_batch_run.argtypes = [vp, vp]
        L.hrt_batch_build.argtypes = [vp, vp]
        L.hrt_batch_trace.argtypes = [vp, vp]
        L.hrt_batch_candidate_count.argtypes = []
        L.hrt_batch_candidate_info.argtypes = [i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        L.hrt_batch_set_candidate.argtypes = [vp, i32]
        L.hrt_batch_get_candidate.argtypes = [vp]
        L.hrt_batch_autotune.argtypes = [vp, i32, C.POINTER(i32), vp]
        L.hrt_batch_dump.argtypes = [vp, u32, i32, vp, sz]
        L.hrt_batch_status.argtypes = [vp, C.POINTER(u32)]
        L.hrt_issue_probe.argtypes = [vp, vp]
        L.hrt_host_alloc.argtypes = [sz, C.POINTER(vp)]
        L.hrt_host_free.argtypes = [vp]
        L.hrt_host_free.restype = None
        L.hrt_server_create.argtypes = [vp, C.c_uint64, u32, C.POINTER(vp)]
        L.hrt_server_destroy.argtypes = [vp]
        L.hrt_server_destroy.restype = None
        L.hrt_server_add_scene.argtypes = [vp, vp, C.POINTER(u32)]
        L.hrt_server_submit.argtypes = [vp, u32, i32, vp, sz, vp, C.POINTER(C.c_uint64)]
        L.hrt_server_poll.argtypes = [vp, C.c_uint64]
        L.hrt_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hrt_server_stats.argtypes = [vp, vp]
        L.hrt_server_tree_rays.argtypes = [vp, vp]
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise HrtError(status, lib().hrt_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(x) -> int:
    """Address of a torch tensor, a numpy array (host) or an int address."""
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return int(x.ctypes.data)
    return int(x.data_ptr())


class Context:
    """One device and one HIP stream (hrt_init / hrt_shutdown)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().hrt_init(device, C.byref(self._h)))
        self.device = device
        _live_contexts.add(self)

    @property
    def stream(self) -> int:
        return int(lib().hrt_ctx_stream(self._h) or 0)

    def eye_make_rays(self, proj_inv, world_view_inv, width: int, height: int, out_rays, out_xy=None,
                      stream: int = 0):
        """Eye rays of a pinhole camera in the 01_eye_rays renderer's superblock order
        (hrt_eye_make_rays); matrices are 16 column-major floats."""
        pi = np.ascontiguousarray(proj_inv, dtype=np.float32).reshape(16)
        wi = np.ascontiguousarray(world_view_inv, dtype=np.float32).reshape(16)
        _check(lib().hrt_eye_make_rays(self._h, _ptr(pi), _ptr(wi), width, height, _dev_ptr(out_rays),
                                       _dev_ptr(out_xy) if out_xy is not None else None, stream or None))

    def issue_probe(self, stream: int = 0):
        """Launches the VALU issue-rate calibration kernel (hrt_issue_probe)."""
        _check(lib().hrt_issue_probe(self._h, stream or None))

    def close(self):
        """Destroys the context, and first every scene still open on it (a scene
        must not outlive the context it was created on)."""
        if self._h:
            for s in _close_order(list(_live_scenes)):
                if s.ctx is self:
                    s.close()
            lib().hrt_shutdown(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """ISceneObject over the GPU engine (one hrt_scene)."""

    def __init__(self, ctx: Context, keep_build_arrays: bool = False):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(lib().hrt_scene_create(ctx._h, C.byref(self._h)))
        _live_scenes.add(self)
        if keep_build_arrays:
            _check(lib().hrt_scene_set_keep_build_arrays(self._h, 1))

    def close(self):
        if self._h:
            lib().hrt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- AddGeom_Triangles3f / AddInstance / CommitScene --------------------
    def clear_geom(self):
        _check(lib().hrt_scene_clear_geom(self._h))

    def clear_instances(self):
        _check(lib().hrt_scene_clear_instances(self._h))

    def add_triangles(self, verts: np.ndarray, idx: np.ndarray, quality: int = 2) -> int:
        v = np.ascontiguousarray(verts, dtype=np.float32)
        stride = 4 * v.shape[1]
        i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        gid = C.c_uint32()
        _check(lib().hrt_add_triangles(self._h, _ptr(v), v.shape[0], stride, _ptr(i), i.size, quality,
                                       C.byref(gid)))
        return gid.value

    def add_instance(self, geom_id: int, matrix_col_major: np.ndarray | None = None) -> int:
        m = np.eye(4, dtype=np.float32) if matrix_col_major is None else np.asarray(matrix_col_major, np.float32)
        m = np.ascontiguousarray(m.reshape(16))
        iid = C.c_uint32()
        _check(lib().hrt_add_instance(self._h, geom_id, _ptr(m), C.byref(iid)))
        return iid.value

    def update_triangles(self, geom_id: int, verts: np.ndarray, idx: np.ndarray | None = None,
                         quality: int = 3):
        """UpdateGeom_Triangles3f: quality 3 (BUILD_REFIT) keeps the hierarchy,
        other qualities rebuild; commit again before tracing."""
        v = np.ascontiguousarray(verts, dtype=np.float32)
        i = None if idx is None else np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        _check(lib().hrt_update_triangles(self._h, geom_id, _ptr(v), v.shape[0], 4 * v.shape[1],
                                          None if i is None else _ptr(i), 0 if i is None else i.size, quality))

    def update_instance(self, inst_id: int, matrix_col_major: np.ndarray):
        m = np.ascontiguousarray(np.asarray(matrix_col_major, np.float32).reshape(16))
        _check(lib().hrt_update_instance(self._h, inst_id, _ptr(m)))

    def set_triangle_test(self, test: int):
        """TRI_MT (the reference's test) or TRI_WATERTIGHT, from the next commit on."""
        _check(lib().hrt_scene_set_triangle_test(self._h, test))

    def commit(self, quality: int = 1, layout: int = LAYOUT_BVH2):
        _check(lib().hrt_commit(self._h, quality, layout))

    @property
    def geom_count(self) -> int:
        return int(lib().hrt_geom_count(self._h))

    @property
    def instance_count(self) -> int:
        return int(lib().hrt_instance_count(self._h))

    # -- batched queries over device buffers ------------------------------
    def trace_nearest(self, rays, n: int, hits, fmt: int = HIT_COMPACT, stream: int = 0, stats=None):
        if stats is None:
            _check(lib().hrt_trace_nearest(self._h, _dev_ptr(rays), n, _dev_ptr(hits), fmt, stream or None))
        else:
            _check(lib().hrt_trace_nearest_stats(self._h, _dev_ptr(rays), n, _dev_ptr(hits), fmt,
                                                 _dev_ptr(stats), stream or None))

    def trace_any(self, rays, n: int, out, stream: int = 0, stats=None):
        if stats is None:
            _check(lib().hrt_trace_any(self._h, _dev_ptr(rays), n, _dev_ptr(out), stream or None))
        else:
            _check(lib().hrt_trace_any_stats(self._h, _dev_ptr(rays), n, _dev_ptr(out), _dev_ptr(stats),
                                             stream or None))

    def trace_nearest_host(self, rays: np.ndarray, stats: bool = False):
        """Host-memory batch (hrt_trace_nearest_host): (n, 8) uint32 hit records[, (n, 2) counters]."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((r.shape[0], 8), dtype=np.uint32)
        st = np.zeros((r.shape[0], 2), dtype=np.uint32) if stats else None
        _check(lib().hrt_trace_nearest_host(self._h, _ptr(r), r.shape[0], _ptr(out), _ptr(st) if stats else None))
        return (out, st) if stats else out

    def trace_any_host(self, rays: np.ndarray) -> np.ndarray:
        """Host-memory batch (hrt_trace_any_host): one bool per ray."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros(r.shape[0], dtype=np.uint8)
        _check(lib().hrt_trace_any_host(self._h, _ptr(r), r.shape[0], _ptr(out)))
        return out.astype(bool)

    def query_nearest_one(self, pos_near, dir_far):
        p = np.ascontiguousarray(pos_near, dtype=np.float32)
        d = np.ascontiguousarray(dir_far, dtype=np.float32)
        out = np.zeros(8, dtype=np.uint32)
        _check(lib().hrt_query_nearest_one(self._h, _ptr(p), _ptr(d), _ptr(out)))
        return (float(out[0:1].view(np.float32)[0]), int(out[1]), int(out[2]), int(out[3]),
                float(out[4:5].view(np.float32)[0]), float(out[5:6].view(np.float32)[0]))

    def rtao_make_rays(self, prim_rays, prim_hits_compact, npix: int, spp: int, tmax: float, seed: int, out_rays,
                       stream: int = 0):
        _check(lib().hrt_rtao_make_rays(self._h, _dev_ptr(prim_rays), _dev_ptr(prim_hits_compact), npix, spp,
                                        float(tmax), seed & 0xFFFFFFFFFFFFFFFF, _dev_ptr(out_rays), stream or None))

    def rtao_trace(self, prim_rays, prim_hits_compact, npix: int, spp: int, tmax: float, seed: int, out_hits,
                   stream: int = 0):
        """rtao_make_rays + trace_any in one call (hrt_rtao_trace): one byte per AO ray in out_hits."""
        _check(lib().hrt_rtao_trace(self._h, _dev_ptr(prim_rays), _dev_ptr(prim_hits_compact), npix, spp,
                                    float(tmax), seed & 0xFFFFFFFFFFFFFFFF, _dev_ptr(out_hits), stream or None))

    def serialize(self) -> bytes:
        """The committed scene with its trees as bytes (hrt_scene_serialize), for a warm start."""
        n = C.c_size_t(0)
        _check(lib().hrt_scene_serialize(self._h, None, 0, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        _check(lib().hrt_scene_serialize(self._h, _ptr(buf), buf.size, C.byref(n)))
        return buf.tobytes()

    @classmethod
    def deserialize(cls, ctx: "Context", data: bytes) -> "Scene":
        """A committed scene from serialize()'s bytes (hrt_scene_deserialize); no tree is rebuilt."""
        s = cls.__new__(cls)
        s.ctx = ctx
        s._h = C.c_void_p()
        buf = np.frombuffer(data, dtype=np.uint8)
        _check(lib().hrt_scene_deserialize(ctx._h, _ptr(buf), buf.size, C.byref(s._h)))
        _live_scenes.add(s)
        return s

    def rtao_render(self, proj_inv, world_view_inv, width: int, height: int, spp: int, tmax: float, seed: int,
                    out_color, power: float = 1.0, stream: int = 0):
        """The 02_rtao frame on the device (hrt_rtao_render): out_color is a device buffer of
        width * height words."""
        pi = np.ascontiguousarray(proj_inv, dtype=np.float32).reshape(16)
        wi = np.ascontiguousarray(world_view_inv, dtype=np.float32).reshape(16)
        _check(lib().hrt_rtao_render(self._h, _ptr(pi), _ptr(wi), width, height, spp, float(tmax),
                                     seed & 0xFFFFFFFFFFFFFFFF, float(power), _dev_ptr(out_color), stream or None))

    def eye_render(self, proj_inv, world_view_inv, width: int, height: int, palette, out_color, stream: int = 0):
        """The 01_eye_rays frame on the device (hrt_eye_render): out_color is a device buffer of
        width * height words."""
        pi = np.ascontiguousarray(proj_inv, dtype=np.float32).reshape(16)
        wi = np.ascontiguousarray(world_view_inv, dtype=np.float32).reshape(16)
        pal = np.ascontiguousarray(palette, dtype=np.uint32).reshape(-1)
        _check(lib().hrt_eye_render(self._h, _ptr(pi), _ptr(wi), width, height, _ptr(pal), pal.size,
                                    _dev_ptr(out_color), stream or None))

    # -- parity hooks -------------------------------------------------------
    def dump(self, geom_id: int, what: int) -> np.ndarray:
        flag = what & DUMP_ANY_TREE
        what &= ~DUMP_ANY_TREE
        sizes = np.zeros(6, dtype=np.uint32)
        _check(lib().hrt_debug_dump(self._h, geom_id, DUMP_SIZES | flag, _ptr(sizes), sizes.nbytes))
        if what == DUMP_SIZES:
            return sizes
        n, L, nq = int(sizes[0]), int(sizes[1]), int(sizes[4])
        if what == DUMP_QUADS32:
            out = np.zeros((nq, 32), dtype=np.uint32)
        elif what == DUMP_QUADS16:
            out = np.zeros((nq, 16), dtype=np.uint32)
        elif what == DUMP_SCENE_BOX:
            out = np.zeros(12, dtype=np.float32)
        elif what == DUMP_PAIRS:
            out = np.zeros((int(sizes[3]), 16), dtype=np.uint32)
        elif what in (DUMP_RUN_START, DUMP_UNIQUE_CODES):
            out = np.zeros(L, dtype=np.uint32)
        elif what == DUMP_LR_NODES:
            out = np.zeros((2 * L - 1, 8), dtype=np.uint32)
        else:
            out = np.zeros(n, dtype=np.uint32)
        _check(lib().hrt_debug_dump(self._h, geom_id, what | flag, _ptr(out), out.nbytes))
        return out


def batch_candidates():
    """[(waves per workgroup, LDS stack entries, refill threshold, claim)] of
    the batched trace kernel (LDS 0 = one ray per lane without refill)."""
    L = lib()
    out = []
    for i in range(L.hrt_batch_candidate_count()):
        w, s, r, c = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(L.hrt_batch_candidate_info(i, C.byref(w), C.byref(s), C.byref(r), C.byref(c)))
        out.append((w.value, s.value, r.value, c.value))
    return out


_hip_rt = None


def _hip():
    """The HIP runtime (already loaded by libhrt.so) for plain host<->device copies."""
    global _hip_rt
    if _hip_rt is None:
        lib()
        h = C.CDLL("libamdhip64.so")
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip_rt = h
    return _hip_rt


def _copy(dst, src, nbytes: int, kind: int) -> None:
    if nbytes:
        rc = _hip().hipMemcpy(dst, src, nbytes, kind)
        if rc != 0:
            raise HrtError(2, f"hipMemcpy failed ({rc})")


class Batch:
    """Many one-geometry scenes built and traced together (hrt_batch_*)."""

    def __init__(self, ctx: Context, n_verts, n_tris, n_rays):
        nv = np.ascontiguousarray(n_verts, dtype=np.uint32)
        nt = np.ascontiguousarray(n_tris, dtype=np.uint32)
        nr = np.ascontiguousarray(n_rays, dtype=np.uint32)
        assert nv.size == nt.size == nr.size
        self.ctx = ctx
        self.n_scenes = int(nv.size)
        self.n_verts, self.n_tris, self.n_rays = nv, nt, nr
        self.vert_off = np.concatenate([[0], np.cumsum(nv, dtype=np.int64)])
        self.tri_off = np.concatenate([[0], np.cumsum(nt, dtype=np.int64)])
        self.ray_off = np.concatenate([[0], np.cumsum(nr, dtype=np.int64)])
        self._h = C.c_void_p()
        _check(lib().hrt_batch_create(ctx._h, self.n_scenes, _ptr(nv), _ptr(nt), _ptr(nr), C.byref(self._h)))
        _live_scenes.add(self)
        v, i, r, h = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().hrt_batch_buffers(self._h, C.byref(v), C.byref(i), C.byref(r), C.byref(h)))
        self.d_verts, self.d_idx, self.d_rays, self.d_hits = v.value, i.value, r.value, h.value

    def close(self):
        if self._h:
            lib().hrt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scenes) -> None:
        """scenes: [(verts (n,3|4), faces (m,3), rays (k,8))] with the sizes given
        at creation.  Waits for the device first (a running batch may still be
        reading its inputs)."""
        assert len(scenes) == self.n_scenes
        v = np.concatenate([np.hstack([np.asarray(x[0], np.float32)[:, :3], np.ones((len(x[0]), 1), np.float32)])
                            for x in scenes])
        f = np.concatenate([np.asarray(x[1], np.uint32).reshape(-1) for x in scenes])
        r = np.concatenate([np.asarray(x[2], np.float32).reshape(-1, 8) for x in scenes])
        assert v.shape[0] == self.vert_off[-1] and f.size == 3 * self.tri_off[-1] and r.shape[0] == self.ray_off[-1]
        for s, x in enumerate(scenes):
            fi = np.asarray(x[1])
            if fi.size and int(fi.max()) >= int(self.n_verts[s]):
                raise ValueError(f"scene {s}: vertex index out of range")
        _hip().hipDeviceSynchronize()
        for host, dev in ((v, self.d_verts), (f, self.d_idx), (r, self.d_rays)):
            host = np.ascontiguousarray(host)
            _copy(dev, host.ctypes.data, host.nbytes, 1)

    def hits(self) -> np.ndarray:
        """(total rays, 2) uint32: t bits and scene-local primId per ray (waits for the device)."""
        n = int(self.ray_off[-1])
        out = np.zeros((n, 2), dtype=np.uint32)
        _hip().hipDeviceSynchronize()
        _copy(out.ctypes.data, self.d_hits, out.nbytes, 2)
        return out

    def run(self, stream: int = 0):
        _check(lib().hrt_batch_run(self._h, stream or None))

    def build(self, stream: int = 0):
        _check(lib().hrt_batch_build(self._h, stream or None))

    def trace(self, stream: int = 0):
        _check(lib().hrt_batch_trace(self._h, stream or None))

    def set_candidate(self, i: int):
        _check(lib().hrt_batch_set_candidate(self._h, i))

    @property
    def candidate(self) -> int:
        return int(lib().hrt_batch_get_candidate(self._h))

    def autotune(self, reps: int = 3):
        """Returns (chosen candidate, ms per candidate)."""
        n = lib().hrt_batch_candidate_count()
        ms = np.zeros(n, dtype=np.float32)
        ch = C.c_int()
        _check(lib().hrt_batch_autotune(self._h, reps, C.byref(ch), _ptr(ms)))
        return ch.value, ms

    def status(self) -> None:
        """Raises HrtError (with .scene set) when the last build could not store some scene's tree."""
        bad = C.c_uint32(0xFFFFFFFF)
        st = lib().hrt_batch_status(self._h, C.byref(bad))
        if st != 0:
            err = HrtError(st, lib().hrt_last_error().decode(errors="replace"))
            err.scene = None if bad.value == 0xFFFFFFFF else int(bad.value)
            raise err

    def dump(self, scene: int, what: int) -> np.ndarray:
        sizes = np.zeros(5, dtype=np.uint32)
        _check(lib().hrt_batch_dump(self._h, scene, DUMP_SIZES, _ptr(sizes), sizes.nbytes))
        if what == DUMP_SIZES:
            return sizes
        if what == DUMP_PAIRS:
            out = np.zeros((int(sizes[3]), 16), dtype=np.uint32)
        else:
            out = np.zeros(int(sizes[0]), dtype=np.uint32)
        _check(lib().hrt_batch_dump(self._h, scene, what, _ptr(out), out.nbytes))
        return out


class _PinnedHolder:
    def __init__(self, ptr: int):
        self.ptr = ptr

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.hrt_host_free(self.ptr)
            self.ptr = 0


class PinnedArray(np.ndarray):
    """numpy array in pinned host memory (hrt_host_alloc); freed with its last view."""


def pinned_array(shape, dtype=np.float32) -> np.ndarray:
    dtype = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(count * dtype.itemsize, 16)
    p = C.c_void_p()
    _check(lib().hrt_host_alloc(nbytes, C.byref(p)))
    holder = _PinnedHolder(p.value)
    buf = (C.c_byte * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=dtype, count=count).reshape(shape).view(PinnedArray)
    arr._holder = holder
    return arr


class Server:
    """Resident scenes and a continuous-batching request loop (hrt_server_*)."""

    def __init__(self, ctx: Context, budget_bytes: int, steps_in_flight: int = 2):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(lib().hrt_server_create(ctx._h, int(budget_bytes), steps_in_flight, C.byref(self._h)))
        _live_scenes.add(self)
        self._keep = {}

    def close(self):
        if self._h:
            lib().hrt_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: "Scene") -> int:
        sid = C.c_uint32()
        _check(lib().hrt_server_add_scene(self._h, scene._h, C.byref(sid)))
        return sid.value

    def submit(self, scene_id: int, rays, out, any_hit: bool = False) -> int:
        """rays: (n, 8) float32; out: (n, 2) uint32 (nearest) or (n,) uint8 (any-hit); each a pinned host
        array (pinned_array) or a device tensor on the server's device."""
        n = rays.shape[0]
        rid = C.c_uint64()
        _check(lib().hrt_server_submit(self._h, scene_id, 1 if any_hit else 0, _dev_ptr(rays), n, _dev_ptr(out),
                                       C.byref(rid)))
        self._keep[rid.value] = (rays, out)
        return rid.value

    def poll(self, rid: int) -> int:
        return int(lib().hrt_server_poll(self._h, rid))

    def wait(self, rid: int):
        """Returns (latency ms: submit -> results on the host, queue ms: submit -> admission)."""
        lat, q = C.c_double(), C.c_double()
        try:
            _check(lib().hrt_server_wait(self._h, rid, C.byref(lat), C.byref(q)))
        finally:
            self._keep.pop(rid, None)
        return lat.value, q.value

    def stats(self) -> dict:
        o = np.zeros(6, dtype=np.uint64)
        _check(lib().hrt_server_stats(self._h, _ptr(o)))
        return {"completed": int(o[0]), "rays": int(o[1]), "steps": int(o[2]),
                "mean_requests_per_step": int(o[3]) / 1000.0, "peak_bytes_in_flight": int(o[4]),
                "budget_bytes": int(o[5])}

    def tree_rays(self) -> dict:
        """Rays launched so far per kind of tree."""
        o = np.zeros(3, dtype=np.uint64)
        _check(lib().hrt_server_tree_rays(self._h, _ptr(o)))
        return {"bvh2": int(o[0]), "bvh4_fp16": int(o[1]), "bvh4_fp32": int(o[2])}

