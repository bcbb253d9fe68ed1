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
    "hrt_add_triangles", "hrt_add_instance", "hrt_commit", "hrt_trace_nearest", "hrt_trace_any",
    "hrt_trace_nearest_stats", "hrt_trace_any_stats", "hrt_query_nearest_one", "hrt_geom_count",
    "hrt_instance_count", "hrt_debug_dump", "hrt_last_error", "hrt_rtao_make_rays",
    "hrt_trace_nearest_host", "hrt_trace_any_host",
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


@atexit.register
def _close_all() -> None:
    for s in list(_live_scenes):
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
        L.hrt_trace_nearest.argtypes = [vp, vp, sz, vp, i32, vp]
        L.hrt_trace_any.argtypes = [vp, vp, sz, vp, vp]
        L.hrt_trace_nearest_stats.argtypes = [vp, vp, sz, vp, i32, vp, vp]
        L.hrt_trace_any_stats.argtypes = [vp, vp, sz, vp, vp, vp]
        L.hrt_query_nearest_one.argtypes = [vp, vp, vp, vp]
        L.hrt_geom_count.argtypes = [vp]
        L.hrt_geom_count.restype = u32
        L.hrt_instance_count.argtypes = [vp]
        L.hrt_instance_count.restype = u32
        L.hrt_debug_dump.argtypes = [vp, u32, i32, vp, sz]
        L.hrt_rtao_make_rays.argtypes = [vp, vp, vp, sz, u32, C.c_float, C.c_uint64, vp, vp]
        L.hrt_trace_nearest_host.argtypes = [vp, vp, sz, vp, vp]
        L.hrt_trace_any_host.argtypes = [vp, vp, sz, vp]
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise HrtError(status, lib().hrt_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(x) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(x, int):
        return x
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

    def close(self):
        """Destroys the context, and first every scene still open on it (a scene
        must not outlive the context it was created on)."""
        if self._h:
            for s in list(_live_scenes):
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

    # -- parity hooks -------------------------------------------------------
    def dump(self, geom_id: int, what: int) -> np.ndarray:
        sizes = np.zeros(5, dtype=np.uint32)
        _check(lib().hrt_debug_dump(self._h, geom_id, DUMP_SIZES, _ptr(sizes), sizes.nbytes))
        if what == DUMP_SIZES:
            return sizes
        n, L, nq = int(sizes[0]), int(sizes[1]), int(sizes[4])
        if what == DUMP_QUADS32:
            out = np.zeros((nq, 32), dtype=np.uint32)
        elif what == DUMP_QUADS16:
            out = np.zeros((nq, 16), dtype=np.uint32)
        elif what == DUMP_SCENE_BOX:
            out = np.zeros(12, dtype=np.float32)
        elif what in (DUMP_RUN_START, DUMP_UNIQUE_CODES):
            out = np.zeros(L, dtype=np.uint32)
        elif what == DUMP_LR_NODES:
            out = np.zeros((2 * L - 1, 8), dtype=np.uint32)
        else:
            out = np.zeros(n, dtype=np.uint32)
        _check(lib().hrt_debug_dump(self._h, geom_id, what, _ptr(out), out.nbytes))
        return out
