"""Python view of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker.  The engine never calls it.

The oracle is a C++ restatement of the reference's LBVH builder and
BVH2CommonRT / BruteForceRT traversal (see hrt_oracle.cpp for the file:line
map).  This wrapper loads ``oracle/_build/libhrt_oracle.so`` (built by
``make -C oracle``) and returns numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhrt_oracle.so")
_lib = None

MODE_BVH2COMMON = 0
MODE_LR = 1
MODE_BRUTE = 2

MODE4_FP32 = 0
MODE4_FP16 = 1
MODE4_BRUTE = 2
MODE4_FP16_FUSED = 3  # FP16 nodes, box tests in the fused form of the engine's any-hit calls

TRI_MT = 0          # the reference's Moller-Trumbore (1e-6 slack)
TRI_WATERTIGHT = 1  # Woop, Benthin, Wald 2013, FP32


def f2h(v: float) -> int:
    return int(lib().hrto_f2h(v))


def round_to_neg(v: float) -> int:
    return int(lib().hrto_round_neg(v))


def round_to_pos(v: float) -> int:
    return int(lib().hrto_round_pos(v))


def h2f(h: int) -> float:
    return float(lib().hrto_h2f(h))


def build_library() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        L.hrto_build.restype = vp
        L.hrto_build.argtypes = [vp, u32, vp, u32]
        L.hrto_build_ex.restype = vp
        L.hrto_build_ex.argtypes = [vp, u32, vp, u32, C.c_int]
        L.hrto_free.argtypes = [vp]
        L.hrto_refit.restype = C.c_int
        L.hrto_refit.argtypes = [vp, vp, u32, vp, u32]
        L.hrto_ploc.restype = None
        L.hrto_ploc.argtypes = [vp, u32]
        L.hrto_superblock_offset.restype = u32
        L.hrto_superblock_offset.argtypes = [u32, u32, u32]
        L.hrto_eye_rays.restype = None
        L.hrto_eye_rays.argtypes = [vp, vp, u32, u32, vp, vp]
        L.hrto_sizes.argtypes = [vp, vp]
        L.hrto_scene_box.argtypes = [vp, vp]
        L.hrto_codes.argtypes = [vp, vp, vp, vp]
        L.hrto_runs.argtypes = [vp, vp, vp]
        L.hrto_lr_nodes.argtypes = [vp, vp]
        L.hrto_peach_nodes.argtypes = [vp, vp]
        L.hrto_perm.argtypes = [vp, vp]
        L.hrto_temp_count.argtypes = [vp, vp]
        L.hrto_trace.argtypes = [vp, C.c_int, vp, u64, vp, vp, C.c_int]
        L.hrto_t_fp64.argtypes = [vp, vp, vp, u64, vp]
        L.hrto_zindex3d.restype = u32
        L.hrto_zindex3d.argtypes = [C.c_int, C.c_int, C.c_int]
        L.hrto_pack_leaf.restype = u32
        L.hrto_pack_leaf.argtypes = [u32, u32]
        L.hrto_safe_inverse.argtypes = [vp, vp]
        L.hrto_slab.argtypes = [vp, vp, vp, vp, vp]
        L.hrto_pcg32.restype = None
        L.hrto_pcg32.argtypes = [u64, u64, u32, vp]
        L.hrto_ao_rays.restype = None
        L.hrto_ao_rays.argtypes = [vp, vp, C.c_uint32, vp, vp, u64, C.c_uint32, C.c_float, C.c_uint64, vp]
        L.hrto_slab_fused.argtypes = [vp, vp, vp, vp, vp]
        L.hrto_slab_robust.argtypes = [vp, vp, vp, vp, vp]
        L.hrto_quad_count.restype = u64
        L.hrto_quad_count.argtypes = [vp]
        L.hrto_quads.argtypes = [vp, C.c_int, vp]
        L.hrto_trace4.argtypes = [vp, C.c_int, vp, u64, vp, vp, C.c_int]
        L.hrto_trace_metrics.restype = None
        L.hrto_trace_metrics.argtypes = [vp, vp, u64, vp, vp]
        L.hrto_set_triangle_test.restype = None
        L.hrto_set_triangle_test.argtypes = [C.c_int]
        L.hrto_tri_eval.restype = None
        L.hrto_tri_eval.argtypes = [C.c_int, vp, vp, C.c_float, vp, vp, vp, vp]
        L.hrto_tri_hit_counts.restype = None
        L.hrto_tri_hit_counts.argtypes = [C.c_int, vp, u32, vp, u32, vp, u64, vp, C.c_int]
        for name in ("hrto_f2h", "hrto_round_neg", "hrto_round_pos"):
            getattr(L, name).restype = u32
            getattr(L, name).argtypes = [C.c_float]
        L.hrto_h2f.restype = C.c_float
        L.hrto_h2f.argtypes = [u32]
        L.hrto_scene_new.restype = vp
        L.hrto_scene_free.argtypes = [vp]
        L.hrto_scene_add_geom.restype = C.c_int
        L.hrto_scene_add_geom.argtypes = [vp, vp, u32, vp, u32]
        L.hrto_scene_add_instance.restype = C.c_int
        L.hrto_scene_add_instance.argtypes = [vp, u32, vp]
        L.hrto_scene_commit.restype = vp
        L.hrto_scene_commit.argtypes = [vp]
        L.hrto_scene_inst_info.argtypes = [vp, vp, vp]
        L.hrto_scene_trace.argtypes = [vp, C.c_int, vp, u64, vp, vp, C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class triangle_test:
    """Context manager: the oracle's traversals use `test` (TRI_MT or TRI_WATERTIGHT) inside the block."""

    def __init__(self, test: int):
        self.test = int(test)

    def __enter__(self):
        lib().hrto_set_triangle_test(self.test)
        return self

    def __exit__(self, *exc):
        lib().hrto_set_triangle_test(TRI_MT)
        return False


def tri_eval(test: int, o, d, tnear: float, A, B, C):
    """One triangle test: (t, u, v, accepted)."""
    args = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, A, B, C)]
    out = np.zeros(4, dtype=np.float32)
    lib().hrto_tri_eval(int(test), _ptr(args[0]), _ptr(args[1]), float(tnear), _ptr(args[2]), _ptr(args[3]),
                        _ptr(args[4]), _ptr(out))
    return float(out[0]), float(out[1]), float(out[2]), bool(out[3])


def tri_hit_counts(test: int, verts, idx, rays, threads: int = 0) -> np.ndarray:
    """Per ray, the number of triangles the test accepts with t in (tNear, |tFar|) (brute force)."""
    v4 = as_vert4(verts)
    f = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    out = np.zeros(r.shape[0], dtype=np.uint32)
    lib().hrto_tri_hit_counts(int(test), _ptr(v4), v4.shape[0], _ptr(f), f.size // 3, _ptr(r), r.shape[0],
                              _ptr(out), threads)
    return out


def as_vert4(verts: np.ndarray) -> np.ndarray:
    """float32 (n,3) or (n,4) -> contiguous (n,4) with w = 1."""
    v = np.asarray(verts, dtype=np.float32)
    out = np.ones((v.shape[0], 4), dtype=np.float32)
    out[:, :3] = v[:, :3]
    return out


class OracleBVH:
    """One LBVH build of one triangle mesh, with the reference's traversals."""

    def __init__(self, verts: np.ndarray, idx: np.ndarray, cubic: bool = False):
        """cubic=True: the engine's any-hit tree (one Morton scale for all axes)."""
        self.v4 = as_vert4(verts)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        L = lib()
        self._owned = True
        self._h = L.hrto_build_ex(_ptr(self.v4), self.v4.shape[0], _ptr(self.idx), self.idx.size, 1 if cubic else 0)
        if not self._h:
            raise ValueError("oracle build needs at least one triangle")
        self._read_sizes()

    @classmethod
    def _view(cls, handle, owner) -> "OracleBVH":
        """A build owned by another oracle object (the top-level tree of a scene)."""
        self = cls.__new__(cls)
        self._owned = False
        self._owner = owner
        self._h = handle
        self._read_sizes()
        return self

    def _read_sizes(self):
        s = np.zeros(5, dtype=np.uint64)
        lib().hrto_sizes(self._h, _ptr(s))
        self.ntris, self.leaves, self.n_lr, self.n_peach, self.max_run = (int(x) for x in s)

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None) and _lib is not None:
            _lib.hrto_free(self._h)
            self._h = None

    def refit(self, verts: np.ndarray, idx: np.ndarray | None = None):
        """BUILD_REFIT: new positions (and optionally indices), same hierarchy."""
        self.v4 = as_vert4(verts)
        if idx is not None:
            self.idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        r = lib().hrto_refit(self._h, _ptr(self.v4), self.v4.shape[0], None if idx is None else _ptr(self.idx),
                             0 if idx is None else self.idx.size)
        if r != 0:
            raise ValueError("a refit keeps the triangle count")

    def ploc(self, radius: int = 16):
        """Replaces the hierarchy above the leaves with the agglomerative (PLOC) one; leaves stay."""
        lib().hrto_ploc(self._h, radius)

    def scene_box(self):
        out = np.zeros(12, dtype=np.float32)
        lib().hrto_scene_box(self._h, _ptr(out))
        return out[0:4].copy(), out[4:8].copy(), out[8:12].copy()

    def codes(self):
        n = self.ntris
        a, b, c = (np.zeros(n, dtype=np.uint32) for _ in range(3))
        lib().hrto_codes(self._h, _ptr(a), _ptr(b), _ptr(c))
        return a, b, c

    def runs(self):
        a = np.zeros(self.leaves, dtype=np.uint32)
        b = np.zeros(self.leaves, dtype=np.uint32)
        lib().hrto_runs(self._h, _ptr(a), _ptr(b))
        return a, b

    def lr_nodes(self) -> np.ndarray:
        """(2L-1, 8) uint32 words: bmin xyz, left, bmax xyz, right."""
        out = np.zeros((self.n_lr, 8), dtype=np.uint32)
        lib().hrto_lr_nodes(self._h, _ptr(out))
        return out

    def peach_nodes(self) -> np.ndarray:
        out = np.zeros((self.n_peach, 8), dtype=np.uint32)
        lib().hrto_peach_nodes(self._h, _ptr(out))
        return out

    def perm(self) -> np.ndarray:
        out = np.zeros(self.ntris, dtype=np.uint32)
        lib().hrto_perm(self._h, _ptr(out))
        return out

    def temp_count(self) -> np.ndarray:
        out = np.zeros(2 * self.ntris - 1, dtype=np.uint32)
        lib().hrto_temp_count(self._h, _ptr(out))
        return out

    def trace(self, rays: np.ndarray, mode: int = MODE_BVH2COMMON, counters: bool = False, threads: int = 0):
        """rays: (n,8) float32 = posNear, dirFar.  Returns (t, primId, u, v[, pairs, tris])."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 6), dtype=np.uint32)
        cnt = np.zeros((n, 2), dtype=np.uint32) if counters else None
        lib().hrto_trace(self._h, mode, _ptr(rays), n, _ptr(hits), _ptr(cnt) if counters else None, threads)
        t = hits[:, 0].view(np.float32).copy()
        prim = hits[:, 1].copy()
        u = hits[:, 4].view(np.float32).copy()
        v = hits[:, 5].view(np.float32).copy()
        if counters:
            return t, prim, u, v, cnt[:, 0].copy(), cnt[:, 1].copy()
        return t, prim, u, v

    def trace_metrics(self, rays: np.ndarray):
        """BVH2CommonRT traversal on one thread with the reference's metrics on (as shipped).
        Returns (t, primId, metrics dict)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 6), dtype=np.uint32)
        m = np.zeros(6, dtype=np.uint64)
        lib().hrto_trace_metrics(self._h, _ptr(rays), rays.shape[0], _ptr(hits), _ptr(m))
        names = ("NC", "LC", "TC", "BLB", "SOC", "WSS128")
        return hits[:, 0].view(np.float32).copy(), hits[:, 1].copy(), {k: int(x) for k, x in zip(names, m)}

    # -- BVH4 (collapse of the same LBVH) ---------------------------------
    def quads(self, fp16: bool = False) -> np.ndarray:
        """(nquads, 4, 8) uint32: per child bmin xyz, ref, bmax xyz, 0."""
        n = int(lib().hrto_quad_count(self._h))
        out = np.zeros((n, 4, 8), dtype=np.uint32)
        lib().hrto_quads(self._h, 1 if fp16 else 0, _ptr(out))
        return out

    def trace4(self, rays: np.ndarray, mode: int = MODE4_FP32, counters: bool = False, threads: int = 0):
        """BVH4 traversal (MODE4_FP32 / MODE4_FP16) or brute force (MODE4_BRUTE),
        nearest hit by (t, primId).  Returns (t, primId, u, v[, quads, tris])."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 6), dtype=np.uint32)
        cnt = np.zeros((n, 2), dtype=np.uint32) if counters else None
        lib().hrto_trace4(self._h, mode, _ptr(rays), n, _ptr(hits), _ptr(cnt) if counters else None, threads)
        t = hits[:, 0].view(np.float32).copy()
        prim = hits[:, 1].copy()
        u = hits[:, 4].view(np.float32).copy()
        v = hits[:, 5].view(np.float32).copy()
        if counters:
            return t, prim, u, v, cnt[:, 0].copy(), cnt[:, 1].copy()
        return t, prim, u, v

    def t_fp64(self, rays: np.ndarray, prim: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        prim = np.ascontiguousarray(prim, dtype=np.uint32)
        out = np.zeros(rays.shape[0], dtype=np.float64)
        lib().hrto_t_fp64(self._h, _ptr(rays), _ptr(prim), rays.shape[0], _ptr(out))
        return out


SCENE_TWO_LEVEL = 0
SCENE_PER_INSTANCE = 1
SCENE_BRUTE = 2


class OracleScene:
    """Geometries, instances and the top-level tree over instance boxes, with
    the two-level traversal (BVH2CommonRT.cpp:143-279), its per-instance closed
    form and brute force."""

    def __init__(self):
        self._h = lib().hrto_scene_new()
        self._keep = []
        self.tlas = None
        self.ninst = 0

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hrto_scene_free(self._h)
            self._h = None

    def add_geom(self, verts: np.ndarray, idx: np.ndarray) -> int:
        v4 = as_vert4(verts)
        i = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        g = lib().hrto_scene_add_geom(self._h, _ptr(v4), v4.shape[0], _ptr(i), i.size)
        if g < 0:
            raise ValueError("geometry needs at least one triangle")
        return g

    def add_instance(self, geom: int, matrix_col_major: np.ndarray) -> int:
        m = np.ascontiguousarray(np.asarray(matrix_col_major, dtype=np.float32).reshape(16))
        r = lib().hrto_scene_add_instance(self._h, geom, _ptr(m))
        if r < 0:
            raise ValueError("unknown geometry or singular matrix")
        self.ninst = r + 1
        return r

    def commit(self):
        h = lib().hrto_scene_commit(self._h)
        self.tlas = OracleBVH._view(h, self) if h else None

    def inst_info(self):
        """(inverse rows (n,3,4) float32, world boxes (n,2,4) float32)."""
        inv = np.zeros((self.ninst, 3, 4), dtype=np.float32)
        box = np.zeros((self.ninst, 2, 4), dtype=np.float32)
        lib().hrto_scene_inst_info(self._h, _ptr(inv), _ptr(box))
        return inv, box

    def trace(self, rays: np.ndarray, mode: int = SCENE_TWO_LEVEL, counters: bool = False, threads: int = 0):
        """Returns (t, primId, instId, geomId, u, v[, pairs, tris])."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 6), dtype=np.uint32)
        cnt = np.zeros((n, 2), dtype=np.uint32) if counters else None
        lib().hrto_scene_trace(self._h, mode, _ptr(rays), n, _ptr(hits), _ptr(cnt) if counters else None, threads)
        out = (hits[:, 0].view(np.float32).copy(), hits[:, 1].copy(), hits[:, 2].copy(), hits[:, 3].copy(),
               hits[:, 4].view(np.float32).copy(), hits[:, 5].view(np.float32).copy())
        if counters:
            return out + (cnt[:, 0].copy(), cnt[:, 1].copy())
        return out


def zindex3d(x: int, y: int, z: int) -> int:
    return int(lib().hrto_zindex3d(x, y, z))


def pack_leaf(start: int, size: int) -> int:
    return int(lib().hrto_pack_leaf(start, size))


def safe_inverse(d) -> np.ndarray:
    d = np.asarray(d, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    lib().hrto_safe_inverse(_ptr(d), _ptr(out))
    return out


def slab(o, inv, bmin, bmax) -> np.ndarray:
    args = [np.asarray(a, dtype=np.float32) for a in (o, inv, bmin, bmax)]
    out = np.zeros(2, dtype=np.float32)
    lib().hrto_slab(*(_ptr(a) for a in args), _ptr(out))
    return out


def slab_robust(o, d, bmin, bmax) -> np.ndarray:
    """The watertight mode's conservative box test: (entry, exit * (1 + 4 * 2^-23)), parallel axes exact."""
    args = [np.asarray(a, dtype=np.float32) for a in (o, d, bmin, bmax)]
    out = np.zeros(2, dtype=np.float32)
    lib().hrto_slab_robust(*(_ptr(a) for a in args), _ptr(out))
    return out


def slab_fused(o, inv, bmin, bmax) -> np.ndarray:
    args = [np.asarray(a, dtype=np.float32) for a in (o, inv, bmin, bmax)]
    out = np.zeros(2, dtype=np.float32)
    lib().hrto_slab_fused(*(_ptr(a) for a in args), _ptr(out))
    return out


def superblock_offset(x: int, y: int, width: int) -> int:
    """SuperBlockIndex2DOpt (examples/01_eye_rays/eye_ray.cpp:114-131): pixel -> thread index."""
    return int(lib().hrto_superblock_offset(x, y, width))


def pcg32(seed: int, seq: int, n: int) -> np.ndarray:
    """n PCG32 draws for (initstate seed, initseq seq), the generator behind ao_rays."""
    out = np.zeros(n, dtype=np.uint32)
    lib().hrto_pcg32(int(seed) & 0xFFFFFFFFFFFFFFFF, int(seq) & 0xFFFFFFFFFFFFFFFF, int(n), _ptr(out))
    return out


def ao_rays(verts: np.ndarray, idx: np.ndarray, prim_rays: np.ndarray, hits_compact: np.ndarray, spp: int,
            tmax: float, seed: int) -> np.ndarray:
    """The engine's AO ray generator (ao.h) restated: (npix * spp, 8) rays from primary rays and
    their compact hits (t bits, primId)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
    pr = np.ascontiguousarray(prim_rays, dtype=np.float32).reshape(-1, 8)
    h = np.ascontiguousarray(hits_compact, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros((pr.shape[0] * spp, 8), dtype=np.float32)
    lib().hrto_ao_rays(_ptr(v), _ptr(f), f.size // 3, _ptr(pr), _ptr(h), pr.shape[0], spp, float(tmax),
                       seed & 0xFFFFFFFFFFFFFFFF, _ptr(out))
    return out


def eye_rays(proj_inv_col_major: np.ndarray, world_view_inv_col_major: np.ndarray, width: int, height: int):
    """kernel_PackXY + kernel_InitEyeRay (eye_ray.cpp:26-47, 133-138): rays (n, 8) in thread order and
    the packed y << 16 | x of each thread."""
    pi = np.ascontiguousarray(proj_inv_col_major, dtype=np.float32).reshape(16)
    wi = np.ascontiguousarray(world_view_inv_col_major, dtype=np.float32).reshape(16)
    n = width * height
    rays = np.zeros((n, 8), dtype=np.float32)
    xy = np.zeros(n, dtype=np.uint32)
    lib().hrto_eye_rays(_ptr(pi), _ptr(wi), width, height, _ptr(rays), _ptr(xy))
    return rays, xy
