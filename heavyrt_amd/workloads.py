"""Synthetic scenes and rays for the five workload shapes (numpy, host side).

All randomness is PCG32 (M. O'Neill, "PCG: a family of simple fast
space-efficient statistically good algorithms for random number generation",
2014; the pcg32_srandom / pcg32_random_r steps of the reference C
implementation).  To stay vectorised, lane j of a draw uses its own PCG32
sequence: initstate = seed, initseq = (stream << 32) + j.  Every generator
below takes an explicit stream number so that each workload is reproducible.

Geometry follows the workload definitions:
  eye rays   icosphere (20 faces subdivided 6x, 81,920 tris, radius 1, centre
             (0,1,0)) on a 2-triangle ground quad, 512x512 pinhole camera
  RTAO       1024x512-quad heightfield over [0,2]x[0,1] (value noise, amp 0.2)
             + 2,000 randomly rotated boxes; 1024x1024 primary rays
  soups      random triangle soups: centres U[0,1]^3, each vertex = centre +
             e*u, e ~ LogNormal(-4, 0.5), u uniform on the sphere
"""
from __future__ import annotations

import math

import numpy as np

PCG_SEED = 0x853C49E6748FEA9B
PCG_MULT = 6364136223846793005
FLT_MAX = np.float32(3.4028234663852886e38)

# stream numbers (one per workload)
STREAM_HEIGHTFIELD = 2
STREAM_BOXES = 3
STREAM_SOUP = 4
STREAM_RAYS = 5
STREAM_TEST = 9
STREAM_BATCH = 10        # batched build+trace: sizes, then soups and rays at 1000 + 2 * scene
STREAM_SERVE = 20        # serving: scene sizes; soups at +10+k; request draws at +100/+200 per rate


def pcg32(n: int, draws: int, stream: int, seed: int = PCG_SEED) -> np.ndarray:
    """(n, draws) uint32: lane j yields `draws` outputs of PCG32(seed, (stream << 32) + j)."""
    M = np.uint64(PCG_MULT)
    with np.errstate(over="ignore"):
        seq = (np.uint64(stream) << np.uint64(32)) + np.arange(n, dtype=np.uint64)
        inc = (seq << np.uint64(1)) | np.uint64(1)
        state = np.zeros(n, dtype=np.uint64)
        state = state * M + inc
        state = state + np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        state = state * M + inc
        out = np.empty((n, draws), dtype=np.uint32)
        for k in range(draws):
            old = state
            state = old * M + inc
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
            rot = (old >> np.uint64(59)).astype(np.uint32)
            out[:, k] = (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))
    return out


def uniform01(bits: np.ndarray) -> np.ndarray:
    """uint32 -> float64 in [0, 1) with 24 bits (exact in float32)."""
    return (bits >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def unit_vectors(u1: np.ndarray, u2: np.ndarray) -> np.ndarray:
    z = 1.0 - 2.0 * u1
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * math.pi * u2
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def icosphere(levels: int = 6, radius: float = 1.0, centre=(0.0, 1.0, 0.0)):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(levels):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mids], axis=0)
        nf = f.shape[0]
        m01, m12, m20 = (inv[0:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:3 * nf] + base)
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=0)
    verts = (v * radius + np.asarray(centre)).astype(np.float32)
    return verts, f.astype(np.uint32)


def ground_quad(half: float = 10.0, y: float = 0.0):
    v = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], dtype=np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.uint32)
    return v, f


def merge(*meshes):
    vs, fs, off = [], [], 0
    for v, f in meshes:
        vs.append(v)
        fs.append(f + off)
        off += v.shape[0]
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.uint32)


def eye_scene():
    """Icosphere (6 levels, 81,920 tris) + ground quad (2 tris) = 81,922 tris."""
    return merge(icosphere(6), ground_quad())


def _value_noise(x: np.ndarray, z: np.ndarray, lattice: np.ndarray, freq: float) -> np.ndarray:
    gx, gz = x * freq, z * freq
    ix, iz = np.floor(gx).astype(np.int64), np.floor(gz).astype(np.int64)
    fx, fz = gx - ix, gz - iz
    sx, sz = fx * fx * (3.0 - 2.0 * fx), fz * fz * (3.0 - 2.0 * fz)
    n = lattice.shape[0]
    a = lattice[ix % n, iz % n]
    b = lattice[(ix + 1) % n, iz % n]
    c = lattice[ix % n, (iz + 1) % n]
    d = lattice[(ix + 1) % n, (iz + 1) % n]
    return (a * (1 - sx) + b * sx) * (1 - sz) + (c * (1 - sx) + d * sx) * sz


def height_fn(seed_stream: int = STREAM_HEIGHTFIELD, amp: float = 0.2):
    lat = uniform01(pcg32(64 * 64, 1, seed_stream)[:, 0]).reshape(64, 64)
    octaves = [(4.0, 1.0), (8.0, 0.5), (16.0, 0.25), (32.0, 0.125)]
    norm = sum(w for _, w in octaves)

    def h(x, z):
        acc = np.zeros(np.broadcast(x, z).shape, dtype=np.float64)
        for f, w in octaves:
            acc += w * _value_noise(x, z, lat, f)
        return amp * acc / norm

    return h


def heightfield(nx: int = 1024, nz: int = 512, size_x: float = 2.0, size_z: float = 1.0, amp: float = 0.2):
    h = height_fn(amp=amp)
    xs = np.linspace(0.0, size_x, nx + 1)
    zs = np.linspace(0.0, size_z, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = h(X, Z)
    verts = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(nx)[:, None]
    j = np.arange(nz)[None, :]
    v00 = (i * (nz + 1) + j).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + (nz + 1)
    v11 = v10 + 1
    t1 = np.stack([v00, v01, v10], axis=1)
    t2 = np.stack([v10, v01, v11], axis=1)
    faces = np.empty((2 * t1.shape[0], 3), dtype=np.uint32)
    faces[0::2] = t1
    faces[1::2] = t2
    return verts, faces, h


_CUBE_V = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                    [-1, 1, 1]], dtype=np.float64) * 0.5
_CUBE_F = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 6, 2], [3, 7, 6],
                    [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.int64)


def _box_frames(count: int, hfn, size_x=2.0, size_z=1.0, stream: int = STREAM_BOXES):
    """Edge length, rotation and centre of each box (same draws for the merged
    mesh and for the instanced form)."""
    r = uniform01(pcg32(count, 6, stream))
    edge = 0.01 + 0.04 * r[:, 0]
    # uniform random rotation (Shoemake's method)
    u1, u2, u3 = r[:, 1], r[:, 2], r[:, 3]
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * math.pi * u2), np.sqrt(1 - u1) * np.cos(2 * math.pi * u2),
                  np.sqrt(u1) * np.sin(2 * math.pi * u3), np.sqrt(u1) * np.cos(2 * math.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], axis=1)
    cx = size_x * r[:, 4]
    cz = size_z * r[:, 5]
    cy = hfn(cx, cz)
    c = np.stack([cx, cy, cz], axis=1)
    return edge, R, c


def boxes(count: int, hfn, size_x=2.0, size_z=1.0, stream: int = STREAM_BOXES):
    edge, R, c = _box_frames(count, hfn, size_x, size_z, stream)
    local = _CUBE_V[None, :, :] * edge[:, None, None]
    world = np.einsum("bij,bkj->bki", R, local) + c[:, None, :]
    verts = world.reshape(-1, 3).astype(np.float32)
    faces = (_CUBE_F[None, :, :] + 8 * np.arange(count)[:, None, None]).reshape(-1, 3).astype(np.uint32)
    return verts, faces


def unit_cube():
    """The 12-triangle cube of edge 1 centred at the origin."""
    return _CUBE_V.astype(np.float32), _CUBE_F.astype(np.uint32)


def box_instance_matrices(count: int, hfn, size_x=2.0, size_z=1.0, stream: int = STREAM_BOXES) -> np.ndarray:
    """The boxes of `boxes` as instances of `unit_cube`: (count, 16) float32
    column-major matrices (rotation x edge, then the centre)."""
    edge, R, c = _box_frames(count, hfn, size_x, size_z, stream)
    M = np.zeros((count, 4, 4), dtype=np.float64)
    M[:, :3, :3] = R * edge[:, None, None]
    M[:, :3, 3] = c
    M[:, 3, 3] = 1.0
    return np.ascontiguousarray(M.transpose(0, 2, 1).reshape(count, 16)).astype(np.float32)


def rtao_instanced_scene(nx: int = 1024, nz: int = 512, nboxes: int = 2000):
    """The RTAO scene with the boxes as instances: geometries [heightfield,
    unit cube], instances [(0, identity)] + [(1, box matrix)] * nboxes."""
    hv, hf, hfn = heightfield(nx, nz)
    cv, cf = unit_cube()
    ident = np.eye(4, dtype=np.float32).reshape(16)
    inst = [(0, ident)] + [(1, m) for m in box_instance_matrices(nboxes, hfn)]
    return [(hv, hf), (cv, cf)], inst


def rtao_scene(nx: int = 1024, nz: int = 512, nboxes: int = 2000):
    """Heightfield (2*nx*nz tris) + nboxes rotated boxes (12 tris each)."""
    hv, hf, hfn = heightfield(nx, nz)
    bv, bf = boxes(nboxes, hfn)
    return merge((hv, hf), (bv, bf))


def soup(ntris: int, stream: int = STREAM_SOUP, mu: float = -4.0, sigma: float = 0.5):
    r = uniform01(pcg32(ntris, 11, stream))
    centres = r[:, 0:3]
    # Box-Muller for the log-normal edge scale
    g = np.sqrt(-2.0 * np.log(np.maximum(r[:, 3], 1e-12))) * np.cos(2 * math.pi * r[:, 4])
    e = np.exp(mu + sigma * g)
    verts = np.empty((ntris, 3, 3), dtype=np.float64)
    for k in range(3):
        u = unit_vectors(r[:, 5 + 2 * k], r[:, 6 + 2 * k])
        verts[:, k, :] = centres + e[:, None] * u
    return verts.reshape(-1, 3).astype(np.float32), np.arange(3 * ntris, dtype=np.uint32).reshape(-1, 3)


# ---------------------------------------------------------------------------
# rays (8 float32 per ray: pos xyz, tNear, dir xyz, tFar)
# ---------------------------------------------------------------------------
def pinhole_rays(width: int, height: int, eye, target, up=(0.0, 1.0, 0.0), fov_y_deg: float = 45.0,
                 tnear: float = 0.0, tfar: float = float(FLT_MAX)) -> np.ndarray:
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    th = math.tan(math.radians(fov_y_deg) * 0.5)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width
    py = (np.arange(height) + 0.5) / height
    X, Y = np.meshgrid(px, py, indexing="xy")
    sx = (2.0 * X - 1.0) * th * aspect
    sy = (1.0 - 2.0 * Y) * th
    d = f[None, None, :] + sx[..., None] * r[None, None, :] + sy[..., None] * u[None, None, :]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    n = width * height
    rays = np.empty((n, 8), dtype=np.float32)
    rays[:, 0:3] = eye.astype(np.float32)
    rays[:, 3] = tnear
    rays[:, 4:7] = d.reshape(n, 3).astype(np.float32)
    rays[:, 7] = tfar
    return rays


def eye_rays(width: int = 512, height: int = 512) -> np.ndarray:
    return pinhole_rays(width, height, eye=(0.0, 1.5, 4.0), target=(0.0, 1.0, 0.0), fov_y_deg=45.0)


EYE_POS = (0.0, 1.5, 4.0)
EYE_TARGET = (0.0, 1.0, 0.0)

# EyeRayCaster::m_palette (examples/01_eye_rays/eye_ray.h:106-113)
EYE_PALETTE = np.array([0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0,
                        0xfff032e6, 0xffd2f53c, 0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8,
                        0xff800000, 0xffaaffc3, 0xff808000, 0xffffd8b1, 0xff000080, 0xff808080], dtype=np.uint32)


def perspective(fov_y_deg: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """OpenGL-style projection (the role of LiteMath::perspectiveMatrix, eye_ray_host.cpp:314), float64."""
    f = 1.0 / math.tan(math.radians(fov_y_deg) * 0.5)
    return np.array([[f / aspect, 0, 0, 0], [0, f, 0, 0],
                     [0, 0, (zfar + znear) / (znear - zfar), 2 * zfar * znear / (znear - zfar)],
                     [0, 0, -1, 0]], dtype=np.float64)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """World -> view matrix looking down -z (the role of LiteMath::lookAt), float64."""
    eye = np.asarray(eye, dtype=np.float64)
    z = eye - np.asarray(target, dtype=np.float64)
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, dtype=np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = x, y, z
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def eye_camera(width: int = 512, height: int = 512, eye=EYE_POS, target=EYE_TARGET, fov_y_deg: float = 45.0):
    """(inverse projection, inverse world-view) as 16 column-major float32 each, as
    EyeRayCaster keeps them in m_projInv / m_worldViewInv (eye_ray_host.cpp:312-318)."""
    proj = perspective(fov_y_deg, width / height, 0.01, 100.0)
    view = look_at(eye, target)
    pi = np.linalg.inv(proj).astype(np.float32)
    wi = np.linalg.inv(view).astype(np.float32)
    return pi.reshape(4, 4).flatten(order="F"), wi.reshape(4, 4).flatten(order="F")


RTAO_EYE = (1.0, 1.1, 0.15)
RTAO_TARGET = (1.0, 0.0, 0.45)
RTAO_FOV = 44.0


def superblock_order(width: int, height: int) -> np.ndarray:
    """Row-major pixel index (y * width + x) of each thread in the renderers' 8x8-superblock
    packing (SuperBlockIndex2DOpt, examples/01_eye_rays/eye_ray.cpp:114-131): superblocks of four
    4x4 blocks, in rows of superblocks.  width and height must be multiples of 8."""
    tid = np.arange(width * height, dtype=np.int64)
    sb, in_h, in_b = tid >> 6, (tid >> 4) & 3, tid & 15
    wb = width >> 3
    bx = (sb % wb) * 2 + (in_h & 1)
    by = (sb // wb) * 2 + (in_h >> 1)
    return ((by * 4 + (in_b >> 2)) * width + bx * 4 + (in_b & 3)).astype(np.int64)


def xcd_stripe_order(width: int, height: int, group: int = 256, stripe: int = 4, xcds: int = 8) -> np.ndarray:
    """Row-major pixel index of each ray in the order hrt_eye_render traces (eye.hip,
    xcd_stripe_pixel): workgroup b of `group` rays runs on XCD b % xcds and gets pixels of
    the rows r with (r // stripe) % xcds == b % xcds, in row order.  Row order when the
    height is not a multiple of stripe * xcds or the width not of group // stripe."""
    tid = np.arange(width * height, dtype=np.int64)
    if height % (stripe * xcds) or width % (group // stripe):
        return tid
    b = tid // group
    x = b % xcds
    k = (b // xcds) * group + tid % group
    rk = k // width
    row = (rk // stripe) * (stripe * xcds) + x * stripe + rk % stripe
    return row * width + k % width


def rtao_primary_rays(width: int = 1024, height: int = 1024, order: str = "rows") -> np.ndarray:
    """RTAO eye rays, one per pixel: in row-major order, or ("superblock") in the renderers'
    8x8-superblock order."""
    rays = pinhole_rays(width, height, eye=RTAO_EYE, target=RTAO_TARGET, fov_y_deg=RTAO_FOV)
    return rays[superblock_order(width, height)] if order == "superblock" else rays


def incoherent_rays(n: int, stream: int = STREAM_RAYS, tfar: float = float(FLT_MAX)) -> np.ndarray:
    r = uniform01(pcg32(n, 5, stream))
    rays = np.empty((n, 8), dtype=np.float32)
    rays[:, 0:3] = r[:, 0:3]
    rays[:, 3] = 0.0
    rays[:, 4:7] = unit_vectors(r[:, 3], r[:, 4])
    rays[:, 7] = tfar
    return rays


def scene_diagonal(verts: np.ndarray) -> float:
    v = np.asarray(verts, dtype=np.float64)
    return float(np.linalg.norm(v.max(axis=0) - v.min(axis=0)))


# ---------------------------------------------------------------------------
# VSGF meshes (header {u64 size; u32 nVert, nIdx, nMat, flags}, cmesh.cpp:86-93, 138-173)
# ---------------------------------------------------------------------------
def load_vsgf(path: str):
    with open(path, "rb") as fh:
        data = fh.read()
    hdr = np.frombuffer(data[:24], dtype=np.uint32)
    nvert, nidx, flags = int(hdr[2]), int(hdr[3]), int(hdr[5])
    off = 24
    pos = np.frombuffer(data[off:off + 16 * nvert], dtype=np.float32).reshape(nvert, 4)
    off += 16 * nvert
    if not (flags & 8):  # HAS_NO_NORMALS
        off += 16 * nvert
    if flags & 1:  # HAS_TANGENT
        off += 16 * nvert
    off += 8 * nvert  # texcoords
    idx = np.frombuffer(data[off:off + 4 * nidx], dtype=np.uint32)
    return pos[:, :3].copy(), idx.reshape(-1, 3).copy()


def log_uniform_int(n: int, lo: float, hi: float, stream: int) -> np.ndarray:
    u = uniform01(pcg32(n, 1, stream))[:, 0]
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batched_workload(n_scenes: int = 256, tris=(1e3, 2e5), rays=(4096, 262144), stream: int = STREAM_BATCH):
    """Batched build+trace workload: n_scenes random soups with triangle counts
    log-uniform in `tris` and ray counts log-uniform in `rays` (incoherent
    rays: origins uniform in the unit cube, directions uniform on the sphere,
    tFar = FLT_MAX).  Returns [(verts, faces, rays)] per scene."""
    nt = log_uniform_int(n_scenes, tris[0], tris[1], stream)
    nr = log_uniform_int(n_scenes, rays[0], rays[1], stream + 1)
    out = []
    for s in range(n_scenes):
        v, f = soup(int(nt[s]), stream=1000 + 2 * s)
        out.append((v, f, incoherent_rays(int(nr[s]), stream=1001 + 2 * s)))
    return out



def serving_scenes(count: int = 16, tris=(5e4, 2e6), stream: int = STREAM_SERVE):
    """Serving workload: `count` resident scenes with triangle counts
    log-uniform in `tris`; even ids are random soups, odd ids heightfields of
    2nz x nz quads with nz = round(sqrt(n / 4)).  Returns [(verts, faces)]."""
    nt = log_uniform_int(count, tris[0], tris[1], stream)
    out = []
    for k in range(count):
        if k % 2 == 0:
            out.append(soup(int(nt[k]), stream=stream + 10 + k))
        else:
            nz = max(8, int(round(math.sqrt(nt[k] / 4.0))))
            v, f, _ = heightfield(2 * nz, nz)
            out.append((v, f))
    return out


def serving_camera(verts: np.ndarray):
    """Eye point, target and AO radius of a serving scene's 1024x1024 request pool."""
    lo, hi = verts.min(0), verts.max(0)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    eye = c + np.array([0.35, 0.8, 1.6]) * ext
    return tuple(eye), tuple(c), 0.1 * ext


def serving_requests(count: int, qps: float, n_scenes: int = 16, pool: int = 1 << 20, sizes=(1e3, 1e6),
                     stream: int = STREAM_SERVE + 100):
    """A Poisson request stream: sizes log-uniform in `sizes`, exponential
    gaps at rate `qps`, scene ids uniform, 50/50 nearest-hit / any-hit, and a
    start offset into the scene's ray pool of `pool` rays.  Returns a dict of
    arrays (size, arrival_s, scene, any_hit, start)."""
    sz = log_uniform_int(count, sizes[0], sizes[1], stream)
    u = uniform01(pcg32(count, 4, stream + 100))
    gaps = -np.log(np.maximum(u[:, 0], 1e-12)) / qps
    return {"size": sz, "arrival_s": np.cumsum(gaps), "scene": np.minimum((u[:, 1] * n_scenes).astype(np.int64),
                                                                          n_scenes - 1),
            "any_hit": u[:, 2] < 0.5, "start": (u[:, 3] * (pool - sz)).astype(np.int64)}
