"""Synthetic scenes for the frame path: geometry, clusters, height maps and
cameras.

The reference loads glTF assets and clusterizes them with meshoptimizer
(``src/scene.cpp:365-552``).  Here geometry is generated: regular grid
patches and icospheres whose triangles are grouped into clusters of at most
128 triangles / 128 vertices along the meshes' own structure, with the
reference's per-cluster vertex duplication (``scene.cpp:519-552``): every
cluster owns a contiguous vertex range and 8-bit local triangle indices.
Height maps are value-noise fBm from an integer hash, so they are identical
on every machine.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import abi

# --------------------------------------------------------------- geometry


@dataclass
class Geometry:
    positions: np.ndarray        # (V, 3) f32, per-cluster vertex ranges
    normals: np.ndarray          # (V, 3) f32
    texcoords: np.ndarray        # (V, 2) f32
    clusters: np.ndarray         # (C,) CLUSTER_DT
    local_triangles: np.ndarray  # (3T,) u8
    bboxes: np.ndarray           # (C,) BBOX_DT
    lo: np.ndarray = field(default=None)
    hi: np.ndarray = field(default=None)

    @property
    def num_triangles(self):
        return int(self.local_triangles.size // 3)

    @property
    def num_vertices(self):
        return int(self.positions.shape[0])

    @property
    def num_clusters(self):
        return int(self.clusters.size)


def _finish(pos_list, nrm_list, uv_list, tri_list):
    """Concatenates per-cluster vertex / triangle blocks into a Geometry."""
    nc = len(pos_list)
    clusters = np.zeros(nc, abi.CLUSTER_DT)
    bboxes = np.zeros(nc, abi.BBOX_DT)
    v0 = t0 = 0
    for c in range(nc):
        nv, nt = pos_list[c].shape[0], tri_list[c].shape[0]
        assert nv <= 128 and nt <= 128
        clusters[c] = (nv, nt, t0, v0, 3 * t0)
        p = pos_list[c]
        tri = tri_list[c]
        e = np.concatenate([np.linalg.norm(p[tri[:, 1]] - p[tri[:, 0]], axis=1),
                            np.linalg.norm(p[tri[:, 2]] - p[tri[:, 1]], axis=1),
                            np.linalg.norm(p[tri[:, 0]] - p[tri[:, 2]], axis=1)])
        bboxes[c] = (p.min(0), p.max(0), e.min(), e.max())
        v0 += nv
        t0 += nt
    pos = np.ascontiguousarray(np.concatenate(pos_list).astype(np.float32))
    g = Geometry(pos, np.ascontiguousarray(np.concatenate(nrm_list).astype(np.float32)),
                 np.ascontiguousarray(np.concatenate(uv_list).astype(np.float32)), clusters,
                 np.ascontiguousarray(np.concatenate(tri_list).astype(np.uint8).reshape(-1)), bboxes)
    g.lo, g.hi = pos.min(0), pos.max(0)
    return g


def grid_patch(nx: int, nz: int, size_x: float, size_z: float, tile: int = 8) -> Geometry:
    """nx x nz quads in the XZ plane, normal +Y, uv in [0,1]^2; clusters are
    tile x tile quad blocks (8 x 8 -> 128 triangles, 81 vertices)."""
    dx, dz = size_x / nx, size_z / nz
    P, N, U, T = [], [], [], []
    for cz in range(0, nz, tile):
        for cx in range(0, nx, tile):
            tx, tz = min(tile, nx - cx), min(tile, nz - cz)
            ii, jj = np.meshgrid(np.arange(tx + 1), np.arange(tz + 1), indexing="xy")
            gi, gj = (cx + ii).reshape(-1), (cz + jj).reshape(-1)
            p = np.stack([gi * dx, np.zeros_like(gi, dtype=np.float64), gj * dz], 1)
            uv = np.stack([gi / nx, gj / nz], 1)
            n = np.tile([0.0, 1.0, 0.0], (p.shape[0], 1))
            a = (jj[:-1, :-1] * (tx + 1) + ii[:-1, :-1]).reshape(-1)  # v00
            b = a + 1                                               # v10
            c = a + (tx + 1)                                        # v01
            d = c + 1                                               # v11
            tri = np.stack([np.stack([a, c, b], 1), np.stack([b, c, d], 1)], 1).reshape(-1, 3)
            P.append(p), N.append(n), U.append(uv), T.append(tri)
    return _finish(P, N, U, T)


def vertex_normals(positions: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (FP64 accumulation) for meshes that
    carry none."""
    p = positions.astype(np.float64)
    t = indices.reshape(-1, 3).astype(np.int64)
    fn = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, t[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 0, n / np.where(ln > 0, ln, 1.0), np.array([0.0, 1.0, 0.0]))
    return n.astype(np.float32)


def mesh_geometry(positions, indices, normals=None, texcoords=None, max_triangles: int = 128,
                  max_vertices: int = 128) -> Geometry:
    """Geometry of an indexed triangle mesh: clusters from the library's
    host clusterizer (tc_clusterize_mesh, the counterpart of the reference's
    meshoptimizer step, scene.cpp:396-552) with per-cluster vertex ranges and
    8-bit local indices.  Missing normals are area-weighted vertex normals,
    missing texture coordinates zero."""
    import ctypes as C
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    if idx.size % 3:
        raise ValueError("indices must hold whole triangles")
    nt, nv = idx.size // 3, pos.shape[0]
    nrm = vertex_normals(pos, idx) if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    uv = np.zeros((nv, 2), np.float32) if texcoords is None else np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
    if nrm.shape[0] != nv or uv.shape[0] != nv:
        raise ValueError("normals / texcoords must have one entry per position")
    clusters = np.zeros(max(nt, 1), abi.CLUSTER_DT)
    bboxes = np.zeros(max(nt, 1), abi.BBOX_DT)
    lverts = np.zeros(max(3 * nt, 1), np.uint32)
    ltris = np.zeros(max(3 * nt, 1), np.uint8)
    ncl, ncv = C.c_uint32(0), C.c_uint32(0)
    abi.check(abi.lib().tc_clusterize_mesh(pos.ctypes.data, nv, idx.ctypes.data, nt, max_triangles, max_vertices,
                                           clusters.ctypes.data, lverts.ctypes.data, ltris.ctypes.data,
                                           bboxes.ctypes.data, C.byref(ncl), C.byref(ncv)), "tc_clusterize_mesh")
    nc, ncv = ncl.value, ncv.value
    lv = lverts[:ncv]
    g = Geometry(np.ascontiguousarray(pos[lv]), np.ascontiguousarray(nrm[lv]), np.ascontiguousarray(uv[lv]),
                 clusters[:nc].copy(), np.ascontiguousarray(ltris[:3 * nt]), bboxes[:nc].copy())
    g.lo, g.hi = pos.min(0), pos.max(0)
    g.source_vertices = lv.copy()  # mesh vertex behind every cluster vertex
    return g


_ICO_V = None


def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def icosphere(segments: int, radius: float = 1.0, block: int = 8) -> Geometry:
    """Icosahedron with every face split into a triangular lattice of
    `segments` per edge (20 * segments^2 triangles), projected on the sphere.
    Clusters are lattice sub-triangles of `block` segments (block^2
    triangles, (block+1)(block+2)/2 vertices)."""
    assert segments % block == 0
    V, F = _icosahedron()
    m = block
    # canonical lattice triangle with m segments: vertices (a, b), a + b <= m
    loc = [(a, b) for b in range(m + 1) for a in range(m + 1 - b)]
    idx = {ab: i for i, ab in enumerate(loc)}
    tris = []
    for b in range(m):
        for a in range(m - b):
            tris.append((idx[(a, b)], idx[(a + 1, b)], idx[(a, b + 1)]))
            if a + b < m - 1:
                tris.append((idx[(a + 1, b)], idx[(a + 1, b + 1)], idx[(a, b + 1)]))
    tris = np.array(tris)
    loc = np.array(loc, np.float64)
    k = segments // m
    P, N, U, T = [], [], [], []
    for A, B, Cc in F:
        pa, pb, pc = V[A], V[B], V[Cc]
        for J in range(k):
            for I in range(k - J):
                for inv in (False, True):
                    if inv and I + J >= k - 1:
                        continue
                    if not inv:
                        li = I * m + loc[:, 0]
                        lj = J * m + loc[:, 1]
                    else:  # point-reflected block keeps orientation
                        li = (I + 1) * m - loc[:, 0]
                        lj = (J + 1) * m - loc[:, 1]
                    p = pa + np.outer(li / segments, pb - pa) + np.outer(lj / segments, pc - pa)
                    n = p / np.linalg.norm(p, axis=1, keepdims=True)
                    u = 0.5 + np.arctan2(n[:, 2], n[:, 0]) / (2 * math.pi)
                    v = np.arccos(np.clip(n[:, 1], -1, 1)) / math.pi
                    P.append(n * radius), N.append(n), U.append(np.stack([u, v], 1)), T.append(tris.copy())
    return _finish(P, N, U, T)


# ------------------------------------------------------------ height maps

def _hash2(x, y, seed):
    h = (x.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (y.astype(np.uint32) * np.uint32(0xD8163841)) ^ np.uint32(seed)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h


def fbm_height(size: int, seed: int, octaves: int, base_freq: int = 8) -> np.ndarray:
    """Tileable value-noise fBm in [0, 1] on a size x size grid (float32).
    FP64 arithmetic in a fixed order; tc_fbm_height computes the same texels
    on the device."""
    coords = (np.arange(size, dtype=np.float64) + 0.5) / size
    acc = np.zeros((size, size), np.float64)
    amp_sum = 0.0
    for o in range(octaves):
        f = base_freq << o
        amp = 0.5 ** o
        x = coords * f
        xi = np.floor(x)
        xf = x - xi
        sx = xf * xf * (3.0 - 2.0 * xf)
        x0 = xi.astype(np.uint64) % f
        x1 = (xi.astype(np.uint64) + 1) % f
        s = np.uint32((seed * 0x27D4EB2F + o * 0x165667B1) & 0xFFFFFFFF)
        # rows are y, columns are x; the lattice is square so both axes share x0 / x1
        v00 = _hash2(x0[None, :], x0[:, None], s) / 4294967296.0
        v10 = _hash2(x1[None, :], x0[:, None], s) / 4294967296.0
        v01 = _hash2(x0[None, :], x1[:, None], s) / 4294967296.0
        v11 = _hash2(x1[None, :], x1[:, None], s) / 4294967296.0
        SX = sx[None, :]
        SY = sx[:, None]
        top = v00 * (1.0 - SX) + v10 * SX
        bot = v01 * (1.0 - SX) + v11 * SX
        acc = acc + amp * (top * (1.0 - SY) + bot * SY)
        amp_sum = amp_sum + amp
    return (acc / amp_sum).astype(np.float32)


def to_r16(h: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(h * 65535.0), 0, 65535).astype(np.uint16)


# ----------------------------------------------------------------- cameras


def perspective_rh_zo(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    """glm::perspectiveRH_ZO with the reference's Y flip (tessellatedclusters.cpp:577-579),
    returned as a row-major 4x4 (math layout)."""
    t = math.tan(fovy / 2)
    M = np.zeros((4, 4))
    M[0, 0] = 1.0 / (aspect * t)
    M[1, 1] = -1.0 / t
    M[2, 2] = far / (near - far)
    M[3, 2] = -1.0
    M[2, 3] = -(far * near) / (far - near)
    return M


def look_at_rh(eye, center, up) -> np.ndarray:
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    M = np.eye(4)
    M[0, :3], M[1, :3], M[2, :3] = s, u, -f
    M[0, 3], M[1, 3], M[2, 3] = -s @ eye, -u @ eye, f @ eye
    return M


def colmajor(M: np.ndarray):
    return (abi.f32 * 16)(*np.asarray(M, np.float32).T.reshape(-1).tolist())


@dataclass
class Camera:
    eye: tuple
    center: tuple
    up: tuple = (0.0, 1.0, 0.0)
    fovy_deg: float = 60.0
    width: int = 1920
    height: int = 1080
    near: float = 0.01
    far: float = 1000.0


RIPPLE_DEFAULT = (100.0, 0.005, 3.7)  # frequency, amplitude, speed (tessellatedclusters.cpp:75-78)


def frame_constants(cam: Camera, tess_rate_pixels: float, sun_dir=(0.3, 0.9, 0.2), disp_scale: float = 1.0,
                    disp_offset: float = 0.0, animation_time: float = None,
                    ripple=RIPPLE_DEFAULT) -> abi.FrameConstants:
    """FrameConstants as tessellatedclusters.cpp:526-613 derives them.
    `animation_time` (seconds of animation clock) turns the ripple state on;
    it only deforms vertices when the frame runs with FLAG_ANIMATION."""
    fc = abi.FrameConstants()
    proj = perspective_rh_zo(math.radians(cam.fovy_deg), cam.width / cam.height, cam.near, cam.far)
    view = look_at_rh(cam.eye, cam.center, cam.up)
    viewI = np.linalg.inv(view)
    vp = proj @ view
    fc.projMatrix = colmajor(proj)
    fc.projMatrixI = colmajor(np.linalg.inv(proj))
    fc.viewProjMatrix = colmajor(vp)
    fc.viewProjMatrixI = colmajor(np.linalg.inv(vp))
    fc.viewMatrix = colmajor(view)
    fc.viewMatrixI = colmajor(viewI)
    fc.viewPos = (abi.f32 * 4)(*[float(x) for x in viewI[:3, 3]], 1.0)
    d = -viewI[:3, 2]
    fc.viewDir = (abi.f32 * 4)(*[float(x) for x in d], 0.0)
    fc.viewport = (abi.i32 * 2)(cam.width, cam.height)
    fc.viewportf = (abi.f32 * 2)(float(cam.width), float(cam.height))
    fc.tessRate = 1.0 / tess_rate_pixels if tess_rate_pixels else 0.0
    fc.displacementScale = disp_scale
    fc.displacementOffset = disp_offset
    fc.nearPlane = cam.near
    fc.farPlane = cam.far
    fc.supersample = 1
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    fc.skyParams.sunDirection = (abi.f32 * 3)(*[float(x) for x in s])
    fc.doShadow = 1
    fc.animationRippleEnabled = 1
    fc.animationRippleFrequency, fc.animationRippleAmplitude, fc.animationRippleSpeed = ripple
    if animation_time is not None:
        fc.doAnimation = 1
        fc.animationState = animation_time
    return fc


# -------------------------------------------------------------- instances


def rigid_transform(yaw: float, translation, pitch: float = 0.0, roll: float = 0.0) -> np.ndarray:
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    cr, sr = math.cos(roll), math.sin(roll)
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    Rz = np.array([[cr, -sr, 0], [sr, cr, 0], [0, 0, 1]])
    M = np.eye(4)
    M[:3, :3] = Ry @ Rx @ Rz
    M[:3, 3] = translation
    return M


@dataclass
class Instance:
    geometry: int
    world: np.ndarray            # 4x4 row-major
    displacement_index: int = -1
    displacement_scale: float = 1.0
    displacement_offset: float = 0.0


@dataclass
class HostScene:
    geometries: list
    instances: list
    textures: list               # (array, abi.TEXTURE_*) pairs or FbmTexture descriptors
    camera: Camera
    tess_rate_pixels: float
    name: str = ""

    @property
    def num_base_triangles(self):
        return sum(self.geometries[i.geometry].num_triangles for i in self.instances)

    @property
    def num_clusters(self):
        return sum(self.geometries[i.geometry].num_clusters for i in self.instances)

    def frame(self, **kw):
        return frame_constants(self.camera, self.tess_rate_pixels, **kw)


# ------------------------------------------------------------ workloads

# the reference renderer's default flags (renderer.hpp:43-48): displacement,
# culling, PN base surface and both transient build forms
REFERENCE_DEFAULT_FLAGS = (abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_PN | abi.FLAG_1X_TRANSIENT |
                           abi.FLAG_2X_TRANSIENT)


def workload_config(large: bool = False, flags: int = None) -> abi.Config:
    """Capacities of the benchmark frames: 1920 x 1080 workloads, or
    (large) the 3840 x 2160 frame of about 70M displaced triangles."""
    if large:
        c = abi.default_config(maxVisibleClusters=1 << 18, maxPartTriangles=1 << 23, maxSplitTriangles=1 << 20,
                               maxGeneratedClusters=1 << 23, maxGeneratedClusterMegs=12288,
                               maxGeneratedVertices=1 << 27)
    else:
        c = abi.default_config(maxVisibleClusters=1 << 20, maxPartTriangles=1 << 22, maxSplitTriangles=1 << 16,
                               maxGeneratedClusters=1 << 21, maxGeneratedClusterMegs=4096,
                               maxGeneratedVertices=1 << 25)
    if flags is not None:
        c.flags = flags
    return c



def scene_icosphere(segments=8, tex_size=512, amplitude=0.05, seed=0, octaves=4, width=1280, height=720,
                    tess_rate_pixels=8.0, distance=3.0):
    """Icosphere of 20*segments^2 triangles (segments=8: 1280), radius 1,
    displaced by 4-octave value-noise fBm along the normal, camera at
    distance 3 with a 60 degree field of view."""
    geo = icosphere(segments)
    h = fbm_height(tex_size, seed, octaves)
    inst = Instance(0, np.eye(4), displacement_index=0, displacement_scale=amplitude)
    cam = Camera(eye=(0.0, 0.0, distance), center=(0.0, 0.0, 0.0), width=width, height=height, near=0.01, far=100.0)
    return HostScene([geo], [inst], [(h, abi.TEXTURE_R32_FLOAT)], cam, tess_rate_pixels, "icosphere")


def scene_grid_instances(num_instances=64, quads=256, patch_size=16.0, tex_size=2048, seed=1, octaves=6,
                         width=1920, height=1080, tess_rate_pixels=4.0, disp_scale=0.6, cam_height=2.5):
    """num_instances grid patches of quads x quads quads laid out on a
    square lattice with seeded yaw / jitter, one R16 fBm height map, camera
    above the field looking across it (a share of the instances falls
    outside the frustum)."""
    geo = grid_patch(quads, quads, patch_size, patch_size)
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(num_instances)))
    insts = []
    for i in range(num_instances):
        gx, gz = i % side, i // side
        yaw = float(rng.uniform(0, 2 * math.pi))
        jitter = rng.uniform(-0.1, 0.1, 3) * patch_size
        c = np.array([gx * patch_size * 1.05, 0.0, gz * patch_size * 1.05]) + jitter * np.array([1, 0.05, 1])
        # rotate about the patch centre
        Mc = rigid_transform(0.0, [-patch_size / 2, 0, -patch_size / 2])
        M = rigid_transform(yaw, c) @ Mc
        insts.append(Instance(0, M, displacement_index=0, displacement_scale=disp_scale, displacement_offset=0.0))
    h = to_r16(fbm_height(tex_size, seed, octaves))
    span = side * patch_size * 1.05
    cam = Camera(eye=(span * 0.5, cam_height, -patch_size * 0.3), center=(span * 0.5, 0.0, span * 0.45),
                 width=width, height=height, near=0.05, far=4 * span)
    return HostScene([geo], insts, [(h, abi.TEXTURE_R16_UNORM)], cam, tess_rate_pixels, "grid_instances")


# ---------------------------------------------------- batched 500-mesh frame


@dataclass
class FbmTexture:
    """A height map generated on the device by tc_fbm_height (numpy's
    fbm_height gives the same texels on the host)."""
    size: int
    seed: int
    octaves: int
    format: int

    def host(self) -> np.ndarray:
        h = fbm_height(self.size, self.seed, self.octaves)
        return to_r16(h) if self.format == abi.TEXTURE_R16_UNORM else h

    @property
    def nbytes(self):
        return self.size * self.size * (2 if self.format == abi.TEXTURE_R16_UNORM else 4)


def _grid_sizes():
    out, n = [], 10.0
    while n < 330:
        out.append(int(round(n)))
        n *= 1.2
    return out


def scene_batched(num_meshes=500, seed=2, tex_size=1024, width=1920, height=1080, tess_rate_pixels=4.0,
                  field=40.0):
    """num_meshes meshes with triangle counts log-uniform over 200..200000:
    grids (nearest of a 1.2-ratio ladder of quad resolutions) and
    icospheres (20 * (8q)^2 triangles), random rigid transforms with a
    uniform scale, each with its own tex_size^2 R16 fBm height map."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(200), math.log(200000), num_meshes))
    kinds = rng.random(num_meshes)
    ladder = _grid_sizes()
    geos, key_of, insts, texs = [], {}, [], []
    for i in range(num_meshes):
        T = counts[i]
        if T >= 1280 and kinds[i] < 0.5:
            key = ("ico", max(1, int(round(math.sqrt(T / 1280.0)))))
        else:
            n = math.sqrt(T / 2.0)
            key = ("grid", min(ladder, key=lambda q: abs(q - n)))
        if key not in key_of:
            key_of[key] = len(geos)
            geos.append(icosphere(8 * key[1], radius=0.5) if key[0] == "ico" else grid_patch(key[1], key[1], 1.0, 1.0))
        g = key_of[key]
        s = float(rng.uniform(1.0, 3.0))
        pos = [float(rng.uniform(-field / 2, field / 2)), float(rng.uniform(0.0, 4.0)), float(rng.uniform(0, field))]
        M = rigid_transform(float(rng.uniform(0, 2 * math.pi)), pos, float(rng.uniform(-0.5, 0.5)),
                            float(rng.uniform(-0.5, 0.5)))
        M[:3, :3] *= s
        insts.append(Instance(g, M, displacement_index=i, displacement_scale=0.05 * s))
        texs.append(FbmTexture(tex_size, 1000 * seed + i, 6, abi.TEXTURE_R16_UNORM))
    cam = Camera(eye=(0.0, 8.0, -12.0), center=(0.0, 1.0, field * 0.4), width=width, height=height, near=0.05,
                 far=10 * field)
    return HostScene(geos, insts, texs, cam, tess_rate_pixels, "batched")


# ------------------------------------------------------- large 4K frame


def scene_large(patches=8, quads_x=512, quads_z=256, quad_size=0.25, tex_size=4096, seed=5, octaves=6,
                width=3840, height=2160, tess_rate_pixels=2.0, disp_scale=4.0):
    """patches x patches instances of a quads_x x quads_z-quad grid patch
    (512 x 256 quads: 262,144 triangles each, 16.8M base triangles for the
    8 x 8 field) on a square lattice, one tex_size^2 R16 fBm height map,
    3840 x 2160 at 2 px per segment.  The camera follows camera_spline()."""
    sx, sz = quads_x * quad_size, quads_z * quad_size
    geo = grid_patch(quads_x, quads_z, sx, sz)
    insts = []
    for j in range(patches):
        for i in range(patches):
            M = rigid_transform(0.0, [i * sx, 0.0, j * sz])
            insts.append(Instance(0, M, displacement_index=0, displacement_scale=disp_scale))
    tex = FbmTexture(tex_size, seed, octaves, abi.TEXTURE_R16_UNORM)
    span_x, span_z = patches * sx, patches * sz
    eye, center = camera_spline(0, 100, span_x, span_z, seed)
    cam = Camera(eye=eye, center=center, width=width, height=height, near=0.05, far=4 * max(span_x, span_z))
    host = HostScene([geo], insts, [tex], cam, tess_rate_pixels, "large")
    host.spline = (span_x, span_z, seed)
    return host


def camera_spline(frame: int, frames: int, span_x: float, span_z: float, seed: int = 5):
    """Eye and look-at point of frame `frame` of `frames` on a closed
    Catmull-Rom spline through 6 seeded control points above the field
    (heights 3..8, looking 12 units ahead along the path, slightly down)."""
    rng = np.random.default_rng(seed)
    ctrl = np.stack([rng.uniform(0.15, 0.85, 6) * span_x, rng.uniform(3.0, 8.0, 6), rng.uniform(0.15, 0.85, 6) * span_z],
                    1)

    def at(u):
        u = u % 1.0
        k = u * len(ctrl)
        i = int(math.floor(k))
        t = k - i
        p0, p1, p2, p3 = (ctrl[(i + d) % len(ctrl)] for d in (-1, 0, 1, 2))
        return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t +
                      (-p0 + 3 * p1 - 3 * p2 + p3) * t * t * t)

    u = frame / float(frames)
    eye = at(u)
    ahead = at(u + 0.02)
    d = ahead - eye
    d[1] = 0.0
    n = np.linalg.norm(d)
    d = d / n if n > 0 else np.array([0.0, 0.0, 1.0])
    center = eye + 12.0 * d
    center[1] = 0.0
    return tuple(float(x) for x in eye), tuple(float(x) for x in center)
