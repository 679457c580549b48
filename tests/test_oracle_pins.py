"""Pins for oracle functions that the stage-level tests do not fix by
themselves: the PN base surface, BLAS setup and insertion, primary-ray
generation, the host instance transforms the caster uses, and the CLAS
blocks the caster reads.  Each check is chosen so that a plausible
transcription mistake (a dropped term, a swapped corner or operand, a
wrong sign, an off-by-one) fails it.

References: shaders/displacement.glsl:47-104 (PN),
blas_setup_insertion.comp.glsl:100-116, blas_clusters_insert.comp.glsl:97-135,
render_raytrace.rgen.glsl:52-86, tessellatedclusters.cpp:569-581 (camera).
"""
import ctypes as C
import math

import numpy as np
import pytest

import harness
from harness import oracle
from vk_tessellated_clusters_amd import abi, scene

# ---------------------------------------------------------------- PN


def pn(verts, normals, bary):
    v = np.ascontiguousarray(verts, np.float32).reshape(9)
    n = np.ascontiguousarray(normals, np.float32).reshape(9)
    b = np.ascontiguousarray(bary, np.float32)
    out = np.zeros(3, np.float32)
    oracle().oracle_pn(v.ctypes.data, n.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def pn_fp64(P, N, bary):
    """Curved PN triangle of Vlachos, Peters, Boyd and Mitchell (I3D 2001)
    written from the paper in FP64, with the reference's corner assignment:
    P[0] at u = 1, P[1] at v = 1, P[2] at w = 1, and
    b(u, v, w) = sum over i + j + k = 3 of 3!/(i! j! k!) b_ijk w^i u^j v^k,
    where b_ijk has i/3 of its weight at P[2], j/3 at P[0], k/3 at P[1]."""
    P = np.asarray(P, np.float64)
    N = np.asarray(N, np.float64)
    u, v, w = (float(x) for x in bary)
    corner = {(0, 3, 0): 0, (0, 0, 3): 1, (3, 0, 0): 2}

    def edge_point(a, b):
        # control point 2/3 of the way from corner a toward b, projected
        # into the tangent plane of a: (2 Pa + Pb - ((Pb - Pa).Na) Na) / 3
        wab = (P[b] - P[a]) @ N[a]
        return (2 * P[a] + P[b] - wab * N[a]) / 3.0

    B = {}
    for ijk, c in corner.items():
        B[ijk] = P[c]
    # index order (i, j, k) = (w, u, v) exponents
    B[(0, 2, 1)] = edge_point(0, 1)
    B[(0, 1, 2)] = edge_point(1, 0)
    B[(1, 0, 2)] = edge_point(1, 2)
    B[(2, 0, 1)] = edge_point(2, 1)
    B[(2, 1, 0)] = edge_point(2, 0)
    B[(1, 2, 0)] = edge_point(0, 2)
    E = sum(B[k] for k in ((0, 2, 1), (0, 1, 2), (1, 0, 2), (2, 0, 1), (2, 1, 0), (1, 2, 0))) / 6.0
    V = (P[0] + P[1] + P[2]) / 3.0
    B[(1, 1, 1)] = E + (E - V) / 2.0
    out = np.zeros(3)
    for (i, j, k), b in B.items():
        coef = math.factorial(3) / (math.factorial(i) * math.factorial(j) * math.factorial(k))
        out += coef * b * (w ** i) * (u ** j) * (v ** k)
    return out


def random_triangle(rng):
    P = rng.normal(size=(3, 3))
    N = rng.normal(size=(3, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    return P.astype(np.float32), N.astype(np.float32)


def random_bary(rng):
    b = rng.random(3)
    return (b / b.sum()).astype(np.float32)


def test_pn_matches_fp64_restatement_of_the_paper():
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(400):
        P, N = random_triangle(rng)
        b = random_bary(rng)
        got = pn(P, N, b).astype(np.float64)
        exp = pn_fp64(P.astype(np.float64), N.astype(np.float64), b.astype(np.float64))
        scale = np.abs(P).max() + 1.0
        worst = max(worst, float(np.abs(got - exp).max() / scale))
    assert worst < 2e-6, worst


def test_pn_reproduces_corners_exactly():
    rng = np.random.default_rng(2)
    for _ in range(100):
        P, N = random_triangle(rng)
        for c in range(3):
            b = np.zeros(3, np.float32)
            b[c] = 1.0
            assert np.array_equal(pn(P, N, b), P[c])


def test_pn_with_face_normals_is_the_flat_triangle():
    """Normals equal to the face normal put every control point in the
    triangle's plane on the linear lattice: the patch is the triangle and
    the parametrisation is linear interpolation."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        P, _ = random_triangle(rng)
        fn = np.cross(P[1] - P[0], P[2] - P[0]).astype(np.float64)
        fn /= np.linalg.norm(fn)
        N = np.tile(fn.astype(np.float32), (3, 1))
        b = random_bary(rng)
        exp = b.astype(np.float64) @ P.astype(np.float64)
        got = pn(P, N, b).astype(np.float64)
        assert np.abs(got - exp).max() <= 4e-6 * (np.abs(P).max() + 1.0)


def test_pn_is_invariant_under_corner_rotation():
    """Rotating the corners and normals together with the barycentrics
    describes the same surface point."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        P, N = random_triangle(rng)
        b = random_bary(rng)
        a = pn(P, N, b).astype(np.float64)
        r = pn(np.roll(P, 1, 0), np.roll(N, 1, 0), np.roll(b, 1)).astype(np.float64)
        assert np.abs(a - r).max() <= 4e-6 * (np.abs(P).max() + 1.0)


def test_pn_bulges_along_the_normals():
    """A flat triangle whose corner normals tilt outward bulges toward the
    normals' side at the centroid (a sign error in the projection or in
    b111's offset flips it)."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    N = np.array([[-1, 1, -1], [1, 1, -0.5], [-0.5, 1, 1]], np.float32)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    c = pn(P, N, np.full(3, 1 / 3, np.float32))
    assert c[1] > 0.02
    exp = pn_fp64(P, N, np.full(3, 1 / 3))
    assert np.abs(c - exp).max() < 1e-6


# ------------------------------------------------ BLAS setup / insertion


def frame_results(host, cfg):
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    return of, of.results()


@pytest.mark.parametrize("flags", [abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING,
                                   abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT |
                                   abi.FLAG_2X_TRANSIENT])
def test_blas_lists_are_each_instances_clas_in_order(flags):
    """Each instance's reference list is exactly the CLAS carrying its
    instance ID, template CLAS first in CLAS order then transient CLAS in
    CLAS order; list offsets are the exclusive prefix sum of the counts
    (blas_setup_insertion.comp.glsl:100-116, blas_clusters_insert.comp.glsl:
    97-135, made deterministic)."""
    host = scene.scene_grid_instances(num_instances=6, quads=48, patch_size=6.0, tex_size=128, tess_rate_pixels=8.0)
    of, res = frame_results(host, harness.small_config(flags=flags))
    n = len(host.instances)
    tid = res.arrays["tempInstanceIDs"]
    toff = res.arrays["tempClusterOffsets"]
    if flags & abi.TRANSIENT_FLAGS:
        rid = res.arrays["transInstanceIDs"]
        roff = res.arrays["transClusterOffsets"]
        assert rid.size > 0
    else:
        rid = np.zeros(0, np.uint32)
        roff = np.zeros(0, np.uint64)
    infos = res.arrays["blasBuildInfos"]
    counts = infos["clusterReferencesCount"].astype(np.int64)
    expected_counts = np.bincount(tid, minlength=n) + np.bincount(rid, minlength=n)
    assert np.array_equal(counts, expected_counts)
    starts = infos["clusterReferences"].astype(np.int64) // 8
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    refs = res.arrays["blasClusterOffsets"]
    assert refs.size == counts.sum() == res.counters["blasClusterCounter"]
    for i in range(n):
        want = np.concatenate([toff[tid == i], roff[rid == i]])
        assert np.array_equal(refs[starts[i]:starts[i] + counts[i]], want), i


# ------------------------------------------------------ primary rays


def primary_ray(fc, x, y, W, H):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    oracle().oracle_primary_ray(C.addressof(fc), x, y, W, H, o.ctypes.data, d.ctypes.data)
    return o, d


def pinhole(cam, x, y):
    """FP64 pinhole camera from its parameters alone: pixel centres on the
    image plane at distance 1 along the view direction, x to the right,
    y downward (the reference flips Y in its projection)."""
    eye, center, up = (np.asarray(v, np.float64) for v in (cam.eye, cam.center, cam.up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    t = math.tan(math.radians(cam.fovy_deg) / 2)
    aspect = cam.width / cam.height
    nx = 2.0 * (x + 0.5) / cam.width - 1.0
    ny = 2.0 * (y + 0.5) / cam.height - 1.0
    d = f + nx * t * aspect * s - ny * t * u
    return eye, d / np.linalg.norm(d)


@pytest.mark.parametrize("W,H,fov", [(101, 51, 60.0), (1920, 1080, 60.0), (64, 200, 35.0)])
def test_primary_rays_are_a_pinhole_camera(W, H, fov):
    cam = scene.Camera(eye=(3.0, 2.0, -4.0), center=(0.5, -0.25, 1.0), width=W, height=H, fovy_deg=fov, near=0.05,
                       far=500.0)
    fc = scene.frame_constants(cam, 4.0)
    rng = np.random.default_rng(W)
    pix = [(0, 0), (W - 1, 0), (0, H - 1), (W - 1, H - 1)] + [tuple(p) for p in
                                                             rng.integers(0, (W, H), size=(64, 2))]
    for x, y in pix:
        o, d = primary_ray(fc, int(x), int(y), W, H)
        eo, ed = pinhole(cam, x, y)
        assert np.abs(o - eo).max() < 1e-5
        assert np.abs(d - ed).max() < 2e-6, (x, y, d, ed)
        assert abs(float(np.linalg.norm(d.astype(np.float64))) - 1.0) < 1e-6
    if W % 2 and H % 2:
        # the centre pixel looks along the view direction
        o, d = primary_ray(fc, W // 2, H // 2, W, H)
        f = np.asarray(cam.center) - np.asarray(cam.eye)
        assert np.abs(d - f / np.linalg.norm(f)).max() < 1e-6


# ------------------------------------------------ instance transforms


@pytest.mark.parametrize("make", [scene.scene_grid_instances, scene.scene_batched])
def test_instance_transforms_invert(make):
    """The caster's per-instance transforms: objectToWorld is the
    instance's world matrix as stored, and objectToWorld x worldToObject is
    the identity in FP64 to single-precision rounding."""
    host = make()
    n = len(host.instances)
    recs = (abi.RenderInstance * n)()
    for i, inst in enumerate(host.instances):
        recs[i].worldMatrix = (abi.f32 * 16)(*np.asarray(inst.world, np.float32).T.reshape(-1).tolist())
    xf = np.zeros(24 * n, np.float32)
    abi.check(abi.lib().tc_instance_xforms_host(C.addressof(recs), n, xf.ctypes.data))
    xf = xf.reshape(n, 2, 3, 4).astype(np.float64)
    for i, inst in enumerate(host.instances):
        W = np.asarray(inst.world, np.float32).astype(np.float64)
        assert np.array_equal(xf[i, 0], W[:3, :4])
        F = np.vstack([xf[i, 0], [0, 0, 0, 1]])
        I = np.vstack([xf[i, 1], [0, 0, 0, 1]])
        scale = np.abs(F).max()
        assert np.abs(F @ I - np.eye(4)).max() < 4e-7 * scale * max(1.0, np.abs(I).max())


# ---------------------------------------------------------- CLAS blocks


def test_clas_blocks_bound_their_triangles():
    """Every CLAS block of a frame (template-instantiated and transient):
    triangle words index inside the block's vertices, each group box bounds
    its (up to four) triangles' vertices exactly, the header box is the
    union of the group boxes, and the primitive IDs are a permutation of the
    cluster's or pattern's triangle indices (templates regroup triangles
    spatially, four per group box)."""
    host = scene.scene_grid_instances(num_instances=4, quads=48, patch_size=6.0, tex_size=128, tess_rate_pixels=8.0)
    flags = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT | abi.FLAG_2X_TRANSIENT
    of, res = frame_results(host, harness.small_config(flags=flags))
    pool = of.buf["genClusterData"]
    offs = np.concatenate([res.arrays["tempClusterOffsets"], res.arrays["transClusterOffsets"]]).astype(np.int64)
    checked = 0
    for off in offs:
        h = pool[off:off + 48].view(abi.CLAS_HEADER_DT)[0]
        nt, nv, ng = int(h["counts"]) & 0xFF, (int(h["counts"]) >> 8) & 0xFF, (int(h["counts"]) >> 16) & 0xFF
        assert ng == (nt + 3) // 4 and nt > 0 and nv > 0
        g0 = off + 48
        t0 = g0 + ((ng * 24 + 15) & ~15)
        v0 = t0 + ((nt * 4 + 15) & ~15)
        groups = pool[g0:g0 + 24 * ng].view(np.float32).reshape(ng, 6)
        words = pool[t0:t0 + 4 * nt].view(np.uint32)
        verts = pool[v0:v0 + 12 * nv].view(np.float32).reshape(nv, 3)
        idx = np.stack([words & 0xFF, (words >> 8) & 0xFF, (words >> 16) & 0xFF], 1)
        assert idx.max() < nv
        assert np.array_equal(np.sort(words >> 24), np.arange(nt))
        for g in range(ng):
            p = verts[idx[4 * g:4 * g + 4].reshape(-1)]
            assert np.array_equal(groups[g, :3], p.min(0)) and np.array_equal(groups[g, 3:], p.max(0))
        assert np.array_equal(h["lo"], groups[:, :3].min(0)) and np.array_equal(h["hi"], groups[:, 3:].max(0))
        checked += 1
    assert checked > 100
