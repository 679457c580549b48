"""Pins the FP64 depth-adjoint checker: closed forms of the frozen forward
map and agreement of its analytic adjoint with central differences of the
L2 depth loss (two independent evaluations of the same derivative)."""
import numpy as np
import pytest

from adjoint_oracle import Problem, identity_xforms
from vk_tessellated_clusters_amd import abi


def _tape(n):
    return np.zeros(n, abi.VERTEX_TAPE_DT)


def test_flat_plane_closed_form():
    # one triangle in the plane y = 0, displaced along +Y by s * h0 + c with a
    # constant height map: a ray from y = 5 with any downward direction meets
    # the plane y = s * h0 + c at t = (5 - y) / -d_y
    tape = _tape(3)
    tape["base"] = [[0, 0, 0], [0, 0, 1], [1, 0, 0]]
    tape["normal"] = [0, 1, 0]
    tape["u"], tape["v"] = [0.1, 0.2, 0.7], [0.3, 0.9, 0.5]
    tape["heightScale"], tape["heightOffset"], tape["texture"] = 0.25, 0.1, 0
    h0 = 0.6
    hmap = np.full((4, 4), h0, np.float32)
    rays = np.zeros(2, abi.RAY_DT)
    rays["origin"] = [[0.2, 5.0, 0.3], [0.4, 5.0, 0.2]]
    rays["direction"] = [[0, -1, 0], [0.1, -2.0, 0.05]]
    hits = np.zeros(2, abi.HIT_DT)
    hits["t"] = 1.0
    hits["instanceID"] = 0
    hv = np.tile(np.array([0, 1, 2, 0], np.uint32), 2)
    p = Problem(rays, hits, hv, tape, [hmap], identity_xforms())
    y = 0.25 * float(np.float32(h0)) + float(np.float32(0.1))
    t = p.t64()
    np.testing.assert_allclose(t, [(5.0 - y) / 1.0, (5.0 - y) / 2.0], rtol=1e-12)
    # a miss stays -1
    hits["t"][1] = -1.0
    assert Problem(rays, hits, hv, tape, [hmap], identity_xforms()).t64()[1] == -1.0


def test_instance_transform_closed_form():
    # object space plane y = c; the instance maps object to world by scaling
    # y by 2 and translating by (0, 1, 0): the world plane is y = 1 + 2c
    tape = _tape(3)
    tape["base"] = [[-5, 0, -5], [-5, 0, 5], [5, 0, 0]]
    tape["normal"] = [0, 1, 0]
    tape["heightScale"], tape["heightOffset"], tape["texture"] = 1.0, 0.3, 0xFFFFFFFF
    o2w = np.array([[1, 0, 0, 0], [0, 2, 0, 1], [0, 0, 1, 0]], np.float64)
    w2o = np.array([[1, 0, 0, 0], [0, 0.5, 0, -0.5], [0, 0, 1, 0]], np.float64)
    xf = np.concatenate([o2w.reshape(-1), w2o.reshape(-1)]).astype(np.float32)
    rays = np.zeros(1, abi.RAY_DT)
    rays["origin"] = [0.1, 7.0, 0.2]
    rays["direction"] = [0, -1, 0]
    hits = np.zeros(1, abi.HIT_DT)
    hv = np.array([0, 1, 2, 0], np.uint32)
    # texture ~0: not displaced, the vertex is its base (y = 0)
    t = Problem(rays, hits, hv, tape, [None], xf).t64()
    np.testing.assert_allclose(t, [7.0 - 1.0], rtol=1e-12)
    tape["texture"] = 0
    hmap = np.zeros((2, 2), np.float32)
    t = Problem(rays, hits, hv, tape, [hmap], xf).t64()
    np.testing.assert_allclose(t, [7.0 - (1.0 + 2 * float(np.float32(0.3)))], rtol=1e-7)


def _random_problem(seed, ntri=24, size=8, tilt=True):
    rng = np.random.default_rng(seed)
    tape = _tape(3 * ntri)
    # triangles roughly facing the rays (no grazing incidence)
    base = rng.uniform(-1, 1, (ntri, 3, 3))
    base[:, :, 1] *= 0.1
    tape["base"] = base.reshape(-1, 3)
    nrm = rng.normal(size=(3 * ntri, 3)) * 0.3 + np.array([0, 1, 0])
    tape["normal"] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    tape["u"], tape["v"] = rng.uniform(0, 1, 3 * ntri), rng.uniform(0, 1, 3 * ntri)
    tape["heightScale"] = rng.uniform(0.05, 0.3, 3 * ntri)
    tape["heightOffset"] = rng.uniform(-0.1, 0.1, 3 * ntri)
    tape["texture"] = 0
    hmap = rng.uniform(0, 1, (size, size)).astype(np.float32)
    # an affine instance transform (rotation about y, anisotropic scale, shift)
    a = 0.7
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]) @ np.diag([1.3, 0.8, 1.1])
    T = np.array([0.2, -0.4, 0.5])
    o2w = np.concatenate([R, T[:, None]], 1)
    Ri = np.linalg.inv(R)
    w2o = np.concatenate([Ri, (-Ri @ T)[:, None]], 1)
    xf = np.concatenate([o2w.reshape(-1), w2o.reshape(-1)]).astype(np.float32)
    probe = Problem(np.zeros(1, abi.RAY_DT), np.zeros(1, abi.HIT_DT), np.zeros(4, np.uint32), tape, [hmap], xf)
    # rays aimed at random interior points of each triangle (world space)
    rays = np.zeros(4 * ntri, abi.RAY_DT)
    hv = np.zeros((4 * ntri, 4), np.uint32)
    hits = np.zeros(4 * ntri, abi.HIT_DT)
    for r in range(4 * ntri):
        tri = r % ntri
        hv[r, :3] = 3 * tri + np.arange(3)
        # displaced FP64 vertices via t64 machinery: reconstruct here
        P = []
        for k in range(3):
            t = tape[3 * tri + k]
            # bilinear of the FP32 sampler coordinates
            x = np.float32(np.float32(t["u"]) * np.float32(size) - np.float32(0.5))
            y = np.float32(np.float32(t["v"]) * np.float32(size) - np.float32(0.5))
            x0, y0 = np.floor(x), np.floor(y)
            fx, fy = float(x - x0), float(y - y0)
            xa, xb = int(x0) % size, (int(x0) + 1) % size
            ya, yb = int(y0) % size, (int(y0) + 1) % size
            h = ((1 - fx) * (1 - fy) * hmap[ya, xa] + fx * (1 - fy) * hmap[ya, xb] + (1 - fx) * fy * hmap[yb, xa] +
                 fx * fy * hmap[yb, xb])
            P.append(t["base"].astype(np.float64) + t["normal"].astype(np.float64) *
                     (h * float(t["heightScale"]) + float(t["heightOffset"])))
        b = rng.dirichlet([2, 2, 2])
        q = b[0] * P[0] + b[1] * P[1] + b[2] * P[2]
        qw = R @ q + T
        d = np.array([rng.uniform(-0.3, 0.3), -1.0, rng.uniform(-0.3, 0.3)]) if tilt else np.array([0, -1.0, 0])
        rays[r]["origin"] = qw - 3.0 * d
        rays[r]["direction"] = d
        hits[r]["t"] = 3.0
    return rays, hits, hv.reshape(-1), tape, hmap, xf, probe


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ray_targets_reproduced(seed):
    rays, hits, hv, tape, hmap, xf, _ = _random_problem(seed)
    t = Problem(rays, hits, hv, tape, [hmap], xf).t64()
    # rays were aimed 3 units before a point of each triangle; origins and
    # transforms are stored in FP32, which grazing triangles amplify
    np.testing.assert_allclose(t, 3.0, rtol=0, atol=3e-5)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_adjoint_matches_central_differences(seed):
    rays, hits, hv, tape, hmap, xf, _ = _random_problem(seed)
    p = Problem(rays, hits, hv, tape, [hmap], xf)
    rng = np.random.default_rng(100 + seed)
    target = (p.t64() + rng.normal(0, 0.05, p.n)).astype(np.float32)
    scale = 1.0 / p.n
    t = p.t64()
    grad_t = 2.0 * scale * (t - target.astype(np.float64))
    (g,) = p.adjoint64(grad_t)
    texels = np.arange(hmap.size, dtype=np.uint64)
    fd = p.loss_fd64(target, scale, 0, texels, 1e-4)
    assert np.abs(g).max() > 0
    np.testing.assert_allclose(g.reshape(-1), fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max())


def test_adjoint_detects_wrong_texel_weights():
    # a planted mistake in the checker's input: moving one vertex's uv
    # (the adjoint then scatters to other texels than the forward samples)
    # must break the agreement with differences of the original problem
    rays, hits, hv, tape, hmap, xf, _ = _random_problem(5)
    p = Problem(rays, hits, hv, tape, [hmap], xf)
    target = (p.t64() + 0.05).astype(np.float32)
    scale = 1.0 / p.n
    grad_t = 2.0 * scale * (p.t64() - target.astype(np.float64))
    fd = p.loss_fd64(target, scale, 0, np.arange(hmap.size, dtype=np.uint64), 1e-4)
    bad = tape.copy()
    bad["u"][0] = (bad["u"][0] + 0.37) % 1.0
    (g,) = Problem(rays, hits, hv, bad, [hmap], xf).adjoint64(grad_t)
    assert not np.allclose(g.reshape(-1), fd, rtol=1e-3, atol=1e-6 * np.abs(fd).max())
