"""Pins the oracle's ripple animation (shaders/displacement.glsl:106-120,
applied after displacement at cluster_classify.comp.glsl:483-485 and
triangle_tess_template_instantiate.comp.glsl:366-368).

The ripple's sin / cos are our own (FP64 range reduction, FP32 Taylor
series) so that the GPU and the oracle agree bit for bit; they are checked
against libm in FP64.  The deformation itself is checked through frames:
with the animation on, every generated vertex is the animation-off vertex
plus the ripple offset evaluated in FP64 from it (seed = instance index,
size = geoHi.w), which a swapped component, a wrong seed or a ripple
applied before displacement breaks.
"""
import math

import numpy as np
import pytest

import harness
from harness import oracle
from vk_tessellated_clusters_amd import abi, scene


def test_sincos_against_libm():
    O = oracle()
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-8.0, 8.0, 4000), rng.uniform(-1e4, 1e4, 4000),
                         np.arange(-64, 65) * (math.pi / 4)]).astype(np.float32)
    err_s = max(abs(O.oracle_ripple_sin(float(x)) - math.sin(float(x))) for x in xs)
    err_c = max(abs(O.oracle_ripple_cos(float(x)) - math.cos(float(x))) for x in xs)
    assert err_s < 1e-6 and err_c < 1e-6, (err_s, err_c)


def test_sincos_special_values_and_symmetry():
    O = oracle()
    assert O.oracle_ripple_sin(0.0) == 0.0
    assert O.oracle_ripple_cos(0.0) == 1.0
    for x in np.linspace(-50.0, 50.0, 997).astype(np.float32):
        x = float(x)
        assert O.oracle_ripple_sin(-x) == -O.oracle_ripple_sin(x)
        assert O.oracle_ripple_cos(-x) == O.oracle_ripple_cos(x)
        # sin^2 + cos^2 = 1 within a few FP32 rounding steps
        s, c = O.oracle_ripple_sin(x), O.oracle_ripple_cos(x)
        assert abs(s * s + c * c - 1.0) < 1e-6


def ripple_fp64(p, seed, size, freq, amp, speed, state):
    """FP64 evaluation of the ripple offset for points p (n, 3)."""
    p = p.astype(np.float64)
    m = np.abs(p).max(axis=1)
    f = freq / size
    ph = state * speed
    wave = np.stack([np.sin(m * f + seed + ph), np.cos(m * f * 3 + seed + ph), np.sin(m * f * 1.2 + seed + ph)], 1)
    zyx = p[:, ::-1]
    n = zyx / np.linalg.norm(zyx, axis=1, keepdims=True)
    return n * (amp * size) * wave


def vertex_owner(res, nv):
    """Instance index of every generated vertex: the vertex ranges of the
    template instantiations tile [0, nv) (without transient builds every
    generated vertex belongs to one CLAS), so sorted by offset each range
    ends where the next starts."""
    ti = res.arrays["tempInstantiations"]
    ids = res.arrays["tempInstanceIDs"]
    off = ti["vertexBufferAddress"].astype(np.int64) // 12
    order = np.argsort(off, kind="stable")
    starts = off[order]
    assert starts[0] == 0 and len(np.unique(starts)) == len(starts)
    ends = np.append(starts[1:], nv)
    owner = np.full(nv, -1, np.int64)
    for s0, e0, k in zip(starts, ends, order):
        owner[s0:e0] = int(ids[k])
    return owner


FREQ, AMP, SPEED = 100.0, 0.005, 3.7


@pytest.fixture(scope="module")
def frames():
    host = scene.scene_grid_instances(num_instances=4, quads=32, patch_size=4.0, tex_size=128, cam_height=1.2)
    base = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING)
    anim = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_ANIMATION)
    out = {}
    for name, cfg, t in (("off", base, 1.3), ("on", anim, 1.3), ("period", anim, 1.3 + 2 * math.pi / SPEED)):
        of = harness.OracleFrame(host, cfg)
        of.run(host.frame(animation_time=t, ripple=(FREQ, AMP, SPEED)))
        out[name] = of.results()
    out["host"] = host
    return out


def test_animation_keeps_tessellation(frames):
    off, on = frames["off"], frames["on"]
    assert off.counters == on.counters
    assert off.readback == on.readback
    assert np.array_equal(off.arrays["partTriangles"].view(np.uint8), on.arrays["partTriangles"].view(np.uint8))


def test_ripple_offset_fp64(frames):
    host, off, on = frames["host"], frames["off"], frames["on"]
    nv = off.counters["genVertexCounter"]
    p0 = off.arrays["genVertices"][:nv]
    p1 = on.arrays["genVertices"][:nv]
    owner = vertex_owner(off, nv)
    assert (owner >= 0).all()
    checked = 0
    for i, inst in enumerate(host.instances):
        geo = host.geometries[inst.geometry]
        size = float(np.float32(np.linalg.norm(geo.hi.astype(np.float64) - geo.lo)))
        sel = owner == i
        exp = p0[sel].astype(np.float64) + ripple_fp64(p0[sel], i, size, FREQ, AMP, SPEED, 1.3)
        # FP32 argument rounding (|arg| < 300) bounds the error by about
        # 1e-4 of the amplitude; a wrong component, seed or order is O(1)
        tol = 1e-3 * AMP * size + 4 * np.spacing(np.abs(p0[sel]).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(p1[sel] - exp) <= tol), np.abs(p1[sel] - exp).max()
        checked += sel.sum()
    assert checked == nv and nv > 1000
    # the deformation is visible at the amplitude's scale
    assert np.abs(p1 - p0).max() > 0.2 * AMP * 4.0


def test_ripple_is_periodic_in_time(frames):
    on, per = frames["on"], frames["period"]
    nv = on.counters["genVertexCounter"]
    d = np.abs(on.arrays["genVertices"][:nv] - per.arrays["genVertices"][:nv])
    assert d.max() < 1e-3 * AMP * 8.0


def test_zero_amplitude_is_identity():
    host = scene.scene_grid_instances(num_instances=2, quads=16, patch_size=2.0, tex_size=64, cam_height=0.8)
    outs = []
    for flags in (abi.FLAG_DISPLACEMENT, abi.FLAG_DISPLACEMENT | abi.FLAG_ANIMATION):
        of = harness.OracleFrame(host, harness.small_config(flags=flags))
        of.run(host.frame(animation_time=0.5, ripple=(FREQ, 0.0, SPEED)))
        outs.append(of.results())
    nv = outs[0].counters["genVertexCounter"]
    assert np.array_equal(outs[0].arrays["genVertices"][:nv], outs[1].arrays["genVertices"][:nv])
