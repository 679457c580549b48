"""HiZ occlusion culling in the oracle (nvhiz_vk.cpp:278-320,
nvhiz-update.comp.glsl, culling.glsl:96-114, instances_classify.comp.glsl:
120).  Pinned by the reference's layout formulas, a numpy max-pool of the
depth image, the conservativeness every HiZ test must have (a rectangle is
reported occluded only if every depth pixel it covers is nearer than its
nearest depth), and a closed form for a constant depth image."""
import ctypes as C

import numpy as np
import pytest

import harness
from harness import oracle
from vk_tessellated_clusters_amd import abi, scene


def setup(w, h):
    hz, nb, f = abi.Hiz(), C.c_uint64(), (C.c_float * 4)()
    smax = C.c_float()
    abi.check(abi.lib().tc_hiz_setup(w, h, C.byref(hz), C.byref(nb), f, C.byref(smax)), "tc_hiz_setup")
    return hz, nb.value, list(f), smax.value


@pytest.mark.parametrize("w,h,size,mips", [(1920, 1080, 1024, 11), (3840, 2160, 2048, 12), (9, 8, 4, 3), (64, 48, 32, 6)])
def test_setup_matches_reference_formulas(w, h, size, mips):
    hz, nbytes, f, smax = setup(w, h)
    assert (hz.size, hz.mipLevels, hz.usedWidth, hz.usedHeight) == (size, mips, w // 2, h // 2)
    assert nbytes == 4 * sum((size >> k) ** 2 for k in range(mips))
    want = [np.float32(w // 2) / np.float32(size), np.float32(h // 2) / np.float32(size),
            np.float32(w // 2 - 2) / np.float32(size), np.float32(h // 2 - 2) / np.float32(size)]
    assert f == [float(x) for x in want] and smax == float(size)
    assert abi.lib().tc_hiz_setup(5, 3, C.byref(abi.Hiz()), None, None, None) == abi.TC_ERROR_INVALID_ARGUMENT


def numpy_pyramid(depth, size, mips):
    h, w = depth.shape
    ys = np.minimum(np.arange(2 * size), h - 1)
    xs = np.minimum(np.arange(2 * size), w - 1)
    img = depth[np.ix_(ys, xs)]
    levels = [img.reshape(size, 2, size, 2).max(axis=(1, 3))]
    for k in range(1, mips):
        d = size >> k
        levels.append(levels[-1].reshape(d, 2, d, 2).max(axis=(1, 3)))
    return np.concatenate([l.reshape(-1) for l in levels])


@pytest.mark.parametrize("w,h", [(37, 23), (64, 48), (7, 6)])
def test_pyramid_is_max_pool(w, h):
    rng = np.random.default_rng(w * h)
    depth = rng.random((h, w), np.float32)
    hz, nbytes, _, _ = setup(w, h)
    tex = np.zeros(nbytes // 4, np.float32)
    hz.texels = tex.ctypes.data
    # strided source: depth inside 48-byte records (tc_hit_decoded)
    rec = np.zeros((h * w, 12), np.float32)
    rec[:, 7] = depth.reshape(-1)
    oracle().oracle_hiz_build(C.byref(hz), rec.ctypes.data + 28, 48)
    assert np.array_equal(tex, numpy_pyramid(depth, hz.size, hz.mipLevels))


def hiz_view(w, h):
    fc = abi.FrameConstants()
    hz, nbytes, f, smax = setup(w, h)
    fc.hizSizeFactors = (abi.f32 * 4)(*f)
    fc.hizSizeMax = smax
    fc.viewportf = (abi.f32 * 2)(w, h)
    return fc, hz, nbytes


def test_hiz_test_is_conservative():
    """Occluded only if every depth pixel under the rectangle is nearer than
    the rectangle's nearest depth; and not trivially always visible."""
    w, h = 160, 96
    rng = np.random.default_rng(1)
    # blocky depth: a few near walls over a far background
    depth = np.full((h, w), 0.9, np.float32)
    for _ in range(12):
        x0, y0 = rng.integers(0, w - 8), rng.integers(0, h - 8)
        depth[y0:y0 + rng.integers(4, 60), x0:x0 + rng.integers(4, 80)] = rng.uniform(0.1, 0.6)
    fc, hz, nbytes = hiz_view(w, h)
    tex = np.zeros(nbytes // 4, np.float32)
    hz.texels = tex.ctypes.data
    oracle().oracle_hiz_build(C.byref(hz), depth.ctypes.data, 4)
    occluded = 0
    for _ in range(3000):
        c = np.sort(rng.uniform(-1, 1, (2, 2)), axis=0).astype(np.float32)
        z = np.float32(rng.uniform(0.05, 0.95))
        cmin = np.array([c[0, 0], c[0, 1], z, 1], np.float32)
        cmax = np.array([c[1, 0], c[1, 1], z, 1], np.float32)
        vis = oracle().oracle_intersect_hiz(C.byref(hz), C.addressof(fc), cmin.ctypes.data, cmax.ctypes.data)
        if vis:
            continue
        occluded += 1
        px0 = int(np.floor((c[0, 0] * 0.5 + 0.5) * w)) - 1
        px1 = int(np.ceil((c[1, 0] * 0.5 + 0.5) * w)) + 1
        py0 = int(np.floor((c[0, 1] * 0.5 + 0.5) * h)) - 1
        py1 = int(np.ceil((c[1, 1] * 0.5 + 0.5) * h)) + 1
        region = depth[max(py0, 0):min(py1, h), max(px0, 0):min(px1, w)]
        assert region.max() + 2.0 / (1 << 24) < z, (c, z, region.max())
    assert occluded > 100


def test_constant_depth_closed_form():
    """Depth 0.25 everywhere: any rectangle is occluded iff its nearest
    depth exceeds 0.25 + 2^-23."""
    w, h = 64, 48
    fc, hz, nbytes = hiz_view(w, h)
    tex = np.zeros(nbytes // 4, np.float32)
    hz.texels = tex.ctypes.data
    depth = np.full((h, w), 0.25, np.float32)
    oracle().oracle_hiz_build(C.byref(hz), depth.ctypes.data, 4)
    nudge = np.float32(2.0 / (1 << 24))
    for z in (0.2, 0.25, np.float32(0.25) + nudge, 0.2500001, 0.26, 0.9):
        for rect in ((-1, -1, 1, 1), (-0.1, -0.1, 0.1, 0.1), (0.3, -0.9, 0.31, -0.89)):
            cmin = np.array([rect[0], rect[1], z, 1], np.float32)
            cmax = np.array([rect[2], rect[3], z, 1], np.float32)
            vis = oracle().oracle_intersect_hiz(C.byref(hz), C.addressof(fc), cmin.ctypes.data, cmax.ctypes.data)
            assert bool(vis) == (np.float32(z) <= np.float32(0.25) + nudge), (z, rect)


def test_instances_behind_a_wall_are_culled():
    """A depth image of a wall at depth 0.3: instances in the frustum whose
    nearest clip depth lies beyond it lose the visible bit (closed form in
    FP64, instances within 1e-4 of the threshold skipped); frustum bits are
    unchanged, and without TC_FLAG_HIZ nothing is occluded."""
    host = scene.scene_grid_instances(num_instances=16, quads=16, patch_size=4.0, tex_size=64, cam_height=1.0)
    fc = host.frame()
    W, H = host.camera.width, host.camera.height
    _, _, f, smax = setup(W, H)
    fc.hizSizeFactors = (abi.f32 * 4)(*f)
    fc.hizSizeMax = smax
    states = {}
    for flags in (abi.FLAG_CULLING, abi.FLAG_CULLING | abi.FLAG_HIZ):
        of = harness.OracleFrame(host, harness.small_config(flags=flags))
        of.set_culling_history(fc, np.full((H, W), 0.3, np.float32))
        of.run(fc, stages="tess")
        states[flags] = of.results().arrays["instanceStates"].copy()
    plain, hiz = states[abi.FLAG_CULLING], states[abi.FLAG_CULLING | abi.FLAG_HIZ]
    assert np.array_equal(plain & abi.INSTANCE_FRUSTUM_BIT, hiz & abi.INSTANCE_FRUSTUM_BIT)
    vp = np.array(fc.viewProjMatrix, np.float64).reshape(4, 4).T
    culled = checked = 0
    for i, inst in enumerate(host.instances):
        if not plain[i] & abi.INSTANCE_VISIBLE_BIT:
            assert not hiz[i] & abi.INSTANCE_VISIBLE_BIT
            continue
        g = host.geometries[inst.geometry]
        corners = np.array([[x, y, z, 1.0] for z in (g.lo[2], g.hi[2]) for y in (g.lo[1], g.hi[1])
                            for x in (g.lo[0], g.hi[0])])
        hp = corners @ (vp @ np.asarray(inst.world, np.float64)).T
        zmin = (hp[:, 2] / np.abs(hp[:, 3])).min()
        if abs(zmin - 0.3) < 1e-4:
            continue
        checked += 1
        expect = zmin <= 0.3
        assert bool(hiz[i] & abi.INSTANCE_VISIBLE_BIT) == expect, (i, zmin)
        culled += not expect
    assert culled > 0 and checked > culled
