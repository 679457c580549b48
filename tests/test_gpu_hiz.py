"""GPU HiZ occlusion culling (TC_FLAG_HIZ) through the C-ABI: the far HiZ
pyramid equals the oracle's bit for bit, and a renderer frame culled against
the previous frame's traced depth equals the oracle's frame with the same
depth and viewLast."""
import ctypes as C

import numpy as np
import pytest

import harness
from harness import oracle
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu


def hiz_setup(w, h):
    hz, nb, f, smax = abi.Hiz(), C.c_uint64(), (C.c_float * 4)(), C.c_float()
    abi.check(abi.lib().tc_hiz_setup(w, h, C.byref(hz), C.byref(nb), f, C.byref(smax)), "tc_hiz_setup")
    return hz, nb.value, list(f), smax.value


@pytest.mark.parametrize("w,h,stride", [(37, 23, 48), (160, 96, 4), (1920, 1080, 48)])
def test_hiz_build_matches_oracle(w, h, stride):
    import torch
    print("case", w, h, stride, flush=True)
    rng = np.random.default_rng(w)
    depth = rng.random((h, w), np.float32)
    rec = np.zeros((h * w, stride // 4), np.float32)
    col = 7 if stride == 48 else 0
    rec[:, col] = depth.reshape(-1)
    hz, nbytes, _, _ = hiz_setup(w, h)
    dev = torch.from_numpy(rec.view(np.uint8).reshape(-1).copy()).cuda()
    tex = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    hz.texels = tex.data_ptr()
    L = abi.lib()
    abi.check(L.tc_hiz_clear(C.byref(hz), None), "tc_hiz_clear")
    torch.cuda.synchronize()
    assert bool((tex == 1.0).all())
    abi.check(L.tc_hiz_build(C.byref(hz), C.c_void_p(dev.data_ptr() + 4 * col), stride, None), "tc_hiz_build")
    torch.cuda.synchronize()
    ref = np.zeros(nbytes // 4, np.float32)
    hr = abi.Hiz.from_buffer_copy(hz)
    hr.texels = ref.ctypes.data
    oracle().oracle_hiz_build(C.byref(hr), depth.ctypes.data, 4)
    assert np.array_equal(tex.cpu().numpy(), ref)
    # argument checks: misaligned depth, bad stride, no pyramid
    assert L.tc_hiz_build(C.byref(hz), C.c_void_p(dev.data_ptr() + 2), stride, None) == abi.TC_ERROR_INVALID_ARGUMENT
    assert L.tc_hiz_build(C.byref(hz), C.c_void_p(dev.data_ptr()), 6, None) == abi.TC_ERROR_INVALID_ARGUMENT
    assert L.tc_hiz_build(C.byref(abi.Hiz()), C.c_void_p(dev.data_ptr()), 4, None) == abi.TC_ERROR_INVALID_ARGUMENT


def wall_scene(width=320, height=180):
    """A vertical wall 8 units in front of the camera and grid patches
    behind it (hidden), plus one patch in front of it (seen)."""
    wall = scene.grid_patch(16, 16, 24.0, 24.0)
    patch = scene.grid_patch(16, 16, 4.0, 4.0)
    W = np.array([[1, 0, 0, -8.0], [0, 0, 1, -8.0], [0, -1, 0, 2.0], [0, 0, 0, 1]], np.float64)  # XZ plane -> XY at z=2
    # instance bounds are the undisplaced ones (renderer.cpp:209-210), so a
    # surface displaced toward the camera would hide its own instance; the
    # wall is displaced away from the camera
    insts = [scene.Instance(0, W, displacement_index=0, displacement_scale=-0.2)]
    for k in range(6):
        insts.append(scene.Instance(1, scene.rigid_transform(0.0, [-4.0 + 4.0 * (k % 3), -1.0, 5.0 + 4.0 * (k // 3)]),
                                    displacement_index=0, displacement_scale=0.2))
    insts.append(scene.Instance(1, scene.rigid_transform(0.0, [2.0, -1.0, -3.0]), displacement_index=0,
                                displacement_scale=0.2))
    h = scene.to_r16(scene.fbm_height(64, 7, 3))
    cam = scene.Camera(eye=(4.0, 2.0, -6.0), center=(4.0, 2.0, 10.0), width=width, height=height, near=0.05, far=100.0)
    return scene.HostScene([wall, patch], insts, [(h, abi.TEXTURE_R16_UNORM)], cam, 4.0, "wall")


def test_renderer_hiz_frames_match_oracle():
    host = wall_scene()
    W, H = host.camera.width, host.camera.height
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_HIZ)
    fc1 = host.frame()
    host.camera.eye = (4.2, 2.1, -6.0)  # the second frame moves a little
    fc2 = host.frame()
    _, _, f, smax = hiz_setup(W, H)
    for fc in (fc1, fc2):  # the renderer writes the HiZ factors into its constants
        fc.hizSizeFactors = (abi.f32 * 4)(*f)
        fc.hizSizeMax = smax
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(fc1, stages="accel")
        first = g.results().arrays["instanceStates"].copy()
        _, dec = g.trace(W, H, abi.TRACE_CULL_BACK)  # builds the HiZ of frame 1
        g.run(fc2)
        res = g.results()
    finally:
        g.close()
    depth = dec["depth"].reshape(H, W).astype(np.float32)
    assert (depth < 1.0).mean() > 0.9  # the wall fills the view
    # frame 1 had no HiZ: nothing was occluded
    assert np.array_equal((first & abi.INSTANCE_VISIBLE_BIT) != 0, (first & abi.INSTANCE_FRUSTUM_BIT) != 0)
    st = res.arrays["instanceStates"]
    hidden = (st & abi.INSTANCE_FRUSTUM_BIT).astype(bool) & ~(st & abi.INSTANCE_VISIBLE_BIT).astype(bool)
    assert hidden[1:7].sum() >= 3 and (st[0] & abi.INSTANCE_VISIBLE_BIT) and (st[7] & abi.INSTANCE_VISIBLE_BIT)
    of = harness.OracleFrame(host, cfg)
    of.set_culling_history(fc1, depth)
    of.run(fc2)
    harness.compare(res, of.results())
