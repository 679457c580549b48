"""tc_frame_tessellate_noalloc through the C-ABI: the tessellation chain
with its look-back state in the library's own device array instead of the
caller's workspace (the bindings carry no workspace).  Two calls in a row,
the second on a smaller input, each against tc_frame_tessellate on the same
renderer and frame: every counter, list, record and generated vertex
bit-identical, and both equal to the oracle."""
import ctypes as C

import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu


def tessellate(g, fc, noalloc):
    L = abi.lib()
    r, cfg = g.r, g.cfg
    b = C.byref(r.bindings)
    abi.check(L.tc_renderer_begin_frame(r.handle, C.byref(fc), None))
    abi.check(L.tc_instances_classify(b, cfg, None))
    abi.check(L.tc_clusters_cull(b, cfg, None))
    if noalloc:
        nb = abi.Bindings.from_buffer_copy(bytes(r.bindings))
        nb.workspace, nb.workspaceBytes = None, 0
        abi.check(L.tc_frame_tessellate_noalloc(C.byref(nb), cfg, 0, None), "tc_frame_tessellate_noalloc")
    else:
        abi.check(L.tc_frame_tessellate(b, cfg, 0, None), "tc_frame_tessellate")
    r.clas_and_blas()
    abi.check(L.tc_device_synchronize())
    return g.results()


def test_noalloc_frame_matches_workspace_frame():
    base = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING
    cases = [
        # split passes, 1X subsets and 2X batches (the batch vertex kernel
        # reads the counter snapshot of the array's header)
        (scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, tess_rate_pixels=2.0),
         harness.small_config(flags=base | abi.FLAG_1X_TRANSIENT | abi.FLAG_2X_TRANSIENT)),
        # smaller: every split pass in one persistent launch (its pass
        # table lives in the array's header)
        (scene.scene_grid_instances(num_instances=2, quads=32, patch_size=4.0, tex_size=128, tess_rate_pixels=1.0),
         harness.small_config(flags=base | abi.FLAG_PERSISTENT_SPLIT)),
    ]
    sizes = []
    for host, cfg in cases:
        print("case:", len(host.instances), "instances", host.num_base_triangles, "tris, flags", hex(cfg.flags),
              flush=True)
        of = harness.OracleFrame(host, cfg)
        of.run(host.frame())
        ref = of.results()
        g = harness.GpuFrame(host, cfg)
        try:
            want = tessellate(g, host.frame(), False)
            got = tessellate(g, host.frame(), True)
        finally:
            g.close()
        assert ref.counters["splitWriteCounter"] > 0
        harness.compare(got, want, max_ulp=0)
        harness.compare(got, ref)
        sizes.append((host.num_base_triangles, ref.readback["numVisibleClusters"]))
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]
