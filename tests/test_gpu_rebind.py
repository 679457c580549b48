"""Frames over rebound instance records, and the prebuilt cluster list.

In ray-tracing mode clusters_cull lists every cluster of every instance
(clusters_cull.comp.glsl:112-161), so a renderer builds the list of its own
instances once, at creation, in a buffer of its own, and its frames start
from it.  Frames over instance records bound in their place with
tc_renderer_rebind_instances cull into the renderer's list buffer instead.
The meshes here differ in cluster count, so a frame that read the wrong list
would differ from the oracle's frame.
"""
import ctypes as C

import numpy as np
import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu


def test_rebound_instances_and_prebuilt_cluster_list():
    import torch

    host = scene.scene_batched(num_meshes=5, tex_size=64, width=320, height=180)
    rev = scene.HostScene(host.geometries, host.instances[::-1], host.textures, host.camera, host.tess_rate_pixels,
                          "batched, instances reversed")
    counts = [host.geometries[i.geometry].num_clusters for i in host.instances]
    assert len(set(counts)) > 1, counts
    cfg = harness.small_config()
    fc = host.frame()
    refs = []
    for h in (host, rev):
        of = harness.OracleFrame(h, cfg)
        of.run(fc)
        refs.append(of.results())

    L = abi.lib()
    g = harness.GpuFrame(host, cfg)
    try:
        # the renderer's instance records (device addresses of its own
        # scene) in reverse order, and their transforms
        n = len(host.instances)
        recs = (abi.RenderInstance * n)(*[g.ds.instances[n - 1 - i] for i in range(n)])
        xf = np.zeros(24 * n, np.float32)
        abi.check(L.tc_instance_xforms_host(C.addressof(recs), n, xf.ctypes.data))
        inst_dev = torch.from_numpy(np.frombuffer(bytes(recs), np.uint8).copy()).cuda()
        xf_dev = torch.from_numpy(xf.view(np.uint8).copy()).cuda()
        torch.cuda.synchronize()
        out = []
        for step, bound in enumerate((False, True, False, True, False)):
            print("frame", step, "rebound" if bound else "own instances", flush=True)
            if bound:
                abi.check(L.tc_renderer_rebind_instances(g.r.handle, C.c_void_p(inst_dev.data_ptr()),
                                                         C.c_void_p(xf_dev.data_ptr()), n))
            else:
                abi.check(L.tc_renderer_rebind_instances(g.r.handle, None, None, n))
            g.run(fc)
            out.append((bound, g.results()))
    finally:
        g.close()
    for bound, res in out:
        harness.compare(res, refs[1] if bound else refs[0])
