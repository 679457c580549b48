"""GPU parity of the tessellation stages through the C-ABI: the HIP
library and the CPU oracle run the same seeded frames and must agree bit
for bit on every counter, list, record, CLAS block and BLAS reference list,
and within 2 ULP on generated positions (the kernels reproduce the
oracle's float evaluation order, so the observed difference is 0 ULP)."""
import ctypes as C

import numpy as np
import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu


def run_both(host, cfg, **frame_kw):
    print("case:", host.name, len(host.instances), "instances", host.num_base_triangles, "tris", flush=True)
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame(**frame_kw))
    ref = of.results()
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(**frame_kw))
        res = g.results()
    finally:
        g.close()
    return res, ref


def small_grid(**kw):
    return scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, **kw)


def test_icosphere_parity():
    res, ref = run_both(scene.scene_icosphere(), harness.small_config())
    assert ref.readback["numTotalTriangles"] > 1280
    harness.compare(res, ref)


def test_grid_split_parity():
    res, ref = run_both(small_grid(), harness.small_config())
    assert ref.counters["splitWriteCounter"] > 0
    harness.compare(res, ref)


def test_pn_parity():
    res, ref = run_both(scene.scene_icosphere(segments=16), harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_PN))
    harness.compare(res, ref)


def test_animation_parity():
    """Ripple animation after displacement (full clusters and part
    templates), after PN, and on 1X transient clusters."""
    anim = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_ANIMATION
    res, ref = run_both(small_grid(), harness.small_config(flags=anim), animation_time=2.25)
    assert ref.counters["partTriangleCounter"] > 0
    harness.compare(res, ref)
    res, ref = run_both(scene.scene_icosphere(segments=16), harness.small_config(flags=anim | abi.FLAG_PN),
                        animation_time=0.75)
    harness.compare(res, ref)
    res, ref = run_both(small_grid(tess_rate_pixels=16.0), harness.small_config(flags=anim | abi.FLAG_1X_TRANSIENT),
                        animation_time=7.5)
    assert ref.counters["transBuildCounter"] > 0
    harness.compare(res, ref)


def test_culling_off_and_invisible_instances():
    host = small_grid()
    host.camera.center = (200.0, 0.0, -100.0)
    res, ref = run_both(host, harness.small_config())
    assert ref.readback["numFullClusters"] > 0
    harness.compare(res, ref)
    res, ref = run_both(small_grid(), harness.small_config(flags=abi.FLAG_DISPLACEMENT))
    harness.compare(res, ref)


def test_capacity_overflow_parity():
    cfg = harness.small_config(maxSplitTriangles=512, maxPartTriangles=4096, maxGeneratedVertices=40000,
                               maxGeneratedClusters=3000)
    res, ref = run_both(small_grid(), cfg)
    assert ref.readback["numPartTriangles"] > 4096
    harness.compare(res, ref)


def test_transient_1x_parity():
    """1X transient clusters: records, part slots at the top of the part
    list, index and mapping bytes, CLAS blocks and both insert passes."""
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT)
    res, ref = run_both(small_grid(tess_rate_pixels=16.0), cfg)
    assert len(ref.arrays["transBuilds"]) > 10
    harness.compare(res, ref)


def test_transient_2x_parity():
    """2X mini batches: batch records and part slots, member vertices,
    index and u16 mapping bytes, CLAS blocks and both insert passes; alone,
    with 1X subsets, with PN and with the ripple."""
    base = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_2X_TRANSIENT
    for flags, kw in ((base, {}), (base | abi.FLAG_1X_TRANSIENT, {}), (base | abi.FLAG_PN, {}),
                      (base | abi.FLAG_1X_TRANSIENT | abi.FLAG_ANIMATION, {"animation_time": 3.5})):
        print("flags", hex(flags), flush=True)
        res, ref = run_both(small_grid(tess_rate_pixels=8.0), harness.small_config(flags=flags), **kw)
        modes = ref.arrays["transBuilds"]["clusterID"] >> 30
        assert (modes == abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED).sum() > 100
        harness.compare(res, ref)


def test_transient_2x_capacity_overflow_parity():
    # batch slots and parts share the part list; vertices run out part way
    flags = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT | abi.FLAG_2X_TRANSIENT
    for kw in ({"maxPartTriangles": 12000}, {"maxGeneratedVertices": 60000}, {"maxGeneratedClusters": 1500},
               {"maxGeneratedClusterMegs": 1}):
        print("capacity", kw, flush=True)
        res, ref = run_both(small_grid(tess_rate_pixels=8.0), harness.small_config(flags=flags, **kw))
        harness.compare(res, ref)


def test_loaded_gltf_scene_parity(tmp_path):
    """A glTF scene (clusters from tc_clusterize_mesh, R16 height maps
    from PNG files, node transforms) through the GPU frame and the oracle."""
    from test_assets import gltf_doc
    from vk_tessellated_clusters_amd import assets
    path, _ = gltf_doc(tmp_path)
    host = assets.load_gltf_scene(str(path), width=640, height=360, tess_rate_pixels=2.0)
    res, ref = run_both(host, harness.small_config())
    assert ref.readback["numTotalTriangles"] > host.num_base_triangles
    harness.compare(res, ref)


def split_queue_error(g):
    """The persistent split's error word (a wait hit its bound)."""
    ws = g.r.bindings.workspace
    return int(g.r.copy(ws + 256 + 12, 4).view(np.uint32)[0])


def test_persistent_split_wait_bound_is_reported():
    """A persistent-split wait that reaches its bound ends every wait of the
    launch and flags the frame in tc_readback.errorFlags.  The bound is
    lowered from the host (tc_tuning.splitWaitLimit = 1 sleep), so every
    wait of the launch still ends at once.  Only classify and the split
    stage run on the flagged frame (the split list is reset to all-ones at
    frame start, so every item the launch reads is either valid or
    rejected); the next frame, with the default bound, is clean and matches
    the oracle."""
    host = small_grid(tess_rate_pixels=1.0)
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_PERSISTENT_SPLIT,
                               maxSplitFactor=2)
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    ref = of.results()
    L = abi.lib()
    fc = host.frame()
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(fc)
        t = abi.Tuning()
        abi.check(L.tc_get_tuning(C.byref(t)))
        t.splitWaitLimit = 1
        abi.check(L.tc_set_tuning(C.byref(t)))
        r = g.r
        b = C.byref(r.bindings)
        print("forced wait bound: classify + persistent split only", flush=True)
        abi.check(L.tc_renderer_begin_frame(r.handle, C.byref(fc), None))
        abi.check(L.tc_instances_classify(b, cfg, None))
        abi.check(L.tc_clusters_cull(b, cfg, None))
        abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_CLASSIFY, None))
        abi.check(L.tc_cluster_classify(b, cfg, None))
        abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_SPLIT, None))
        abi.check(L.tc_triangle_split(b, cfg, None))
        abi.check(L.tc_device_synchronize())
        flagged = r.readback().errorFlags
        assert split_queue_error(g) == 1
        assert flagged == abi.READBACK_ERROR_SPLIT_WAIT
        abi.check(L.tc_set_tuning(None))
        g.run(fc)
        res = g.results()
    finally:
        abi.check(L.tc_set_tuning(None))
        g.close()
    assert res.readback["errorFlags"] == 0
    harness.compare(res, ref)


def test_persistent_split_parity():
    """TC_FLAG_PERSISTENT_SPLIT runs every split pass in one launch; the
    frame equals the multipass frame and the oracle bit for bit (lists,
    counters and SPLIT_PASS state), for 4 and 14 passes, with the split list
    overflowing, and with transient builds."""
    base = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING
    cases = [(4.0, {}, base), (1.0, {}, base), (1.0, {"maxSplitFactor": 2}, base),
             (1.0, {"maxSplitTriangles": 300}, base), (1.0, {}, base | abi.FLAG_1X_TRANSIENT | abi.FLAG_2X_TRANSIENT)]
    for rate, kw, flags in cases:
        host = small_grid(tess_rate_pixels=rate)
        multi = harness.small_config(flags=flags, **kw)
        pers = harness.small_config(flags=flags | abi.FLAG_PERSISTENT_SPLIT, **kw)
        print("case", rate, kw, hex(flags), flush=True)
        of = harness.OracleFrame(host, multi)
        of.run(host.frame())
        ref = of.results()
        out = []
        for cfg in (multi, pers):
            g = harness.GpuFrame(host, cfg)
            try:
                g.run(host.frame())
                g.run(host.frame())  # a second frame reuses the queue state
                out.append(g.results())
                if cfg is pers:
                    assert split_queue_error(g) == 0
            finally:
                g.close()
        assert ref.counters["splitWriteCounter"] > 0
        harness.compare(out[1], out[0], max_ulp=0)
        harness.compare(out[1], ref)


def test_transient_1x_capacity_overflow_parity():
    # the part list is shared by parts (from the bottom) and transient slots
    # (from the top); acceptance stays a prefix of the canonical order
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT,
                               maxPartTriangles=6000, maxGeneratedVertices=200000)
    res, ref = run_both(small_grid(tess_rate_pixels=16.0), cfg)
    assert ref.readback["numPartTriangles"] > 6000
    harness.compare(res, ref)


def test_many_instances_parity():
    host = scene.scene_grid_instances(num_instances=16, quads=32, patch_size=4.0, tex_size=128, cam_height=1.5)
    res, ref = run_both(host, harness.small_config())
    harness.compare(res, ref)


def test_repeat_frames_are_identical():
    host = small_grid()
    cfg = harness.small_config()
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame())
        r1 = g.results()
        g.run(host.frame())
        r2 = g.results()
    finally:
        g.close()
    harness.compare(r2, r1, max_ulp=0)


def test_launch_shapes_do_not_change_results():
    """Work-queue kernels give identical frames, BVHs and hits for every
    launch shape the tuner may pick."""
    import numpy as np
    from vk_tessellated_clusters_amd import autotune
    host = small_grid()
    cfg = harness.small_config()
    shapes = [None, dict(classifyBlocksPerCU=1, splitBlocksPerCU=1, instantiateBlocksPerCU=1, clasWavesPerCU=4,
                         traceBlocksPerCU=2, traceThreads=64, bvhBlocksPerCU=1),
              dict(classifyBlocksPerCU=12, splitBlocksPerCU=16, instantiateBlocksPerCU=8, clasWavesPerCU=128,
                   traceBlocksPerCU=128, traceThreads=128, bvhBlocksPerCU=16)]
    out = []
    try:
        for shape in shapes:
            autotune.set_tuning(shape)
            g = harness.GpuFrame(host, cfg)
            try:
                g.run(host.frame(), stages="accel")
                res = g.results()
                arr = g.accel_arrays()
                hits, dec = g.trace(96, 54, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
            finally:
                g.close()
            out.append((res, arr, hits, dec))
    finally:
        autotune.set_tuning(None)
    def written(a):
        """Node records that the build writes (n - 1 internal nodes per
        instance) and leaves as offsets into the CLAS pool."""
        sb, n = a["sb"], len(a["infos"])
        keep = np.zeros(len(a["blasNodes"]), bool)
        for info in a["infos"]:
            first = int(info["clusterReferences"] - sb.blasClusterAddresses) // 8
            keep[first:first + max(0, int(info["clusterReferencesCount"]) - 1)] = True
        return (a["blasNodes"][keep].view(np.uint8), a["blasNodes4"][keep].view(np.uint8),
                a["tlasNodes"][:n - 1].view(np.uint8), a["tlasNodes4"][:n - 1].view(np.uint8),
                a["blasLeaves"] - np.uint64(sb.genClusterData))

    r0, a0, h0, d0 = out[0]
    for r, a, h, d in out[1:]:
        harness.compare(r, r0, max_ulp=0)
        for x, y in zip(written(a), written(a0)):
            assert np.array_equal(x, y)
        assert np.array_equal(h.view(np.uint8), h0.view(np.uint8))
        assert np.array_equal(d.view(np.uint8), d0.view(np.uint8))


def test_graph_replays_match_eager_frames():
    """Graph replays with new frame constants (queued back to back, more
    than the pinned ring holds) end in the same state as an eager frame
    with the last constants."""
    import ctypes as C
    import numpy as np
    import torch
    host = small_grid()
    cfg = harness.small_config()
    frames = []
    for k in range(20):
        host.camera.eye = (8.0 + 0.3 * k, 2.5, -2.4 + 0.1 * k)
        frames.append(host.frame())
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(frames[-1], stages="accel")
        eager = g.results()
        eh, ed = g.trace(64, 36, abi.TRACE_CULL_BACK)
        stream = torch.cuda.Stream()
        s = C.c_void_p(stream.cuda_stream)
        L = g.r.L
        abi.check(L.tc_renderer_graph_capture(g.r.handle, C.byref(frames[0]), None, s))
        for fc in frames:
            abi.check(L.tc_renderer_graph_launch(g.r.handle, C.byref(fc), s))
        stream.synchronize()
        replay = g.results()
        rh, rd = g.trace(64, 36, abi.TRACE_CULL_BACK)
    finally:
        g.close()
    ce, cr = eager.arrays["clasBlocks"], replay.arrays["clasBlocks"]
    if ce.shape == cr.shape and not np.array_equal(ce, cr):
        # name the first differing block and the part of it that differs
        off = 0
        for i, sz in enumerate(eager.arrays["tempClusterSizes"]):
            a, b = ce[off:off + int(sz)], cr[off:off + int(sz)]
            if not np.array_equal(a, b):
                d = np.nonzero(a != b)[0]
                print("block", i, "of", len(eager.arrays["tempClusterSizes"]), "size", int(sz), "differs at bytes",
                      d[:8], "count", d.size, "headers", a[:48].view(abi.CLAS_HEADER_DT), b[:48].view(abi.CLAS_HEADER_DT))
                break
            off += int(sz)
    harness.compare(replay, eager, max_ulp=0)
    assert np.array_equal(rh.view(np.uint8), eh.view(np.uint8))


def test_device_fbm_matches_numpy():
    """The device height-map generator reproduces the numpy generator texel
    for texel (R32 and R16)."""
    import numpy as np
    import torch
    L = abi.lib()
    for size, seed, octaves, fmt in ((64, 1, 6, abi.TEXTURE_R32_FLOAT), (256, 7, 4, abi.TEXTURE_R16_UNORM),
                                     (100, 3, 5, abi.TEXTURE_R32_FLOAT)):
        ref = scene.fbm_height(size, seed, octaves)
        if fmt == abi.TEXTURE_R16_UNORM:
            ref = scene.to_r16(ref)
        out = torch.zeros(size * size * ref.itemsize, dtype=torch.uint8, device="cuda")
        abi.check(L.tc_fbm_height(out.data_ptr(), size, seed, octaves, 8, fmt, None))
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(ref.dtype).reshape(size, size)
        assert np.array_equal(got, ref)
