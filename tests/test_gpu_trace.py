"""Cluster acceleration structures and ray traversal on the GPU.

The BVH is validated on the host first (every CLAS reachable exactly once,
boxes equal to the unions of their subtrees, Morton order inside each
instance); only then are rays traced, and hits are compared with the CPU
oracle's hierarchy-free ray caster (closest hit over every CLAS, same
watertight test and tie-break): hit IDs must agree on >= 99.99 % of rays
and t within 1e-5 relative."""
import numpy as np
import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu


def small_scene():
    return scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256)


def clas_boxes(g, addrs):
    out = {}
    for a in np.unique(addrs):
        h = g.r.copy(int(a), 48).view(abi.CLAS_HEADER_DT)[0]
        out[int(a)] = (h["lo"].copy(), h["hi"].copy(), int(h["counts"]) | int(h["clusterID"]) << 32)
    return out


def check_blas(arr, boxes, inst):
    info = arr["infos"][inst]
    n = int(info["clusterReferencesCount"])
    first = int(info["clusterReferences"] - arr["sb"].blasClusterAddresses) // 8
    leaves = arr["blasLeaves"][first:first + n]
    assert sorted(leaves.tolist()) == sorted(arr["refs"][first:first + n].tolist())
    # the leaf info repeats each leaf's header counts and cluster ID
    assert [int(x) for x in arr["blasLeafInfo"][first:first + n]] == [boxes[int(a)][2] for a in leaves]
    if n == 1:
        return
    nodes = arr["blasNodes"][first:first + n - 1]
    seen_nodes, seen_leaves = set(), []

    def visit(word):
        if word & abi.BVH_LEAF:
            k = word & ~abi.BVH_LEAF
            assert k < n
            seen_leaves.append(k)
            return boxes[int(leaves[k])]
        assert word < n - 1 and word not in seen_nodes
        seen_nodes.add(word)
        nd = nodes[word]
        b0 = visit(int(nd["child0"]))
        b1 = visit(int(nd["child1"]))
        assert np.array_equal(nd["lo0"], b0[0]) and np.array_equal(nd["hi0"], b0[1])
        assert np.array_equal(nd["lo1"], b1[0]) and np.array_equal(nd["hi1"], b1[1])
        return np.minimum(b0[0], b1[0]), np.maximum(b0[1], b1[1])

    visit(0)
    assert len(seen_nodes) == n - 1 and sorted(seen_leaves) == list(range(n))
    check_wide(nodes, arr["blasNodes4"][first:first + n - 1])


def check_wide(nodes, wide):
    """Every four-wide record lists the node's leaf children and the
    children of its internal children, with the binary tree's boxes."""
    for i, nd in enumerate(nodes):
        want = []
        for c in (0, 1):
            ch, lo, hi = int(nd["child%d" % c]), nd["lo%d" % c], nd["hi%d" % c]
            if ch & abi.BVH_LEAF:
                want.append((ch, tuple(lo), tuple(hi)))
            else:
                g = nodes[ch]
                want += [(int(g["child0"]), tuple(g["lo0"]), tuple(g["hi0"])),
                         (int(g["child1"]), tuple(g["lo1"]), tuple(g["hi1"]))]
        w = wide[i]
        got = [(int(w["child"][k]), tuple(w["lo"][:, k]), tuple(w["hi"][:, k])) for k in range(4)
               if int(w["child"][k]) != abi.BVH4_EMPTY]
        assert got == want, i
        assert all(int(w["child"][k]) == abi.BVH4_EMPTY for k in range(len(want), 4))


def test_accel_structure():
    host = small_scene()
    g = harness.GpuFrame(host, harness.small_config())
    try:
        g.run(host.frame(), stages="accel")
        arr = g.accel_arrays()
        boxes = clas_boxes(g, arr["refs"])
        for i in range(len(host.instances)):
            check_blas(arr, boxes, i)
        assert harness.check_blas_key_order(g, arr) == len(arr["refs"])
        # TLAS reaches every instance once
        n = len(host.instances)
        seen = []

        def tvisit(word, depth=0):
            assert depth < 64
            if word & abi.BVH_LEAF:
                seen.append(word & ~abi.BVH_LEAF)
                return
            nd = arr["tlasNodes"][word]
            tvisit(int(nd["child0"]), depth + 1)
            tvisit(int(nd["child1"]), depth + 1)

        tvisit(arr["tlasRoot"])
        assert sorted(seen) == list(range(n))
        check_wide(arr["tlasNodes"][:n - 1], arr["tlasNodes4"][:n - 1])
    finally:
        g.close()


compare_hits = harness.compare_hits


def test_trace_parity_primary_and_shadow():
    host = small_scene()
    cfg = harness.small_config()
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    W, H = 160, 90
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
        # the same traversal with the whole stack in private memory, and the
        # wave-wide packet traversal
        ph, pd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW | abi.TRACE_PRIVATE_STACK)
        kh, kd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW | abi.TRACE_PACKETS)
    finally:
        g.close()
    assert np.array_equal(ph.view(np.uint8), gh.view(np.uint8))
    assert np.array_equal(pd.view(np.uint8), gd.view(np.uint8))
    assert np.array_equal(kh.view(np.uint8), gh.view(np.uint8))
    assert np.array_equal(kd.view(np.uint8), gd.view(np.uint8))
    assert (oh["t"] > 0).mean() > 0.3, "camera must see geometry"
    agree = compare_hits(gh, oh, gd, od)
    print("hit agreement", agree, "hits", (oh["t"] > 0).sum(), "shadowed", od["shadowed"].sum())


def test_trace_parity_icosphere():
    host = scene.scene_icosphere()
    cfg = harness.small_config()
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    W, H = 128, 72
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    finally:
        g.close()
    # a radius-1 sphere at distance 3 under a 60 degree, 16:9 frustum covers
    # about 17% of the pixels
    assert 0.12 < (oh["t"] > 0).mean() < 0.25
    compare_hits(gh, oh, gd, od)


def test_trace_parity_transient_1x():
    """Hits on 1X transient CLAS decode to the base cluster and triangle."""
    host = scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, tess_rate_pixels=16.0)
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT)
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    W, H = 160, 90
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    finally:
        g.close()
    assert ((oh["clusterID"] >> 30) == abi.RT_CLUSTER_MODE_1X_SUBSET_CLUSTER).sum() > 20
    agree = compare_hits(gh, oh, gd, od)
    print("hit agreement", agree)


def test_trace_parity_transient_2x():
    """Hits on 2X batch CLAS decode to the base cluster, triangle and
    sub-triangle (with 1X subsets in the same frame)."""
    host = scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, tess_rate_pixels=8.0)
    cfg = harness.small_config(flags=abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_1X_TRANSIENT |
                               abi.FLAG_2X_TRANSIENT)
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    W, H = 160, 90
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    finally:
        g.close()
    assert ((oh["clusterID"] >> 30) == abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED).sum() > 50
    agree = compare_hits(gh, oh, gd, od)
    print("hit agreement", agree)


def test_tile_row_order_keeps_hits():
    """A renderer's camera traces time their tile rows, the next frame start
    ranks the rows (longest wave first), and that frame's traces of the same
    height take the rows in that order.  Hits and decoded records do not
    depend on the order: ordered traces, traces after a frame that traced
    another height (tile order again), and the 128- and 256-thread workgroup
    forms all equal the first trace byte for byte."""
    import ctypes as C
    host = small_scene()
    cfg = harness.small_config()
    flags = abi.TRACE_CULL_BACK | abi.TRACE_SHADOW
    W, H = 160, 90
    L = abi.lib()
    g = harness.GpuFrame(host, cfg)

    def frame_and_trace(w, h):
        g.run(host.frame(), stages="accel")
        return g.trace(w, h, flags)

    try:
        first = frame_and_trace(W, H)
        later = [frame_and_trace(W, H), frame_and_trace(W, H)]
        small = frame_and_trace(96, 54)
        small2 = frame_and_trace(96, 54)
        later += [frame_and_trace(W, H), frame_and_trace(W, H)]
        for threads in (128, 256):
            t = abi.Tuning()
            abi.check(L.tc_get_tuning(C.byref(t)))
            t.traceThreads = threads
            abi.check(L.tc_set_tuning(C.byref(t)))
            later += [frame_and_trace(W, H), frame_and_trace(W, H)]
    finally:
        abi.check(L.tc_set_tuning(None))
        g.close()
    assert (first[0]["t"] > 0).mean() > 0.3, "camera must see geometry"
    for h, d in later:
        assert np.array_equal(h.view(np.uint8), first[0].view(np.uint8))
        assert np.array_equal(d.view(np.uint8), first[1].view(np.uint8))
    assert np.array_equal(small2[0].view(np.uint8), small[0].view(np.uint8))
    assert np.array_equal(small2[1].view(np.uint8), small[1].view(np.uint8))


def test_packet_traversal_matches_per_ray():
    """Packet traversal gives the per-ray traversal's hits byte for byte,
    for camera rays (ragged tiles) and explicit rays in any directions."""
    host = scene.scene_icosphere()
    cfg = harness.small_config()
    g = harness.GpuFrame(host, cfg)
    rng = np.random.default_rng(7)
    rays = np.zeros(5000, abi.RAY_DT)
    rays["origin"] = rng.normal(0, 0.3, (5000, 3)) + np.array([0.0, 0.0, 3.0])
    dirs = rng.normal(size=(5000, 3)) + np.array([0.0, 0.0, -2.0])
    rays["direction"] = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    rays["tMin"], rays["tMax"] = 0.0, 100.0
    try:
        g.run(host.frame(), stages="accel")
        for W, H in ((100, 60), (64, 64)):
            a = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
            b = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW | abi.TRACE_PACKETS)
            assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
            assert np.array_equal(a[1].view(np.uint8), b[1].view(np.uint8))
        a = g.trace(0, 0, 0, rays=rays)
        b = g.trace(0, 0, abi.TRACE_PACKETS, rays=rays)
        assert (a[0]["t"] > 0).sum() > 500
        assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8))
    finally:
        g.close()


def lattice_scene(n=1100, width=96, height=54):
    """n one-cluster patches (2 x 2 quads) on a square lattice with seeded
    yaw, one displacement map, and a camera looking down across the field:
    more instances than the one-workgroup TLAS build takes, so the TLAS is
    built by separate launches with a radix sort over instances."""
    geo = scene.grid_patch(2, 2, 1.4, 1.4)
    side = int(np.ceil(np.sqrt(n)))
    rng = np.random.default_rng(5)
    insts = []
    for k in range(n):
        x, z = 1.0 * (k % side - side / 2), 1.0 * (k // side - side / 2)
        insts.append(scene.Instance(0, scene.rigid_transform(float(rng.uniform(0, 6.28)), [x, 0.0, z]),
                                    displacement_index=0, displacement_scale=0.1))
    h = scene.to_r16(scene.fbm_height(64, 3, 3))
    cam = scene.Camera(eye=(0.0, 14.0, -22.0), center=(0.0, 0.0, 0.0), width=width, height=height, near=0.05,
                       far=500.0)
    return scene.HostScene([geo], insts, [(h, abi.TEXTURE_R16_UNORM)], cam, 8.0, "lattice")


def test_large_tlas_structure_and_trace():
    host = lattice_scene()
    n = len(host.instances)
    cfg = harness.small_config()
    W, H = host.camera.width, host.camera.height
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        arr = g.accel_arrays()
        # more than 256 instances: the key sort runs its instance digits
        assert harness.check_blas_key_order(g, arr) == len(arr["refs"])
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    finally:
        g.close()
    seen = []

    def tvisit(word, depth=0):
        assert depth < 64
        if word & abi.BVH_LEAF:
            seen.append(word & ~abi.BVH_LEAF)
            return
        nd = arr["tlasNodes"][word]
        tvisit(int(nd["child0"]), depth + 1)
        tvisit(int(nd["child1"]), depth + 1)

    tvisit(arr["tlasRoot"])
    assert sorted(seen) == list(range(n))
    check_wide(arr["tlasNodes"][:n - 1], arr["tlasNodes4"][:n - 1])
    assert (oh["t"] > 0).mean() > 0.3, "camera must see geometry"
    assert len(np.unique(oh["instanceID"][oh["t"] > 0])) > 200
    compare_hits(gh, oh, gd, od)


def test_large_tlas_with_clas_overflow():
    """The TLAS has key space of its own, so a scene with more instances
    (1100) than maxGeneratedClusters (1024) builds: the frame overflows its
    CLAS capacity exactly as the oracle's does, instances without accepted
    CLAS stay in the TLAS with empty BLAS, and hits match the oracle."""
    host = lattice_scene()
    n = len(host.instances)
    cfg = harness.small_config(maxGeneratedClusters=1024)
    W, H = host.camera.width, host.camera.height
    of = harness.OracleFrame(host, cfg)
    of.run(host.frame())
    ref = of.results()
    oh, od = of.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(host.frame(), stages="accel")
        res = g.results()
        arr = g.accel_arrays()
        assert harness.check_blas_key_order(g, arr) == len(arr["refs"])
        gh, gd = g.trace(W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW)
    finally:
        g.close()
    harness.compare(res, ref)
    assert ref.counters["tempInstantiateCounter"] <= 1024 < n
    seen = []

    def tvisit(word, depth=0):
        assert depth < 64
        if word & abi.BVH_LEAF:
            seen.append(word & ~abi.BVH_LEAF)
            return
        nd = arr["tlasNodes"][word]
        tvisit(int(nd["child0"]), depth + 1)
        tvisit(int(nd["child1"]), depth + 1)

    tvisit(arr["tlasRoot"])
    assert sorted(seen) == list(range(n))
    compare_hits(gh, oh, gd, od)
