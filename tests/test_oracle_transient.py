"""1X transient cluster builds in the oracle (TC_FLAG_1X_TRANSIENT).

The reference has no golden data for them, so the oracle is pinned by what
the transient path must leave unchanged: the frame's triangle total, the
set of untessellated triangles (moved from (1,1,1) part instantiations
into transient CLAS blocks), the order of every other part, and the
traced surface (the same displaced triangles, decoded to the same base
cluster and triangle)."""
import numpy as np
import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

CFG_111 = 0  # lookup index of factors (1, 1, 1)


def small_scene():
    return scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, tess_rate_pixels=16.0)


@pytest.fixture(scope="module")
def frames():
    host = small_scene()
    out = {}
    for name, flags in (("off", 0), ("on", abi.FLAG_1X_TRANSIENT)):
        cfg = harness.small_config()
        cfg.flags |= flags
        f = harness.OracleFrame(host, cfg)
        f.run(host.frame())
        out[name] = (f, f.results(), cfg)
    return host, out


def _parts(res):
    p = res.arrays["partTriangles"]
    key = np.stack([p["instanceID"], p["clusterID"], p["triangleID_config"] & 0xFFFF], 1)
    cfg = p["triangleID_config"] >> 16
    return key, cfg, p


def _trans_triangles(f, res):
    """(instance, cluster, triangle) of every triangle in a transient CLAS."""
    out = []
    buf = f.buf["partTriangles"]
    for r in res.arrays["transBuilds"]:
        slot = int(r["clusterID"]) & 0x3FFFFFFF
        n = int(r["packed"]) & 511
        raw = buf[24 * slot: 24 * slot + 8 + n]
        inst, cl = raw[:8].view(np.uint32)
        out += [(int(inst), int(cl), int(t)) for t in raw[8:8 + n]]
    return out


def test_transient_clusters_exist(frames):
    _, fr = frames
    res = fr["on"][1]
    assert len(res.arrays["transBuilds"]) > 10
    assert fr["on"][1].readback["numTransBuilds"] == len(res.arrays["transBuilds"])


def test_triangle_total_unchanged(frames):
    _, fr = frames
    assert fr["on"][1].readback["numTotalTriangles"] == fr["off"][1].readback["numTotalTriangles"]
    assert fr["on"][1].readback["numFullClusters"] == fr["off"][1].readback["numFullClusters"]


def test_simple_triangles_move_into_transient_clas(frames):
    _, fr = frames
    koff, coff, _ = _parts(fr["off"][1])
    kon, con, _ = _parts(fr["on"][1])
    f_on, r_on, _ = fr["on"]
    moved = _trans_triangles(f_on, r_on)
    assert len(moved) == len(set(moved))
    simple_off = {tuple(k) for k in koff[coff == CFG_111].tolist()}
    simple_on = {tuple(k) for k in kon[con == CFG_111].tolist()}
    assert simple_on.isdisjoint(moved)
    assert simple_on | set(moved) == simple_off
    # every other part is unchanged and in the same order
    keep = np.array([tuple(k) not in set(moved) for k in koff.tolist()], bool)
    assert np.array_equal(koff[keep], kon)
    assert np.array_equal(coff[keep], con)
    # each transient cluster moved more than one triangle (threshold 1)
    slots = {}
    for t in moved:
        slots.setdefault(t[:2], 0)
        slots[t[:2]] += 1
    assert min(slots.values()) >= 2


def test_transient_blocks_reference_cluster_vertices(frames):
    host, fr = frames
    f, res, cfg = fr["on"]
    gv = f.buf["genVertices"].view(np.float32).reshape(-1, 3)
    blocks = res.arrays["transClasBlocks"]
    off = 0
    for r, size in zip(res.arrays["transBuilds"], res.arrays["transClusterSizes"]):
        blk = blocks[off: off + int(size)]
        off += int(size)
        h = blk[:48].view(abi.CLAS_HEADER_DT)[0]
        nt, nv, ng = int(h["counts"]) & 0xFF, (int(h["counts"]) >> 8) & 0xFF, (int(h["counts"]) >> 16) & 0xFF
        assert nt == (int(r["packed"]) & 511) and nv == ((int(r["packed"]) >> 9) & 511)
        assert int(h["clusterID"]) == int(r["clusterID"])
        vb = int(r["vertexBuffer"]) // 12
        assert int(h["vertexBase"]) == vb
        r16 = lambda v: (v + 15) & ~15
        tris = blk[48 + r16(24 * ng): 48 + r16(24 * ng) + 4 * nt].view(np.uint32)
        verts = blk[48 + r16(24 * ng) + r16(4 * nt): 48 + r16(24 * ng) + r16(4 * nt) + 12 * nv].view(np.float32).reshape(-1, 3)
        assert np.array_equal(verts, gv[vb: vb + nv])
        idx = f.buf["genVertices"][12 * (vb + nv): 12 * (vb + nv) + 3 * nt].reshape(-1, 3)
        assert np.array_equal(tris & 0xFF, idx[:, 0]) and np.array_equal((tris >> 8) & 0xFF, idx[:, 1])
        assert np.array_equal(tris >> 24, np.arange(nt))
    assert off == len(blocks)


def test_traced_surface_unchanged(frames):
    _, fr = frames
    W, H = 96, 54
    flags = abi.TRACE_CULL_BACK | abi.TRACE_SHADOW
    h0, d0 = fr["off"][0].trace(W, H, flags)
    h1, d1 = fr["on"][0].trace(W, H, flags)
    hit = h0["t"] > 0
    assert hit.mean() > 0.3
    # a (1,1,1) part re-normalises the interpolated normal, the cluster path
    # does not: vertices may differ in the last bit, so hits agree up to
    # rays through shared edges and t to rounding
    same = np.ones(len(h0), bool)
    for k in ("instanceID", "clusterID", "triangleID"):
        same &= d0[k] == d1[k]
    assert same.mean() >= 0.999
    both = same & hit
    assert np.abs(h0["t"][both] - h1["t"][both]).max() <= 1e-5 * np.abs(h0["t"][both]).max()
    modes = h1["clusterID"][hit] >> 30
    assert (modes == abi.RT_CLUSTER_MODE_1X_SUBSET_CLUSTER).sum() > 20
