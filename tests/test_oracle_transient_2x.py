"""2X transient mini batches in the oracle (TC_FLAG_2X_TRANSIENT,
cluster_classify.comp.glsl:667-905, hit decode
render_raytrace_clusters.rchit.glsl:152-170).

The reference has no golden data for them.  The oracle is pinned by what
the batches must leave unchanged against a frame without them: the frame's
triangle total; the base triangles, which move from the part list into
batches (every other part keeps its order); each batch member's vertices,
bit for bit the vertices the template path instantiates for the same
(factors, base triangle); the batch layout (8 consecutive mini triangles of
one 32-triangle group); and the traced surface, decoded to the same base
cluster, triangle and sub-triangle."""
import numpy as np
import pytest

import harness
from harness import oracle
from vk_tessellated_clusters_amd import abi, scene

FULL = {0, 0x8000, 0x8000 << 16}  # encoded corners of a whole base triangle


def small_scene():
    return scene.scene_grid_instances(num_instances=4, quads=64, patch_size=8.0, tex_size=256, tess_rate_pixels=8.0)


def cfg_factors(cfg):
    idx = (cfg & 0x7FFF) + 273
    return idx & 15, (idx >> 4) & 15, (idx >> 8) & 15


def get_config(f):
    vtx = np.array([0, 0x8000, 0x8000 << 16], np.uint32)
    c = oracle().oracle_get_config(np.asarray(f, np.uint32).ctypes.data, vtx.ctypes.data)
    return int(c), vtx


@pytest.fixture(scope="module")
def frames():
    host = small_scene()
    out = {}
    for name, flags in (("off", 0), ("2x", abi.FLAG_2X_TRANSIENT),
                        ("1x2x", abi.FLAG_1X_TRANSIENT | abi.FLAG_2X_TRANSIENT)):
        cfg = harness.small_config()
        cfg.flags |= flags
        f = harness.OracleFrame(host, cfg)
        f.run(host.frame())
        out[name] = (f, f.results(), cfg)
    return host, out


def batches(f, res):
    """Per 2X record: (record, instance, cluster, [(triangle, sub, factors)])."""
    lk = np.zeros(4096, abi.TABLE_ENTRY_DT)
    oracle().oracle_table_lookup(lk.ctypes.data)
    buf = f.buf["partTriangles"]
    out = []
    for k, r in enumerate(res.arrays["transBuilds"]):
        if int(r["clusterID"]) >> 30 != abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED:
            continue
        slot = int(r["clusterID"]) & 0x3FFFFFFF
        n = int(r["packed"]) & 511
        raw = buf[24 * slot: 24 * slot + 8 + 2 * n]
        inst, cl = raw[:8].view(np.uint32)
        m = raw[8:].view(np.uint16).astype(np.uint32)
        ents = [(int(x & 0xFF), int((x >> 8) & 0xF), (1 + int((x >> 12) & 1), 1 + int((x >> 13) & 1), 1 + int(x >> 14)))
                for x in m]
        out.append((k, r, int(inst), int(cl), ents))
    return out, lk


def classify_parts(res):
    """Parts covering a whole base triangle (emitted by classify), as
    {(instance, cluster, triangle): (index, config)}."""
    p = res.arrays["partTriangles"]
    out = {}
    for i in range(len(p)):
        if set(int(v) for v in p["vtxEncoded"][i]) == FULL:
            out[(int(p["instanceID"][i]), int(p["clusterID"][i]), int(p["triangleID_config"][i]) & 0xFFFF)] = (
                i, int(p["triangleID_config"][i]) >> 16)
    return out


def test_batches_exist(frames):
    _, fr = frames
    for name in ("2x", "1x2x"):
        f, res, _ = fr[name]
        b, _ = batches(f, res)
        assert len(b) > 20, name
        assert res.readback["numTransBuilds"] == len(res.arrays["transBuilds"])
    # with 1X the simple triangles of mixed clusters go to subset clusters
    modes = fr["1x2x"][1].arrays["transBuilds"]["clusterID"] >> 30
    assert (modes == abi.RT_CLUSTER_MODE_1X_SUBSET_CLUSTER).sum() > 0


def test_triangle_total_unchanged(frames):
    _, fr = frames
    for name in ("2x", "1x2x"):
        assert fr[name][1].readback["numTotalTriangles"] == fr["off"][1].readback["numTotalTriangles"], name
        assert fr[name][1].readback["numFullClusters"] == fr["off"][1].readback["numFullClusters"], name


def test_mini_triangles_move_into_batches(frames):
    _, fr = frames
    off = classify_parts(fr["off"][1])
    f, res, _ = fr["2x"]
    b, lk = batches(f, res)
    moved = {}
    for _, r, inst, cl, ents in b:
        tris = 0
        for t, sub, fac in ents:
            moved.setdefault((inst, cl, t), []).append((sub, fac))
            tris += 1
        assert tris == int(r["packed"]) & 511
    # exactly the whole-triangle parts with every factor <= 2 moved
    mini = {k for k, (_, c) in off.items() if max(cfg_factors(c)) <= 2}
    assert set(moved) == mini
    for key, subs in moved.items():
        c = off[key][1]
        cf, _ = get_config(subs[0][1])
        assert cf == c, key
        assert [s for s, _ in subs] == list(range(int(lk[c & 0x7FFF]["numTriangles"])))
    # every other part is unchanged and in the same order
    p0, p1 = fr["off"][1].arrays["partTriangles"], res.arrays["partTriangles"]
    key0 = [(int(a), int(c), int(t) & 0xFFFF) for a, c, t in zip(p0["instanceID"], p0["clusterID"], p0["triangleID_config"])]
    keep = np.array([not (k in moved and set(int(v) for v in p0["vtxEncoded"][i]) == FULL) for i, k in enumerate(key0)])
    assert np.array_equal(p0[keep].view(np.uint8), p1.view(np.uint8))


def test_batch_layout(frames):
    """Batches are 8 consecutive mini triangles of one 32-triangle group of
    a cluster (the last batch of a group may be short)."""
    _, fr = frames
    f, res, _ = fr["2x"]
    b, _ = batches(f, res)
    groups = {}
    for k, r, inst, cl, ents in b:
        members = sorted({t for t, _, _ in ents}, key=[t for t, _, _ in ents].index)
        assert len({t // 32 for t in members}) == 1
        assert 1 <= len(members) <= 8
        groups.setdefault((inst, cl, members[0] // 32), []).append(members)
    for key, lists in groups.items():
        flat = [t for ms in lists for t in ms]
        assert flat == sorted(flat)
        assert all(len(ms) == 8 for ms in lists[:-1])


def test_batch_vertices_equal_instantiated_vertices(frames):
    """Member vertices are bit for bit the template path's vertices of the
    same part; unused slots, and index bytes past the batch's triangles,
    are zero; the u8 indices name the member's config triangles."""
    _, fr = frames
    f0, r0, _ = fr["off"]
    f1, r1, _ = fr["2x"]
    gv0 = f0.buf["genVertices"].view(np.float32).reshape(-1, 3)
    gv1 = f1.buf["genVertices"].view(np.float32).reshape(-1, 3)
    off = classify_parts(r0)
    ti = r0.arrays["tempInstantiations"]
    vb_of_part = {int(c) & 0x3FFFFFFF: int(v) // 12 for c, v in zip(ti["clusterIdOffset"], ti["vertexBufferAddress"])
                  if int(c) >> 30 == abi.RT_CLUSTER_MODE_SINGLE_TESSELLATED}
    b, lk = batches(f1, r1)
    checked = 0
    for _, r, inst, cl, ents in b:
        vb = int(r["vertexBuffer"]) // 12  # relative to genVertices
        members = sorted({t for t, _, _ in ents}, key=[t for t, _, _ in ents].index)
        idx = f1.buf["genVertices"][12 * (vb + 48): 12 * (vb + 56)]
        o = 0
        for rel in range(8):
            slots = gv1[vb + 6 * rel: vb + 6 * rel + 6]
            if rel >= len(members):
                assert not slots.any()
                continue
            i, c = off[(inst, cl, members[rel])]
            nv, nt = int(lk[c & 0x7FFF]["numVertices"]), int(lk[c & 0x7FFF]["numTriangles"])
            src = vb_of_part[i]
            assert np.array_equal(slots[:nv], gv0[src: src + nv])
            assert not slots[nv:].any()
            for tri in range(nt):
                tv = np.zeros(3, np.uint32)
                oracle().oracle_config_triangle_vertices(c, tri, tv.ctypes.data)
                assert list(idx[3 * o: 3 * o + 3]) == [int(v) + 6 * rel for v in tv]
                o += 1
            checked += 1
        assert o == int(r["packed"]) & 511
        assert not idx[3 * o:].any()
    assert checked > 100


def test_transient_blocks(frames):
    _, fr = frames
    f, res, _ = fr["2x"]
    gv = f.buf["genVertices"].view(np.float32).reshape(-1, 3)
    blocks = res.arrays["transClasBlocks"]
    off = 0
    r16 = lambda v: (v + 15) & ~15
    for r, size in zip(res.arrays["transBuilds"], res.arrays["transClusterSizes"]):
        blk = blocks[off: off + int(size)]
        off += int(size)
        h = blk[:48].view(abi.CLAS_HEADER_DT)[0]
        nt, nv, ng = int(h["counts"]) & 0xFF, (int(h["counts"]) >> 8) & 0xFF, (int(h["counts"]) >> 16) & 0xFF
        assert nt == (int(r["packed"]) & 511) and nv == 48
        vb = int(r["vertexBuffer"]) // 12
        verts = blk[48 + r16(24 * ng) + r16(4 * nt): 48 + r16(24 * ng) + r16(4 * nt) + 12 * nv].view(np.float32)
        assert np.array_equal(verts.reshape(-1, 3), gv[vb: vb + nv])
        tris = blk[48 + r16(24 * ng): 48 + r16(24 * ng) + 4 * nt].view(np.uint32)
        assert np.array_equal(tris >> 24, np.arange(nt))
        # group boxes bound their triangles' vertices
        boxes = blk[48: 48 + 24 * ng].view(np.float32).reshape(-1, 6)
        for t in range(nt):
            p = gv[vb + np.array([tris[t] & 0xFF, (tris[t] >> 8) & 0xFF, (tris[t] >> 16) & 0xFF])]
            g = boxes[t // 4]
            assert np.all(p >= g[:3]) and np.all(p <= g[3:])
    assert off == len(blocks)


@pytest.mark.parametrize("name", ["2x", "1x2x"])
def test_traced_surface_unchanged(frames, name):
    _, fr = frames
    W, H = 96, 54
    flags = abi.TRACE_CULL_BACK | abi.TRACE_SHADOW
    h0, d0 = fr["off"][0].trace(W, H, flags)
    h1, d1 = fr[name][0].trace(W, H, flags)
    hit = h0["t"] > 0
    assert hit.mean() > 0.3
    same = np.ones(len(h0), bool)
    for k in ("instanceID", "clusterID", "triangleID"):
        same &= d0[k] == d1[k]
    # a 2X member has exactly its part's vertices; only 1X (1,1,1) subsets
    # can differ in the last bit (no re-normalised interpolated normal)
    assert same.mean() >= (1.0 if name == "2x" else 0.999)
    both = same & hit
    assert np.abs(h0["t"][both] - h1["t"][both]).max() <= 1e-5 * np.abs(h0["t"][both]).max()
    m3 = (h1["clusterID"] >> 30) == abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED
    assert (m3 & hit).sum() > 20
    sel = m3 & both
    assert np.array_equal(d0["subTriangleID"][sel], d1["subTriangleID"][sel])
    assert np.abs(d0["baryBase"][sel] - d1["baryBase"][sel]).max() <= 1e-6
    if name == "2x":
        assert np.array_equal(h0["t"], h1["t"])
