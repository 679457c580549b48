"""Pins the tessellation table and its 16^3 lookup expansion.

References: src/tessellation_table_nv_raw.hpp (data), README.md:30-92
(format), src/tessellation_table.cpp:36-100 (lookup), shaders/shaderio_scene.h:
25-51 (limits).
"""
import ctypes as C
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from harness import oracle, ROOT
from vk_tessellated_clusters_amd import abi

GOLDEN = Path(__file__).resolve().parent / "golden" / "table_digest.json"


def table_arrays():
    O = oracle()
    nv, nt, nc = O.oracle_table_num_vertices(), O.oracle_table_num_triangles(), O.oracle_table_num_configs()
    v = np.ctypeslib.as_array(O.oracle_table_vertices(), (nv,)).copy()
    t = np.ctypeslib.as_array(O.oracle_table_triangles(), (nt,)).copy()
    c = np.ctypeslib.as_array(O.oracle_table_configs(), (nc * 4,)).copy().reshape(nc, 4)
    return v, t, c


def raw_configs():
    return [(x, y, z) for x in range(1, 12) for y in range(1, x + 1) for z in range(1, y + 1)]


def test_table_matches_reference_digest():
    """The build's copy of the table equals the upstream header word for word
    (digest produced from oracle/_ref/table_dump, compiled from that header)."""
    g = json.loads(GOLDEN.read_text())
    v, t, c = table_arrays()
    assert g["max_edge_segments"] == 11
    assert (g["num_vertices"], g["num_triangles"], g["num_configs"]) == (v.size, t.size, c.shape[0])
    assert hashlib.sha256(v.astype(np.uint32).tobytes()).hexdigest() == g["vertices_sha256"]
    assert hashlib.sha256(t.astype(np.uint32).tobytes()).hexdigest() == g["triangles_sha256"]
    assert hashlib.sha256(c.astype(np.uint16).reshape(-1).tobytes()).hexdigest() == g["configs_sha256"]


def test_table_matches_reference_build_when_present():
    dump = ROOT / "oracle" / "_ref" / "table_dump"
    if not dump.exists():
        pytest.skip("oracle/_ref not built (upstream tree absent)")
    out = subprocess.run([str(dump)], capture_output=True, text=True, check=True).stdout.split("\n")
    v, t, c = table_arrays()
    assert [int(l[2:], 16) for l in out if l.startswith("V ")] == v.tolist()
    assert [int(l[2:], 16) for l in out if l.startswith("T ")] == t.tolist()
    assert [int(l[2:]) for l in out if l.startswith("C ")] == c.reshape(-1).tolist()


def test_table_invariants():
    v, t, c = table_arrays()
    assert c.shape[0] == 286 == len(raw_configs())
    # ranges are contiguous and in bounds
    assert c[0, 0] == 0 and c[0, 1] == 0
    assert np.all(c[1:, 0] == c[:-1, 0] + c[:-1, 2])
    assert np.all(c[1:, 1] == c[:-1, 1] + c[:-1, 3])
    assert c[-1, 0] + c[-1, 2] == t.size and c[-1, 1] + c[-1, 3] == v.size
    # u + v <= 1.0 in fixed point, top triangle byte unused
    assert np.all((v & 0xFFFF) + (v >> 16) <= 32768)
    assert np.all(t >> 24 == 0)
    for i, (x, y, z) in enumerate(raw_configs()):
        ft, fv, nt, nv = (int(k) for k in c[i])
        idx = t[ft:ft + nt]
        loc = np.stack([idx & 0xFF, (idx >> 8) & 0xFF, (idx >> 16) & 0xFF], 1)
        assert loc.max() < nv
        # every pattern vertex is used
        assert np.unique(loc).size == nv
        # uniform n: n^2 triangles, (n+1)(n+2)/2 vertices (shaderio_scene.h:45-51)
        if x == y == z:
            assert nt == x * x and nv == (x + 1) * (x + 2) // 2
        if x <= 2:
            assert nt <= 4 and nv <= 6
        # the three corners are present
        corners = {0, 0x8000, 0x80000000}
        assert corners <= set(v[fv:fv + nv].tolist())
        # edge segments: vertices on edge w=0 (u+v == 1), v == 0 and u == 0
        vv = v[fv:fv + nv]
        u_, w_ = vv & 0xFFFF, vv >> 16
        assert np.count_nonzero(w_ == 0) == x + 1          # edge 0: corner 0 -> corner 1
        assert np.count_nonzero(u_ + w_ == 32768) == y + 1  # edge 1: corner 1 -> corner 2
        assert np.count_nonzero(u_ == 0) == z + 1          # edge 2: corner 2 -> corner 0
    # (1,1,1): identity triangle (README.md:41-52)
    assert v[:3].tolist() == [0, 0x8000, 0x80000000] and t[0] == 0x00020100


def test_patterns_tile_the_triangle():
    """Pattern triangles are CCW and their areas sum to the unit triangle."""
    v, t, c = table_arrays()
    for i in range(c.shape[0]):
        ft, fv, nt, nv = (int(k) for k in c[i])
        vv = v[fv:fv + nv].astype(np.int64)
        uv = np.stack([vv & 0xFFFF, vv >> 16], 1)
        idx = t[ft:ft + nt]
        a, b, cc = (uv[(idx >> s) & 0xFF] for s in (0, 8, 16))
        area2 = (b[:, 0] - a[:, 0]) * (cc[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (cc[:, 0] - a[:, 0])
        assert np.all(area2 > 0)
        assert area2.sum() == 32768 * 32768


def lookup_oracle():
    out = np.zeros(4096, abi.TABLE_ENTRY_DT)
    oracle().oracle_table_lookup(out.ctypes.data)
    return out


def test_lookup_expansion():
    """tessellation_table.cpp:53-80: (x,y,z) and its alias (x,z,y)."""
    lk = lookup_oracle()
    v, t, c = table_arrays()
    for i, (x, y, z) in enumerate(raw_configs()):
        e = lk[x + 16 * y + 256 * z - 273]
        assert tuple(e) == tuple(int(k) for k in c[i])
        if z != y and x > 1:
            assert tuple(lk[x + 16 * z + 256 * y - 273]) == tuple(int(k) for k in c[i])
    used = np.count_nonzero(lk["numTriangles"])
    assert used == 286 + sum(1 for (x, y, z) in raw_configs() if z != y and x > 1)


def test_product_lookup_equals_oracle():
    L = abi.lib()
    mine = np.zeros(4096, abi.TABLE_ENTRY_DT)
    abi.check(L.tc_table_lookup_host(mine.ctypes.data))
    assert np.array_equal(mine, lookup_oracle())


def test_pattern_templates_are_permutations():
    """Our per-pattern CLAS templates hold every pattern triangle once, with
    the flipped winding for (x,z,y) aliases (tessellation_table.cpp:130-153)."""
    L = abi.lib()
    n = C.c_uint32(0)
    abi.check(L.tc_table_templates_host(None, None, None, C.byref(n)))
    words = np.zeros(n.value, np.uint32)
    recs = np.zeros(4096 * 32, np.uint8)
    sizes = np.zeros(4096, np.uint32)
    abi.check(L.tc_table_templates_host(recs.ctypes.data, sizes.ctypes.data, words.ctypes.data, C.byref(n)))
    r = recs.view(np.dtype([("nt", "<u4"), ("nv", "<u4"), ("ng", "<u4"), ("cid", "<u4"), ("tri", "<u8"),
                            ("pad", "<u8")]))
    lk = lookup_oracle()
    v, t, c = table_arrays()
    for li in range(4096):
        e = lk[li]
        assert r["nt"][li] == e["numTriangles"]
        if not e["numTriangles"]:
            continue
        f = li + 273
        flipped = ((f >> 8) & 15) > ((f >> 4) & 15)
        w = words[int(r["tri"][li]):int(r["tri"][li]) + int(e["numTriangles"])]
        prim = w >> 24
        assert sorted(prim.tolist()) == list(range(int(e["numTriangles"])))
        src = t[int(e["firstTriangle"]) + prim]
        i0, i1, i2 = src & 0xFF, (src >> 8) & 0xFF, (src >> 16) & 0xFF
        if flipped:
            i1, i2 = i2, i1
        assert np.array_equal(w & 0xFFFFFF, i0 | (i1 << 8) | (i2 << 16))
        assert sizes[li] == L.tc_clas_size(int(e["numTriangles"]), int(e["numVertices"]))
