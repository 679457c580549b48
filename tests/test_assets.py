"""Input formats: the host clusterizer (tc_clusterize_mesh) and the OBJ /
glTF / image loaders.

The reference clusterizes with meshoptimizer, which is not in the
reference tree (a dependency fetched at configure time), so the clusters
themselves are ours and "parity unpinned" against meshoptimizer; the
clusterizer is pinned by what every clustering must satisfy (each triangle
once, with its own vertex positions; limits; bounds and edge lengths by
brute force; determinism) and by its fill on regular meshes.  The loaders
are pinned by files written here with known contents.
"""
import ctypes as C
import json
import base64
import struct

import numpy as np
import pytest

from vk_tessellated_clusters_amd import abi, assets, scene


def grid_mesh(nx, nz, size=1.0):
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(nz + 1), indexing="xy")
    pos = np.stack([ii.reshape(-1) * size / nx, np.zeros((nx + 1) * (nz + 1)), jj.reshape(-1) * size / nz], 1)
    a = (jj[:-1, :-1] * (nx + 1) + ii[:-1, :-1]).reshape(-1)
    tri = np.stack([np.stack([a, a + nx + 1, a + 1], 1), np.stack([a + 1, a + nx + 1, a + nx + 2], 1)], 1).reshape(-1)
    return pos.astype(np.float32), tri.astype(np.uint32)


def check_clustering(pos, idx, g, max_t, max_v):
    nt = idx.size // 3
    assert g.num_triangles == nt
    cl = g.clusters
    assert cl["numTriangles"].max() <= max_t and cl["numVertices"].max() <= max_v
    assert cl["numTriangles"].sum() == nt and cl["numVertices"].sum() == g.num_vertices
    assert np.array_equal(cl["firstLocalTriangle"], 3 * cl["firstTriangle"])
    src = g.source_vertices
    seen = []
    for c in range(g.num_clusters):
        n_t, n_v = int(cl["numTriangles"][c]), int(cl["numVertices"][c])
        t0, v0 = int(cl["firstTriangle"][c]), int(cl["firstLocalVertex"][c])
        lt = g.local_triangles[3 * t0: 3 * (t0 + n_t)].reshape(-1, 3).astype(np.int64)
        assert lt.max() < n_v and len(np.unique(lt)) == n_v  # every cluster vertex is used
        verts = src[v0: v0 + n_v]
        assert len(np.unique(verts)) == n_v
        assert np.array_equal(g.positions[v0: v0 + n_v], pos[verts])
        seen += [tuple(sorted(verts[t])) for t in lt]
        p = pos[verts].astype(np.float32)
        bb = g.bboxes[c]
        assert np.array_equal(bb["lo"], p.min(0)) and np.array_equal(bb["hi"], p.max(0))
        e = np.concatenate([np.linalg.norm(p[lt[:, k]] - p[lt[:, (k + 1) % 3]], axis=1) for k in range(3)])
        assert abs(bb["shortestEdge"] - e.min()) <= 1e-6 * max(1.0, e.min())
        assert abs(bb["longestEdge"] - e.max()) <= 1e-6 * max(1.0, e.max())
    want = sorted(tuple(sorted(t)) for t in idx.reshape(-1, 3).tolist())
    assert sorted(seen) == want  # every triangle exactly once


@pytest.mark.parametrize("max_t,max_v", [(128, 128), (64, 64), (128, 64), (32, 256)])
def test_clusterize_grid(max_t, max_v):
    pos, idx = grid_mesh(64, 48)
    g = scene.mesh_geometry(pos, idx, max_triangles=max_t, max_vertices=max_v)
    check_clustering(pos, idx, g, max_t, max_v)
    # fill: close to the minimum cluster count (a grid cluster with V
    # vertices holds at most 2 (sqrt(V) - 1)^2 triangles), compact clusters
    cap = min(max_t, 2 * (int(np.sqrt(max_v)) - 1) ** 2)
    lower = int(np.ceil(idx.size / 3 / cap))
    assert g.num_clusters <= 1.6 * lower + 1
    ext = g.bboxes["hi"] - g.bboxes["lo"]
    area = ext[:, 0] * ext[:, 2]
    ideal = g.clusters["numTriangles"] / 2 / (64 * 48)  # quads per cluster over the unit square
    assert np.median(area / ideal) < 3.0


def test_clusterize_sphere_and_determinism():
    g0 = scene.icosphere(16)
    # one shared-vertex mesh from the icosphere: merge equal positions
    key, inv = np.unique(g0.positions.round(6), axis=0, return_inverse=True)
    pos = key.astype(np.float32)
    idx = np.concatenate([inv[g0.local_triangles.reshape(-1, 3)[t0:t0 + nt].astype(np.int64) + v0].reshape(-1)
                          for nt, t0, v0 in zip(g0.clusters["numTriangles"], g0.clusters["firstTriangle"],
                                                g0.clusters["firstLocalVertex"])]).astype(np.uint32)
    a = scene.mesh_geometry(pos, idx)
    b = scene.mesh_geometry(pos, idx)
    check_clustering(pos, idx, a, 128, 128)
    for k in ("clusters", "local_triangles", "positions", "bboxes"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    # vertex normals of a sphere point outwards
    n = scene.vertex_normals(pos, idx)
    assert (np.sum(n * pos, axis=1) > 0.99).all()


def test_clusterize_argument_errors():
    L = abi.lib()
    pos, idx = grid_mesh(2, 2)
    out = [np.zeros(64, abi.CLUSTER_DT), np.zeros(64, np.uint32), np.zeros(64, np.uint8), np.zeros(64, abi.BBOX_DT)]
    a, b = C.c_uint32(), C.c_uint32()
    ptrs = [o.ctypes.data for o in out]
    assert L.tc_clusterize_mesh(pos.ctypes.data, 9, idx.ctypes.data, 8, 128, 128, *ptrs, C.byref(a), C.byref(b)) == 0
    assert a.value == 1 and b.value == 9
    assert L.tc_clusterize_mesh(pos.ctypes.data, 8, idx.ctypes.data, 8, 128, 128, *ptrs, C.byref(a), C.byref(b)) == 1
    assert L.tc_clusterize_mesh(pos.ctypes.data, 9, idx.ctypes.data, 8, 129, 128, *ptrs, C.byref(a), C.byref(b)) == 1
    assert L.tc_clusterize_mesh(pos.ctypes.data, 9, idx.ctypes.data, 8, 128, 2, *ptrs, C.byref(a), C.byref(b)) == 1
    assert L.tc_clusterize_mesh(None, 9, idx.ctypes.data, 8, 128, 128, *ptrs, C.byref(a), C.byref(b)) == 1


def test_load_obj(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# quad and a pentagon\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\nv 3 1 0\nv 2 2 0\n"
                 "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                 "vn 0 0 1\n"
                 "f 1/1/1 2/2/1 3/3/1 4/4/1\n"
                 "f -5/1/1 -3/2/1 -2/3/1 -1/4/1 -4/1/1\n")
    m = assets.load_obj(str(p))
    tris = m.indices.reshape(-1, 3)
    assert len(tris) == 2 + 3
    assert np.array_equal(m.positions[tris[0]], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    assert np.array_equal(m.texcoords[tris[1]], [[0, 0], [1, 1], [0, 1]])
    assert (m.normals == [0, 0, 1]).all()
    # negative indices: the pentagon's corners are v3, v5, v6, v7, v4
    assert np.array_equal(m.positions[tris[2]], [[1, 1, 0], [2, 0, 0], [3, 1, 0]])
    g = scene.mesh_geometry(m.positions, m.indices, m.normals, m.texcoords)
    assert g.num_triangles == 5


def write_png16(path, a):
    from PIL import Image
    Image.fromarray(a.astype(np.uint16)).save(path)  # uint16 -> mode I;16


def gltf_doc(tmp_path, glb=False):
    """Two grid meshes, three nodes (a parent with TRS, a child with a
    matrix, a second root), a displacement material over a 16-bit PNG, an
    8-bit PNG on the other mesh, and a perspective camera."""
    from PIL import Image
    pa, ia = grid_mesh(8, 8, 2.0)
    pb, ib = grid_mesh(4, 4, 1.0)
    uv = np.stack([pa[:, 0] / 2, pa[:, 2] / 2], 1).astype(np.float32)
    blob = b"".join([pa.tobytes(), uv.tobytes(), ia.tobytes(), pb.tobytes(), ib.astype(np.uint16).tobytes()])
    offs = np.cumsum([0, pa.nbytes, uv.nbytes, ia.nbytes, pb.nbytes])
    h16 = (np.arange(16 * 16).reshape(16, 16) * 251 % 65536).astype(np.uint16)
    write_png16(tmp_path / "h16.png", h16)
    h8 = (np.arange(8 * 8).reshape(8, 8) * 3 % 256).astype(np.uint8)
    Image.fromarray(h8).save(tmp_path / "h8.png")  # uint8 -> mode L
    views = [dict(buffer=0, byteOffset=int(offs[i]), byteLength=int(n))
             for i, n in enumerate([pa.nbytes, uv.nbytes, ia.nbytes, pb.nbytes, ib.size * 2])]
    acc = [dict(bufferView=0, componentType=5126, count=len(pa), type="VEC3"),
           dict(bufferView=1, componentType=5126, count=len(uv), type="VEC2"),
           dict(bufferView=2, componentType=5125, count=ia.size, type="SCALAR"),
           dict(bufferView=3, componentType=5126, count=len(pb), type="VEC3"),
           dict(bufferView=4, componentType=5123, count=ib.size, type="SCALAR")]
    doc = {
        "asset": {"version": "2.0"},
        "buffers": [dict(byteLength=len(blob))],
        "bufferViews": views, "accessors": acc,
        "images": [{"uri": "h16.png"}, {"uri": "h8.png"}],
        "textures": [{"source": 1}, {"source": 0}],
        "materials": [{"extensions": {"KHR_materials_displacement": {
            "displacementGeometryFactor": 0.4, "displacementGeometryOffset": -0.2,
            "displacementGeometryTexture": {"index": 1}}}},
            {"extensions": {"KHR_materials_displacement": {"displacementGeometryTexture": {"index": 0}}}}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "indices": 2, "material": 0}]},
                   {"primitives": [{"attributes": {"POSITION": 3}, "indices": 4, "material": 1}]}],
        "nodes": [{"mesh": 0, "translation": [1, 2, 3], "rotation": [0, 0.7071067811865476, 0, 0.7071067811865476],
                   "scale": [2, 2, 2], "children": [1]},
                  {"mesh": 1, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1]},
                  {"mesh": 0, "translation": [-4, 0, 0]},
                  {"camera": 0, "translation": [0, 3, 10]}],
        "cameras": [{"type": "perspective", "perspective": {"yfov": 0.8, "znear": 0.1}}],
        "scenes": [{"nodes": [0, 2, 3]}], "scene": 0,
    }
    if glb:
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        b = blob + b"\0" * (-len(blob) % 4)
        data = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(b))
        data += struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(b), 0x004E4942) + b
        path = tmp_path / "s.glb"
        path.write_bytes(data)
    else:
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        path = tmp_path / "s.gltf"
        path.write_text(json.dumps(doc))
    return path, (pa, ia, pb, ib, h16, h8)


@pytest.mark.parametrize("glb", [False, True])
def test_load_gltf(tmp_path, glb):
    path, (pa, ia, pb, ib, h16, h8) = gltf_doc(tmp_path, glb)
    g = assets.Gltf(str(path))
    m0, m1 = g.mesh(0), g.mesh(1)
    assert np.array_equal(m0.positions, pa) and np.array_equal(m0.indices, ia)
    assert np.array_equal(m1.indices, ib.astype(np.uint32)) and m1.normals is None and m1.texcoords is None
    assert m0.displacement == (1, 0.4, -0.2) and m1.displacement == (0, 1.0, 0.0)
    inst = g.instances()
    assert [mi for mi, _ in inst] == [0, 1, 0]
    # parent: T * R(90 deg about Y) * S(2); child adds a translation of 5 in
    # the parent's space
    R = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float64)
    W0 = np.eye(4)
    W0[:3, :3], W0[:3, 3] = 2 * R, [1, 2, 3]
    assert np.allclose(inst[0][1], W0, atol=1e-12)
    T5 = np.eye(4)
    T5[0, 3] = 5
    assert np.allclose(inst[1][1], W0 @ T5, atol=1e-12)
    sc = assets.load_gltf_scene(str(path), width=640, height=360)
    # textures numbered in first use: mesh 0 uses glTF texture 1 (the 16-bit image)
    assert np.array_equal(sc.textures[0][0], h16) and sc.textures[0][1] == abi.TEXTURE_R16_UNORM
    assert np.array_equal(sc.textures[1][0], h8.astype(np.uint16) * 257)
    assert [i.displacement_index for i in sc.instances] == [0, 1, 0]
    assert sc.instances[0].displacement_scale == 0.4 and sc.instances[0].displacement_offset == -0.2
    assert sc.camera.eye == (0.0, 3.0, 10.0) and abs(sc.camera.fovy_deg - np.degrees(0.8)) < 1e-9
    assert sc.camera.width == 640
    assert sc.num_base_triangles == 2 * ia.size // 3 + ib.size // 3
    sc.frame()  # frame constants from the loaded camera


def test_loaded_scene_runs_in_the_oracle(tmp_path):
    """A loaded glTF scene tessellates in the oracle with the displaced
    vertices of its own height map."""
    import harness
    path, _ = gltf_doc(tmp_path)
    sc = assets.load_gltf_scene(str(path), width=320, height=180, tess_rate_pixels=8.0)
    of = harness.OracleFrame(sc, harness.small_config())
    of.run(sc.frame())
    res = of.results()
    assert res.readback["numTotalTriangles"] >= sc.num_base_triangles
