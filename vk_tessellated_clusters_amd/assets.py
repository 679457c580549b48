"""Mesh and scene files: Wavefront OBJ, glTF 2.0 (.gltf / .glb) with the
unofficial ``KHR_materials_displacement`` extension, and height-map images.

The reference loads glTF through cgltf (``src/scene_gltf.cpp:175-511``):
every triangle primitive of a mesh is appended to one geometry, a
primitive's material with ``KHR_materials_displacement`` gives the geometry
its displacement texture, factor and offset (``:315-320``), every node with
a mesh becomes an instance with the node's world transform (``:136-170``),
and the first perspective camera node gives the camera (``:467-483``).
Images are read with stb_image (``scene.cpp:211-256``) and sampled through
their red channel.  This module restates those rules in Python with its own
parser; clustering goes through the library's host clusterizer
(``scene.mesh_geometry``).  Height maps become R16_UNORM textures (8-bit
values v map exactly to v * 257) or stay R32_FLOAT.
"""
import base64
import json
import math
import os
import struct
from dataclasses import dataclass

import numpy as np

from . import abi
from .scene import Camera, HostScene, Instance, mesh_geometry


@dataclass
class Mesh:
    positions: np.ndarray              # (V, 3) f32
    indices: np.ndarray                # (3T,) u32
    normals: np.ndarray = None         # (V, 3) f32 or None
    texcoords: np.ndarray = None       # (V, 2) f32 or None
    displacement: tuple = None         # (texture index, factor, offset) or None


# ------------------------------------------------------------------ OBJ


def load_obj(path: str) -> Mesh:
    """Positions, texture coordinates and normals of every face of an OBJ
    file as one indexed mesh; polygons become triangle fans and every
    distinct (v, vt, vn) corner becomes one vertex."""
    vs, vts, vns = [], [], []
    corners = {}
    pos, uv, nrm, idx = [], [], [], []
    has_uv = has_n = True

    def corner(tok):
        nonlocal has_uv, has_n
        parts = tok.split("/")
        vi = int(parts[0])
        ti = int(parts[1]) if len(parts) > 1 and parts[1] else None
        ni = int(parts[2]) if len(parts) > 2 and parts[2] else None
        vi = vi - 1 if vi > 0 else len(vs) + vi
        if ti is not None:
            ti = ti - 1 if ti > 0 else len(vts) + ti
        if ni is not None:
            ni = ni - 1 if ni > 0 else len(vns) + ni
        has_uv &= ti is not None
        has_n &= ni is not None
        key = (vi, ti, ni)
        k = corners.get(key)
        if k is None:
            k = corners[key] = len(pos)
            pos.append(vs[vi])
            uv.append(vts[ti] if ti is not None else (0.0, 0.0))
            nrm.append(vns[ni] if ni is not None else (0.0, 0.0, 0.0))
        return k

    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "v":
                vs.append(tuple(float(x) for x in t[1:4]))
            elif t[0] == "vt":
                vts.append((float(t[1]), float(t[2]) if len(t) > 2 else 0.0))
            elif t[0] == "vn":
                vns.append(tuple(float(x) for x in t[1:4]))
            elif t[0] == "f":
                ks = [corner(x) for x in t[1:]]
                for i in range(1, len(ks) - 1):
                    idx += [ks[0], ks[i], ks[i + 1]]
    if not idx:
        raise ValueError("%s holds no faces" % path)
    return Mesh(np.asarray(pos, np.float32), np.asarray(idx, np.uint32),
                np.asarray(nrm, np.float32) if has_n else None, np.asarray(uv, np.float32) if has_uv else None)


# ----------------------------------------------------------------- images


def load_height_image(path_or_bytes) -> tuple:
    """(texels, abi.TEXTURE_*) of an image's red channel: 16-bit images and
    8-bit images as R16_UNORM (8-bit v -> v * 257, the same normalised
    value), float images as R32_FLOAT."""
    from PIL import Image
    import io
    img = Image.open(io.BytesIO(path_or_bytes) if isinstance(path_or_bytes, (bytes, bytearray)) else path_or_bytes)
    if img.mode in ("I;16", "I;16B", "I;16L"):
        a = np.asarray(img, dtype=np.uint16)
        return np.ascontiguousarray(a), abi.TEXTURE_R16_UNORM
    if img.mode == "I":
        a = np.asarray(img, dtype=np.int64)
        return np.ascontiguousarray(np.clip(a, 0, 65535).astype(np.uint16)), abi.TEXTURE_R16_UNORM
    if img.mode == "F":
        return np.ascontiguousarray(np.asarray(img, dtype=np.float32)), abi.TEXTURE_R32_FLOAT
    a = np.asarray(img.convert("RGBA") if img.mode not in ("L", "RGB", "RGBA") else img)
    r = a if a.ndim == 2 else a[..., 0]
    return np.ascontiguousarray(r.astype(np.uint16) * np.uint16(257)), abi.TEXTURE_R16_UNORM


# ------------------------------------------------------------------ glTF

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _node_local(node) -> np.ndarray:
    """Row-major 4x4 local transform of a node (matrix, or T * R * S)."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T
    M = np.eye(4)
    S = np.diag(list(node.get("scale", [1.0, 1.0, 1.0])) + [1.0])
    R = np.eye(4)
    R[:3, :3] = _quat_matrix(node.get("rotation", [0.0, 0.0, 0.0, 1.0]))
    M[:3, 3] = node.get("translation", [0.0, 0.0, 0.0])
    return M @ R @ S


class Gltf:
    """A parsed glTF 2.0 document with its binary buffers."""

    def __init__(self, path: str):
        self.path = path
        self.base = os.path.dirname(os.path.abspath(path))
        with open(path, "rb") as f:
            data = f.read()
        glb_bin = None
        if data[:4] == b"glTF":
            version, length = struct.unpack_from("<II", data, 4)
            if version != 2:
                raise ValueError("only glTF 2.0 binaries are supported")
            off, js = 12, None
            while off < length:
                clen, ctype = struct.unpack_from("<II", data, off)
                chunk = data[off + 8: off + 8 + clen]
                if ctype == 0x4E4F534A:
                    js = json.loads(chunk.decode("utf-8"))
                elif ctype == 0x004E4942:
                    glb_bin = chunk
                off += 8 + clen
            self.doc = js
        else:
            self.doc = json.loads(data.decode("utf-8"))
        self.buffers = []
        for i, b in enumerate(self.doc.get("buffers", [])):
            uri = b.get("uri")
            if uri is None:
                if glb_bin is None or i != 0:
                    raise ValueError("buffer %d has no data" % i)
                self.buffers.append(glb_bin)
            else:
                self.buffers.append(self._read_uri(uri))

    def _read_uri(self, uri: str) -> bytes:
        if uri.startswith("data:"):
            return base64.b64decode(uri.split(",", 1)[1])
        from urllib.parse import unquote
        with open(os.path.join(self.base, unquote(uri)), "rb") as f:
            return f.read()

    def view_bytes(self, vi: int) -> tuple:
        v = self.doc["bufferViews"][vi]
        buf = self.buffers[v["buffer"]]
        o = v.get("byteOffset", 0)
        return buf[o: o + v["byteLength"]], v.get("byteStride")

    def accessor(self, ai: int) -> np.ndarray:
        """Accessor data as (count, width) float32 (normalised integers
        converted as the glTF spec says) or (count,) uint32 for integer
        scalars."""
        a = self.doc["accessors"][ai]
        if "sparse" in a:
            raise ValueError("sparse accessors are not supported")
        dt = np.dtype(_COMPONENT[a["componentType"]])
        w, n = _WIDTH[a["type"]], a["count"]
        if "bufferView" in a:
            raw, stride = self.view_bytes(a["bufferView"])
            off = a.get("byteOffset", 0)
            stride = stride or dt.itemsize * w
            arr = np.ndarray((n, w), dt, raw, off, (stride, dt.itemsize))
        else:
            arr = np.zeros((n, w), dt)
        if a["type"] == "SCALAR" and dt.kind in "ui" and not a.get("normalized", False):
            return arr[:, 0].astype(np.uint32)
        out = arr.astype(np.float32)
        if a.get("normalized", False):
            if dt == np.uint8:
                out /= 255.0
            elif dt == np.uint16:
                out /= 65535.0
            elif dt == np.int8:
                out = np.maximum(out / 127.0, -1.0)
            elif dt == np.int16:
                out = np.maximum(out / 32767.0, -1.0)
        return out

    def mesh(self, mi: int) -> Mesh:
        """All triangle primitives of mesh `mi` as one indexed mesh."""
        P, N, U, I = [], [], [], []
        disp = None
        base = 0
        has_n = has_uv = True
        for prim in self.doc["meshes"][mi]["primitives"]:
            if prim.get("mode", 4) != 4 or "POSITION" not in prim.get("attributes", {}):
                continue
            at = prim["attributes"]
            p = self.accessor(at["POSITION"])[:, :3]
            n = p.shape[0]
            P.append(p)
            if "NORMAL" in at:
                N.append(self.accessor(at["NORMAL"])[:, :3])
            else:
                has_n = False
            if "TEXCOORD_0" in at:
                U.append(self.accessor(at["TEXCOORD_0"])[:, :2])
            else:
                has_uv = False
            idx = self.accessor(prim["indices"]) if "indices" in prim else np.arange(n, dtype=np.uint32)
            I.append(idx[: idx.size - idx.size % 3].astype(np.uint32) + np.uint32(base))
            base += n
            mat = prim.get("material")
            if mat is not None:
                ext = self.doc["materials"][mat].get("extensions", {}).get("KHR_materials_displacement")
                if ext and "displacementGeometryTexture" in ext:
                    disp = (ext["displacementGeometryTexture"]["index"], float(ext.get("displacementGeometryFactor", 1.0)),
                            float(ext.get("displacementGeometryOffset", 0.0)))
        if not P:
            return None
        return Mesh(np.ascontiguousarray(np.concatenate(P), np.float32), np.concatenate(I),
                    np.ascontiguousarray(np.concatenate(N), np.float32) if has_n else None,
                    np.ascontiguousarray(np.concatenate(U), np.float32) if has_uv else None, disp)

    def instances(self) -> list:
        """(mesh index, row-major world matrix) of every node with a mesh,
        depth first from the scene's root nodes."""
        nodes = self.doc.get("nodes", [])
        out = []

        def walk(ni, parent):
            node = nodes[ni]
            world = parent @ _node_local(node)
            if "mesh" in node:
                out.append((node["mesh"], world))
            for c in node.get("children", []):
                walk(c, world)

        scenes = self.doc.get("scenes")
        if scenes:
            roots = scenes[self.doc.get("scene", 0)].get("nodes", [])
        else:
            children = {c for n in nodes for c in n.get("children", [])}
            roots = [i for i in range(len(nodes)) if i not in children]
        for r in roots:
            walk(r, np.eye(4))
        return out

    def camera(self) -> tuple:
        """(yfov radians, row-major world matrix) of the first perspective
        camera node, or None."""
        nodes = self.doc.get("nodes", [])
        cams = self.doc.get("cameras", [])
        parents = {c: i for i, n in enumerate(nodes) for c in n.get("children", [])}
        for ni, node in enumerate(nodes):
            if "camera" in node and cams[node["camera"]].get("type") == "perspective":
                world = _node_local(node)
                p = parents.get(ni)
                while p is not None:
                    world = _node_local(nodes[p]) @ world
                    p = parents.get(p)
                return cams[node["camera"]]["perspective"]["yfov"], world
        return None

    def texture_image(self, ti: int) -> tuple:
        img = self.doc["images"][self.doc["textures"][ti]["source"]]
        if "uri" in img:
            return load_height_image(self._read_uri(img["uri"]))
        raw, _ = self.view_bytes(img["bufferView"])
        return load_height_image(bytes(raw))


def scene_from_meshes(meshes: list, instances: list, textures: list, camera: Camera = None, tess_rate_pixels=4.0,
                      name="meshes", max_triangles=128, max_vertices=128) -> HostScene:
    """HostScene of meshes (Mesh), instances ((mesh index, 4x4 world)) and
    textures ((texels, format)); each mesh's displacement (texture,
    factor, offset) becomes its instances' displacement
    (renderer.cpp:205-207).  Without a camera, one looks at the scene's
    bounds from their diagonal."""
    geos = [mesh_geometry(m.positions, m.indices, m.normals, m.texcoords, max_triangles, max_vertices) for m in meshes]
    insts = []
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for mi, world in instances:
        m = meshes[mi]
        d = m.displacement
        insts.append(Instance(mi, np.asarray(world, np.float64), displacement_index=d[0] if d else -1,
                              displacement_scale=d[1] if d else 1.0, displacement_offset=d[2] if d else 0.0))
        g = geos[mi]
        corners = np.array([[x, y, z, 1.0] for x in (g.lo[0], g.hi[0]) for y in (g.lo[1], g.hi[1])
                            for z in (g.lo[2], g.hi[2])]) @ np.asarray(world, np.float64).T
        lo, hi = np.minimum(lo, corners[:, :3].min(0)), np.maximum(hi, corners[:, :3].max(0))
    if camera is None:
        c = (lo + hi) / 2
        r = float(np.linalg.norm(hi - lo)) / 2 or 1.0
        camera = Camera(eye=tuple(c + np.array([0.6, 0.5, 1.0]) * r * 1.6), center=tuple(c), near=r * 0.001,
                        far=r * 20.0)
    return HostScene(geos, insts, textures, camera, tess_rate_pixels, name)


def load_gltf_scene(path: str, width=1920, height=1080, tess_rate_pixels=4.0, max_triangles=128,
                    max_vertices=128) -> HostScene:
    """HostScene of a glTF file, as the reference's loader builds it."""
    g = Gltf(path)
    meshes, remap = [], {}
    for mi in range(len(g.doc.get("meshes", []))):
        m = g.mesh(mi)
        if m is not None:
            remap[mi] = len(meshes)
            meshes.append(m)
    # textures referenced by displacement, numbered in first use
    tex_ids, textures = {}, []
    for m in meshes:
        if m.displacement is not None:
            t = m.displacement[0]
            if t not in tex_ids:
                tex_ids[t] = len(textures)
                textures.append(g.texture_image(t))
            m.displacement = (tex_ids[t], m.displacement[1], m.displacement[2])
    insts = [(remap[mi], w) for mi, w in g.instances() if mi in remap]
    scene = scene_from_meshes(meshes, insts, textures, None, tess_rate_pixels, os.path.basename(path), max_triangles,
                              max_vertices)
    cam = g.camera()
    if cam is not None:
        yfov, world = cam
        # :476-480 the camera looks at the scene's bounds centre
        c = np.array(scene.camera.center)
        scene.camera = Camera(eye=tuple(world[:3, 3]), center=tuple(c), fovy_deg=math.degrees(yfov),
                              near=scene.camera.near, far=scene.camera.far)
    scene.camera.width, scene.camera.height = width, height
    return scene
