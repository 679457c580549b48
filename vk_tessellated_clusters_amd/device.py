"""Device residency of a synthetic scene and a thin wrapper of the C++
renderer.

PyTorch only provides device memory for the scene arrays and the stream;
every stage runs in ``libtc_hip.so`` through the C ABI.
"""
import ctypes as C

import numpy as np

from . import abi
from .scene import HostScene


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU: the HIP path needs a ROCm device")
    return torch


def cluster_templates(geo, lib=None):
    """Per-cluster template words (group order) and CLAS sizes of a geometry,
    computed by the library's host code."""
    L = lib or abi.lib()
    words = np.zeros(geo.num_triangles, np.uint32)
    sizes = np.zeros(geo.num_clusters, np.uint32)
    for c, cl in enumerate(geo.clusters):
        nt, nv = int(cl["numTriangles"]), int(cl["numVertices"])
        t0, v0 = int(cl["firstTriangle"]), int(cl["firstLocalVertex"])
        lt = np.ascontiguousarray(geo.local_triangles[3 * t0:3 * (t0 + nt)])
        pos = np.ascontiguousarray(geo.positions[v0:v0 + nv])
        out = words[t0:t0 + nt]
        abi.check(L.tc_cluster_template_order(nt, nv, lt.ctypes.data, pos.ctypes.data, out.ctypes.data),
                  "tc_cluster_template_order")
        sizes[c] = L.tc_clas_size(nt, nv)
    return words, sizes


def template_records(geo, words_addr: int):
    """tc_clas_template records of a geometry's clusters; triangles point
    into the word array at `words_addr`."""
    rec = np.zeros(geo.num_clusters, np.dtype([("numTriangles", "<u4"), ("numVertices", "<u4"), ("numGroups", "<u4"),
                                               ("clusterID", "<u4"), ("triangles", "<u8"), ("_pad", "<u8")]))
    nt = geo.clusters["numTriangles"].astype(np.uint32)
    rec["numTriangles"] = nt
    rec["numVertices"] = geo.clusters["numVertices"]
    rec["numGroups"] = (nt + abi.CLAS_GROUP_TRIANGLES - 1) // abi.CLAS_GROUP_TRIANGLES
    rec["clusterID"] = np.arange(geo.num_clusters, dtype=np.uint32)
    rec["triangles"] = np.uint64(words_addr) + geo.clusters["firstTriangle"].astype(np.uint64) * np.uint64(4)
    return rec


def instance_record(inst, geo, addrs: dict) -> abi.RenderInstance:
    r = abi.RenderInstance()
    r.worldMatrix = (abi.f32 * 16)(*np.asarray(inst.world, np.float32).T.reshape(-1).tolist())
    r.geometryID = inst.geometry
    r.numTriangles = geo.num_triangles
    r.numVertices = geo.num_vertices
    r.numClusters = geo.num_clusters
    r.displacementIndex = inst.displacement_index
    r.displacementScale = inst.displacement_scale
    r.displacementOffset = inst.displacement_offset
    r.geoLo = (abi.f32 * 4)(*[float(x) for x in geo.lo], 1.0)
    diag = float(np.linalg.norm(geo.hi.astype(np.float64) - geo.lo))
    r.geoHi = (abi.f32 * 4)(*[float(x) for x in geo.hi], diag)
    for k in ("positions", "normals", "texcoords", "clusters", "clusterLocalTriangles", "clusterBboxes",
              "clusterTemplateAdresses", "clusterTemplateInstantiatonSizes"):
        setattr(r, k, addrs[k])
    return r


class DeviceScene:
    """Uploads a HostScene: geometry arrays, cluster templates, height maps
    and instance / texture records with device addresses."""

    def __init__(self, host: HostScene):
        torch = _torch()
        L = abi.lib()
        self.host = host
        self._keep = []
        dev = torch.device("cuda:0")

        def up(arr):
            t = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1).copy()).to(dev)
            self._keep.append(t)
            return t.data_ptr()

        self.geo_addrs = []
        self.template_words = []
        for geo in host.geometries:
            words, sizes = cluster_templates(geo, L)
            self.template_words.append(words)
            a = {"positions": up(geo.positions), "normals": up(geo.normals), "texcoords": up(geo.texcoords),
                 "clusters": up(geo.clusters), "clusterLocalTriangles": up(geo.local_triangles),
                 "clusterBboxes": up(geo.bboxes), "clusterTemplateInstantiatonSizes": up(sizes)}
            waddr = up(words)
            rec = template_records(geo, waddr)
            raddr = up(rec)
            a["templateRecords"] = raddr
            a["clusterTemplateAdresses"] = up(np.uint64(raddr) + np.arange(geo.num_clusters, dtype=np.uint64) *
                                              np.uint64(32))
            self.geo_addrs.append(a)
        from .scene import FbmTexture
        self.textures = (abi.DisplacementTexture * max(1, len(host.textures)))()
        for i, tex in enumerate(host.textures):
            t = self.textures[i]
            if isinstance(tex, FbmTexture):
                buf = torch.empty(tex.nbytes, dtype=torch.uint8, device=dev)
                self._keep.append(buf)
                abi.check(L.tc_fbm_height(buf.data_ptr(), tex.size, tex.seed, tex.octaves, 8, tex.format, None),
                          "tc_fbm_height")
                t.texels, t.width, t.height, t.format = buf.data_ptr(), tex.size, tex.size, tex.format
            else:
                arr, fmt = tex
                if isinstance(arr, torch.Tensor):
                    # device-resident texels the caller owns (kept alive here)
                    if not arr.is_cuda or not arr.is_contiguous() or arr.dim() != 2:
                        raise ValueError("texture tensors must be contiguous 2-D device tensors")
                    self._keep.append(arr)
                    t.texels = arr.data_ptr()
                else:
                    t.texels = up(arr)
                t.height, t.width = arr.shape
                t.format = fmt
        self.num_textures = len(host.textures)
        self.instances = (abi.RenderInstance * len(host.instances))()
        for i, inst in enumerate(host.instances):
            self.instances[i] = instance_record(inst, host.geometries[inst.geometry], self.geo_addrs[inst.geometry])
        torch.cuda.synchronize()


class Renderer:
    """C++ renderer (tc_renderer_*) over a DeviceScene."""

    def __init__(self, dscene: DeviceScene, config: abi.Config):
        self.L = abi.lib()
        self.scene = dscene
        self.config = config
        h = C.c_void_p()
        abi.check(self.L.tc_renderer_create(C.addressof(dscene.instances), len(dscene.instances),
                                            C.addressof(dscene.textures), dscene.num_textures, C.byref(config),
                                            C.byref(h)), "tc_renderer_create")
        self.handle = h
        self.bindings = abi.Bindings()
        abi.check(self.L.tc_renderer_get_bindings(h, C.byref(self.bindings)))
        self.building = abi.SceneBuilding()
        abi.check(self.L.tc_renderer_get_building(h, C.byref(self.building)))

    def close(self):
        if self.handle:
            self.L.tc_renderer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def stream_ptr(stream=None):
        if stream is None:
            return None
        return C.c_void_p(stream.cuda_stream)

    def tessellate(self, fc: abi.FrameConstants, stream=None):
        abi.check(self.L.tc_renderer_tessellate(self.handle, C.byref(fc), self.stream_ptr(stream)),
                  "tc_renderer_tessellate")

    def build_accel(self, stream=None):
        abi.check(self.L.tc_renderer_build_accel(self.handle, self.stream_ptr(stream)), "tc_renderer_build_accel")

    def accel(self):
        a = abi.Accel()
        xf = C.c_void_p()
        abi.check(self.L.tc_renderer_get_accel(self.handle, C.byref(a), C.byref(xf)))
        return a, xf.value

    def trace(self, params: abi.TraceParams, stream=None):
        abi.check(self.L.tc_renderer_trace(self.handle, C.byref(params), self.stream_ptr(stream)), "tc_renderer_trace")

    def clas_and_blas(self, stream=None):
        s = self.stream_ptr(stream)
        b = C.byref(self.bindings)
        abi.check(self.L.tc_clas_instantiate(b, self.config, s), "tc_clas_instantiate")
        abi.check(self.L.tc_clas_build_transient(b, self.config, s), "tc_clas_build_transient")
        abi.check(self.L.tc_blas_setup_insertion(b, self.config, s), "tc_blas_setup_insertion")
        abi.check(self.L.tc_blas_clusters_insert(b, self.config, 0, s), "tc_blas_clusters_insert")
        abi.check(self.L.tc_blas_clusters_insert(b, self.config, 1, s), "tc_blas_clusters_insert")

    # ------------------------------------------------------------ readback

    def copy(self, addr: int, nbytes: int) -> np.ndarray:
        out = np.empty(max(0, int(nbytes)), np.uint8)
        if nbytes:
            abi.check(self.L.tc_memcpy(out.ctypes.data, C.c_void_p(addr), C.c_uint64(nbytes)), "tc_memcpy")
        return out

    def read_struct(self, ctype, addr):
        raw = self.copy(addr, C.sizeof(ctype))
        return ctype.from_buffer_copy(raw.tobytes())

    def scene_building(self) -> abi.SceneBuilding:
        return self.read_struct(abi.SceneBuilding, self.bindings.build)

    def readback(self) -> abi.Readback:
        return self.read_struct(abi.Readback, self.bindings.readback)
