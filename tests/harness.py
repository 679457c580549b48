"""Runs one frame of the tessellation path through the CPU oracle and through
the HIP library on identical inputs, and returns comparable results.

Device addresses differ between the two runs, so every address-valued field
is turned into an offset from its buffer base (or a template identity)
before comparison.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from vk_tessellated_clusters_amd import abi
from vk_tessellated_clusters_amd.device import cluster_templates, template_records, instance_record

ROOT = Path(__file__).resolve().parent.parent
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

P = C.c_void_p


class OracleCtx(C.Structure):
    _fields_ = [("view", P), ("tessTable", P), ("readback", P), ("instances", P), ("build", P),
                ("displacementTextures", P), ("numRenderInstances", C.c_uint32), ("numDisplacementTextures", C.c_uint32),
                ("genVerticesValue", C.c_uint64), ("viewLast", P), ("hiz", abi.Hiz)]


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        ctx = C.POINTER(OracleCtx)
        for n in ("oracle_instances_classify", "oracle_clusters_cull", "oracle_cluster_classify", "oracle_triangle_split",
                  "oracle_triangle_tess_template_instantiate", "oracle_blas_setup_insertion", "oracle_clas_instantiate",
                  "oracle_clas_build_transient"):
            getattr(L, n).argtypes = [ctx, abi.Config]
            getattr(L, n).restype = None
        L.oracle_build_setup.argtypes = [ctx, abi.Config, C.c_int]
        L.oracle_blas_clusters_insert.argtypes = [ctx, abi.Config, C.c_int]
        L.oracle_frame_tessellate.argtypes = [ctx, abi.Config, C.c_uint32]
        L.oracle_frame_tessellate_mt.argtypes = [ctx, abi.Config, C.c_uint32, C.c_uint32]
        L.oracle_num_split_passes.argtypes = [C.c_uint32]
        L.oracle_num_split_passes.restype = C.c_uint32
        L.oracle_encode_barycentrics.argtypes = [P]
        L.oracle_encode_barycentrics.restype = C.c_uint32
        L.oracle_decode_barycentrics.argtypes = [C.c_uint32, P]
        L.oracle_tess_factors.argtypes = [P, P, P, P, C.c_float, C.c_float, C.c_float, P]
        L.oracle_split_factor.argtypes = [P, C.c_uint32, P]
        L.oracle_get_config.argtypes = [P, P]
        L.oracle_get_config.restype = C.c_uint32
        L.oracle_config_factors.argtypes = [C.c_uint32, P]
        L.oracle_config_triangle_vertices.argtypes = [C.c_uint32, C.c_uint32, P]
        L.oracle_config_vertex_barycentrics.argtypes = [C.c_uint32, C.c_uint32, P]
        L.oracle_sample_height.argtypes = [P, C.c_float, C.c_float]
        L.oracle_sample_height.restype = C.c_float
        L.oracle_hiz_build.argtypes = [C.POINTER(abi.Hiz), P, C.c_uint32]
        L.oracle_hiz_build.restype = None
        L.oracle_intersect_hiz.argtypes = [C.POINTER(abi.Hiz), P, P, P]
        L.oracle_intersect_hiz.restype = C.c_int
        for fn in (L.oracle_ripple_sin, L.oracle_ripple_cos):
            fn.argtypes = [C.c_float]
            fn.restype = C.c_float
        L.oracle_table_lookup.argtypes = [P]
        L.oracle_trace.argtypes = [ctx, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P, C.c_uint32]
        L.oracle_trace.restype = None
        L.oracle_trace_pixels.argtypes = [ctx, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P, C.c_uint32]
        L.oracle_trace_pixels.restype = None
        L.oracle_primary_ray.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P]
        L.oracle_primary_ray.restype = None
        L.oracle_pn.argtypes = [P, P, P, P]
        L.oracle_pn.restype = None
        for n in ("oracle_table_num_vertices", "oracle_table_num_triangles", "oracle_table_num_configs"):
            getattr(L, n).restype = C.c_uint32
        for n, t in (("oracle_table_vertices", C.c_uint32), ("oracle_table_triangles", C.c_uint32),
                     ("oracle_table_configs", C.c_uint16)):
            getattr(L, n).restype = C.POINTER(t)
        _oracle = L
    return _oracle


def small_config(**kw) -> abi.Config:
    c = abi.default_config()
    c.maxVisibleClusters = 1 << 14
    c.maxPartTriangles = 1 << 18
    c.maxSplitTriangles = 1 << 14
    c.maxGeneratedClusters = (1 << 14) + (1 << 18)
    c.maxGeneratedClusterMegs = 512
    c.maxGeneratedVertices = 1 << 22
    for k, v in kw.items():
        setattr(c, k, v)
    return c


ADDR_FIELDS = {"instanceStates", "visibleClusters", "fullClusters", "splitTriangles", "partTriangles",
               "basicClusterSizes", "genClusterData", "genVertices", "tempInstanceIDs", "tempInstantiations",
               "tempClusterAddresses", "tempClusterSizes", "transInstanceIDs", "transBuilds", "transClusterAddresses",
               "transClusterSizes", "transTriMappings", "transTriIndices", "blasBuildInfos", "blasBuildSizes",
               "blasClusterAddresses", "blasBuildData"}


def counters(sb: abi.SceneBuilding) -> dict:
    out = {}
    for name, _ in sb._fields_:
        if name in ADDR_FIELDS or name.startswith("_"):
            continue
        v = getattr(sb, name)
        if isinstance(v, C.Structure):
            v = tuple(getattr(v, f) for f, _ in v._fields_)
        elif isinstance(v, C.Array):
            v = tuple(v)
        out[name] = v
    return out


def readback_dict(rb: abi.Readback) -> dict:
    return {n: (tuple(getattr(rb, n)) if isinstance(getattr(rb, n), C.Array) else getattr(rb, n))
            for n, _ in rb._fields_}


class Results:
    """Normalised frame outputs."""

    def __init__(self):
        self.counters = {}
        self.readback = {}
        self.arrays = {}


def gather_ranges(fetch, base, offsets, sizes) -> np.ndarray:
    """Bytes [base + offsets[i], + sizes[i]) of every range, concatenated
    in list order, from one bulk fetch of the span they cover (frames of a
    few million CLAS would otherwise take one copy per block)."""
    offs = np.asarray(offsets, np.int64)
    lens = np.asarray(sizes, np.int64)
    if offs.size == 0:
        return np.zeros(0, np.uint8)
    lo = int(offs.min())
    hi = int((offs + lens).max())
    pool = fetch(int(base) + lo, hi - lo)
    rel = offs - lo
    if rel[0] == 0 and np.array_equal(rel[1:], (rel + lens)[:-1]):
        return pool[:int(rel[-1] + lens[-1])].copy()  # contiguous in list order
    return np.concatenate([pool[o:o + n] for o, n in zip(rel.tolist(), lens.tolist())])


def _finish_results(sb, rb, fetch, cfg, template_id):
    """fetch(addr, nbytes) -> np.uint8 array.  template_id maps a template
    address to a comparable identity."""
    R = Results()
    R.counters = counters(sb)
    R.readback = readback_dict(rb)
    b = sb
    base = {k: getattr(b, k) for k in ADDR_FIELDS}
    n_inst = b.numRenderInstances
    R.arrays["instanceStates"] = fetch(base["instanceStates"], 4 * n_inst).view(np.uint32)
    nvis = min(b.visibleClusterCounter, cfg.maxVisibleClusters)
    R.arrays["visibleClusters"] = fetch(base["visibleClusters"], 8 * nvis).view(np.uint32)
    R.arrays["splitTriangles"] = fetch(base["splitTriangles"], 24 * cfg.maxSplitTriangles).view(abi.TESS_TRIANGLE_INFO)
    npart = min(b.partTriangleCounter, cfg.maxPartTriangles)
    R.arrays["partTriangles"] = fetch(base["partTriangles"], 24 * npart).view(abi.TESS_TRIANGLE_INFO)
    ntemp = min(b.tempInstantiateCounter, cfg.maxGeneratedClusters)
    ti = fetch(base["tempInstantiations"], 32 * ntemp).view(abi.TEMPLATE_INSTANTIATE_INFO).copy()
    tmpl = [template_id(int(a)) for a in ti["clusterTemplateAddress"]]
    R.arrays["templateIdentity"] = tmpl
    vb = ti["vertexBufferAddress"].astype(np.int64) - np.int64(base["genVertices"])
    ti["vertexBufferAddress"] = vb.astype(np.uint64)
    ti["clusterTemplateAddress"] = 0
    R.arrays["tempInstantiations"] = ti
    R.arrays["tempInstanceIDs"] = fetch(base["tempInstanceIDs"], 4 * ntemp).view(np.uint32)
    ca = fetch(base["tempClusterAddresses"], 8 * ntemp).view(np.uint64)
    R.arrays["tempClusterOffsets"] = ca - np.uint64(base["genClusterData"])
    sizes = fetch(base["tempClusterSizes"], 4 * ntemp).view(np.uint32)
    R.arrays["tempClusterSizes"] = sizes
    nvtx = b.genVertexCounter
    R.genVerticesPrefix = False
    if nvtx > cfg.maxGeneratedVertices:
        # overflow: the accepted vertex allocations form a prefix of the
        # list, so every vertex below the last accepted record's start is
        # written (by records of either kind)
        starts = [int(vb.max()) if ntemp else 0]
        if cfg.flags & abi.TRANSIENT_FLAGS:
            ntr = min(b.transBuildCounter, cfg.maxGeneratedClusters)
            tvb = fetch(base["transBuilds"], 64 * ntr).view(abi.CLAS_BUILD_INFO_DT)["vertexBuffer"]
            if ntr:
                starts.append(int(tvb.max()) - int(base["genVertices"]))
        nvtx = max(starts) // 12
        R.genVerticesPrefix = True
    R.arrays["genVertices"] = fetch(base["genVertices"], 12 * nvtx).view(np.float32).reshape(-1, 3)
    # CLAS blocks, concatenated in CLAS order
    R.arrays["clasBlocks"] = gather_ranges(fetch, base["genClusterData"], R.arrays["tempClusterOffsets"], sizes)
    bi = fetch(base["blasBuildInfos"], 16 * n_inst).view(abi.BLAS_BUILD_INFO).copy()
    bi["clusterReferences"] = (bi["clusterReferences"].astype(np.int64) - np.int64(base["blasClusterAddresses"])
                               ).astype(np.uint64)
    R.arrays["blasBuildInfos"] = bi
    nref = b.blasClusterCounter
    refs = fetch(base["blasClusterAddresses"], 8 * nref).view(np.uint64)
    R.arrays["blasClusterOffsets"] = refs - np.uint64(base["genClusterData"])
    if cfg.flags & abi.TRANSIENT_FLAGS:
        ntr = min(b.transBuildCounter, cfg.maxGeneratedClusters)
        tb = fetch(base["transBuilds"], 64 * ntr).view(abi.CLAS_BUILD_INFO_DT).copy()
        for f in ("vertexBuffer", "indexBuffer"):
            tb[f] = (tb[f].astype(np.int64) - np.int64(base["genVertices"])).astype(np.uint64)
        R.arrays["transBuilds"] = tb
        R.arrays["transInstanceIDs"] = fetch(base["transInstanceIDs"], 4 * ntr).view(np.uint32)
        ta = fetch(base["transClusterAddresses"], 8 * ntr).view(np.uint64)
        R.arrays["transClusterOffsets"] = ta - np.uint64(base["genClusterData"])
        tsz = fetch(base["transClusterSizes"], 4 * ntr).view(np.uint32)
        R.arrays["transClusterSizes"] = tsz
        R.arrays["transClasBlocks"] = gather_ranges(fetch, base["genClusterData"], R.arrays["transClusterOffsets"], tsz)
        # written bytes of every transient part slot: ClusterInfo + mappings
        # (u8 per triangle for 1X subsets, u16 for 2X batches)
        def slot_bytes(r):
            width = 2 if int(r["clusterID"]) >> 30 == abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED else 1
            return 8 + width * (int(r["packed"]) & 511)

        slot_offs = 24 * (tb["clusterID"].astype(np.int64) & 0x3FFFFFFF)
        slot_lens = np.array([slot_bytes(r) for r in tb], np.int64)
        R.arrays["transSlots"] = gather_ranges(fetch, base["partTriangles"], slot_offs, slot_lens)
    return R


# ----------------------------------------------------------------- oracle


class OracleFrame:
    """Host buffers of one frame and the oracle chain over them."""

    def __init__(self, host_scene, cfg: abi.Config, textures=None):
        """textures: (array, format) pairs that replace the scene's height
        maps (e.g. the texels a GpuFrame generated on the device)."""
        L = abi.lib()
        self.cfg = cfg
        self.keep = []
        self.tmpl_ids = {}

        def keep(a):
            a = np.ascontiguousarray(a)
            self.keep.append(a)
            return a

        # table templates with host addresses
        nwords = C.c_uint32(0)
        abi.check(L.tc_table_templates_host(None, None, None, C.byref(nwords)))
        words = keep(np.zeros(max(1, nwords.value), np.uint32))
        recs = keep(np.zeros(4096 * 32, np.uint8))
        sizes = keep(np.zeros(4096, np.uint32))
        abi.check(L.tc_table_templates_host(recs.ctypes.data, sizes.ctypes.data, words.ctypes.data, C.byref(nwords)))
        r = recs.view(np.dtype([("numTriangles", "<u4"), ("numVertices", "<u4"), ("numGroups", "<u4"),
                                ("clusterID", "<u4"), ("triangles", "<u8"), ("_pad", "<u8")]))
        r["triangles"] = np.uint64(words.ctypes.data) + r["triangles"] * np.uint64(4)
        taddr = keep(np.where(r["numTriangles"] > 0,
                              np.uint64(recs.ctypes.data) + np.arange(4096, dtype=np.uint64) * np.uint64(32),
                              np.uint64(0)).astype(np.uint64))
        for li in range(4096):
            if taddr[li]:
                self.tmpl_ids[int(taddr[li])] = ("pattern", li)
        self.table = abi.TessellationTable(0, 0, 0, taddr.ctypes.data, sizes.ctypes.data)

        # geometry with host addresses
        self.geo_addrs = []
        for gi, geo in enumerate(host_scene.geometries):
            w, s = cluster_templates(geo, L)
            w, s = keep(w), keep(s)
            rec = keep(template_records(geo, w.ctypes.data))
            ra = keep(np.uint64(rec.ctypes.data) + np.arange(geo.num_clusters, dtype=np.uint64) * np.uint64(32))
            for c in range(geo.num_clusters):
                self.tmpl_ids[int(ra[c])] = ("cluster", gi, c)
            self.geo_addrs.append({"positions": keep(geo.positions).ctypes.data, "normals": keep(geo.normals).ctypes.data,
                                   "texcoords": keep(geo.texcoords).ctypes.data, "clusters": keep(geo.clusters).ctypes.data,
                                   "clusterLocalTriangles": keep(geo.local_triangles).ctypes.data,
                                   "clusterBboxes": keep(geo.bboxes).ctypes.data, "clusterTemplateAdresses": ra.ctypes.data,
                                   "clusterTemplateInstantiatonSizes": s.ctypes.data})
        from vk_tessellated_clusters_amd.scene import FbmTexture
        self.textures = (abi.DisplacementTexture * max(1, len(host_scene.textures)))()
        for i, tex in enumerate(textures if textures is not None else host_scene.textures):
            arr, fmt = (tex.host(), tex.format) if isinstance(tex, FbmTexture) else tex
            a = keep(arr)
            self.textures[i].texels = a.ctypes.data
            self.textures[i].height, self.textures[i].width = a.shape
            self.textures[i].format = fmt
        self.instances = (abi.RenderInstance * len(host_scene.instances))()
        for i, inst in enumerate(host_scene.instances):
            self.instances[i] = instance_record(inst, host_scene.geometries[inst.geometry],
                                                self.geo_addrs[inst.geometry])
        # frame buffers
        n = len(host_scene.instances)
        self.buf = {}

        def alloc(name, nbytes):
            a = keep(np.zeros(max(16, int(nbytes)), np.uint8))
            self.buf[name] = a
            return a.ctypes.data

        sb = abi.SceneBuilding()
        sb.numRenderInstances = n
        sb.instanceStates = alloc("instanceStates", 4 * n)
        sb.visibleClusters = alloc("visibleClusters", 8 * cfg.maxVisibleClusters)
        sb.splitTriangles = alloc("splitTriangles", 24 * cfg.maxSplitTriangles)
        sb.partTriangles = alloc("partTriangles", 24 * cfg.maxPartTriangles)
        sb.genVertices = alloc("genVertices", 12 * cfg.maxGeneratedVertices)
        sb.tempInstanceIDs = alloc("tempInstanceIDs", 4 * cfg.maxGeneratedClusters)
        sb.tempInstantiations = alloc("tempInstantiations", 32 * cfg.maxGeneratedClusters)
        sb.tempClusterAddresses = alloc("tempClusterAddresses", 8 * cfg.maxGeneratedClusters)
        sb.tempClusterSizes = alloc("tempClusterSizes", 4 * cfg.maxGeneratedClusters)
        sb.blasBuildInfos = alloc("blasBuildInfos", 16 * n)
        sb.blasBuildSizes = alloc("blasBuildSizes", 4 * n)
        sb.blasClusterAddresses = alloc("blasClusterAddresses", 8 * cfg.maxGeneratedClusters)
        if cfg.flags & abi.TRANSIENT_FLAGS:
            sb.transBuilds = alloc("transBuilds", 64 * cfg.maxGeneratedClusters)
            sb.transInstanceIDs = alloc("transInstanceIDs", 4 * cfg.maxGeneratedClusters)
            sb.transClusterAddresses = alloc("transClusterAddresses", 8 * cfg.maxGeneratedClusters)
            sb.transClusterSizes = alloc("transClusterSizes", 4 * cfg.maxGeneratedClusters)
            sb.transTriIndices = sb.genVertices
            sb.transTriMappings = sb.partTriangles
        # the oracle fills up to maxGeneratedClusterMegs (its overflow test);
        # np.zeros pages are only backed once written
        self.clas_bytes = cfg.maxGeneratedClusterMegs << 20
        sb.genClusterData = alloc("genClusterData", self.clas_bytes)
        self.sb = sb
        self.rb = abi.Readback()
        self.view = abi.FrameConstants()
        self.view_last = abi.FrameConstants()
        self.ctx = OracleCtx(C.addressof(self.view), C.addressof(self.table), C.addressof(self.rb),
                             C.addressof(self.instances), C.addressof(self.sb), C.addressof(self.textures), n,
                             len(host_scene.textures), 0)
        self.hiz_texels = None

    def set_culling_history(self, fc_last: abi.FrameConstants = None, depth: np.ndarray = None):
        """Culls against `fc_last` (culling.glsl's viewLast) and, with a
        depth image (H, W float32), the far HiZ the oracle builds from it."""
        if fc_last is None:
            self.ctx.viewLast = None
        else:
            C.memmove(C.addressof(self.view_last), C.addressof(fc_last), C.sizeof(fc_last))
            self.ctx.viewLast = C.addressof(self.view_last)
        if depth is None:
            self.ctx.hiz = abi.Hiz()
            self.hiz_texels = None
            return
        h, w = depth.shape
        hz = abi.Hiz()
        nbytes = C.c_uint64()
        abi.check(abi.lib().tc_hiz_setup(w, h, C.byref(hz), C.byref(nbytes), None, None), "tc_hiz_setup")
        self.hiz_texels = np.zeros(nbytes.value // 4, np.float32)
        hz.texels = self.hiz_texels.ctypes.data
        self.hiz_depth = np.ascontiguousarray(depth, np.float32)
        oracle().oracle_hiz_build(C.byref(hz), self.hiz_depth.ctypes.data, 4)
        self.ctx.hiz = hz

    def run(self, fc: abi.FrameConstants, nthreads: int = 1, stages: str = "all"):
        O = oracle()
        C.memmove(C.addressof(self.view), C.addressof(fc), C.sizeof(fc))
        # frame reset (renderer_raytrace_clusters_tess.cpp:412-419)
        for name, _ in self.sb._fields_:
            if name not in ADDR_FIELDS and name != "numRenderInstances" and not name.startswith("_"):
                v = getattr(self.sb, name)
                if isinstance(v, C.Structure):
                    C.memset(C.addressof(v), 0, C.sizeof(v))
                elif isinstance(v, C.Array):
                    C.memset(C.addressof(v), 0, C.sizeof(v))
                else:
                    setattr(self.sb, name, 0)
        self.sb.viewPos = (abi.f32 * 3)(*fc.viewPos[:3])
        C.memset(C.addressof(self.rb), 0, C.sizeof(self.rb))
        self.buf["splitTriangles"][:] = 0xFF
        ctx = C.byref(self.ctx)
        O.oracle_instances_classify(ctx, self.cfg)
        O.oracle_clusters_cull(ctx, self.cfg)
        if nthreads > 1:
            O.oracle_frame_tessellate_mt(ctx, self.cfg, 0, nthreads)
        else:
            O.oracle_frame_tessellate(ctx, self.cfg, 0)
        if stages == "tess":
            return
        O.oracle_clas_instantiate(ctx, self.cfg)
        O.oracle_clas_build_transient(ctx, self.cfg)
        O.oracle_blas_setup_insertion(ctx, self.cfg)
        O.oracle_blas_clusters_insert(ctx, self.cfg, 0)
        O.oracle_blas_clusters_insert(ctx, self.cfg, 1)

    def trace(self, width, height, flags, nthreads=8, rays=None):
        xf = np.zeros(24 * len(self.instances), np.float32)
        abi.check(abi.lib().tc_instance_xforms_host(C.addressof(self.instances), len(self.instances), xf.ctypes.data))
        n = len(rays) if rays is not None else width * height
        hits = np.zeros(n, abi.HIT_DT)
        dec = np.zeros(n, abi.HIT_DECODED_DT)
        rp = rays.ctypes.data if rays is not None else None
        oracle().oracle_trace(C.byref(self.ctx), xf.ctypes.data, rp, n if rays is not None else 0, width, height, flags,
                              hits.ctypes.data, dec.ctypes.data, nthreads)
        return hits, dec

    def trace_pixels(self, width, height, flags, pixels, nthreads=8):
        """Camera rays of the listed pixels (y * width + x) only."""
        xf = np.zeros(24 * len(self.instances), np.float32)
        abi.check(abi.lib().tc_instance_xforms_host(C.addressof(self.instances), len(self.instances), xf.ctypes.data))
        px = np.ascontiguousarray(pixels, np.uint32)
        hits = np.zeros(px.size, abi.HIT_DT)
        dec = np.zeros(px.size, abi.HIT_DECODED_DT)
        oracle().oracle_trace_pixels(C.byref(self.ctx), xf.ctypes.data, px.ctypes.data, px.size, width, height, flags,
                                     hits.ctypes.data, dec.ctypes.data, nthreads)
        return hits, dec

    def results(self) -> Results:
        def fetch(addr, nbytes):
            return np.frombuffer(C.string_at(addr, nbytes), np.uint8).copy() if nbytes else np.zeros(0, np.uint8)

        return _finish_results(self.sb, self.rb, fetch, self.cfg, lambda a: self.tmpl_ids.get(a, ("unknown", a)))


# -------------------------------------------------------------------- GPU


class GpuFrame:
    def __init__(self, host_scene, cfg: abi.Config):
        from vk_tessellated_clusters_amd.device import DeviceScene, Renderer
        self.cfg = cfg
        self.ds = DeviceScene(host_scene)
        self.r = Renderer(self.ds, cfg)
        L = abi.lib()
        tb = self.r.read_struct(abi.TessellationTable, self.r.bindings.tessTable)
        taddr = self.r.copy(tb.templateAddresses, 8 * 4096).view(np.uint64)
        self.tmpl_ids = {int(a): ("pattern", li) for li, a in enumerate(taddr) if a}
        for gi, a in enumerate(self.ds.geo_addrs):
            geo = host_scene.geometries[gi]
            for c in range(geo.num_clusters):
                self.tmpl_ids[int(a["templateRecords"] + 32 * c)] = ("cluster", gi, c)

    def run(self, fc: abi.FrameConstants, stages: str = "all"):
        self.r.tessellate(fc)
        if stages == "accel":
            self.r.build_accel()
        elif stages != "tess":
            self.r.clas_and_blas()
        abi.check(abi.lib().tc_device_synchronize(), "synchronize")

    def trace(self, width, height, flags, rays=None):
        import torch
        n = len(rays) if rays is not None else width * height
        hits = torch.zeros(n * abi.HIT_DT.itemsize, dtype=torch.uint8, device="cuda")
        dec = torch.zeros(n * abi.HIT_DECODED_DT.itemsize, dtype=torch.uint8, device="cuda")
        p = abi.TraceParams()
        if rays is not None:
            rd = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
            p.rays, p.numRays = rd.data_ptr(), n
        p.width, p.height, p.flags = width, height, flags
        p.hits, p.decoded = hits.data_ptr(), dec.data_ptr()
        self.r.trace(p)
        torch.cuda.synchronize()
        return hits.cpu().numpy().view(abi.HIT_DT), dec.cpu().numpy().view(abi.HIT_DECODED_DT)

    def device_textures(self):
        """Host copies (array, format) of the height maps as the device
        holds them."""
        out = []
        for i in range(self.ds.num_textures):
            t = self.ds.textures[i]
            dt = np.uint16 if t.format == abi.TEXTURE_R16_UNORM else np.float32
            n = t.width * t.height * np.dtype(dt).itemsize
            out.append((self.r.copy(t.texels, n).view(dt).reshape(t.height, t.width), t.format))
        return out

    def accel_arrays(self):
        """Host copies of the built acceleration structure."""
        a, _ = self.r.accel()
        sb = self.r.scene_building()
        n = a.numInstances
        out = {"accel": a, "sb": sb}
        out["tlasInstances"] = self.r.copy(a.tlasInstances, 128 * n).view(abi.TLAS_INSTANCE_DT)
        out["tlasNodes"] = self.r.copy(a.tlasNodes, 64 * n).view(abi.BVH_NODE_DT)
        out["tlasRoot"] = int(self.r.copy(a.tlasRoot, 4).view(np.uint32)[0])
        total = sb.blasClusterCounter
        out["blasNodes"] = self.r.copy(a.blasNodes, 64 * total).view(abi.BVH_NODE_DT)
        out["blasNodes4"] = self.r.copy(a.blasNodes4, 128 * total).view(abi.BVH4_NODE_DT)
        out["tlasNodes4"] = self.r.copy(a.tlasNodes4, 128 * n).view(abi.BVH4_NODE_DT)
        out["blasLeaves"] = self.r.copy(a.blasLeaves, 8 * total).view(np.uint64)
        out["blasLeafInfo"] = self.r.copy(a.blasLeafInfo, 8 * total).view(np.uint64)
        out["refs"] = self.r.copy(sb.blasClusterAddresses, 8 * total).view(np.uint64)
        out["infos"] = self.r.copy(sb.blasBuildInfos, 16 * n).view(abi.BLAS_BUILD_INFO)
        return out

    def results(self) -> Results:
        sb = self.r.scene_building()
        rb = self.r.readback()
        return _finish_results(sb, rb, self.r.copy, self.cfg, lambda a: self.tmpl_ids.get(a, ("unknown", a)))

    def close(self):
        self.r.close()


def compare_hits(gh, oh, gd, od, min_agree=0.9999):
    """Closest-hit identities (instance, cluster, primitive) agree on at
    least min_agree of the rays, t within 1e-5 relative where they agree,
    and the decoded base triangle, sub-triangle, cluster and shadow flag on
    at least min_agree of those rays.  Returns the agreement."""
    gid = np.stack([gh["instanceID"], gh["clusterID"], gh["primitiveID"]], 1)
    oid = np.stack([oh["instanceID"], oh["clusterID"], oh["primitiveID"]], 1)
    same = np.all(gid == oid, axis=1)
    agree = same.mean()
    assert agree >= min_agree, "hit agreement %.6f" % agree
    hit = same & (oh["t"] > 0)
    rel = np.abs(gh["t"][hit] - oh["t"][hit]) / np.maximum(np.abs(oh["t"][hit]), 1e-30)
    assert rel.max(initial=0) <= 1e-5, "hit t differs by %g relative" % rel.max(initial=0)
    for k in ("triangleID", "subTriangleID", "clusterID", "shadowed"):
        assert (gd[k][same] == od[k][same]).mean() >= min_agree, k
    return agree


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|ULP distance| between float32 arrays of equal shape."""
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-2 ** 31) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2 ** 31) - ib, ib)
    return np.abs(ia - ib)


def compare(gpu: Results, ref: Results, max_ulp: int = 2):
    """Raises AssertionError naming the first differing output."""
    assert gpu.readback["errorFlags"] == 0, "frame flagged invalid: errorFlags 0x%x" % gpu.readback["errorFlags"]
    assert gpu.counters == ref.counters, _diff_dict("SceneBuilding", gpu.counters, ref.counters)
    assert gpu.readback == ref.readback, _diff_dict("Readback", gpu.readback, ref.readback)
    keys = ["instanceStates", "visibleClusters", "splitTriangles", "partTriangles", "tempInstantiations",
            "tempInstanceIDs", "tempClusterOffsets", "tempClusterSizes", "blasBuildInfos", "blasClusterOffsets"]
    keys += [k for k in ("transBuilds", "transInstanceIDs", "transClusterOffsets", "transClusterSizes", "transClasBlocks",
                         "transSlots") if k in ref.arrays]
    for k in keys:
        a, b = gpu.arrays[k], ref.arrays[k]
        assert a.shape == b.shape, "%s: shape %s vs %s" % (k, a.shape, b.shape)
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0]
            raise AssertionError("%s differs at %d entries, first %d: gpu %s oracle %s" % (
                k, bad.size, bad[0], a[bad[0]], b[bad[0]]))
    assert gpu.arrays["templateIdentity"] == ref.arrays["templateIdentity"], "template identities differ"
    gv, rv = gpu.arrays["genVertices"], ref.arrays["genVertices"]
    assert gv.shape == rv.shape, "genVertices shape %s vs %s" % (gv.shape, rv.shape)
    if not np.array_equal(gv.view(np.uint32), rv.view(np.uint32)):
        d = ulp_diff(gv, rv)
        assert d.max(initial=0) <= max_ulp, "genVertices differ by %d ULP" % d.max()
    cb, rbk = gpu.arrays["clasBlocks"], ref.arrays["clasBlocks"]
    assert cb.shape == rbk.shape, "CLAS blocks size"
    assert np.array_equal(cb, rbk), "CLAS blocks differ at %d bytes" % int((cb != rbk).sum())


def _diff_dict(name, a, b):
    d = {k: (a.get(k), b.get(k)) for k in set(a) | set(b) if a.get(k) != b.get(k)}
    return "%s differs: %s" % (name, dict(list(d.items())[:12]))


# ------------------------------------------------------- BVH key order


def _spread8(q: np.ndarray) -> np.ndarray:
    """Bits 0..7 of q moved to bits 0, 3, 6, ... 21."""
    q = q.astype(np.uint32)
    q = (q | (q << np.uint32(8))) & np.uint32(0x0300F00F)
    q = (q | (q << np.uint32(4))) & np.uint32(0x030C30C3)
    q = (q | (q << np.uint32(2))) & np.uint32(0x09249249)
    return q


def blas_keys_host(lo: np.ndarray, hi: np.ndarray, geo_lo: np.ndarray, geo_hi: np.ndarray,
                   inst: np.ndarray) -> np.ndarray:
    """64-bit BLAS sort keys recomputed on the host in FP32: the instance
    index above a 24-bit Morton code (8 bits per axis) of each CLAS box
    centre normalised to its instance's undisplaced geometry bounds, with
    the device's operation order ((lo + hi) * 0.5, (c - lo) / extent,
    clamp, truncation of n * 255 + 0.5)."""
    f = np.float32
    c = (lo.astype(f) + hi.astype(f)) * f(0.5)
    ext = (geo_hi - geo_lo).astype(f)
    safe = np.where(ext > 0, ext, f(1))
    n = np.where(ext > 0, (c - geo_lo.astype(f)) / safe, f(0.5)).astype(f)
    n = np.minimum(np.maximum(n, f(0)), f(1))
    q = (n * f(255) + f(0.5)).astype(np.uint32)
    code = _spread8(q[:, 0]) | (_spread8(q[:, 1]) << np.uint32(1)) | (_spread8(q[:, 2]) << np.uint32(2))
    return (inst.astype(np.uint64) << np.uint64(32)) | code.astype(np.uint64)


def check_blas_key_order(g: "GpuFrame", arr: dict) -> int:
    """Asserts that the BVH leaves of every instance are the stable sort of
    the inserted CLAS references by their recomputed 64-bit keys (one sort
    over all instances, as the device runs it); returns the reference
    count."""
    refs = arr["refs"].astype(np.uint64)
    total = refs.size
    if total == 0:
        return 0
    infos = arr["infos"]
    base = int(arr["sb"].blasClusterAddresses)
    counts = infos["clusterReferencesCount"].astype(np.int64)
    starts = (infos["clusterReferences"].astype(np.int64) - base) // 8
    inst = np.zeros(total, np.int64)
    for i in np.nonzero(counts)[0]:
        inst[starts[i]:starts[i] + counts[i]] = i
    # CLAS headers of every reference from one bulk copy of the span
    lo_a = int(refs.min())
    span = g.r.copy(lo_a, int(refs.max()) - lo_a + 48).view(np.uint32)  # blocks are 16-byte aligned
    rel = (refs - np.uint64(lo_a)).astype(np.int64) // 4
    hdr = span[rel[:, None] + np.arange(12)].reshape(-1).view(abi.CLAS_HEADER_DT)
    ri = g.ds.instances
    glo = np.array([list(ri[i].geoLo)[:3] for i in range(len(ri))], np.float32)
    ghi = np.array([list(ri[i].geoHi)[:3] for i in range(len(ri))], np.float32)
    keys = blas_keys_host(hdr["lo"], hdr["hi"], glo[inst], ghi[inst], inst)
    order = np.argsort(keys, kind="stable")
    want = refs[order]
    got = arr["blasLeaves"][:total]
    if not np.array_equal(got, want):
        bad = np.nonzero(got != want)[0]
        raise AssertionError("BVH leaf order differs from the key sort at %d of %d positions (first %d)"
                             % (bad.size, total, bad[0]))
    return total
