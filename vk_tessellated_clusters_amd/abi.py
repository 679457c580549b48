"""ctypes mirror of ``include/tc_abi.h`` and the loader of ``libtc_hip.so``.

The structures here are byte-for-byte the C records (checked against the C
sizes in the tests).  Record arrays read back from device memory are viewed
through the numpy dtypes at the bottom.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtc_hip.so"

u8, u16, u32, i32, u64, f32 = C.c_uint8, C.c_uint16, C.c_uint32, C.c_int32, C.c_uint64, C.c_float

TC_OK = 0
TC_ERROR_INVALID_ARGUMENT = 1
TC_ERROR_CAPACITY = 6
TC_STATUS = {0: "TC_OK", 1: "TC_ERROR_INVALID_ARGUMENT", 2: "TC_ERROR_LAUNCH", 3: "TC_ERROR_OUT_OF_MEMORY",
             4: "TC_ERROR_NO_DEVICE", 5: "TC_ERROR_UNSUPPORTED", 6: "TC_ERROR_CAPACITY"}

TESSTABLE_COORD_MAX = 1 << 15
TESSTABLE_SIZE = 11
INSTANCE_FRUSTUM_BIT = 1
INSTANCE_VISIBLE_BIT = 2

BUILD_SETUP_CLASSIFY = 0
BUILD_SETUP_INSTANTIATE_TESS = 1
BUILD_SETUP_DRAW_TESS = 2
BUILD_SETUP_SPLIT = 3
BUILD_SETUP_SPLIT_PASS = 4
BUILD_SETUP_BUILD_BLAS = 5

RT_CLUSTER_MODE_FULL_CLUSTER = 0
RT_CLUSTER_MODE_SINGLE_TESSELLATED = 1
RT_CLUSTER_MODE_1X_SUBSET_CLUSTER = 2
RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED = 3

READBACK_ERROR_SPLIT_WAIT = 1

TEXTURE_R16_UNORM = 0
TEXTURE_R32_FLOAT = 1

FLAG_DISPLACEMENT = 1 << 0
FLAG_CULLING = 1 << 1
FLAG_ANIMATION = 1 << 2
FLAG_PN = 1 << 3
FLAG_1X_TRANSIENT = 1 << 4
FLAG_2X_TRANSIENT = 1 << 5
TRANSIENT_FLAGS = FLAG_1X_TRANSIENT | FLAG_2X_TRANSIENT  # TESS_USE_TRANSIENTBUILDS
FLAG_PERSISTENT_SPLIT = 1 << 6
FLAG_VERTEX_TAPE = 1 << 7
FLAG_HIZ = 1 << 8

CLAS_GROUP_TRIANGLES = 4


class DispatchIndirect(C.Structure):
    _fields_ = [("gridX", u32), ("gridY", u32), ("gridZ", u32)]


class DrawMeshTasksIndirect(C.Structure):
    _fields_ = [("count", u32), ("first", u32)]


class TessTableEntry(C.Structure):
    _fields_ = [("firstTriangle", u16), ("firstVertex", u16), ("numTriangles", u16), ("numVertices", u16)]


class TessellationTable(C.Structure):
    _fields_ = [("vertices", u64), ("triangles", u64), ("entries", u64), ("templateAddresses", u64),
                ("templateInstantiationSizes", u64)]


class Cluster(C.Structure):
    _fields_ = [("numVertices", u16), ("numTriangles", u16), ("firstTriangle", u32), ("firstLocalVertex", u32),
                ("firstLocalTriangle", u32)]


class RenderInstance(C.Structure):
    _fields_ = [("worldMatrix", f32 * 16), ("geometryID", u32), ("numTriangles", u32), ("numVertices", u32),
                ("numClusters", u32), ("displacementIndex", i32), ("displacementScale", f32),
                ("displacementOffset", f32), ("_pad", f32), ("geoLo", f32 * 4), ("geoHi", f32 * 4),
                ("positions", u64), ("normals", u64), ("texcoords", u64), ("clusters", u64),
                ("clusterLocalTriangles", u64), ("clusterBboxes", u64), ("clusterTemplateAdresses", u64),
                ("clusterTemplateInstantiatonSizes", u64)]


class ClusterInfo(C.Structure):
    _fields_ = [("instanceID", u32), ("clusterID", u32)]


class SceneBuilding(C.Structure):
    _fields_ = [
        ("viewPos", f32 * 3), ("_pad", u32),
        ("numRenderInstances", u32), ("visibleClusterCounter", u32),
        ("fullClusterCounter", u32), ("partTriangleCounter", u32),
        ("dualPartTriangleCounter", u64),
        ("splitTriangleCounter", i32), ("splitReadCounter", u32), ("splitWriteCounter", u32), ("splitPass", u32),
        ("splitPassStart", u32), ("splitPassEnd", u32),
        ("genVertexCounter", u32), ("genClusterCounter", u32), ("genClusterDataCounter", u64),
        ("dispatchClassify", DispatchIndirect), ("dispatchTriangleSplit", DispatchIndirect),
        ("instanceStates", u64), ("visibleClusters", u64), ("fullClusters", u64), ("splitTriangles", u64),
        ("partTriangles", u64),
        ("drawFullClusters", DrawMeshTasksIndirect), ("drawPartTriangles", DrawMeshTasksIndirect),
        ("dispatchClusterInstantiate", DispatchIndirect), ("dispatchTriangleInstantiate", DispatchIndirect),
        ("dispatchBlasTempInsert", DispatchIndirect), ("dispatchBlasTransInsert", DispatchIndirect),
        ("positionTruncateBitCount", u32), ("blasClusterCounter", u32), ("tempInstantiateCounter", u32),
        ("transBuildCounter", u32),
        ("basicClusterSizes", u64), ("genClusterData", u64), ("genVertices", u64),
        ("tempInstanceIDs", u64), ("tempInstantiations", u64), ("tempClusterAddresses", u64),
        ("tempClusterSizes", u64),
        ("transInstanceIDs", u64), ("transBuilds", u64), ("transClusterAddresses", u64), ("transClusterSizes", u64),
        ("transTriMappings", u64), ("transTriIndices", u64),
        ("blasBuildInfos", u64), ("blasBuildSizes", u64), ("blasClusterAddresses", u64), ("blasBuildData", u64),
        ("numBlasReservedSizes", u32),
    ]


class Readback(C.Structure):
    _fields_ = [
        ("numVisibleClusters", u32), ("numFullClusters", u32), ("numSplitTriangles", u32), ("numPartTriangles", u32),
        ("numTotalTriangles", u32), ("numTempInstantiations", u32), ("numGenVertices", u32), ("numBlasClusters", u32),
        ("numTransBuilds", u32), ("numTransPartTriangles", u32), ("numActualTransBuilds", u32),
        ("numActualTempInstantiations", u32), ("numGenDatas", u64), ("numGenActualDatas", u64),
        ("numBlasReservedSizes", u32), ("numBlasActualSizes", u32), ("debugU64", u64),
        ("clusterTriangleId", u32), ("_packedDepth0", u32), ("instanceId", u32), ("_packedDepth1", u32),
        ("debugI", i32), ("debugUI", u32), ("debugF", u32),
        ("debugA", u32 * 64), ("debugB", u32 * 64), ("debugC", u32 * 64), ("errorFlags", u32),
    ]


class SkyParams(C.Structure):
    _fields_ = [("sunDirection", f32 * 3), ("_pad", f32)]


class FrameConstants(C.Structure):
    _fields_ = [
        ("projMatrix", f32 * 16), ("projMatrixI", f32 * 16), ("viewProjMatrix", f32 * 16),
        ("viewProjMatrixI", f32 * 16), ("viewMatrix", f32 * 16), ("viewMatrixI", f32 * 16),
        ("viewPos", f32 * 4), ("viewDir", f32 * 4), ("viewPlane", f32 * 4), ("skyProjMatrixI", f32 * 16),
        ("viewport", i32 * 2), ("viewportf", f32 * 2), ("viewPixelSize", f32 * 2), ("viewClipSize", f32 * 2),
        ("wLightPos", f32 * 3), ("tessRate", f32), ("displacementScale", f32), ("displacementOffset", f32),
        ("lightMixer", f32), ("doShadow", u32), ("wUpDir", f32 * 3), ("sceneSize", f32), ("bgColor", f32 * 4),
        ("lodScale", f32), ("animationState", f32), ("ambientOcclusionRadius", f32),
        ("ambientOcclusionSamples", i32), ("animationRippleEnabled", i32), ("animationRippleFrequency", f32),
        ("animationRippleAmplitude", f32), ("animationRippleSpeed", f32), ("_pad", u32 * 3), ("visualize", u32),
        ("doAnimation", u32), ("flipWinding", u32), ("nearPlane", f32), ("farPlane", f32),
        ("hizSizeFactors", f32 * 4), ("nearSizeFactors", f32 * 4), ("hizSizeMax", f32), ("facetShading", i32),
        ("supersample", i32), ("colorXor", u32), ("dbgUint", u32), ("dbgFloat", f32), ("time", f32), ("frame", u32),
        ("mousePosition", u32 * 2), ("wireThickness", f32), ("wireSmoothing", f32), ("wireColor", f32 * 3),
        ("wireStipple", u32), ("wireBackfaceColor", f32 * 3), ("wireStippleRepeats", f32),
        ("wireStippleLength", f32), ("doWireframe", u32), ("visFilterInstanceID", u32),
        ("visFilterClusterID", u32), ("skyParams", SkyParams),
    ]


class DisplacementTexture(C.Structure):
    _fields_ = [("texels", u64), ("width", u32), ("height", u32), ("format", u32), ("_pad", u32)]


class ClasTemplate(C.Structure):
    _fields_ = [("numTriangles", u32), ("numVertices", u32), ("numGroups", u32), ("clusterID", u32),
                ("triangles", u64), ("_pad", u64)]


class Ray(C.Structure):
    _fields_ = [("origin", f32 * 3), ("tMin", f32), ("direction", f32 * 3), ("tMax", f32)]


class Hit(C.Structure):
    _fields_ = [("t", f32), ("u", f32), ("v", f32), ("instanceID", u32), ("clusterID", u32), ("primitiveID", u32)]


class HitDecoded(C.Structure):
    _fields_ = [("instanceID", u32), ("clusterID", u32), ("triangleID", u32), ("subTriangleID", u32),
                ("baryBase", f32 * 3), ("depth", f32), ("wPos", f32 * 3), ("shadowed", u32)]


class Accel(C.Structure):
    _fields_ = [("blasNodes", C.c_void_p), ("blasLeaves", C.c_void_p), ("tlasInstances", C.c_void_p),
                ("tlasNodes", C.c_void_p), ("tlasRoot", C.c_void_p), ("numInstances", u32), ("capacity", u32),
                ("memory", C.c_void_p), ("memoryBytes", u64), ("blasNodes4", C.c_void_p), ("tlasNodes4", C.c_void_p),
                ("blasLeafInfo", C.c_void_p)]


class TraceParams(C.Structure):
    _fields_ = [("rays", C.c_void_p), ("numRays", u32), ("width", u32), ("height", u32), ("flags", u32),
                ("hits", C.c_void_p), ("decoded", C.c_void_p), ("hitVertices", C.c_void_p)]


TRACE_CULL_BACK = 1
TRACE_SHADOW = 2
TRACE_PRIVATE_STACK = 4
TRACE_PACKETS = 8

RAY_DT = np.dtype([("origin", "<f4", (3,)), ("tMin", "<f4"), ("direction", "<f4", (3,)), ("tMax", "<f4")])
HIT_DT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instanceID", "<u4"), ("clusterID", "<u4"),
                   ("primitiveID", "<u4")])
HIT_DECODED_DT = np.dtype([("instanceID", "<u4"), ("clusterID", "<u4"), ("triangleID", "<u4"),
                           ("subTriangleID", "<u4"), ("baryBase", "<f4", (3,)), ("depth", "<f4"),
                           ("wPos", "<f4", (3,)), ("shadowed", "<u4")])
BVH_NODE_DT = np.dtype([("lo0", "<f4", (3,)), ("child0", "<u4"), ("hi0", "<f4", (3,)), ("child1", "<u4"),
                        ("lo1", "<f4", (3,)), ("pad0", "<u4"), ("hi1", "<f4", (3,)), ("pad1", "<u4")])
CLAS_BUILD_INFO_DT = np.dtype([("clusterID", "<u4"), ("clusterFlags", "<u4"), ("packed", "<u4"),
                               ("baseGeometryIndexAndFlags", "<u4"), ("indexBufferStride", "<u2"),
                               ("vertexBufferStride", "<u2"), ("geometryIndexAndFlagsBufferStride", "<u2"),
                               ("opacityMicromapIndexBufferStride", "<u2"), ("indexBuffer", "<u8"), ("vertexBuffer", "<u8"),
                               ("geometryIndexAndFlagsBuffer", "<u8"), ("opacityMicromapArray", "<u8"),
                               ("opacityMicromapIndexBuffer", "<u8")])
BVH4_NODE_DT = np.dtype([("lo", "<f4", (3, 4)), ("hi", "<f4", (3, 4)), ("child", "<u4", (4,)), ("pad", "<u4", (4,))])
BVH4_EMPTY = 0xFFFFFFFF
TLAS_INSTANCE_DT = np.dtype([("objectToWorld", "<f4", (12,)), ("worldToObject", "<f4", (12,)), ("blasNodes", "<u8"),
                             ("blasLeaves", "<u8"), ("blasRoot", "<u4"), ("numClusters", "<u4"), ("flags", "<u4"),
                             ("pad", "<u4")])
CLAS_HEADER_DT = np.dtype([("lo", "<f4", (3,)), ("clusterID", "<u4"), ("hi", "<f4", (3,)), ("counts", "<u4"),
                           ("vertexBase", "<u4"), ("pad", "<u4", (3,))])
BVH_LEAF = 0x80000000


class Config(C.Structure):
    _fields_ = [("flags", u32), ("maxVisibleClusters", u32), ("maxPartTriangles", u32), ("maxSplitTriangles", u32),
                ("maxGeneratedClusters", u32), ("maxGeneratedClusterMegs", u32), ("maxGeneratedVertices", u32),
                ("clusterVertexCount", u32), ("clusterTriangleCount", u32), ("maxSplitFactor", u32)]


class Hiz(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("size", u32), ("mipLevels", u32), ("usedWidth", u32), ("usedHeight", u32),
                ("width", u32), ("height", u32)]


class Bindings(C.Structure):
    _fields_ = [("view", C.c_void_p), ("tessTable", C.c_void_p), ("readback", C.c_void_p), ("instances", C.c_void_p),
                ("build", C.c_void_p), ("displacementTextures", C.c_void_p), ("numRenderInstances", u32),
                ("numDisplacementTextures", u32), ("workspace", C.c_void_p), ("workspaceBytes", u64),
                ("vertexTape", C.c_void_p), ("viewLast", C.c_void_p), ("hiz", Hiz)]


class Tuning(C.Structure):
    _fields_ = [("classifyBlocksPerCU", u32), ("splitBlocksPerCU", u32), ("instantiateBlocksPerCU", u32),
                ("clasWavesPerCU", u32), ("traceBlocksPerCU", u32), ("traceThreads", u32), ("bvhBlocksPerCU", u32),
                ("splitWaitLimit", u32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if not k.startswith("_")}


class AdjointParams(C.Structure):
    _fields_ = [("rays", C.c_void_p), ("hits", C.c_void_p), ("hitVertices", C.c_void_p), ("gradT", C.c_void_p),
                ("numRays", u32), ("numTextures", u32), ("genVertices", C.c_void_p), ("tape", C.c_void_p),
                ("tlasInstances", C.c_void_p), ("textures", C.c_void_p), ("gradTexels", C.c_void_p)]


VERTEX_TAPE_DT = np.dtype([("base", "<f4", (3,)), ("normal", "<f4", (3,)), ("u", "<f4"), ("v", "<f4"),
                           ("heightScale", "<f4"), ("heightOffset", "<f4"), ("texture", "<u4"), ("pad", "<u4")])


EXPECTED_SIZES = {"TessTableEntry": 8, "Cluster": 16, "RenderInstance": 192, "ClusterInfo": 8,
                  "SceneBuilding": 368, "FrameConstants": 800, "Readback": 880, "DisplacementTexture": 24,
                  "ClasTemplate": 32, "Config": 40}

# ------------------------------------------------------------- numpy views

TESS_TRIANGLE_INFO = np.dtype([("instanceID", "<u4"), ("clusterID", "<u4"), ("vtxEncoded", "<u4", (3,)),
                               ("triangleID_config", "<u4")])
TEMPLATE_INSTANTIATE_INFO = np.dtype([("clusterIdOffset", "<u4"), ("geometryIndexOffset", "<u4"),
                                      ("clusterTemplateAddress", "<u8"), ("vertexBufferAddress", "<u8"),
                                      ("vertexBufferStride", "<u8")])
BLAS_BUILD_INFO = np.dtype([("clusterReferencesCount", "<u4"), ("clusterReferencesStride", "<u4"),
                            ("clusterReferences", "<u8")])
CLUSTER_DT = np.dtype([("numVertices", "<u2"), ("numTriangles", "<u2"), ("firstTriangle", "<u4"),
                       ("firstLocalVertex", "<u4"), ("firstLocalTriangle", "<u4")])
BBOX_DT = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("shortestEdge", "<f4"), ("longestEdge", "<f4")])
TABLE_ENTRY_DT = np.dtype([("firstTriangle", "<u2"), ("firstVertex", "<u2"), ("numTriangles", "<u2"),
                           ("numVertices", "<u2")])

# ----------------------------------------------------------------- loader

_lib = None

EXPORTS = [
    "tc_device_check", "tc_workspace_size", "tc_workspace_init", "tc_table_upload", "tc_table_free",
    "tc_table_lookup_host", "tc_clas_size", "tc_cluster_template_order", "tc_instances_classify", "tc_clusters_cull",
    "tc_build_setup", "tc_cluster_classify", "tc_triangle_split", "tc_triangle_tess_template_instantiate",
    "tc_blas_setup_insertion", "tc_blas_clusters_insert", "tc_frame_tessellate", "tc_frame_tessellate_noalloc",
    "tc_renderer_create",
    "tc_renderer_destroy", "tc_renderer_get_bindings", "tc_renderer_get_building", "tc_renderer_begin_frame",
    "tc_renderer_tessellate", "tc_num_split_passes", "tc_memcpy", "tc_device_synchronize", "tc_clas_instantiate",
    "tc_table_templates_host", "tc_accel_size", "tc_accel_bind", "tc_accel_build", "tc_trace_rays",
    "tc_renderer_get_accel", "tc_renderer_build_accel", "tc_renderer_trace", "tc_renderer_frame",
    "tc_instance_xforms_host", "tc_renderer_graph_capture", "tc_renderer_graph_launch", "tc_renderer_graph_capture_meshwise",
    "tc_fbm_height",
    "tc_renderer_rebind_instances", "tc_depth_adjoint", "tc_get_tuning", "tc_set_tuning", "tc_clas_build_transient",
    "tc_clusterize_mesh", "tc_hiz_setup", "tc_hiz_clear", "tc_hiz_build",
]


def lib():
    """Loads ``libtc_hip.so``; raises if the HIP library was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError("libtc_hip.so is missing at %s: run __graft_entry__.build() first" % LIB_PATH)
    # PyTorch-ROCm ships its own copy of the HIP runtime (same soname as
    # the system one this library links; the first one loaded serves the
    # process).  When this library is loaded first and torch initialises
    # the runtime afterwards, every later allocation made through the
    # library returns hipErrorNoDevice; with torch's runtime loaded and
    # started first, both share the device.  So start it first whenever
    # torch is present.  (Without torch the system runtime works alone.)
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    sig = {
        "tc_device_check": (i32, [C.POINTER(C.c_int)]),
        "tc_workspace_size": (u64, [C.POINTER(Config)]),
        "tc_workspace_init": (i32, [C.POINTER(Bindings), C.POINTER(Config), P]),
        "tc_table_upload": (i32, [C.POINTER(P)]),
        "tc_table_free": (i32, [P]),
        "tc_table_lookup_host": (i32, [P]),
        "tc_clas_size": (u32, [u32, u32]),
        "tc_cluster_template_order": (i32, [u32, u32, P, P, P]),
        "tc_clusterize_mesh": (i32, [P, u32, P, u32, u32, u32, P, P, P, P, C.POINTER(u32), C.POINTER(u32)]),
        "tc_hiz_setup": (i32, [u32, u32, C.POINTER(Hiz), C.POINTER(u64), P, C.POINTER(f32)]),
        "tc_hiz_clear": (i32, [C.POINTER(Hiz), P]),
        "tc_hiz_build": (i32, [C.POINTER(Hiz), P, u32, P]),
        "tc_instances_classify": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_clusters_cull": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_build_setup": (i32, [C.POINTER(Bindings), Config, C.c_int, P]),
        "tc_cluster_classify": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_triangle_split": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_triangle_tess_template_instantiate": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_blas_setup_insertion": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_blas_clusters_insert": (i32, [C.POINTER(Bindings), Config, C.c_int, P]),
        "tc_frame_tessellate": (i32, [C.POINTER(Bindings), Config, u32, P]),
        "tc_frame_tessellate_noalloc": (i32, [C.POINTER(Bindings), Config, u32, P]),
        "tc_renderer_create": (i32, [P, u32, P, u32, C.POINTER(Config), C.POINTER(P)]),
        "tc_renderer_destroy": (i32, [P]),
        "tc_renderer_get_bindings": (i32, [P, C.POINTER(Bindings)]),
        "tc_renderer_get_building": (i32, [P, C.POINTER(SceneBuilding)]),
        "tc_renderer_begin_frame": (i32, [P, C.POINTER(FrameConstants), P]),
        "tc_renderer_tessellate": (i32, [P, C.POINTER(FrameConstants), P]),
        "tc_num_split_passes": (u32, [u32]),
        "tc_memcpy": (i32, [P, P, u64]),
        "tc_device_synchronize": (i32, []),
        "tc_clas_instantiate": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_table_templates_host": (i32, [P, P, P, C.POINTER(u32)]),
        "tc_accel_size": (u64, [u32, u32]),
        "tc_accel_bind": (i32, [P, u64, u32, u32, C.POINTER(Accel)]),
        "tc_accel_build": (i32, [C.POINTER(Bindings), Config, C.POINTER(Accel), P, P]),
        "tc_trace_rays": (i32, [C.POINTER(Bindings), Config, C.POINTER(Accel), C.POINTER(TraceParams), P]),
        "tc_renderer_get_accel": (i32, [P, C.POINTER(Accel), C.POINTER(P)]),
        "tc_renderer_build_accel": (i32, [P, P]),
        "tc_renderer_trace": (i32, [P, C.POINTER(TraceParams), P]),
        "tc_renderer_frame": (i32, [P, C.POINTER(FrameConstants), C.POINTER(TraceParams), P]),
        "tc_instance_xforms_host": (i32, [P, u32, P]),
        "tc_renderer_graph_capture": (i32, [P, C.POINTER(FrameConstants), C.POINTER(TraceParams), P]),
        "tc_renderer_graph_launch": (i32, [P, C.POINTER(FrameConstants), P]),
        "tc_renderer_graph_capture_meshwise": (i32, [P, P, P, u32, C.POINTER(FrameConstants), P]),
        "tc_fbm_height": (i32, [P, u32, u32, u32, u32, u32, P]),
        "tc_depth_adjoint": (i32, [C.POINTER(AdjointParams), P]),
        "tc_clas_build_transient": (i32, [C.POINTER(Bindings), Config, P]),
        "tc_get_tuning": (i32, [C.POINTER(Tuning)]),
        "tc_set_tuning": (i32, [C.POINTER(Tuning)]),
        "tc_renderer_rebind_instances": (i32, [P, P, P, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, what: str = "call"):
    if status != TC_OK:
        raise RuntimeError("%s failed: %s" % (what, TC_STATUS.get(status, status)))


def default_config(**kw) -> Config:
    """The reference's RendererConfig capacities (renderer.hpp:35-68) with
    clusters of up to 128 triangles / vertices."""
    c = Config()
    c.flags = FLAG_DISPLACEMENT | FLAG_CULLING
    c.maxVisibleClusters = 1 << 20
    c.maxPartTriangles = 1 << 20
    c.maxSplitTriangles = 1 << 16
    c.maxGeneratedClusters = (1 << 20) + (1 << 20)
    c.maxGeneratedClusterMegs = 1024
    c.maxGeneratedVertices = 1 << 24
    c.clusterVertexCount = 128
    c.clusterTriangleCount = 128
    c.maxSplitFactor = 8
    for k, v in kw.items():
        setattr(c, k, v)
    return c
