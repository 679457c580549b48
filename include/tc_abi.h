/*
 * tc_abi.h -- C ABI of the MI355X tessellated-cluster frame path.
 *
 * Every record below mirrors, field for field and byte for byte, the
 * host/device contract of nvpro-samples/vk_tessellated_clusters
 * (shaders/shaderio_core.h, shaderio_scene.h, shaderio_building.h,
 * shaderio.h).  Device addresses are uint64_t exactly as the reference's
 * C++ side declares BUFFER_REF (shaderio_core.h:29).  Sizes and offsets are
 * pinned with static asserts at the end of this file.
 *
 * Entry points take plain pointers: a host-side tc_bindings (the
 * reference's single descriptor set, shaderio.h:48-58) whose members are
 * device pointers, a tc_config by value (the reference's compile-time
 * macros, renderer_raytrace_clusters_tess.cpp:117-138), and a HIP stream.
 * Calls are asynchronous on the stream.  They return TC_OK or an argument /
 * launch error; data-dependent capacity overflow is never an error return: as
 * in the reference it clamps counters and drops work, visible through
 * tc_readback (build_setup.comp.glsl:105-264).
 */
#ifndef TC_ABI_H
#define TC_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef enum tc_status {
  TC_OK = 0,
  TC_ERROR_INVALID_ARGUMENT = 1,
  TC_ERROR_LAUNCH = 2,
  TC_ERROR_OUT_OF_MEMORY = 3,
  TC_ERROR_NO_DEVICE = 4,
  TC_ERROR_UNSUPPORTED = 5,
  TC_ERROR_CAPACITY = 6
} tc_status;

/* HIP stream handle (hipStream_t), kept opaque so this header needs no HIP. */
typedef void* tc_stream;

/* --------------------------------------------------------------- constants */

#define TC_TESSTABLE_COORD_MAX (1u << 15) /* shaderio_scene.h:25 */
#define TC_TESSTABLE_SIZE 11u             /* shaderio_scene.h:26 */
#define TC_TESSTABLE_LOOKUP_SIZE 16u      /* tessellation_table.cpp:39-40 */
#define TC_TESSTABLE_NUM_LOOKUP (16u * 16u * 16u)

#define TC_INSTANCE_FRUSTUM_BIT 1u /* shaderio_building.h:54 */
#define TC_INSTANCE_VISIBLE_BIT 2u /* shaderio_building.h:55 */

/* build_setup modes, shaderio.h:67-72 */
#define TC_BUILD_SETUP_CLASSIFY 0
#define TC_BUILD_SETUP_INSTANTIATE_TESS 1
#define TC_BUILD_SETUP_DRAW_TESS 2
#define TC_BUILD_SETUP_SPLIT 3
#define TC_BUILD_SETUP_SPLIT_PASS 4
#define TC_BUILD_SETUP_BUILD_BLAS 5

/* cluster id mode bits, shaderio.h:93-99 */
#define TC_RT_CLUSTER_MODE_FULL_CLUSTER 0u
#define TC_RT_CLUSTER_MODE_SINGLE_TESSELLATED 1u
#define TC_RT_CLUSTER_MODE_1X_SUBSET_CLUSTER 2u
#define TC_RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED 3u

/* ------------------------------------------------------- records (L1 mirror) */

/* shaderio_core.h:96-101 */
typedef struct tc_dispatch_indirect {
  uint32_t gridX, gridY, gridZ;
} tc_dispatch_indirect;

/* shaderio_core.h:103-107 */
typedef struct tc_draw_mesh_tasks_indirect {
  uint32_t count, first;
} tc_draw_mesh_tasks_indirect;

/* shaderio_scene.h:65-71 */
typedef struct tc_tess_table_entry {
  uint16_t firstTriangle;
  uint16_t firstVertex;
  uint16_t numTriangles;
  uint16_t numVertices;
} tc_tess_table_entry;

/* shaderio_scene.h:75-87: device addresses of the table arrays */
typedef struct tc_tessellation_table {
  uint64_t vertices;                   /* u32[]: u | v << 16                   */
  uint64_t triangles;                  /* u32[]: 3 x u8 CCW indices            */
  uint64_t entries;                    /* tc_tess_table_entry[4096] lookup     */
  uint64_t templateAddresses;          /* u64[4096] -> tc_clas_template        */
  uint64_t templateInstantiationSizes; /* u32[4096] worst-case CLAS bytes      */
} tc_tessellation_table;

/* shaderio_scene.h:89-95 */
typedef struct tc_bbox {
  float lo[3];
  float hi[3];
  float shortestEdge;
  float longestEdge;
} tc_bbox;

/* shaderio_scene.h:101-111 */
typedef struct tc_cluster {
  uint16_t numVertices;
  uint16_t numTriangles;
  uint32_t firstTriangle;
  uint32_t firstLocalVertex;
  uint32_t firstLocalTriangle;
} tc_cluster;

/* shaderio_scene.h:116-156 (column-major world matrix, glm layout) */
typedef struct tc_render_instance {
  float worldMatrix[16];
  uint32_t geometryID;
  uint32_t numTriangles;
  uint32_t numVertices;
  uint32_t numClusters;
  int32_t displacementIndex;
  float displacementScale;
  float displacementOffset;
  float _pad;
  float geoLo[4];
  float geoHi[4]; /* w = length of the geometry bbox diagonal */
  uint64_t positions;                        /* float3[]                      */
  uint64_t normals;                          /* float3[]                      */
  uint64_t texcoords;                        /* float2[]                      */
  uint64_t clusters;                         /* tc_cluster[]                  */
  uint64_t clusterLocalTriangles;            /* u8[3 * T]                     */
  uint64_t clusterBboxes;                    /* tc_bbox[]                     */
  uint64_t clusterTemplateAdresses;          /* u64[] -> tc_clas_template     */
  uint64_t clusterTemplateInstantiatonSizes; /* u32[] worst-case CLAS bytes   */
} tc_render_instance;

/* shaderio_building.h:62-66 */
typedef struct tc_cluster_info {
  uint32_t instanceID;
  uint32_t clusterID;
} tc_cluster_info;

/* shaderio_building.h:71-82 */
typedef struct tc_sub_triangle_info {
  uint32_t vtxEncoded[3];     /* 2 x 16-bit barycentrics per corner     */
  uint32_t triangleID_config; /* [15:0] triangle, [31:16] config+flip   */
} tc_sub_triangle_info;

/* shaderio_building.h:85-89 */
typedef struct tc_tess_triangle_info {
  tc_cluster_info cluster;
  tc_sub_triangle_info subTriangle;
} tc_tess_triangle_info;

/* shaderio_building.h:130-137 */
typedef struct tc_template_instantiate_info {
  uint32_t clusterIdOffset;
  uint32_t geometryIndexOffset;
  uint64_t clusterTemplateAddress;
  uint64_t vertexBufferAddress;
  uint64_t vertexBufferStride;
} tc_template_instantiate_info;

/* shaderio_building.h:95-126 (ClasBuildInfo): one transient CLAS build.
 * packed = triangleCount[8:0] | vertexCount[17:9] |
 * positionTruncateBitCount[23:18] | indexType[27:24] (1: u8). */
typedef struct tc_clas_build_info {
  uint32_t clusterID;
  uint32_t clusterFlags;
  uint32_t packed;
  uint32_t baseGeometryIndexAndFlags;
  uint16_t indexBufferStride;
  uint16_t vertexBufferStride;
  uint16_t geometryIndexAndFlagsBufferStride;
  uint16_t opacityMicromapIndexBufferStride;
  uint64_t indexBuffer;
  uint64_t vertexBuffer;
  uint64_t geometryIndexAndFlagsBuffer;
  uint64_t opacityMicromapArray;
  uint64_t opacityMicromapIndexBuffer;
} tc_clas_build_info;
#define TC_CLAS_GEOMETRY_OPAQUE (4u << 29) /* ClasGeometryFlag_OPAQUE_BIT_NV */

/* shaderio_building.h:141-149 */
typedef struct tc_blas_build_info {
  uint32_t clusterReferencesCount;
  uint32_t clusterReferencesStride;
  uint64_t clusterReferences;
} tc_blas_build_info;

/* shaderio_building.h:156-263, every field in order */
typedef struct tc_scene_building {
  float viewPos[3];
  uint32_t _pad;

  uint32_t numRenderInstances;
  uint32_t visibleClusterCounter;

  uint32_t fullClusterCounter;
  uint32_t partTriangleCounter;

  uint64_t dualPartTriangleCounter;

  int32_t splitTriangleCounter;
  uint32_t splitReadCounter;
  uint32_t splitWriteCounter;
  uint32_t splitPass;
  uint32_t splitPassStart;
  uint32_t splitPassEnd;

  uint32_t genVertexCounter;
  uint32_t genClusterCounter;
  uint64_t genClusterDataCounter;

  tc_dispatch_indirect dispatchClassify;
  tc_dispatch_indirect dispatchTriangleSplit;

  uint64_t instanceStates;  /* u32[]                   */
  uint64_t visibleClusters; /* tc_cluster_info[]       */
  uint64_t fullClusters;    /* tc_cluster_info[]       */
  uint64_t splitTriangles;  /* tc_tess_triangle_info[] */
  uint64_t partTriangles;   /* tc_tess_triangle_info[] */

  tc_draw_mesh_tasks_indirect drawFullClusters;
  tc_draw_mesh_tasks_indirect drawPartTriangles;

  tc_dispatch_indirect dispatchClusterInstantiate;
  tc_dispatch_indirect dispatchTriangleInstantiate;
  tc_dispatch_indirect dispatchBlasTempInsert;
  tc_dispatch_indirect dispatchBlasTransInsert;

  uint32_t positionTruncateBitCount;

  uint32_t blasClusterCounter;
  uint32_t tempInstantiateCounter;
  uint32_t transBuildCounter;

  uint64_t basicClusterSizes;

  uint64_t genClusterData;
  uint64_t genVertices; /* float3[] */

  uint64_t tempInstanceIDs;      /* u32[]                          */
  uint64_t tempInstantiations;   /* tc_template_instantiate_info[] */
  uint64_t tempClusterAddresses; /* u64[]                          */
  uint64_t tempClusterSizes;     /* u32[]                          */

  uint64_t transInstanceIDs;
  uint64_t transBuilds;
  uint64_t transClusterAddresses;
  uint64_t transClusterSizes;

  uint64_t transTriMappings;
  uint64_t transTriIndices;

  uint64_t blasBuildInfos;       /* tc_blas_build_info[] per instance */
  uint64_t blasBuildSizes;       /* u32[] per instance                */
  uint64_t blasClusterAddresses; /* u64[] per-instance regions        */
  uint64_t blasBuildData;

  uint32_t numBlasReservedSizes;
} tc_scene_building;

/* shaderio.h:263-309 (C++ side layout) */
typedef struct tc_readback {
  uint32_t numVisibleClusters;
  uint32_t numFullClusters;
  uint32_t numSplitTriangles;
  uint32_t numPartTriangles;
  uint32_t numTotalTriangles;
  uint32_t numTempInstantiations;
  uint32_t numGenVertices;
  uint32_t numBlasClusters;
  uint32_t numTransBuilds;
  uint32_t numTransPartTriangles;
  uint32_t numActualTransBuilds;
  uint32_t numActualTempInstantiations;
  uint64_t numGenDatas;
  uint64_t numGenActualDatas;
  uint32_t numBlasReservedSizes;
  uint32_t numBlasActualSizes;
  uint64_t debugU64;
  uint32_t clusterTriangleId;
  uint32_t _packedDepth0;
  uint32_t instanceId;
  uint32_t _packedDepth1;
  int32_t debugI;
  uint32_t debugUI;
  uint32_t debugF;
  uint32_t debugA[64];
  uint32_t debugB[64];
  uint32_t debugC[64];
  /* Not in the reference's Readback: it takes the record's tail padding,
   * so the size and every field above are unchanged.  TC_READBACK_ERROR_*
   * bits of this frame, cleared with the rest of the readback at frame
   * start. */
  uint32_t errorFlags;
} tc_readback;

/* tc_readback.errorFlags: a persistent split (TC_FLAG_PERSISTENT_SPLIT)
 * wait reached its bound, so that launch gave up and its split and part
 * lists and counters are invalid for this frame. */
#define TC_READBACK_ERROR_SPLIT_WAIT 1u

/* Sun direction of the reference's SkySimpleParameters (an nvpro_core2 type
 * that is not part of the reference tree); the path reads only this member
 * (render_raytrace_clusters.rchit.glsl:287-289). */
typedef struct tc_sky_params {
  float sunDirection[3];
  float _pad;
} tc_sky_params;

/* shaderio.h:180-261; all matrices column-major (glm).  The trailing
 * SkySimpleParameters is replaced by tc_sky_params. */
typedef struct tc_frame_constants {
  float projMatrix[16];
  float projMatrixI[16];
  float viewProjMatrix[16];
  float viewProjMatrixI[16];
  float viewMatrix[16];
  float viewMatrixI[16];
  float viewPos[4];
  float viewDir[4];
  float viewPlane[4];
  float skyProjMatrixI[16];
  int32_t viewport[2];
  float viewportf[2];
  float viewPixelSize[2];
  float viewClipSize[2];
  float wLightPos[3];
  float tessRate;
  float displacementScale;
  float displacementOffset;
  float lightMixer;
  uint32_t doShadow;
  float wUpDir[3];
  float sceneSize;
  float bgColor[4];
  float lodScale;
  float animationState;
  float ambientOcclusionRadius;
  int32_t ambientOcclusionSamples;
  int32_t animationRippleEnabled;
  float animationRippleFrequency;
  float animationRippleAmplitude;
  float animationRippleSpeed;
  uint32_t _pad[3];
  uint32_t visualize;
  uint32_t doAnimation;
  uint32_t flipWinding;
  float nearPlane;
  float farPlane;
  float hizSizeFactors[4];
  float nearSizeFactors[4];
  float hizSizeMax;
  int32_t facetShading;
  int32_t supersample;
  uint32_t colorXor;
  uint32_t dbgUint;
  float dbgFloat;
  float time;
  uint32_t frame;
  uint32_t mousePosition[2];
  float wireThickness;
  float wireSmoothing;
  float wireColor[3];
  uint32_t wireStipple;
  float wireBackfaceColor[3];
  float wireStippleRepeats;
  float wireStippleLength;
  uint32_t doWireframe;
  uint32_t visFilterInstanceID;
  uint32_t visFilterClusterID;
  tc_sky_params skyParams;
} tc_frame_constants;

/* ------------------------------------------------ inputs without a Vulkan type */

/* Stands for one entry of `sampler2D displacementTextures[]`
 * (cluster_classify.comp.glsl:122).  Sampled with FP32 bilinear weights,
 * REPEAT addressing, texel centres at (i + 0.5) / size. */
#define TC_TEXTURE_R16_UNORM 0u
#define TC_TEXTURE_R32_FLOAT 1u
typedef struct tc_displacement_texture {
  uint64_t texels; /* device address, row-major, width * height texels */
  uint32_t width;
  uint32_t height;
  uint32_t format; /* TC_TEXTURE_* */
  uint32_t _pad;
} tc_displacement_texture;

/* ----------------------------------------------- own cluster acceleration data */

/* Replaces an NV cluster template (tessellation_table.cpp:102-404,
 * raytracing_cluster_data.cpp:29-280): the triangle order in which a
 * cluster's triangles are grouped for its small BVH.  `triangles` holds
 * numTriangles packed words v0 | v1 << 8 | v2 << 16 | primitiveID << 24,
 * already in group order; groups are runs of TC_CLAS_GROUP_TRIANGLES. */
#define TC_CLAS_GROUP_TRIANGLES 4u
typedef struct tc_clas_template {
  uint32_t numTriangles;
  uint32_t numVertices;
  uint32_t numGroups;
  uint32_t clusterID; /* added to clusterIdOffset when instantiated */
  uint64_t triangles;
  uint64_t _pad;
} tc_clas_template;

/* Header of one instantiated cluster (CLAS) at tempClusterAddresses[i]:
 *   tc_clas_header | groups: numGroups x tc_aabb | tris: u32[numTriangles]
 *   | vertices: float3[numVertices]   (sizes rounded to 16 bytes) */
typedef struct tc_aabb {
  float lo[3];
  float hi[3];
} tc_aabb;

typedef struct tc_clas_header {
  float lo[3];
  uint32_t clusterID; /* mode << 30 | id, the value of gl_ClusterIDNV */
  float hi[3];
  uint32_t counts;     /* numTriangles | numVertices << 8 | numGroups << 16 */
  uint32_t vertexBase; /* genVertices index of the source of vertex 0 */
  uint32_t _pad[3];
} tc_clas_header;

/* Two-child BVH node used by the per-instance BLAS and the TLAS.  A child
 * index with the top bit set is a leaf: bits [30:0] index the leaf list. */
typedef struct tc_bvh_node {
  float lo0[3];
  uint32_t child0;
  float hi0[3];
  uint32_t child1;
  float lo1[3];
  uint32_t _pad0;
  float hi1[3];
  uint32_t _pad1;
} tc_bvh_node;
#define TC_BVH_LEAF 0x80000000u

/* Four-wide view of the same tree, one record per internal node of the
 * two-child tree with the same index: the slots are the node's children,
 * with every internal child replaced by that child's two children (leaf
 * children stay leaves).  Traversal visits every other level of the binary
 * tree.  Unused slots hold TC_BVH4_EMPTY. */
typedef struct tc_bvh4_node {
  float lo[3][4]; /* [axis][slot] */
  float hi[3][4];
  uint32_t child[4]; /* TC_BVH_LEAF | leaf, internal node index, or TC_BVH4_EMPTY */
  uint32_t _pad[4];
} tc_bvh4_node;
#define TC_BVH4_EMPTY 0xFFFFFFFFu

/* One TLAS instance (replaces VkAccelerationStructureInstanceKHR,
 * renderer.cpp:240-266). */
typedef struct tc_tlas_instance {
  float objectToWorld[12]; /* 3x4 row-major */
  float worldToObject[12]; /* 3x4 row-major */
  uint64_t blasNodes;      /* tc_bvh_node[] of this instance's BLAS       */
  uint64_t blasLeaves;     /* u64[] CLAS addresses (clusterReferences)    */
  uint32_t blasRoot;       /* root child word (leaf flag if single CLAS)  */
  uint32_t numClusters;
  uint32_t flags; /* bit 0: force opaque, bit 1: front CCW flip */
  uint32_t _pad;
} tc_tlas_instance;

/* Ray and hit records of the traversal kernel. */
typedef struct tc_ray {
  float origin[3];
  float tMin;
  float direction[3];
  float tMax;
} tc_ray;

typedef struct tc_hit {
  float t;              /* < 0 : miss                                    */
  float u, v;           /* hit barycentrics (gl_HitAttribute)            */
  uint32_t instanceID;  /* gl_InstanceID                                 */
  uint32_t clusterID;   /* gl_ClusterIDNV incl. mode bits                */
  uint32_t primitiveID; /* gl_PrimitiveID within the CLAS                */
} tc_hit;

/* Closest-hit decode output (render_raytrace_clusters.rchit.glsl:128-236,
 * 338-339). */
typedef struct tc_hit_decoded {
  uint32_t instanceID;
  uint32_t clusterID;     /* base cluster of the instance's geometry  */
  uint32_t triangleID;    /* base triangle within that cluster        */
  uint32_t subTriangleID; /* micro triangle within the tessellation   */
  float baryBase[3];      /* barycentrics on the base triangle        */
  float depth;            /* projPos.z / projPos.w, 1.0 on miss       */
  float wPos[3];
  uint32_t shadowed;      /* 1 if the shadow ray hit something        */
} tc_hit_decoded;

/* --------------------------------------------------------- configuration */

#define TC_FLAG_DISPLACEMENT (1u << 0)     /* HAS_DISPLACEMENT_TEXTURES */
#define TC_FLAG_CULLING (1u << 1)          /* DO_CULLING               */
#define TC_FLAG_ANIMATION (1u << 2)        /* DO_ANIMATION             */
#define TC_FLAG_PN (1u << 3)               /* TESS_USE_PN              */
#define TC_FLAG_1X_TRANSIENT (1u << 4)     /* TESS_USE_1X_TRANSIENTBUILDS */
#define TC_FLAG_2X_TRANSIENT (1u << 5)     /* TESS_USE_2X_TRANSIENTBUILDS */
#define TC_FLAG_PERSISTENT_SPLIT (1u << 6) /* TESS_USE_PERSISTENT_KERNEL */
/* No reference counterpart: cluster_classify and
   triangle_tess_template_instantiate also write one tc_vertex_tape record
   per generated vertex (tc_bindings.vertexTape), the input of
   tc_depth_adjoint. */
#define TC_FLAG_VERTEX_TAPE (1u << 7)
/* Occlusion culling of instances against the far HiZ of the previous
   frame's traced depth (instances_classify.comp.glsl:120 with
   culling.glsl:96-114), bound as tc_bindings.hiz; the renderer builds the
   HiZ after every camera trace.  The reference does this whenever culling
   is on; here it is a flag so frames without a HiZ stay reproducible. */
#define TC_FLAG_HIZ (1u << 8)

/* The reference's specialisation macros
 * (renderer_raytrace_clusters_tess.cpp:117-138). */
typedef struct tc_config {
  uint32_t flags;
  uint32_t maxVisibleClusters;      /* MAX_VISIBLE_CLUSTERS        */
  uint32_t maxPartTriangles;        /* MAX_PART_TRIANGLES          */
  uint32_t maxSplitTriangles;       /* MAX_SPLIT_TRIANGLES         */
  uint32_t maxGeneratedClusters;    /* MAX_GENERATED_CLUSTERS      */
  uint32_t maxGeneratedClusterMegs; /* MAX_GENERATED_CLUSTER_MEGS  */
  uint32_t maxGeneratedVertices;    /* MAX_GENERATED_VERTICES      */
  uint32_t clusterVertexCount;      /* CLUSTER_VERTEX_COUNT  (<=255) */
  uint32_t clusterTriangleCount;    /* CLUSTER_TRIANGLE_COUNT (<=128) */
  uint32_t maxSplitFactor;          /* TESS_MAX_SPLIT_FACTOR       */
} tc_config;

/* The reference's descriptor set (shaderio.h:48-58) as device pointers,
 * plus the caller-owned workspace that replaces dispatch-indirect state. */
/* How a generated vertex was displaced (own record, no reference
   counterpart): position = base + normal * (H(u, v) * heightScale +
   heightOffset) with H the bilinear sample of displacementTextures[texture]
   (displacement.glsl, cluster_classify.comp.glsl:477-479 and
   triangle_tess_template_instantiate.comp.glsl, interpolated or cluster
   attributes).  texture is ~0u for a vertex that was not displaced. */
typedef struct tc_vertex_tape {
  float base[3];
  float normal[3]; /* normalised, object space */
  float u, v;
  float heightScale;  /* instance scale * frame scale */
  float heightOffset; /* instance offset + frame offset */
  uint32_t texture;
  uint32_t _pad;
} tc_vertex_tape;

/* Far HiZ pyramid (nvhiz_vk.cpp:278-320 with hizFarLevel 0,
 * nvhiz-update.comp.glsl): level 0 holds the maximum depth of each 2 x 2
 * block of the width x height depth image, every further level the maximum
 * of 2 x 2 texels of the level before, in a size x size allocation (size the
 * power of two >= max(width, height) / 2); levels are stored back to back,
 * level k being (size >> k)^2 floats. */
typedef struct tc_hiz {
  float* texels;       /* device memory, tc_hiz_setup's byte count */
  uint32_t size;
  uint32_t mipLevels;
  uint32_t usedWidth;  /* width / 2 */
  uint32_t usedHeight; /* height / 2 */
  uint32_t width;      /* source depth image */
  uint32_t height;
} tc_hiz;

typedef struct tc_bindings {
  const tc_frame_constants* view;                      /* FRAME_UBO        */
  const tc_tessellation_table* tessTable;              /* TESSTABLE_UBO    */
  tc_readback* readback;                               /* READBACK_SSBO    */
  const tc_render_instance* instances;                 /* RENDERINSTANCES  */
  tc_scene_building* build;                            /* SCENEBUILDING    */
  const tc_displacement_texture* displacementTextures; /* DISPLACED_TEXTURES */
  uint32_t numRenderInstances;
  uint32_t numDisplacementTextures;
  void* workspace;         /* tc_workspace_size() bytes, device memory */
  uint64_t workspaceBytes;
  struct tc_vertex_tape* vertexTape; /* maxGeneratedVertices records, only read
                                        with TC_FLAG_VERTEX_TAPE */
  /* culling.glsl's viewLast: the frame constants the HiZ depth was traced
   * with (NULL: view); used by tc_instances_classify */
  const tc_frame_constants* viewLast;
  tc_hiz hiz; /* texels NULL: no occlusion test (with TC_FLAG_HIZ) */
} tc_bindings;

/* ------------------------------------------------------------ entry points */

/* Device count / basic probe.  Returns TC_ERROR_NO_DEVICE without a GPU. */
tc_status tc_device_check(int* deviceCount);

/* Launch shapes of the work-queue kernels (no reference counterpart: the
 * reference's dispatch sizes are fixed by its shaders and indirect
 * arguments).  Grids are given in workgroups per CU; the work-queue kernels
 * take their tiles by ticket, so every setting produces identical results
 * and only the speed changes.  traceThreads is the traversal workgroup
 * size (64, 128 or 256).  Settings are process-wide and apply to later
 * calls (a captured graph keeps the shapes it was captured with); a NULL
 * argument to tc_set_tuning restores the defaults. */
typedef struct tc_tuning {
  uint32_t classifyBlocksPerCU;    /* cluster_classify, 256-thread workgroups */
  uint32_t splitBlocksPerCU;       /* triangle_split, 64-thread workgroups */
  uint32_t instantiateBlocksPerCU; /* template instantiate, 256-thread workgroups */
  uint32_t clasWavesPerCU;         /* CLAS build, one wave per CLAS */
  uint32_t traceBlocksPerCU;       /* traversal workgroups per CU */
  uint32_t traceThreads;           /* traversal workgroup size */
  uint32_t bvhBlocksPerCU;         /* BVH keys / hierarchy / collapse */
  /* Sleeps before a wait of the persistent split gives up and flags the
   * frame (TC_READBACK_ERROR_SPLIT_WAIT); default 2^24, and 0 in
   * tc_set_tuning restores it.  Not a performance knob: only failure-path
   * tests lower it. */
  uint32_t splitWaitLimit;
} tc_tuning;

tc_status tc_get_tuning(tc_tuning* out);
tc_status tc_set_tuning(const tc_tuning* tuning);

/* Bytes of workspace the stage kernels need for `config`. */
uint64_t tc_workspace_size(const tc_config* config);

/* One-time workspace initialisation (clears tile status words). */
tc_status tc_workspace_init(const tc_bindings* bindings, const tc_config* config, tc_stream stream);

/* Uploads the built-in table: vertices, triangles, the 16^3 lookup
 * (tessellation_table.cpp:36-100) and per-pattern templates with their
 * instantiation sizes.  Allocates device memory owned by the library until
 * tc_table_free.  *tableDev receives a device pointer to a
 * tc_tessellation_table. */
tc_status tc_table_upload(tc_tessellation_table** tableDev);
tc_status tc_table_free(tc_tessellation_table* tableDev);

/* Host copy of the expanded 16^3 lookup (4096 entries) for inspection. */
tc_status tc_table_lookup_host(tc_tess_table_entry* out4096);

/* Host copy of the per-pattern templates: out4096[i].triangles holds the
 * word offset of lookup entry i's triangle words in `words` (*numWords
 * words; pass words = NULL to query the count), sizes4096 the CLAS sizes. */
tc_status tc_table_templates_host(tc_clas_template* out4096, uint32_t* sizes4096, uint32_t* words, uint32_t* numWords);

/* Synchronous copy between any host / device pointers (hipMemcpyDefault),
 * after all device work queued so far on any stream has finished. */
tc_status tc_memcpy(void* dst, const void* src, uint64_t bytes);
tc_status tc_device_synchronize(void);

/* Worst-case CLAS bytes of a cluster with the given counts, the quantity the
 * reference takes from its template builds (raytracing_cluster_data.cpp:362-388). */
uint32_t tc_clas_size(uint32_t numTriangles, uint32_t numVertices);

/* Host-side mesh clustering (stands for meshopt_buildMeshletsSpatial and
 * the cluster bounds / per-cluster vertices of scene.cpp:396-552):
 * clusters of at most maxTriangles (<= 128) triangles and maxVertices
 * (<= 256) distinct vertices by recursive bisection of the triangle
 * centroids, deterministic.  Outputs, sized by the caller: clusters and
 * clusterBboxes numTriangles entries, clusterLocalVertices and
 * clusterLocalTriangles 3 x numTriangles entries.  clusterLocalVertices[i]
 * is the mesh vertex behind cluster vertex i (the caller gathers the
 * per-cluster vertex arrays from it); clusterLocalTriangles holds 3 u8
 * cluster-local indices per triangle.  Returns the cluster and cluster
 * vertex counts. */
tc_status tc_clusterize_mesh(const float* positions, uint32_t numVertices, const uint32_t* indices,
                             uint32_t numTriangles, uint32_t maxTriangles, uint32_t maxVertices, tc_cluster* clusters,
                             uint32_t* clusterLocalVertices, uint8_t* clusterLocalTriangles, tc_bbox* clusterBboxes,
                             uint32_t* numClusters, uint32_t* numClusterVertices);

/* Builds a per-cluster template in host memory: writes the packed triangle
 * words in group order into trianglesOut (numTriangles u32).
 * localTriangles: 3 x u8 per triangle; positions: float3 per vertex. */
tc_status tc_cluster_template_order(uint32_t numTriangles, uint32_t numVertices, const uint8_t* localTriangles,
                                    const float* positions, uint32_t* trianglesOut);

/* instances_classify.comp.glsl:100-127: per-instance frustum, size and
 * (with TC_FLAG_HIZ and a bound HiZ) occlusion state bits against viewLast. */
tc_status tc_instances_classify(const tc_bindings* b, tc_config config, tc_stream stream);

/* Host: the HiZ layout for a width x height depth image (6 .. 65536 per
 * axis; nvhiz_vk.cpp:
 * 278-320): size, levels and used texels into *hiz (texels untouched), its
 * byte count, and the frame constants' hizSizeFactors / hizSizeMax
 * (nvhiz_vk.cpp:29-40; tessellatedclusters.cpp:615-619). */
tc_status tc_hiz_setup(uint32_t width, uint32_t height, tc_hiz* hiz, uint64_t* bytes, float sizeFactors[4],
                       float* sizeMax);

/* Fills every HiZ texel with the far depth 1.0 (an unbuilt HiZ occludes
 * nothing). */
tc_status tc_hiz_clear(const tc_hiz* hiz, tc_stream stream);

/* nvhiz-update.comp.glsl: builds every level of the far HiZ from a depth
 * image of hiz->width x hiz->height floats, pixel p at
 * depth + p * pixelStride bytes (row-major; e.g. tc_hit_decoded.depth with
 * pixelStride sizeof(tc_hit_decoded)); source reads clamp to the image. */
tc_status tc_hiz_build(const tc_hiz* hiz, const void* depth, uint32_t pixelStride, tc_stream stream);

/* clusters_cull.comp.glsl:112-161 in ray-tracing mode: appends every
 * cluster of every instance, instance by instance, to build.visibleClusters. */
tc_status tc_clusters_cull(const tc_bindings* b, tc_config config, tc_stream stream);

/* build_setup.comp.glsl:102-264, one device thread. */
tc_status tc_build_setup(const tc_bindings* b, tc_config config, int mode, tc_stream stream);

/* cluster_classify.comp.glsl:150-907 over build.visibleClusters. */
tc_status tc_cluster_classify(const tc_bindings* b, tc_config config, tc_stream stream);

/* triangle_split.comp.glsl:598-632, one pass over
 * [build.splitPassStart, build.splitPassEnd); with TC_FLAG_PERSISTENT_SPLIT
 * (:476-594) every pass, SPLIT_PASS setups included, in one launch. */
tc_status tc_triangle_split(const tc_bindings* b, tc_config config, tc_stream stream);

/* triangle_tess_template_instantiate.comp.glsl:126-373 over build.partTriangles. */
tc_status tc_triangle_tess_template_instantiate(const tc_bindings* b, tc_config config, tc_stream stream);

/* blas_setup_insertion.comp.glsl:100-116 (exclusive scan over instances).
 * Lays every instance's reference list out after build.blasClusterCounter
 * and advances it, as the shader does: run it once per frame, after the
 * frame's reset.  tc_accel_build needs the lists to tile
 * [0, blasClusterCounter) exactly, which a second run in the same frame
 * breaks (the first lists then belong to no instance). */
tc_status tc_blas_setup_insertion(const tc_bindings* b, tc_config config, tc_stream stream);

/* blas_clusters_insert.comp.glsl:97-135; transient = push constant 1. */
tc_status tc_blas_clusters_insert(const tc_bindings* b, tc_config config, int transient, tc_stream stream);

/* Runs the reference's per-frame chain classify .. insert
 * (renderer_raytrace_clusters_tess.cpp:410-686) with numSplitPasses split
 * passes (0: derive from maxSplitFactor as the reference does). */
tc_status tc_frame_tessellate(const tc_bindings* b, tc_config config, uint32_t numSplitPasses, tc_stream stream);

/* No counterpart in the reference: tc_frame_tessellate for callers that
 * cannot allocate device memory or bring a workspace (code under graph
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * look-back state lives in a library-owned device array of 1 MiB + 1 KiB
 * per device (b->workspace is ignored and may be null), so it calls no
 * allocator.  Limits: ceil(maxVisibleClusters / 64), ceil(maxSplitTriangles
 * / 64) and ceil(maxPartTriangles / 1024) must each be at most 16384, else
 * TC_ERROR_CAPACITY.  Calls that overlap on one device race on that array;
 * calls in stream order do not. */
tc_status tc_frame_tessellate_noalloc(const tc_bindings* b, tc_config config, uint32_t numSplitPasses,
                                      tc_stream stream);

/* Our CLAS build (replaces the NV template instantiation,
 * renderer_raytrace_clusters_tess.cpp:595-623): one block per
 * tempInstantiations entry at tempClusterAddresses[i]; writes
 * tempClusterSizes[i]. */
tc_status tc_clas_instantiate(const tc_bindings* b, tc_config config, tc_stream stream);

/* Our CLAS build from transient build records (replaces the NV transient
 * CLAS build, renderer_raytrace_clusters_tess.cpp:625-655; with
 * TC_FLAG_1X_TRANSIENT, else a no-op): one block per transBuilds entry at
 * transClusterAddresses[i]; writes transClusterSizes[i]. */
tc_status tc_clas_build_transient(const tc_bindings* b, tc_config config, tc_stream stream);

/* ---- cluster acceleration structures (replace the NV BLAS build and the
 *      TLAS build, renderer_raytrace_clusters_tess.cpp:695-745) ---- */

/* Device arrays of the two-level structure built each frame.  blasNodes and
 * blasLeaves are indexed per instance from the instance's first reference
 * (tc_tlas_instance.blasNodes / blasLeaves point at that offset). */
typedef struct tc_accel {
  tc_bvh_node* blasNodes;
  uint64_t* blasLeaves; /* CLAS address per sorted reference */
  tc_tlas_instance* tlasInstances;
  tc_bvh_node* tlasNodes;
  uint32_t* tlasRoot; /* [0] root child word */
  uint32_t numInstances;
  uint32_t capacity; /* CLAS references (>= config.maxGeneratedClusters) */
  void* memory;
  uint64_t memoryBytes;
  tc_bvh4_node* blasNodes4; /* four-wide view of blasNodes, same indices */
  tc_bvh4_node* tlasNodes4; /* four-wide view of tlasNodes */
  uint64_t* blasLeafInfo;   /* per sorted reference: its CLAS header's counts
                               | clusterID << 32 (read with blasLeaves, so the
                               traversal need not wait for the header) */
} tc_accel;

/* Bytes of device memory for tc_accel_bind. */
uint64_t tc_accel_size(uint32_t capacity, uint32_t numInstances);
/* Carves the arrays of a tc_accel out of caller-owned device memory. */
tc_status tc_accel_bind(void* memory, uint64_t bytes, uint32_t capacity, uint32_t numInstances, tc_accel* out);
/* Per-instance LBVH over the CLAS of build.blasBuildInfos (after
 * tc_blas_clusters_insert) and the TLAS over the instances.  The instances'
 * reference lists must tile [0, build.blasClusterCounter), as one
 * tc_blas_setup_insertion per frame leaves them.
 * instanceXforms: device array, per instance objectToWorld then
 * worldToObject, each 3x4 row-major (24 floats). */
tc_status tc_accel_build(const tc_bindings* b, tc_config config, const tc_accel* accel, const float* instanceXforms,
                         tc_stream stream);

/* ---- ray tracing (vkCmdTraceRaysKHR with render_raytrace.rgen.glsl,
 *      render_raytrace_clusters.rchit.glsl, render_shading.glsl:263-274) ---- */

#define TC_TRACE_CULL_BACK (1u << 0) /* gl_RayFlagsCullBackFacingTrianglesEXT on primary rays */
#define TC_TRACE_SHADOW (1u << 1)    /* one shadow ray toward skyParams.sunDirection per hit */
#define TC_TRACE_PRIVATE_STACK (1u << 2) /* keep the whole traversal stack in private memory
                                            (slower; exercises the overflow part of the stack) */
#define TC_TRACE_PACKETS (1u << 3) /* wave-wide packet traversal (shared node stack); same results,
                                      faster for coherent rays */

/* Rays are either explicit (rays != NULL, numRays) or the primary rays of
 * the frame camera, one per pixel of width x height (rgen:52-86).  Hits are
 * closest hits with ties broken by (instance, cluster, primitive).  Front
 * faces are counter-clockwise seen from the ray origin (right-hand normal
 * toward the origin). */
typedef struct tc_trace_params {
  const tc_ray* rays; /* device, or NULL for camera rays */
  uint32_t numRays;
  uint32_t width;
  uint32_t height;
  uint32_t flags; /* TC_TRACE_* */
  tc_hit* hits;            /* device, per ray (t < 0: miss), may be NULL */
  tc_hit_decoded* decoded; /* device, per ray, may be NULL */
  uint32_t* hitVertices;   /* device, 4 per ray: genVertices indices of the hit
                              triangle's vertices (~0 on a miss), may be NULL */
} tc_trace_params;

tc_status tc_trace_rays(const tc_bindings* b, tc_config config, const tc_accel* accel, const tc_trace_params* params,
                        tc_stream stream);

/* Adjoint of hit distance with respect to height-map texels (no reference
 * counterpart; the reference is not differentiable).  For every ray with
 * a hit and gradT[r] != 0 it adds
 *   gradT[r] * dt/dH(texel)
 * to gradTexels[texture][texel], where t is the distance along the ray to
 * the plane of the hit triangle, the triangle's vertices are the
 * genVertices named by hitVertices, and each vertex moves along its tape
 * normal by heightScale * H(u, v), H the bilinear REPEAT sample.  The
 * tessellation pattern and the hit triangle are those of the forward
 * pass (frozen).  With n the object-space normal of the hit triangle, d
 * the object-space ray direction and b its barycentrics,
 *   dt/dP_k = b_k n / (n . d),  dP_k/dH = heightScale * normal_k.
 * Gradients are accumulated with float atomics (summation order varies). */
typedef struct tc_adjoint_params {
  const tc_ray* rays;          /* device, the traced rays (world space) */
  const tc_hit* hits;          /* device, from tc_trace_rays */
  const uint32_t* hitVertices; /* device, 4 per ray, from tc_trace_rays */
  const float* gradT;          /* device, dL/dt per ray */
  uint32_t numRays;
  uint32_t numTextures;
  const float* genVertices;                      /* device, SceneBuilding.genVertices */
  const tc_vertex_tape* tape;                    /* device, tc_bindings.vertexTape */
  const tc_tlas_instance* tlasInstances;         /* device, tc_accel.tlasInstances */
  const tc_displacement_texture* textures;       /* device, numTextures records */
  float* const* gradTexels; /* device array of numTextures device pointers
                               (width * height floats each, NULL: skip) */
} tc_adjoint_params;

tc_status tc_depth_adjoint(const tc_adjoint_params* params, tc_stream stream);

/* Synthetic workload input: a size x size tileable value-noise fBm height
 * map (octaves from baseFreq, integer lattice hash) written to device
 * memory as R16_UNORM or R32_FLOAT.  No reference counterpart: the
 * reference loads height maps from glTF assets. */
tc_status tc_fbm_height(void* texelsDev, uint32_t size, uint32_t seed, uint32_t octaves, uint32_t baseFreq, uint32_t format,
                        tc_stream stream);

/* ---- frame driver (renderer_raytrace_clusters_tess.cpp:166-310, 410-780) ---- */

typedef struct tc_renderer tc_renderer;

/* Allocates every per-frame buffer for `config` (one device arena), copies
 * the instance and texture records (whose addresses are device memory the
 * caller keeps alive) and uploads the tessellation table. */
tc_status tc_renderer_create(const tc_render_instance* instancesHost, uint32_t numInstances,
                             const tc_displacement_texture* texturesHost, uint32_t numTextures, const tc_config* config,
                             tc_renderer** out);
tc_status tc_renderer_destroy(tc_renderer* renderer);
/* Device pointers of the renderer's bindings, for the stage entry points. */
tc_status tc_renderer_get_bindings(tc_renderer* renderer, tc_bindings* out);
/* Initial per-frame SceneBuilding (device addresses of all buffers). */
tc_status tc_renderer_get_building(tc_renderer* renderer, tc_scene_building* out);
/* Uploads frame constants, resets SceneBuilding / Readback, fills the split
 * list with ~0 (renderer_raytrace_clusters_tess.cpp:412-419). */
tc_status tc_renderer_begin_frame(tc_renderer* renderer, const tc_frame_constants* frameHost, tc_stream stream);
/* begin_frame, instances_classify, clusters_cull and tc_frame_tessellate.
 * In ray-tracing mode clusters_cull lists every cluster of every instance,
 * so over the renderer's own instances that list is built once, at
 * creation, in a buffer of its own: the frame's SceneBuilding then names it
 * (visibleClusters, visibleClusterCounter) and the culling launches are
 * skipped.  Over instances bound with tc_renderer_rebind_instances the frame
 * culls into the buffer tc_renderer_get_building names. */
tc_status tc_renderer_tessellate(tc_renderer* renderer, const tc_frame_constants* frameHost, tc_stream stream);

/* Accel arrays of the renderer and its per-instance transforms (device). */
tc_status tc_renderer_get_accel(tc_renderer* renderer, tc_accel* out, const float** instanceXforms);
/* tc_clas_instantiate, tc_blas_setup_insertion, tc_blas_clusters_insert,
 * tc_accel_build over the renderer's buffers. */
tc_status tc_renderer_build_accel(tc_renderer* renderer, tc_stream stream);
/* tc_trace_rays over the renderer's accel arrays.  Camera traces take their
 * 8 x 8 tile rows in the order the renderer's last frame start ranked from
 * the previous frame's traces of the same height (rows with the longest
 * waves first); hits do not depend on it.  Launches of more than 16 waves
 * per resident wave slot (e.g. 3840 x 2160 on 256 CUs) keep tile order.  A
 * renderer's traces share that state and must not overlap. */
tc_status tc_renderer_trace(tc_renderer* renderer, const tc_trace_params* params, tc_stream stream);
/* tessellate + build_accel + trace (params may be NULL). */
tc_status tc_renderer_frame(tc_renderer* renderer, const tc_frame_constants* frameHost, const tc_trace_params* params,
                            tc_stream stream);
/* Points the renderer at n instance records and 24-float transforms already
 * in device memory (n = the count given at creation); NULL restores its own
 * copies.  Lets one small renderer process many meshes one after another. */
tc_status tc_renderer_rebind_instances(tc_renderer* renderer, const tc_render_instance* devInstances, const float* devXforms,
                                       uint32_t n);

/* Captures one whole frame into a HIP graph on `stream` (not the null
 * stream); tc_renderer_graph_launch replays it with new frame constants. */
tc_status tc_renderer_graph_capture(tc_renderer* renderer, const tc_frame_constants* frameHost,
                                    const tc_trace_params* params, tc_stream stream);
tc_status tc_renderer_graph_launch(tc_renderer* renderer, const tc_frame_constants* frameHost, tc_stream stream);

/* No counterpart in the reference (its frame processes every instance in
 * one dispatch per stage).  The per-mesh form of a multi-mesh frame, for
 * comparison with the batched frame: one graph that runs the chain
 * tessellation .. CLAS .. BLAS insertion .. BLAS / TLAS build once per
 * instance record of devInstances (count records in device memory, with 24
 * floats per record in devXforms as tc_instance_xforms_host writes them),
 * on a renderer created for one instance.  Replay with
 * tc_renderer_graph_launch. */
tc_status tc_renderer_graph_capture_meshwise(tc_renderer* renderer, const tc_render_instance* devInstances,
                                             const float* devXforms, uint32_t count, const tc_frame_constants* frameHost,
                                             tc_stream stream);

/* objectToWorld and worldToObject (3x4 row-major, 24 floats) per instance,
 * the transforms the renderer uploads; worldToObject is inverted in FP64. */
tc_status tc_instance_xforms_host(const tc_render_instance* instances, uint32_t n, float* out);

/* Number of split passes the reference records for a split factor
 * (renderer_raytrace_clusters_tess.cpp:516-543). */
uint32_t tc_num_split_passes(uint32_t splitFactor);

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(tc_cluster_info) == 8, "ClusterInfo_size");            /* shaderio_building.h:68  */
static_assert(sizeof(tc_tess_triangle_info) == 24, "TessTriangleInfo_size"); /* shaderio_building.h:92  */
static_assert(sizeof(tc_template_instantiate_info) == 32, "TemplateInstantiateInfo");
static_assert(sizeof(tc_blas_build_info) == 16, "BlasBuildInfo");
static_assert(sizeof(tc_tess_table_entry) == 8, "TessTableEntry");
static_assert(sizeof(tc_cluster) == 16, "Cluster");
static_assert(sizeof(tc_bbox) == 32, "BBox");
static_assert(sizeof(tc_render_instance) == 192, "RenderInstance");
static_assert(offsetof(tc_scene_building, genClusterDataCounter) == 72, "SceneBuilding layout");
static_assert(offsetof(tc_scene_building, instanceStates) == 104, "SceneBuilding layout");
static_assert(offsetof(tc_scene_building, positionTruncateBitCount) == 208, "SceneBuilding layout");
static_assert(offsetof(tc_scene_building, numBlasReservedSizes) == 360, "SceneBuilding layout");
static_assert(sizeof(tc_scene_building) == 368, "SceneBuilding");
static_assert(offsetof(tc_readback, numGenDatas) == 48, "Readback layout");
static_assert(sizeof(tc_clas_header) == 48, "clas header");
static_assert(sizeof(tc_bvh_node) == 64, "bvh node");
static_assert(sizeof(tc_bvh4_node) == 128, "bvh4 node");
static_assert(sizeof(tc_clas_build_info) == 64, "ClasBuildInfo");
static_assert(sizeof(tc_clas_template) == 32, "clas template");
static_assert(sizeof(tc_vertex_tape) == 48, "vertex tape");
static_assert(sizeof(tc_bindings) == 120, "bindings");
static_assert(sizeof(tc_hiz) == 32, "hiz");
static_assert(sizeof(tc_adjoint_params) == 80, "adjoint params");
#endif

#endif /* TC_ABI_H */
