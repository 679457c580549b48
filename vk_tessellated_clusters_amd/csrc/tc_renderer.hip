// tc_renderer.hip -- C++ frame driver: owns the per-frame buffers and
// records the stage chain (renderer_raytrace_clusters_tess.cpp:166-310 for
// the buffer layout, :410-780 for the frame).
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "tc_kernels.h"
#include "tc_trace.h"

namespace tc {

struct Renderer {
  tc_config cfg{};
  uint32_t numInstances = 0, numTextures = 0;
  // device memory
  void* arena = nullptr;
  uint64_t arenaBytes = 0;
  tc_frame_constants* view = nullptr;
  tc_frame_constants* viewLast = nullptr;  // constants of the frame before (culling)
  tc_readback* readback = nullptr;
  tc_render_instance* instances = nullptr;
  tc_displacement_texture* textures = nullptr;
  tc_scene_building* build = nullptr;
  tc_scene_building* buildTemplate = nullptr;  // per-frame initial state
  tc_tessellation_table* table = nullptr;
  void* workspace = nullptr;
  uint64_t workspaceBytes = 0;
  tc_vertex_tape* tape = nullptr;  // with TC_FLAG_VERTEX_TAPE
  tc_scene_building hostBuild{};
  // The visible-cluster list of the renderer's own instances.  In
  // ray-tracing mode culling appends every cluster of every instance
  // (clusters_cull.comp.glsl:112-161), a function of the instance records
  // alone, so it is built once at creation; frames over these instances
  // start from it and skip the culling launches.
  uint64_t clusterList = 0;      // device tc_cluster_info[maxVisibleClusters]
  uint32_t listedClusters = 0;   // its visibleClusterCounter (unclamped)
  tc::TraceOrder* traceOrder = nullptr;  // tile-row order of the camera traces (device)
  // Frame constants travel through a ring of pinned slots: the host fills
  // slot hostFrame % kViewSlots, k_frame_start (in the stream or the
  // graph) reads the slot named by a device counter that advances once per
  // frame, and an event per slot keeps the host from refilling a slot
  // that a queued frame has not read yet.
  tc_frame_constants* pinnedRing = nullptr;
  uint32_t* frameCounter = nullptr;  // device
  uint64_t hostFrame = 0;
  hipEvent_t slotEvent[8] = {};
  bool slotPending[8] = {};
  bool capturing = false;
  tc_frame_constants lastFrame{};  // constants of the latest frame queued or captured
  uint64_t splitBytes = 0;
  float* xforms = nullptr;  // per instance objectToWorld | worldToObject (3x4 row-major)
  tc_accel accel{};
  tc_render_instance* ownInstances = nullptr;
  float* ownXforms = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  tc_trace_params graphParams{};
  // with TC_FLAG_HIZ: the far HiZ of the latest camera trace, sized for the
  // frame constants' viewport, and its shader factors
  tc_hiz hiz{};
  float hizFactors[4] = {0, 0, 0, 0};
  float hizSizeMax = 0.0f;
};

static uint64_t a256(uint64_t v) { return (v + 255) & ~255ull; }
constexpr uint32_t kViewSlots = 8;

// inverse of a 4x4 row-major matrix in double precision (cofactors)
static bool invert4(const double m[16], double out[16]) {
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.0) return false;
  for (int i = 0; i < 16; i++) out[i] = inv[i] / det;
  return true;
}

// Word w of the frame's initial SceneBuilding: the template's, with the
// eye position of the frame and, when `list` is set, the renderer's
// prebuilt cluster list (its address and count) in place of an empty one.
__device__ __forceinline__ uint32_t building_word(const tc_scene_building* tmpl, uint32_t w, const float* eye,
                                                  uint64_t list, uint32_t listCount) {
  constexpr uint32_t kList = offsetof(tc_scene_building, visibleClusters) / 4;
  constexpr uint32_t kCount = offsetof(tc_scene_building, visibleClusterCounter) / 4;
  static_assert(offsetof(tc_scene_building, viewPos) == 0, "the eye position leads the building record");
  uint32_t v = reinterpret_cast<const uint32_t*>(tmpl)[w];
  if (w < 3) v = __float_as_uint(eye[w]);
  if (list) {
    if (w == kList) v = (uint32_t)list;
    if (w == kList + 1) v = (uint32_t)(list >> 32);
    if (w == kCount) v = listCount;
  }
  return v;
}

// Resets the per-frame state: SceneBuilding from its template (with the
// eye position of the frame), Readback to zero and the split list to ~0
// (renderer_raytrace_clusters_tess.cpp:412-419).
__global__ void k_frame_reset(tc_scene_building* build, const tc_scene_building* tmpl, const tc_frame_constants* view,
                              tc_readback* rb, uint32_t* split, uint64_t splitWords, uint64_t list, uint32_t listCount) {
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (gid < sizeof(tc_scene_building) / 4)
    reinterpret_cast<uint32_t*>(build)[gid] = building_word(tmpl, (uint32_t)gid, view->viewPos, list, listCount);
  for (uint64_t i = gid; i < sizeof(tc_readback) / 4; i += stride) reinterpret_cast<uint32_t*>(rb)[i] = 0;
  for (uint64_t i = gid; i < splitWords; i += stride) split[i] = ~0u;
}

// The frame's start in one launch: workgroup 0 copies the pinned slot the
// device frame counter names into the frame constants (the constants they
// replace become viewLast; the first frame's viewLast is its own view),
// advances the counter and resets SceneBuilding, taking the eye position
// from that slot; workgroup 1 ranks the tile rows of the traces timed since
// the last frame start (the order this frame's camera traces take); every
// workgroup clears Readback and fills the split list with ~0.  No workgroup
// waits on another.
__global__ void k_frame_start(const tc_frame_constants* ring, uint32_t* counter, tc_frame_constants* view,
                              tc_frame_constants* viewLast, tc_scene_building* build, const tc_scene_building* tmpl,
                              tc_readback* rb, uint32_t* split, uint64_t splitWords, uint64_t list, uint32_t listCount,
                              tc::TraceOrder* order) {
  __shared__ uint32_t s_frame;
  __shared__ uint32_t s_rows[tc::kMaxOrderRows];
  if (blockIdx.x == 1) tc::rank_tile_rows(order, s_rows);
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) s_frame = counter[0];
    __syncthreads();
    const uint32_t frame = s_frame;
    const tc_frame_constants* src = ring + frame % kViewSlots;
    const uint32_t* sw = reinterpret_cast<const uint32_t*>(src);
    uint32_t* dst = reinterpret_cast<uint32_t*>(view);
    uint32_t* last = reinterpret_cast<uint32_t*>(viewLast);
    for (uint32_t i = threadIdx.x; i < sizeof(tc_frame_constants) / 4; i += blockDim.x) {
      const uint32_t v = sw[i];
      last[i] = frame ? dst[i] : v;
      dst[i] = v;
    }
    if (threadIdx.x < sizeof(tc_scene_building) / 4)
      reinterpret_cast<uint32_t*>(build)[threadIdx.x] = building_word(tmpl, threadIdx.x, src->viewPos, list, listCount);
    __syncthreads();
    if (threadIdx.x == 0) counter[0] = frame + 1;
  }
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gid; i < sizeof(tc_readback) / 4; i += stride) reinterpret_cast<uint32_t*>(rb)[i] = 0;
  for (uint64_t i = gid; i < splitWords; i += stride) split[i] = ~0u;
}
static_assert(sizeof(tc_scene_building) / 4 <= 256, "the building record is reset by workgroup 0 of k_frame_start");

// Writes the frame constants into a ring slot, first waiting until the
// frame that used the slot last has read it.
static bool fill_slot(Renderer* r, uint32_t slot, const tc_frame_constants* fc) {
  if (r->slotPending[slot]) {
    if (hipEventSynchronize(r->slotEvent[slot]) != hipSuccess) return false;
    r->slotPending[slot] = false;
  }
  memcpy(&r->pinnedRing[slot], fc, sizeof(tc_frame_constants));
  if (r->hiz.texels) {
    // tessellatedclusters.cpp:615-619 the HiZ factors of the frame
    memcpy(r->pinnedRing[slot].hizSizeFactors, r->hizFactors, sizeof(r->hizFactors));
    r->pinnedRing[slot].hizSizeMax = r->hizSizeMax;
  }
  if (fc != &r->lastFrame) r->lastFrame = *fc;
  return true;
}

// With TC_FLAG_HIZ: (re)allocates the HiZ for the viewport of `fc`, filled
// with the far depth (outside graph capture; a captured graph that used an
// older HiZ is dropped).
static bool ensure_hiz(Renderer* r, const tc_frame_constants* fc) {
  if (!(r->cfg.flags & TC_FLAG_HIZ) || !fc) return true;
  const uint32_t w = (uint32_t)fc->viewport[0], h = (uint32_t)fc->viewport[1];
  if (r->hiz.texels && r->hiz.width == w && r->hiz.height == h) return true;
  tc_hiz nh{};
  uint64_t bytes = 0;
  float factors[4], sizeMax;
  if (tc_hiz_setup(w, h, &nh, &bytes, factors, &sizeMax) != TC_OK) return false;
  if (hipDeviceSynchronize() != hipSuccess) return false;
  if (r->hiz.texels) (void)hipFree(r->hiz.texels);
  r->hiz.texels = nullptr;
  if (r->graphExec) (void)hipGraphExecDestroy(r->graphExec), r->graphExec = nullptr;
  if (r->graph) (void)hipGraphDestroy(r->graph), r->graph = nullptr;
  if (hipMalloc((void**)&nh.texels, bytes) != hipSuccess) return false;
  if (tc_hiz_clear(&nh, nullptr) != TC_OK || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(nh.texels);
    return false;
  }
  r->hiz = nh;
  memcpy(r->hizFactors, factors, sizeof(factors));
  r->hizSizeMax = sizeMax;
  return true;
}

// Marks the slot as read by the work just queued on `st`.
static bool queued_slot(Renderer* r, uint32_t slot, hipStream_t st) {
  if (hipEventRecord(r->slotEvent[slot], st) != hipSuccess) return false;
  r->slotPending[slot] = true;
  r->hostFrame++;
  return true;
}

}  // namespace tc

using namespace tc;

extern "C" {

// objectToWorld (the instance's column-major world matrix as 3x4 row-major)
// and its inverse, per instance
tc_status tc_instance_xforms_host(const tc_render_instance* instances, uint32_t n, float* out) {
  if (!instances || !out) return TC_ERROR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; i++) {
    const float* m = instances[i].worldMatrix;  // m[col * 4 + row]
    double rm[16];
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) rm[r * 4 + c] = m[c * 4 + r];
    double inv[16];
    if (!invert4(rm, inv)) return TC_ERROR_INVALID_ARGUMENT;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 4; c++) {
        out[24 * i + r * 4 + c] = (float)rm[r * 4 + c];
        out[24 * i + 12 + r * 4 + c] = (float)inv[r * 4 + c];
      }
  }
  return TC_OK;
}

tc_status tc_table_upload(tc_tessellation_table** tableDev);
tc_status tc_table_free(tc_tessellation_table* tableDev);
tc_status tc_workspace_init(const tc_bindings* b, const tc_config* config, tc_stream stream);
tc_status tc_instances_classify(const tc_bindings* b, tc_config config, tc_stream stream);
tc_status tc_clusters_cull(const tc_bindings* b, tc_config config, tc_stream stream);

static Renderer* R(tc_renderer* r) { return reinterpret_cast<Renderer*>(r); }

tc_status tc_renderer_create(const tc_render_instance* instancesHost, uint32_t numInstances,
                             const tc_displacement_texture* texturesHost, uint32_t numTextures, const tc_config* config,
                             tc_renderer** out) {
  if (!instancesHost || !numInstances || !config || !out) return TC_ERROR_INVALID_ARGUMENT;
  if (numTextures && !texturesHost) return TC_ERROR_INVALID_ARGUMENT;
  uint64_t wsBytes = tc_workspace_size(config);
  if (!wsBytes) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = new Renderer();
  r->cfg = *config;
  r->numInstances = numInstances;
  r->numTextures = numTextures;
  const tc_config& c = *config;

  // carve one arena (renderer_raytrace_clusters_tess.cpp:200-310)
  struct Slot {
    uint64_t* dst;
    uint64_t bytes;
  };
  tc_scene_building& B = r->hostBuild;
  memset(&B, 0, sizeof(B));
  B.numRenderInstances = numInstances;
  uint64_t oView, oViewLast, oRb, oInst, oTex, oBuild, oTmpl, oWs, oXf, oAccel, oTape, oCounter, oList, oOrder;
  std::vector<Slot> slots = {
      {&oView, sizeof(tc_frame_constants)},
      {&oViewLast, sizeof(tc_frame_constants)},
      {&oRb, sizeof(tc_readback)},
      {&oInst, sizeof(tc_render_instance) * (uint64_t)numInstances},
      {&oTex, sizeof(tc_displacement_texture) * (uint64_t)std::max(numTextures, 1u)},
      {&oBuild, sizeof(tc_scene_building)},
      {&oTmpl, sizeof(tc_scene_building)},
      {&B.instanceStates, 4ull * numInstances},
      {&B.visibleClusters, sizeof(tc_cluster_info) * (uint64_t)c.maxVisibleClusters},
      {&B.splitTriangles, sizeof(tc_tess_triangle_info) * (uint64_t)c.maxSplitTriangles},
      {&B.partTriangles, sizeof(tc_tess_triangle_info) * (uint64_t)c.maxPartTriangles},
      {&B.genVertices, 12ull * c.maxGeneratedVertices},
      {&B.tempInstanceIDs, 4ull * c.maxGeneratedClusters},
      {&B.tempInstantiations, sizeof(tc_template_instantiate_info) * (uint64_t)c.maxGeneratedClusters},
      {&B.tempClusterAddresses, 8ull * c.maxGeneratedClusters},
      {&B.tempClusterSizes, 4ull * c.maxGeneratedClusters},
      {&B.blasBuildInfos, sizeof(tc_blas_build_info) * (uint64_t)numInstances},
      {&B.blasBuildSizes, 4ull * numInstances},
      {&B.blasClusterAddresses, 8ull * c.maxGeneratedClusters},
      // renderer_raytrace_clusters_tess.cpp:224-229 (1X transient builds)
      {&B.transBuilds, transient_builds(c.flags) ? sizeof(tc_clas_build_info) * (uint64_t)c.maxGeneratedClusters : 0},
      {&B.transInstanceIDs, transient_builds(c.flags) ? 4ull * c.maxGeneratedClusters : 0},
      {&B.transClusterAddresses, transient_builds(c.flags) ? 8ull * c.maxGeneratedClusters : 0},
      {&B.transClusterSizes, transient_builds(c.flags) ? 4ull * c.maxGeneratedClusters : 0},
      {&B.genClusterData, (uint64_t)c.maxGeneratedClusterMegs * 1024ull * 1024ull},
      {&oWs, wsBytes},
      {&oXf, 24ull * 4 * numInstances},
      {&oAccel, tc_accel_size(c.maxGeneratedClusters, numInstances)},
      {&oTape, (c.flags & TC_FLAG_VERTEX_TAPE) ? sizeof(tc_vertex_tape) * (uint64_t)c.maxGeneratedVertices : 0},
      {&oCounter, 256},
      {&oList, sizeof(tc_cluster_info) * (uint64_t)c.maxVisibleClusters},
      {&oOrder, sizeof(tc::TraceOrder)},
  };
  uint64_t o = 0;
  for (auto& s : slots) {
    *s.dst = o;
    o = a256(o + s.bytes);
  }
  r->arenaBytes = o;
  if (hipMalloc(&r->arena, o) != hipSuccess) {
    delete r;
    return TC_ERROR_OUT_OF_MEMORY;
  }
  const uint64_t base = reinterpret_cast<uint64_t>(r->arena);
  char* mem = static_cast<char*>(r->arena);
  r->view = reinterpret_cast<tc_frame_constants*>(mem + oView);
  r->viewLast = reinterpret_cast<tc_frame_constants*>(mem + oViewLast);
  r->readback = reinterpret_cast<tc_readback*>(mem + oRb);
  r->instances = reinterpret_cast<tc_render_instance*>(mem + oInst);
  r->textures = reinterpret_cast<tc_displacement_texture*>(mem + oTex);
  r->build = reinterpret_cast<tc_scene_building*>(mem + oBuild);
  r->buildTemplate = reinterpret_cast<tc_scene_building*>(mem + oTmpl);
  r->workspace = mem + oWs;
  r->workspaceBytes = wsBytes;
  r->xforms = reinterpret_cast<float*>(mem + oXf);
  if (c.flags & TC_FLAG_VERTEX_TAPE) r->tape = reinterpret_cast<tc_vertex_tape*>(mem + oTape);
  r->frameCounter = reinterpret_cast<uint32_t*>(mem + oCounter);
  r->traceOrder = reinterpret_cast<tc::TraceOrder*>(mem + oOrder);
  r->ownInstances = r->instances;
  r->ownXforms = r->xforms;
  tc_accel_bind(mem + oAccel, tc_accel_size(c.maxGeneratedClusters, numInstances), c.maxGeneratedClusters, numInstances,
                &r->accel);
  std::vector<float> xf(24ull * numInstances);
  tc_instance_xforms_host(instancesHost, numInstances, xf.data());
  for (uint64_t* f : {&B.instanceStates, &B.visibleClusters, &B.splitTriangles, &B.partTriangles, &B.genVertices,
                      &B.tempInstanceIDs, &B.tempInstantiations, &B.tempClusterAddresses, &B.tempClusterSizes,
                      &B.blasBuildInfos, &B.blasBuildSizes, &B.blasClusterAddresses, &B.genClusterData, &B.transBuilds,
                      &B.transInstanceIDs, &B.transClusterAddresses, &B.transClusterSizes})
    *f += base;
  if (!transient_builds(c.flags)) B.transBuilds = B.transInstanceIDs = B.transClusterAddresses = B.transClusterSizes = 0;
  // :251, :293 the transient index and mapping bytes alias these lists
  B.transTriIndices = B.genVertices;
  B.transTriMappings = B.partTriangles;
  r->splitBytes = sizeof(tc_tess_triangle_info) * (uint64_t)c.maxSplitTriangles;

  tc_status s = TC_OK;
  if (hipMemcpy(r->instances, instancesHost, sizeof(tc_render_instance) * numInstances, hipMemcpyHostToDevice) != hipSuccess ||
      (numTextures &&
       hipMemcpy(r->textures, texturesHost, sizeof(tc_displacement_texture) * numTextures, hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(r->buildTemplate, &B, sizeof(B), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(r->xforms, xf.data(), xf.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(r->arena, 0, oInst) != hipSuccess || hipMemset(r->frameCounter, 0, 256) != hipSuccess ||
      hipMemset(r->traceOrder, 0, sizeof(tc::TraceOrder)) != hipSuccess ||
      hipMemset(reinterpret_cast<void*>(B.blasBuildSizes), 0, 4ull * numInstances) != hipSuccess)
    s = TC_ERROR_LAUNCH;
  if (s == TC_OK) s = tc_table_upload(&r->table);
  if (s == TC_OK &&
      hipHostMalloc((void**)&r->pinnedRing, sizeof(tc_frame_constants) * kViewSlots, hipHostMallocDefault) != hipSuccess)
    s = TC_ERROR_OUT_OF_MEMORY;
  for (uint32_t k = 0; s == TC_OK && k < kViewSlots; k++)
    if (hipEventCreateWithFlags(&r->slotEvent[k], hipEventDisableTiming) != hipSuccess) s = TC_ERROR_LAUNCH;
  if (s == TC_OK) {
    tc_bindings b;
    tc_renderer_get_bindings(reinterpret_cast<tc_renderer*>(r), &b);
    s = tc_workspace_init(&b, config, nullptr);
    if (s == TC_OK && hipDeviceSynchronize() != hipSuccess) s = TC_ERROR_LAUNCH;
  }
  if (s == TC_OK) {
    // the own instances' cluster list, by the culling kernels into the
    // list's own buffer (the building record is reset at every frame start)
    tc_scene_building L = B;
    L.visibleClusters = base + oList;
    L.visibleClusterCounter = 0;
    tc_bindings b;
    tc_renderer_get_bindings(reinterpret_cast<tc_renderer*>(r), &b);
    if (hipMemcpy(r->build, &L, sizeof(L), hipMemcpyHostToDevice) != hipSuccess) s = TC_ERROR_LAUNCH;
    if (s == TC_OK) s = tc_clusters_cull(&b, *config, nullptr);
    if (s == TC_OK && hipMemcpy(&r->listedClusters, &r->build->visibleClusterCounter, 4, hipMemcpyDeviceToHost) != hipSuccess)
      s = TC_ERROR_LAUNCH;
    r->clusterList = base + oList;
  }
  if (s != TC_OK) {
    tc_renderer_destroy(reinterpret_cast<tc_renderer*>(r));
    return s;
  }
  *out = reinterpret_cast<tc_renderer*>(r);
  return TC_OK;
}

tc_status tc_renderer_destroy(tc_renderer* rr) {
  if (!rr) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  if (r->graphExec) (void)hipGraphExecDestroy(r->graphExec);
  if (r->graph) (void)hipGraphDestroy(r->graph);
  if (r->table) tc_table_free(r->table);
  for (uint32_t k = 0; k < kViewSlots; k++)
    if (r->slotEvent[k]) (void)hipEventDestroy(r->slotEvent[k]);
  if (r->pinnedRing) {
    (void)hipDeviceSynchronize();
    (void)hipHostFree(r->pinnedRing);
  }
  if (r->arena) (void)hipFree(r->arena);
  if (r->hiz.texels) (void)hipFree(r->hiz.texels);
  delete r;
  return TC_OK;
}

tc_status tc_renderer_get_bindings(tc_renderer* rr, tc_bindings* out) {
  if (!rr || !out) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  out->view = r->view;
  out->tessTable = r->table;
  out->readback = r->readback;
  out->instances = r->instances;
  out->build = r->build;
  out->displacementTextures = r->numTextures ? r->textures : nullptr;
  out->numRenderInstances = r->numInstances;
  out->numDisplacementTextures = r->numTextures;
  out->workspace = r->workspace;
  out->workspaceBytes = r->workspaceBytes;
  out->vertexTape = r->tape;
  out->viewLast = r->viewLast;
  out->hiz = r->hiz;
  return TC_OK;
}

tc_status tc_renderer_get_accel(tc_renderer* rr, tc_accel* out, const float** instanceXforms) {
  if (!rr || !out) return TC_ERROR_INVALID_ARGUMENT;
  *out = R(rr)->accel;
  if (instanceXforms) *instanceXforms = R(rr)->xforms;
  return TC_OK;
}

// CLAS build, BLAS insertion, BLAS + TLAS build
// (renderer_raytrace_clusters_tess.cpp:592-745).
tc_status tc_renderer_build_accel(tc_renderer* rr, tc_stream stream) {
  if (!rr) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  tc_bindings b;
  tc_renderer_get_bindings(rr, &b);
  tc_status s;
  if ((s = tc_clas_instantiate(&b, r->cfg, stream)) != TC_OK) return s;
  if ((s = tc_clas_build_transient(&b, r->cfg, stream)) != TC_OK) return s;
  if ((s = tc_blas_setup_insertion(&b, r->cfg, stream)) != TC_OK) return s;
  if ((s = tc_blas_clusters_insert(&b, r->cfg, 0, stream)) != TC_OK) return s;
  if ((s = tc_blas_clusters_insert(&b, r->cfg, 1, stream)) != TC_OK) return s;
  return tc_accel_build(&b, r->cfg, &r->accel, r->xforms, stream);
}

tc_status tc_renderer_trace(tc_renderer* rr, const tc_trace_params* params, tc_stream stream) {
  if (!rr) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  tc_bindings b;
  tc_renderer_get_bindings(rr, &b);
  // camera traces take their tile rows longest first (the order ranked at
  // the frame start from the previous frame's traces of the same height)
  tc_status s = tc::trace_rays(&b, r->cfg, &r->accel, params, r->traceOrder, stream);
  if (s != TC_OK) return s;
  // resources.cpp:503-512: the HiZ of this frame's depth, for the next
  // frame's instance culling (camera rays of the HiZ's viewport only)
  if (r->hiz.texels && params && !params->rays && params->decoded && params->width == r->hiz.width &&
      params->height == r->hiz.height)
    s = tc_hiz_build(&r->hiz, reinterpret_cast<const char*>(params->decoded) + offsetof(tc_hit_decoded, depth),
                     sizeof(tc_hit_decoded), stream);
  return s;
}

// One whole frame: tessellation, cluster acceleration structures, rays.
tc_status tc_renderer_frame(tc_renderer* rr, const tc_frame_constants* frameHost, const tc_trace_params* params,
                            tc_stream stream) {
  tc_status s;
  if ((s = tc_renderer_tessellate(rr, frameHost, stream)) != TC_OK) return s;
  if ((s = tc_renderer_build_accel(rr, stream)) != TC_OK) return s;
  return params ? tc_renderer_trace(rr, params, stream) : TC_OK;
}

// Points the renderer at `n` instance records and transforms already in
// device memory (n equal to the count given at creation); NULL arguments
// restore the renderer's own copies.
tc_status tc_renderer_rebind_instances(tc_renderer* rr, const tc_render_instance* devInstances, const float* devXforms,
                                       uint32_t n) {
  if (!rr) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  if (!devInstances || !devXforms) {
    r->instances = r->ownInstances;
    r->xforms = r->ownXforms;
    return TC_OK;
  }
  if (n != r->numInstances) return TC_ERROR_INVALID_ARGUMENT;
  r->instances = const_cast<tc_render_instance*>(devInstances);
  r->xforms = const_cast<float*>(devXforms);
  return TC_OK;
}

// Captures one whole frame (tc_renderer_frame) into a graph on `stream`.
// The frame constants are read at replay time from the renderer's pinned
// buffer, which tc_renderer_graph_launch fills.
tc_status tc_renderer_graph_capture(tc_renderer* rr, const tc_frame_constants* frameHost, const tc_trace_params* params,
                                    tc_stream stream) {
  if (!rr || !frameHost || !stream) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  hipStream_t st = (hipStream_t)stream;
  if (r->graphExec) (void)hipGraphExecDestroy(r->graphExec), r->graphExec = nullptr;
  if (r->graph) (void)hipGraphDestroy(r->graph), r->graph = nullptr;
  if (params) r->graphParams = *params;
  if (!ensure_hiz(r, frameHost)) return TC_ERROR_OUT_OF_MEMORY;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return TC_ERROR_LAUNCH;
  r->capturing = true;
  r->lastFrame = *frameHost;
  tc_status s = tc_renderer_frame(rr, frameHost, params ? &r->graphParams : nullptr, stream);
  r->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (s != TC_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return s != TC_OK ? s : TC_ERROR_LAUNCH;
  }
  r->graph = g;
  if (hipGraphInstantiate(&r->graphExec, g, nullptr, nullptr, 0) != hipSuccess) return TC_ERROR_LAUNCH;
  return TC_OK;
}

// Replays the captured frame with new frame constants.  The pinned buffer
// is rewritten on the host, so a caller changing the camera between
// replays must let the previous replay finish its upload first (the frame
// constants copy is the graph's first node).
tc_status tc_renderer_graph_launch(tc_renderer* rr, const tc_frame_constants* frameHost, tc_stream stream) {
  if (!rr || !R(rr)->graphExec) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  const uint32_t slot = (uint32_t)(r->hostFrame % kViewSlots);
  if (!fill_slot(r, slot, frameHost ? frameHost : &r->lastFrame)) return TC_ERROR_LAUNCH;
  if (hipGraphLaunch(r->graphExec, (hipStream_t)stream) != hipSuccess) return TC_ERROR_LAUNCH;
  return queued_slot(r, slot, (hipStream_t)stream) ? TC_OK : TC_ERROR_LAUNCH;
}

tc_status tc_renderer_get_building(tc_renderer* rr, tc_scene_building* out) {
  if (!rr || !out) return TC_ERROR_INVALID_ARGUMENT;
  *out = R(rr)->hostBuild;
  return TC_OK;
}

}  // extern "C"

// Frame start: frame constants upload (from pinned host memory, unless
// fetchView is false: the frame keeps the constants already fetched) and
// per-frame state reset.
static tc_status begin_frame(Renderer* r, const tc_frame_constants* frameHost, hipStream_t st, bool fetchView,
                             bool listed) {
  const uint64_t list = listed ? r->clusterList : 0;
  if (fetchView) {
    // during graph capture only the fetch is recorded; each replay fills
    // its slot in tc_renderer_graph_launch
    const uint32_t slot = (uint32_t)(r->hostFrame % kViewSlots);
    if (!r->capturing && !ensure_hiz(r, frameHost)) return TC_ERROR_OUT_OF_MEMORY;
    if (!r->capturing && !fill_slot(r, slot, frameHost)) return TC_ERROR_LAUNCH;
    hipLaunchKernelGGL(k_frame_start, dim3(256), dim3(256), 0, st, (const tc_frame_constants*)r->pinnedRing, r->frameCounter,
                       r->view, r->viewLast, r->build, (const tc_scene_building*)r->buildTemplate, r->readback,
                       reinterpret_cast<uint32_t*>(r->hostBuild.splitTriangles), r->splitBytes / 4, list,
                       r->listedClusters, r->traceOrder);
    if (!r->capturing && !queued_slot(r, slot, st)) return TC_ERROR_LAUNCH;
  } else {
    hipLaunchKernelGGL(k_frame_reset, dim3(256), dim3(256), 0, st, r->build, (const tc_scene_building*)r->buildTemplate,
                       (const tc_frame_constants*)r->view, r->readback,
                       reinterpret_cast<uint32_t*>(r->hostBuild.splitTriangles), r->splitBytes / 4, list,
                       r->listedClusters);
  }
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

// The frame's start and its tessellation: instances_classify through the
// template instantiation (renderer_raytrace_clusters_tess.cpp:433-582).
// Over the renderer's own instances the frame starts from their prebuilt
// cluster list; over instance records bound in their place it culls.
static tc_status tessellate(tc_renderer* rr, const tc_frame_constants* frameHost, tc_stream stream, bool fetchView) {
  Renderer* r = R(rr);
  tc_status s;
  const bool listed = r->clusterList && r->instances == r->ownInstances;
  if ((s = begin_frame(r, frameHost, (hipStream_t)stream, fetchView, listed)) != TC_OK) return s;
  tc_bindings b;
  tc_renderer_get_bindings(rr, &b);
  if ((s = tc_instances_classify(&b, r->cfg, stream)) != TC_OK) return s;
  if (!listed && (s = tc_clusters_cull(&b, r->cfg, stream)) != TC_OK) return s;
  return tc_frame_tessellate(&b, r->cfg, 0, stream);
}

extern "C" {

tc_status tc_renderer_begin_frame(tc_renderer* rr, const tc_frame_constants* frameHost, tc_stream stream) {
  if (!rr || !frameHost) return TC_ERROR_INVALID_ARGUMENT;
  return begin_frame(R(rr), frameHost, (hipStream_t)stream, true, false);
}

// Captures the per-mesh form of a multi-mesh frame into the renderer's
// graph: the chain tessellation .. CLAS .. BLAS insertion .. BLAS / TLAS
// build run once for each of `count` instance records (device memory, 192
// bytes each, with 24 transform floats each in devXforms) by a renderer
// created for one instance.  The first chain fetches the frame constants,
// the others reuse them.  tc_renderer_graph_launch replays it.
tc_status tc_renderer_graph_capture_meshwise(tc_renderer* rr, const tc_render_instance* devInstances, const float* devXforms,
                                             uint32_t count, const tc_frame_constants* frameHost, tc_stream stream) {
  if (!rr || !devInstances || !devXforms || !count || !frameHost || !stream) return TC_ERROR_INVALID_ARGUMENT;
  Renderer* r = R(rr);
  if (r->numInstances != 1) return TC_ERROR_INVALID_ARGUMENT;
  hipStream_t st = (hipStream_t)stream;
  if (r->graphExec) (void)hipGraphExecDestroy(r->graphExec), r->graphExec = nullptr;
  if (r->graph) (void)hipGraphDestroy(r->graph), r->graph = nullptr;
  if (!ensure_hiz(r, frameHost)) return TC_ERROR_OUT_OF_MEMORY;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return TC_ERROR_LAUNCH;
  r->capturing = true;
  r->lastFrame = *frameHost;
  tc_status s = TC_OK;
  for (uint32_t i = 0; i < count && s == TC_OK; i++) {
    r->instances = const_cast<tc_render_instance*>(devInstances + i);
    r->xforms = const_cast<float*>(devXforms + 24ull * i);
    s = tessellate(rr, frameHost, stream, i == 0);
    if (s == TC_OK) s = tc_renderer_build_accel(rr, stream);
  }
  r->instances = r->ownInstances;
  r->xforms = r->ownXforms;
  r->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (s != TC_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return s != TC_OK ? s : TC_ERROR_LAUNCH;
  }
  r->graph = g;
  if (hipGraphInstantiate(&r->graphExec, g, nullptr, nullptr, 0) != hipSuccess) return TC_ERROR_LAUNCH;
  return TC_OK;
}

// The frame's start and its tessellation; tc_renderer_build_accel
// continues from the CLAS builds.
tc_status tc_renderer_tessellate(tc_renderer* rr, const tc_frame_constants* frameHost, tc_stream stream) {
  if (!rr || !frameHost) return TC_ERROR_INVALID_ARGUMENT;
  return tessellate(rr, frameHost, stream, true);
}

}  // extern "C"
