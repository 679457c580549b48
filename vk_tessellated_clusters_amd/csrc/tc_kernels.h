// tc_kernels.h -- kernel argument block and kernel declarations.
#pragma once

#include "tc_device.h"

namespace tc {

constexpr uint32_t kMaxClusterVerts = 256;
constexpr uint32_t kMaxClusterTris = 128;

// TESS_USE_TRANSIENTBUILDS (shaderio.h:140-141): 1X or 2X transient builds
__host__ __device__ inline bool transient_builds(uint32_t flags) {
  return (flags & (TC_FLAG_1X_TRANSIENT | TC_FLAG_2X_TRANSIENT)) != 0;
}

// 2X mini batches (shaderio.h:149-158, cluster_classify.comp.glsl:686-737):
// up to 8 base triangles whose factors are all <= 2 share one transient
// CLAS of 6 vertex slots and up to 4 triangles per member; a batch never
// spans two 32-triangle groups of a cluster (the reference's subgroup).
constexpr uint32_t kMiniBatch = 8, kMiniTris = 4, kMiniVerts = 6, kMiniGroup = 32;
constexpr uint32_t kMiniBatchTris = kMiniBatch * kMiniTris, kMiniBatchVerts = kMiniBatch * kMiniVerts;
constexpr uint32_t kMiniVertexSize = kMiniBatchVerts + (kMiniBatchTris * 3 + 11) / 12;  // vertex slots
constexpr uint32_t kMiniPartSize = (8 + kMiniBatchTris * 2 + 24 - 1) / 24;             // part slots

// Everything a stage kernel reads: the bindings (device pointers), the
// configuration and the look-back workspace.
struct KArgs {
  const tc_frame_constants* view;
  const tc_tessellation_table* tessTable;
  tc_readback* readback;
  const tc_render_instance* instances;
  tc_scene_building* build;
  const tc_displacement_texture* textures;
  uint32_t numInstances;
  uint32_t numTextures;
  WsHeader* ws;
  uint64_t* status;  // kMaxCounters x maxTiles words
  uint32_t maxTiles;
  uint32_t _pad;
  tc_vertex_tape* tape;  // non-null with TC_FLAG_VERTEX_TAPE
  tc_config cfg;
  const tc_frame_constants* viewLast;  // culling constants (view when unbound)
  tc_hiz hiz;                          // texels null: no occlusion test
};

// Workspace layout (bytes from the workspace start).
struct WsLayout {
  uint64_t status;       // look-back status words
  uint64_t sortKeys[2];  // radix sort ping-pong keys (u64)
  uint64_t sortVals[2];  // radix sort ping-pong values (u32)
  uint64_t sortStatus;   // per-tile per-digit look-back words of the radix sort
  uint64_t sortState;    // radix sort epoch, tickets and digit bases
  uint64_t sortAux;      // segment starts of the BLAS insertion
  uint64_t cullOffsets;  // per-instance offsets of clusters_cull
  uint64_t total;
  uint32_t maxTiles;
  uint32_t sortCapacity;
  uint32_t sortMaxTiles;
  uint32_t instanceCapacity;
};

constexpr uint32_t kSortThreads = 256;    // threads of a radix-sort pass workgroup
constexpr uint32_t kSortKeysPerThread = 8;
constexpr uint32_t kSortTileKeys = kSortThreads * kSortKeysPerThread;  // keys per pass tile
constexpr uint32_t kSortRadix = 256;


WsLayout ws_layout(const tc_config& cfg, uint32_t numInstances);

tc_status make_args_public(const tc_bindings* b, const tc_config& cfg, KArgs& a);
int num_cu_public();
tc_tuning tuning_snapshot();

__global__ void k_prologue(KArgs a, uint32_t stage);
__global__ void k_build_setup(KArgs a, int mode);
__global__ void k_setup_prologue(KArgs a, int mode, uint32_t stage);
template <bool TWO_X, bool TRANS>
__global__ void k_cluster_classify(KArgs a);
__global__ void k_mini_vertices(KArgs a);
__global__ void k_triangle_split(KArgs a);
__global__ void k_split_queue_init(KArgs a, uint32_t numPasses, uint32_t waitLimit);
__global__ void k_triangle_split_persistent(KArgs a);
__global__ void k_instantiate(KArgs a);
template <bool PN>
__global__ void k_instantiate_vertices(KArgs a);
__global__ void k_blas_setup_insertion(KArgs a);
__global__ void k_instances_classify(KArgs a);
__global__ void k_clusters_cull_scan(KArgs a, uint32_t* offsets);
__global__ void k_clusters_cull_fill(KArgs a, const uint32_t* offsets);

}  // namespace tc
