// tc_trace.h -- traversal kernel arguments.
#pragma once

#include "tc_kernels.h"

namespace tc {

struct TraceArgs {
  const tc_frame_constants* view;
  const tc_tessellation_table* tessTable;
  const tc_scene_building* build;
  const tc_tlas_instance* tlasInstances;
  const tc_bvh_node* tlasNodes;
  const tc_bvh4_node* tlasNodes4;
  const tc_bvh4_node* blasNodes4;
  uint64_t blasNodesBase;  // tc_tlas_instance.blasNodes - this = first node of the instance (x 64 B)
  uint64_t blasLeavesBase;          // tc_tlas_instance.blasLeaves - this = first reference (x 8 B)
  const uint64_t* blasLeafInfo;     // per reference: counts | clusterID << 32
  const uint32_t* tlasRoot;
  const tc_ray* rays;
  uint32_t width, height, numRays, flags;
  uint32_t numItems;  // camera rays: 64 per wave of the strip walk (some waves idle)
  tc_hit* hits;
  tc_hit_decoded* decoded;
  uint32_t* hitVertices;
};

}  // namespace tc
