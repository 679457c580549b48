// tc_bvh.h -- BLAS / TLAS build arguments.
#pragma once

#include "tc_sort.h"

namespace tc {

struct BvhArgs {
  tc_scene_building* build;
  const tc_render_instance* instances;
  uint32_t numInstances;
  uint32_t capacity;  // CLAS references the arrays hold
  // BLAS
  uint64_t* keys;  // sort input
  uint32_t* vals;
  const uint64_t* sortedKeys;
  const uint32_t* sortedVals;
  uint64_t* leafAddr;  // CLAS address per sorted position (blasLeaves)
  uint64_t* leafInfo;  // its header's counts | clusterID << 32 (blasLeafInfo)
  uint32_t* child;     // 2 per internal node, relative to the instance start
  uint32_t* parent;    // [0, capacity) internal nodes, [capacity, 2 capacity) leaves
  uint32_t* counter;   // refit arrival counters
  uint32_t* rangeEnd;  // per internal node: the far end of its key range (global position)
  uint32_t* pendingCount;  // per instance: straddling nodes listed by the chunk refit
  uint32_t* pendingList;   // their node indices, per instance from (first >> 10) + instance
  tc_bvh_node* nodes;  // blasNodes
  tc_bvh4_node* nodes4;  // blasNodes4
  tc_aabb* nodeBox;    // union box of every internal node
  // TLAS
  tc_tlas_instance* tlasInst;
  tc_aabb* tlasLeafBox;
  uint64_t* tlasKeys;
  uint32_t* tlasVals;
  const uint64_t* tlasSortedKeys;
  const uint32_t* tlasSortedVals;
  uint32_t* tlasChild;
  uint32_t* tlasParent;
  uint32_t* tlasCounter;
  tc_bvh_node* tlasNodes;
  tc_bvh4_node* tlasNodes4;
  tc_aabb* tlasNodeBox;
  uint32_t* tlasRoot;  // [0] root word, [1] instance count
};

struct AccelLayout {
  uint64_t blasNodes, blasLeaves, nodeBox, child, parent, counter, rangeEnd, pendingCount, pendingList;
  uint64_t blasNodes4, tlasNodes4, blasLeafInfo;
  uint64_t tlasInstances, tlasNodes, tlasNodeBox, tlasLeafBox, tlasChild, tlasParent, tlasCounter, tlasKeys, tlasVals,
      tlasRoot;
  uint64_t total;
};

AccelLayout accel_layout(uint32_t capacity, uint32_t numInstances);

}  // namespace tc
