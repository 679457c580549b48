// tc_trace.h -- traversal kernel arguments.
#pragma once

#include "tc_kernels.h"

namespace tc {

// Tile-row order of camera traces (a renderer's own traces).  A launch's
// span is bounded by its longest waves (8 x 8 tiles of grazing rays near
// the horizon), and waves are dispatched in tile order.  Each wave times its
// tile and the longest time of each tile row is kept; the renderer's next
// frame start ranks the rows by it, longest first, and the traces of that
// frame with the same height take the rows in that order, so the costly
// rows start first and the cheap ones fill the end.  That tail is a small
// part of a launch whose waves run many rounds per resident wave slot, and
// there the timing atomics cost more than the order saves: launches of more
// than kOrderRounds waves per slot are neither timed nor ordered.  Results
// do not depend on the order.
constexpr uint32_t kOrderRounds = 16;
constexpr uint32_t kMaxOrderRows = 1536;
struct TraceOrder {
  uint32_t rows;         // tile rows `order` ranks (0: none yet)
  uint32_t pendingRows;  // tile rows of the traces timed since the last ranking (0: none)
  uint32_t _pad[2];
  uint32_t rowTime[kMaxOrderRows];  // longest wave of each tile row, 10 ns ticks
  uint32_t order[kMaxOrderRows];    // dispatch position -> tile row
};

// Ranks the timed rows longest first (ties by row) into `order` and clears
// the times; no-op when nothing was timed.  One workgroup; `lds` holds
// kMaxOrderRows words.
__device__ __forceinline__ void rank_tile_rows(TraceOrder* ord, uint32_t* lds) {
  const uint32_t rows = ord->pendingRows;
  if (rows == 0 || rows > kMaxOrderRows) return;
  for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x) lds[r] = ord->rowTime[r];
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < rows; r += blockDim.x) {
    const uint32_t c = lds[r];
    uint32_t rank = 0;
    for (uint32_t q = 0; q < rows; q++) {
      const uint32_t cq = lds[q];
      rank += (cq > c || (cq == c && q < r)) ? 1u : 0u;
    }
    ord->order[rank] = r;
  }
  for (uint32_t r = threadIdx.x; r < kMaxOrderRows; r += blockDim.x) ord->rowTime[r] = 0;
  if (threadIdx.x == 0) ord->rows = rows, ord->pendingRows = 0;
}

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
  TraceOrder* order;  // camera traces of a renderer, else null
};

// tc_trace_rays with the renderer's tile-row order state (null: dispatch in
// tile order).
tc_status trace_rays(const tc_bindings* b, tc_config config, const tc_accel* accel, const tc_trace_params* p,
                     TraceOrder* order, tc_stream stream);

}  // namespace tc
