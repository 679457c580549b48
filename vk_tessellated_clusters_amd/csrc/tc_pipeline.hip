// tc_pipeline.hip -- the tessellation stages of one frame on gfx950.
//
// Stage kernels follow the reference shaders named at each kernel.  The
// reference sub-allocates every list with order-free atomicAdd; here each
// stage runs as a work queue of tiles taken in ticket order, and a
// decoupled look-back over the tiles' counts gives every tile the exact
// offsets a serial execution would produce.  Output order is therefore the
// canonical serial order: (visible cluster, triangle) for classify,
// (split item, child) for split, part index for instantiate.
#include "tc_stage.h"

namespace tc {

// ------------------------------------------------------------ prologue

// Opens a new look-back launch: bumps the epoch (clearing stale status
// words when it wraps), resets the ticket and snapshots the counters the
// stage advances.  One workgroup.
__device__ void prologue_body(const KArgs& a, uint32_t stage) {
  __shared__ uint32_t s_epoch;
  WsHeader* h = a.ws;
  tc_scene_building* b = a.build;
  if (threadIdx.x == 0) {
    uint32_t e = h->epoch + 1;
    s_epoch = e >= (1u << kEpochBits) ? 0u : e;
  }
  __syncthreads();
  uint32_t e = s_epoch;
  if (e == 0) {
    uint64_t n = (uint64_t)a.maxTiles * kMaxCounters;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) a.status[i] = 0;
    __threadfence();
    e = 1;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  h->epoch = e;
  h->ticket = 0;
  if (stage == kStageClassify) {
    const uint32_t n = min(b->visibleClusterCounter, a.cfg.maxVisibleClusters);
    h->numTiles = min((n + kClassifyTile - 1) / kClassifyTile, a.maxTiles);
    h->itemBegin = 0;
    h->itemEnd = n;
    h->base[0] = (uint32_t)b->splitTriangleCounter;
    h->base[1] = b->partTriangleCounter;
    h->base[2] = b->genClusterCounter;
    h->base[3] = b->genVertexCounter;
    h->base[4] = b->genClusterDataCounter;
    h->base[5] = b->tempInstantiateCounter;
    h->base[6] = (uint32_t)(b->dualPartTriangleCounter >> 32);
    h->base[7] = b->transBuildCounter;
  } else if (stage == kStageSplit) {
    uint32_t s = min(b->splitPassStart, a.cfg.maxSplitTriangles), en = min(b->splitPassEnd, a.cfg.maxSplitTriangles);
    uint32_t n = en > s ? en - s : 0;
    h->numTiles = min((n + kWave - 1) / kWave, a.maxTiles);
    h->itemBegin = s;
    h->itemEnd = s + n;
    h->base[0] = b->splitWriteCounter;
    h->base[1] = b->partTriangleCounter;
    h->base[2] = (uint32_t)(b->dualPartTriangleCounter >> 32);
  } else {
    uint32_t n = min(b->partTriangleCounter, a.cfg.maxPartTriangles);
    h->numTiles = min((n + kInstTile - 1) / kInstTile, a.maxTiles);
    h->itemBegin = 0;
    h->itemEnd = n;
    h->base[0] = b->genClusterCounter;
    h->base[1] = b->genVertexCounter;
    h->base[2] = b->genClusterDataCounter;
    h->base[3] = b->tempInstantiateCounter;
  }
}

__global__ __launch_bounds__(256) void k_prologue(KArgs a, uint32_t stage) { prologue_body(a, stage); }

// ---------------------------------------------------------- build_setup

// build_setup.comp.glsl:102-264 (transient builds with TC_FLAG_1X_TRANSIENT
// or TC_FLAG_2X_TRANSIENT).  One thread.
__device__ void build_setup_one(const KArgs& a, int mode) {
  tc_scene_building* b = a.build;
  tc_readback* rb = a.readback;
  const tc_config& cfg = a.cfg;
  if (mode == TC_BUILD_SETUP_CLASSIFY) {
    uint32_t counter = b->visibleClusterCounter;
    rb->numVisibleClusters = counter;
    counter = min(counter, cfg.maxVisibleClusters);
    b->visibleClusterCounter = counter;
    b->dispatchClassify = tc_dispatch_indirect{counter, 1, 1};
  } else if (mode == TC_BUILD_SETUP_DRAW_TESS) {
    uint32_t full = b->fullClusterCounter, part = b->partTriangleCounter;
    rb->numFullClusters = full;
    rb->numPartTriangles = part;
    rb->numSplitTriangles = b->splitWriteCounter;
    full = min(full, cfg.maxVisibleClusters);
    part = min(part, cfg.maxPartTriangles);
    b->drawFullClusters = tc_draw_mesh_tasks_indirect{full, 0};
    b->drawPartTriangles = tc_draw_mesh_tasks_indirect{(part + 31) / 32, 0};
    rb->numBlasClusters = full;
  } else if (mode == TC_BUILD_SETUP_SPLIT) {
    uint32_t count = min((uint32_t)b->splitTriangleCounter, cfg.maxSplitTriangles);
    b->splitWriteCounter = count;
    b->splitTriangleCounter = (int32_t)count;
    b->splitPassStart = 0;
    b->splitPassEnd = count;
    b->dispatchTriangleSplit = tc_dispatch_indirect{(count + 63) / 64, 1, 1};
  } else if (mode == TC_BUILD_SETUP_SPLIT_PASS) {
    b->splitPass = b->splitPass + 1;
    uint32_t start = min(b->splitPassEnd, cfg.maxSplitTriangles);
    uint32_t end = min(b->splitWriteCounter, cfg.maxSplitTriangles);
    b->splitPassStart = start;
    b->splitPassEnd = end;
    b->dispatchTriangleSplit = tc_dispatch_indirect{(end - start + 63) / 64, 1, 1};
  } else if (mode == TC_BUILD_SETUP_BUILD_BLAS) {
    const bool trans = transient_builds(cfg.flags);
    uint32_t temp = b->tempInstantiateCounter;
    uint32_t tr = trans ? b->transBuildCounter : 0u;
    rb->numBlasClusters = temp + tr;
    if (trans) rb->numTransBuilds = tr;
    rb->numTempInstantiations = temp;
    rb->numGenDatas = b->genClusterDataCounter;
    rb->numGenVertices = b->genVertexCounter;
    rb->numBlasReservedSizes = b->numBlasReservedSizes;
    temp = min(cfg.maxGeneratedClusters, temp);
    b->tempInstantiateCounter = temp;
    rb->numActualTempInstantiations = temp;
    b->dispatchBlasTempInsert = tc_dispatch_indirect{(temp + 63) / 64, 1, 1};
    if (trans) {
      tr = min(cfg.maxGeneratedClusters, temp + tr) - temp;
      b->transBuildCounter = tr;
      rb->numActualTransBuilds = tr;
      b->dispatchBlasTransInsert = tc_dispatch_indirect{(tr + 63) / 64, 1, 1};
    }
  } else if (mode == TC_BUILD_SETUP_INSTANTIATE_TESS) {
    uint32_t part = b->partTriangleCounter;
    rb->numSplitTriangles = b->splitWriteCounter;
    if (transient_builds(cfg.flags)) {
      // :240-253: the instantiated parts are the accepted prefix
      // x + y + 2 <= MAX_PART_TRIANGLES of the lower counter
      const uint32_t y = (uint32_t)(b->dualPartTriangleCounter >> 32);
      rb->numPartTriangles = part + y;
      rb->numTransPartTriangles = y;
      const uint32_t cap = cfg.maxPartTriangles > y + 1 ? cfg.maxPartTriangles - y - 1 : 0u;
      part = min(part, cap);
    } else {
      rb->numPartTriangles = part;
      part = min(part, cfg.maxPartTriangles);
    }
    b->partTriangleCounter = part;
    b->dispatchTriangleInstantiate = tc_dispatch_indirect{(part + 31) / 32, 1, 1};
  }
}

__global__ void k_build_setup(KArgs a, int mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  build_setup_one(a, mode);
}

// A build_setup followed by the next stage's prologue in one launch: the
// prologue's snapshot is taken by the same thread after the setup's writes.
__global__ __launch_bounds__(256) void k_setup_prologue(KArgs a, int mode, uint32_t stage) {
  if (threadIdx.x == 0) build_setup_one(a, mode);
  __syncthreads();
  prologue_body(a, stage);
}

// -------------------------------------------------------- triangle_split

struct SplitItem {
  v3 pos[3];   // world-space base triangle (fillBaseVertices)
  v3 bary[3];  // decoded sub-triangle corners
  uint32_t cfg;
  uint32_t firstTriangle;
  uint32_t firstVertex;
  uint32_t start;  // first virtual thread
  uint32_t count;  // children
  tc_cluster_info cluster;
  uint32_t triangleID_config;
};

// triangle_split.comp.glsl:201-252: corners and factors of one child
__device__ __forceinline__ void split_child(const TableView& T, const SplitItem& it, uint32_t child, const ViewParams& vp,
                                            uint32_t vtx[3], uint32_t f[3]) {
  uint32_t vi[3];
  cfg_tri_vertices(T, it.cfg, it.firstTriangle, child, vi);
#pragma unroll
  for (int v = 0; v < 3; v++) {
    v3 w = cfg_vertex_bary(T, it.cfg, it.firstVertex, vi[v]);
    vtx[v] = encode_bary(interp(it.bary[0], it.bary[1], it.bary[2], w));
  }
  v3 wp[3];
#pragma unroll
  for (int v = 0; v < 3; v++) wp[v] = interp(it.pos[0], it.pos[1], it.pos[2], decode_bary(vtx[v]));
  tess_factors(wp[0], wp[1], wp[2], vp, f);
}


// triangle_split.comp.glsl:598-632 + processAllSubTasks :350-472 for one
// tile of 64 split items (one wave); children packed densely over the wave.
// Counters: 0 split, 1 part.  The last tile of a pass writes the pass's
// split / part totals to SceneBuilding.
__device__ __forceinline__ void split_tile(const KArgs& a, const TableView& T, const ViewParams& vp, SplitItem* s_items,
                                           uint32_t* s_start, uint32_t tile, uint32_t numTiles, uint32_t epoch,
                                           uint32_t itemBegin, uint32_t itemEnd, uint64_t B0, uint64_t B1, uint64_t transY,
                                           const uint32_t* abort = nullptr) {
  const uint32_t lane = lane_id();
  tc_scene_building* b = a.build;
  const tc_config& cfg = a.cfg;
  // setupTask + fillBaseVertices (:146-199) once per item
  const uint32_t idx = itemBegin + tile * kWave + lane;
  bool valid = idx < itemEnd;
  uint32_t count = 0;
  {
    SplitItem it;
    tc_tess_triangle_info ti;
    if (valid) {
      ti = ptr<const tc_tess_triangle_info>(b->splitTriangles)[idx];
      valid = ti.cluster.instanceID < a.numInstances && cfg_index(ti.subTriangle.triangleID_config >> 16) < TC_TESSTABLE_NUM_LOOKUP &&
              ti.cluster.clusterID < a.instances[ti.cluster.instanceID].numClusters;
    }
    if (valid) {
      it.cluster = ti.cluster;
      it.triangleID_config = ti.subTriangle.triangleID_config;
      it.cfg = ti.subTriangle.triangleID_config >> 16;
      tc_tess_table_entry e = cfg_entry(T, it.cfg);
      it.firstTriangle = e.firstTriangle;
      it.firstVertex = e.firstVertex;
      count = e.numTriangles;
      const tc_render_instance& inst = a.instances[ti.cluster.instanceID];
      const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ti.cluster.clusterID];
      const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
      uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        it.pos[v] = xform_point(inst.worldMatrix, ld3(inst.positions, lt[tri * 3 + v] + cl.firstLocalVertex));
        it.bary[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
      }
    }
    uint32_t incl = wave_incl_scan(count);
    it.start = incl - count;
    it.count = count;
    s_items[lane] = it;
    s_start[lane] = valid ? it.start : 0xFFFFFFFFu;
  }
  const uint32_t total = shfl(wave_incl_scan(count), kWave - 1);
  __syncthreads();

  // phase A: count children going to the split and part lists
  uint32_t ns = 0, np = 0;
  for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
    const uint32_t t = t0 + lane;
    bool needSplit = false, needPart = false;
    if (t < total) {
      const SplitItem& it = s_items[find_owner(s_start, t)];
      uint32_t vtx[3], f[3];
      split_child(T, it, t - it.start, vp, vtx, f);
      uint32_t mx = umax3(f[0], f[1], f[2]);
      needSplit = mx > TC_TESSTABLE_SIZE;
      needPart = !needSplit;
    }
    ns += __popcll(ballot(needSplit));
    np += __popcll(ballot(needPart));
  }
  uint64_t agg[2] = {ns, np}, excl[2];
  lookback<2>(a.status, a.maxTiles, epoch, tile, agg, excl, abort);

  // phase B: recompute and emit in (item, child) order
  uint32_t runSplit = (uint32_t)(B0 + excl[0]), runPart = (uint32_t)(B1 + excl[1]);
  for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
    const uint32_t t = t0 + lane;
    bool needSplit = false, needPart = false;
    uint32_t vtx[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    const SplitItem* itp = nullptr;
    if (t < total) {
      itp = &s_items[find_owner(s_start, t)];
      split_child(T, *itp, t - itp->start, vp, vtx, f);
      uint32_t mx = umax3(f[0], f[1], f[2]);
      needSplit = mx > TC_TESSTABLE_SIZE;
      needPart = !needSplit;
    }
    uint64_t vs = ballot(needSplit), vpm = ballot(needPart);
    uint32_t offSplit = runSplit + mask_prefix(vs);
    uint32_t offPart = runPart + mask_prefix(vpm);
    runSplit += __popcll(vs);
    runPart += __popcll(vpm);
    if (needSplit || needPart) {
      if (needSplit) {
        f[0] = split_factor(f[0], cfg.maxSplitFactor);
        f[1] = split_factor(f[1], cfg.maxSplitFactor);
        f[2] = split_factor(f[2], cfg.maxSplitFactor);
      }
      uint32_t cf = get_config(f, vtx);
      tc_tess_triangle_info ti;
      ti.cluster = itp->cluster;
      ti.subTriangle.vtxEncoded[0] = vtx[0];
      ti.subTriangle.vtxEncoded[1] = vtx[1];
      ti.subTriangle.vtxEncoded[2] = vtx[2];
      ti.subTriangle.triangleID_config = (itp->triangleID_config & 0xFFFFu) | (cf << 16);
      // with transient builds the part list shares its top with their
      // slots (build.glsl:68-83): x + y + 2 <= MAX_PART_TRIANGLES
      const bool partOk = transient_builds(cfg.flags)
                              ? (uint64_t)offPart + transY + 2 <= cfg.maxPartTriangles
                              : offPart < cfg.maxPartTriangles;
      if (needSplit && offSplit < cfg.maxSplitTriangles)
        ptr<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
      else if (needPart && partOk)
        ptr<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
    }
  }

  if (tile == numTiles - 1 && lane == 0) {
    b->splitWriteCounter = (uint32_t)(B0 + excl[0] + agg[0]);
    b->partTriangleCounter = (uint32_t)(B1 + excl[1] + agg[1]);
    b->dualPartTriangleCounter = (transY << 32) | (uint32_t)(B1 + excl[1] + agg[1]);
  }
  __syncthreads();
}

// One split pass (multipass form, the reference default): tiles by ticket.
__global__ __launch_bounds__(64) void k_triangle_split(KArgs a) {
  __shared__ SplitItem s_items[kWave];
  __shared__ uint32_t s_start[kWave];
  WsHeader* h = a.ws;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, itemBegin = h->itemBegin, itemEnd = h->itemEnd;
  const uint64_t B0 = h->base[0], B1 = h->base[1], transY = h->base[2];
  const ViewParams vp = view_params(a);
  const TableView T = load_table(a.tessTable);
  while (true) {
    uint32_t tile = 0;
    if (lane_id() == 0) tile = take_ticket_below(h, numTiles);
    tile = uniform(tile);
    if (tile >= numTiles) break;
    split_tile(a, T, vp, s_items, s_start, tile, numTiles, epoch, itemBegin, itemEnd, B0, B1, transY);
  }
}

// ---------------------------------------------- persistent triangle_split

// Loads of words other waves write during the launch (never cached in the
// scalar cache).
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// build_setup SPLIT_PASS (build_setup.comp.glsl:168-190) for the passes
// after `p`, publishing the first of them that has items; when none is
// left the queue is finished.  One thread, after pass p completed.
__device__ void publish_split_passes(KArgs a, SplitQueue* q, WsHeader* h, uint32_t p) {
  tc_scene_building* b = a.build;
  const tc_config& cfg = a.cfg;
  uint32_t tileEnd = ld_agent(&q->pass[p].tileBegin) + ld_agent(&q->pass[p].numTiles), epoch = ld_agent(&q->pass[p].epoch);
  const uint32_t numPasses = ld_agent(&q->numPasses);
  for (uint32_t np = p + 1; np < numPasses; np++) {
    b->splitPass = ld_agent(&b->splitPass) + 1;
    const uint32_t start = min(ld_agent(&b->splitPassEnd), cfg.maxSplitTriangles);
    const uint32_t writeCounter = ld_agent(&b->splitWriteCounter);
    const uint32_t end = min(writeCounter, cfg.maxSplitTriangles);
    b->splitPassStart = start;
    b->splitPassEnd = end;
    b->dispatchTriangleSplit = tc_dispatch_indirect{(end - start + 63) / 64, 1, 1};
    const uint32_t n = end > start ? end - start : 0u;
    SplitPass& nx = q->pass[np];
    nx.tileBegin = tileEnd;
    nx.numTiles = min((n + kWave - 1) / kWave, a.maxTiles);
    nx.itemBegin = start;
    nx.itemEnd = start + n;
    nx.done = 0;
    nx.epoch = ++epoch;
    nx.b0 = writeCounter;
    nx.b1 = ld_agent(&b->partTriangleCounter);
    h->epoch = epoch;  // the next stage's prologue continues from here
    if (nx.numTiles) {
      __hip_atomic_store(&q->passReady, np + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
  __hip_atomic_store(&q->passReady, numPasses, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&q->finished, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Pass 0 of the persistent split from the prologue's snapshot (workspace
// header), with look-back epochs for every pass.
__global__ __launch_bounds__(256) void k_split_queue_init(KArgs a, uint32_t numPasses, uint32_t waitLimit) {
  __shared__ uint32_t s_wrap;
  WsHeader* h = a.ws;
  SplitQueue* q = split_queue(h);
  if (threadIdx.x == 0) s_wrap = h->epoch + numPasses >= (1u << kEpochBits);
  __syncthreads();
  if (s_wrap) {
    // the passes' epochs would wrap: clear the status words, restart at 1
    const uint64_t n = (uint64_t)a.maxTiles * kMaxCounters;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) a.status[i] = 0;
    __threadfence();
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (s_wrap) h->epoch = 1;
  q->numPasses = min(numPasses, kMaxSplitPasses);
  q->error = 0;
  q->waitLimit = waitLimit;
  q->finished = 0;
  SplitPass& p0 = q->pass[0];
  p0.tileBegin = 0;
  p0.numTiles = h->numTiles;
  p0.itemBegin = h->itemBegin;
  p0.itemEnd = h->itemEnd;
  p0.done = 0;
  p0.epoch = h->epoch;
  p0.b0 = h->base[0];
  p0.b1 = h->base[1];
  if (q->numPasses == 0) {
    q->passReady = 0;
    q->finished = 1;
  } else if (p0.numTiles) {
    q->passReady = 1;
  } else {
    q->passReady = 1;
    publish_split_passes(a, q, h, 0);
  }
  __threadfence();
}

// triangle_split.comp.glsl:476-594 (TESS_USE_PERSISTENT_KERNEL) as one
// launch over every pass: tickets are global over the passes, a ticket of
// a pass not yet published waits for it, and the wave completing a pass
// publishes the next.  Every ticket before a waiting one is held by a
// running wave that does not wait on anything later, so the waits end; a
// wait is still bounded and sets `error` when the bound is hit.  The split
// and part lists, counters and SPLIT_PASS state equal the multipass form's.
__global__ __launch_bounds__(64) void k_triangle_split_persistent(KArgs a) {
  __shared__ SplitItem s_items[kWave];
  __shared__ uint32_t s_start[kWave];
  WsHeader* h = a.ws;
  SplitQueue* q = split_queue(h);
  const uint32_t lane = lane_id();
  const uint64_t transY = h->base[2];
  const ViewParams vp = view_params(a);
  const TableView T = load_table(a.tessTable);
  const uint32_t waitLimit = ld_agent(&q->waitLimit);
  // p: the pass of this wave's last tile; seen: the passes published when
  // this wave last acquired the pass table (lane 0's copies)
  uint32_t p = 0, seen = 0;
  while (true) {
    // A finished queue has no tile left: leave without an atomic on the one
    // ticket word (every wave of the launch, when nothing splits).
    uint32_t g = 0, fin = 0;
    if (lane == 0) {
      fin = ld_agent(&q->finished);
      if (!fin) g = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (uniform(fin)) break;
    g = uniform(g);
    uint32_t found = 0xFFFFFFFFu;
    if (lane == 0) {
      for (uint32_t spins = 0;; spins++) {
        while (p < seen && g >= ld_agent(&q->pass[p].tileBegin) + ld_agent(&q->pass[p].numTiles)) p++;
        if (p < seen) {
          found = p;
          break;
        }
        // Poll with plain agent-scope loads: an acquire at that scope
        // invalidates the XCD's L2, and thousands of waiting waves would
        // keep clearing it under the waves at work.  The acquire runs once
        // per publication seen.
        const uint32_t ready = ld_agent(&q->passReady);
        if (ready > seen) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          seen = ready;
          continue;
        }
        // passReady reached numPasses before `finished` was set, so a wave
        // that finds the queue finished holds a ticket past every pass
        if (ld_agent(&q->finished)) break;
        // a wait far beyond any frame's split time: flag the launch and end
        // every wait (look-backs included) so the grid drains
        if (spins >= waitLimit || ld_agent(&q->error)) {
          __hip_atomic_store(&q->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    found = uniform(found);
    if (found == 0xFFFFFFFFu) {
      // a launch that gave up reports it with the frame's readback
      if (lane == 0 && ld_agent(&q->error))
        __hip_atomic_fetch_or(&a.readback->errorFlags, TC_READBACK_ERROR_SPLIT_WAIT, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    p = found;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const SplitPass& Q = q->pass[p];
    const uint32_t tileBegin = ld_agent(&Q.tileBegin), numTiles = ld_agent(&Q.numTiles);
    split_tile(a, T, vp, s_items, s_start, g - tileBegin, numTiles, ld_agent(&Q.epoch), ld_agent(&Q.itemBegin),
               ld_agent(&Q.itemEnd), ld_agent(&Q.b0), ld_agent(&Q.b1), transY, &q->error);
    // completion: this tile's lists and counters are visible before it counts
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) {
      const uint32_t old = __hip_atomic_fetch_add(&q->pass[p].done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == numTiles) {
        // every other tile of the pass released before it counted
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        publish_split_passes(a, q, h, p);
      }
    }
  }
}

// ------------------------------------------------- blas_setup_insertion

// blas_setup_insertion.comp.glsl:100-116 as an exclusive scan over
// instances in index order (one workgroup).
__global__ __launch_bounds__(1024) void k_blas_setup_insertion(KArgs a) {
  __shared__ uint32_t s_wave[16];
  __shared__ uint32_t s_carry;
  __shared__ uint32_t s_sizes;
  tc_scene_building* b = a.build;
  tc_blas_build_info* infos = ptr<tc_blas_build_info>(b->blasBuildInfos);
  const uint32_t* sizes = ptr<const uint32_t>(b->blasBuildSizes);
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_carry = b->blasClusterCounter, s_sizes = 0;
  __syncthreads();
  uint32_t sizeSum = 0;
  for (uint32_t i0 = 0; i0 < a.numInstances; i0 += blockDim.x) {
    const uint32_t i = i0 + threadIdx.x;
    const uint32_t cnt = i < a.numInstances ? infos[i].clusterReferencesCount : 0u;
    if (i < a.numInstances) sizeSum += sizes[i];
    const uint32_t incl = wave_incl_scan(cnt);
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t waveBase = 0, blockTotal = 0;
    for (uint32_t w = 0; w < blockDim.x / kWave; w++) {
      if (w < wave) waveBase += s_wave[w];
      blockTotal += s_wave[w];
    }
    const uint32_t off = s_carry + waveBase + incl - cnt;
    if (i < a.numInstances) {
      tc_blas_build_info bi;
      bi.clusterReferencesCount = 0;
      bi.clusterReferencesStride = 8;
      bi.clusterReferences = b->blasClusterAddresses + (uint64_t)off * 8ull;
      infos[i] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_carry += blockTotal;
    __syncthreads();
  }
  sizeSum = wave_sum(sizeSum);
  if (lane == 0) atomicAdd(&s_sizes, sizeSum);
  __syncthreads();
  if (threadIdx.x == 0) {
    b->blasClusterCounter = s_carry;
    a.readback->numBlasActualSizes += s_sizes;
  }
}

// ------------------------------------------ culling front-end (fixture)

__device__ __forceinline__ void xform4(const float* m, const float p[4], float o[4]) {
#pragma unroll
  for (int r = 0; r < 4; r++) o[r] = ((m[r] * p[0] + m[4 + r] * p[1]) + m[8 + r] * p[2]) + m[12 + r] * p[3];
}

// instances_classify.comp.glsl:100-127 with culling.glsl:30-90 (frustum and
// size test; there is no HiZ pyramid, so the occlusion test passes)
__global__ void k_instances_classify(KArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.numInstances) return;
  const tc_render_instance& inst = a.instances[i];
  const tc_frame_constants* vw = a.viewLast;  // culling.glsl reads viewLast
  float wvp[16];
  for (int col = 0; col < 4; col++)
    for (int r = 0; r < 4; r++)
      wvp[col * 4 + r] = ((vw->viewProjMatrix[r] * inst.worldMatrix[col * 4 + 0] + vw->viewProjMatrix[4 + r] * inst.worldMatrix[col * 4 + 1]) +
                          vw->viewProjMatrix[8 + r] * inst.worldMatrix[col * 4 + 2]) +
                         vw->viewProjMatrix[12 + r] * inst.worldMatrix[col * 4 + 3];
  uint32_t bits = 0x7F;
  float cmin[4], cmax[4];
  bool clipValid = true;
  for (int n = 0; n < 8; n++) {
    float p[4] = {(n & 1) ? inst.geoHi[0] : inst.geoLo[0], (n & 2) ? inst.geoHi[1] : inst.geoLo[1],
                  (n & 4) ? inst.geoHi[2] : inst.geoLo[2], 1.0f};
    float hp[4];
    xform4(wvp, p, hp);
    bool valid = !(-1.2e-07f < hp[3] && hp[3] < 1.2e-07f);
    float aw = fabsf(hp[3]);
    float clip[4] = {hp[0] / aw, hp[1] / aw, hp[2] / aw, hp[3]};
    uint32_t cb = 0;
    cb |= hp[0] < -hp[3] ? 1u : 0u;
    cb |= hp[0] > hp[3] ? 2u : 0u;
    cb |= hp[1] < -hp[3] ? 4u : 0u;
    cb |= hp[1] > hp[3] ? 8u : 0u;
    cb |= hp[2] < 0 ? 16u : 0u;
    cb |= hp[2] > hp[3] ? 32u : 0u;
    cb |= hp[3] <= 0 ? 64u : 0u;
    bits &= cb;
    for (int k = 0; k < 4; k++) {
      cmin[k] = n == 0 ? clip[k] : fminf(cmin[k], clip[k]);
      cmax[k] = n == 0 ? clip[k] : fmaxf(cmax[k], clip[k]);
    }
    clipValid = clipValid && valid;
  }
  for (int k = 0; k < 2; k++) {
    cmin[k] = fminf(fmaxf(cmin[k], -1.0f), 1.0f);
    cmax[k] = fminf(fmaxf(cmax[k], -1.0f), 1.0f);
  }
  const bool inFrustum = bits == 0;
  const bool sizeOk = (cmax[0] - cmin[0]) > 2.0f / vw->viewportf[0] || (cmax[1] - cmin[1]) > 2.0f / vw->viewportf[1];
  // :120 the occlusion test against the previous frame's far HiZ
  const bool hizOk = !a.hiz.texels || !(inFrustum && clipValid && sizeOk) || intersect_hiz(a.hiz, vw, cmin, cmax);
  const bool visible = inFrustum && (!clipValid || (sizeOk && hizOk));
  ptr<uint32_t>(a.build->instanceStates)[i] = (inFrustum ? TC_INSTANCE_FRUSTUM_BIT : 0u) | (visible ? TC_INSTANCE_VISIBLE_BIT : 0u);
  ptr<tc_blas_build_info>(a.build->blasBuildInfos)[i].clusterReferencesCount = 0;
}

// clusters_cull.comp.glsl:112-161 in ray-tracing mode (every cluster of
// every instance, instance by instance).  Workgroup i appends instance i at
// the offset of the exclusive scan held in `offsets`.
// One wave: exclusive scan of the instances' cluster counts, 64 instances
// at a time, after the clusters already appended.
__global__ __launch_bounds__(64) void k_clusters_cull_scan(KArgs a, uint32_t* offsets) {
  if (blockIdx.x != 0) return;
  const uint32_t lane = lane_id();
  uint32_t run = a.build->visibleClusterCounter;
  for (uint32_t i0 = 0; i0 < a.numInstances; i0 += kWave) {
    const uint32_t i = i0 + lane;
    const uint32_t n = i < a.numInstances ? a.instances[i].numClusters : 0u;
    const uint32_t incl = wave_incl_scan(n);
    if (i < a.numInstances) offsets[i] = run + incl - n;
    run += shfl(incl, kWave - 1);
  }
  if (lane == 0) a.build->visibleClusterCounter = run;
}

__global__ void k_clusters_cull_fill(KArgs a, const uint32_t* offsets) {
  const uint32_t i = blockIdx.x;
  const uint32_t n = a.instances[i].numClusters, off = offsets[i];
  tc_cluster_info* out = ptr<tc_cluster_info>(a.build->visibleClusters);
  for (uint32_t c = threadIdx.x; c < n; c += blockDim.x)
    if (off + c < a.cfg.maxVisibleClusters) out[off + c] = tc_cluster_info{i, c};
}

}  // namespace tc
