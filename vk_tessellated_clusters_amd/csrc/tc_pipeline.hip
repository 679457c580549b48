// tc_pipeline.hip -- the tessellation stages of one frame on gfx950.
//
// Stage kernels follow the reference shaders named at each kernel.  The
// reference sub-allocates every list with order-free atomicAdd; here each
// stage runs as a work queue of tiles taken in ticket order, and a
// decoupled look-back over the tiles' counts gives every tile the exact
// offsets a serial execution would produce.  Output order is therefore the
// canonical serial order: (visible cluster, triangle) for classify,
// (split item, child) for split, part index for instantiate.
#include "tc_kernels.h"

namespace tc {

// ------------------------------------------------------------ prologue

enum Stage : uint32_t { kStageClassify = 0, kStageSplit = 1, kStageInstantiate = 2 };

// Opens a new look-back launch: bumps the epoch (clearing stale status
// words when it wraps), resets the ticket and snapshots the counters the
// stage advances.  One workgroup.
__global__ __launch_bounds__(256) void k_prologue(KArgs a, uint32_t stage) {
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
    const uint32_t n = min(min(b->visibleClusterCounter, a.cfg.maxVisibleClusters), a.maxTiles);
    h->numTiles = n;
    h->itemBegin = 0;
    h->itemEnd = n;
    h->base[0] = (uint32_t)b->splitTriangleCounter;
    h->base[1] = b->partTriangleCounter;
    h->base[2] = b->genClusterCounter;
    h->base[3] = b->genVertexCounter;
    h->base[4] = b->genClusterDataCounter;
    h->base[5] = b->tempInstantiateCounter;
  } else if (stage == kStageSplit) {
    uint32_t s = min(b->splitPassStart, a.cfg.maxSplitTriangles), en = min(b->splitPassEnd, a.cfg.maxSplitTriangles);
    uint32_t n = en > s ? en - s : 0;
    h->numTiles = (n + kWave - 1) / kWave;
    h->itemBegin = s;
    h->itemEnd = s + n;
    h->base[0] = b->splitWriteCounter;
    h->base[1] = b->partTriangleCounter;
  } else {
    uint32_t n = min(b->partTriangleCounter, a.cfg.maxPartTriangles);
    h->numTiles = (n + kWave - 1) / kWave;
    h->itemBegin = 0;
    h->itemEnd = n;
    h->base[0] = b->genClusterCounter;
    h->base[1] = b->genVertexCounter;
    h->base[2] = b->genClusterDataCounter;
    h->base[3] = b->tempInstantiateCounter;
  }
}

// ---------------------------------------------------------- build_setup

// build_setup.comp.glsl:102-264 (no transient builds)
__global__ void k_build_setup(KArgs a, int mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    uint32_t temp = b->tempInstantiateCounter;
    rb->numBlasClusters = temp;
    rb->numTempInstantiations = temp;
    rb->numGenDatas = b->genClusterDataCounter;
    rb->numGenVertices = b->genVertexCounter;
    rb->numBlasReservedSizes = b->numBlasReservedSizes;
    temp = min(cfg.maxGeneratedClusters, temp);
    b->tempInstantiateCounter = temp;
    rb->numActualTempInstantiations = temp;
    b->dispatchBlasTempInsert = tc_dispatch_indirect{(temp + 63) / 64, 1, 1};
  } else if (mode == TC_BUILD_SETUP_INSTANTIATE_TESS) {
    uint32_t part = b->partTriangleCounter;
    rb->numPartTriangles = part;
    rb->numSplitTriangles = b->splitWriteCounter;
    part = min(part, cfg.maxPartTriangles);
    b->partTriangleCounter = part;
    b->dispatchTriangleInstantiate = tc_dispatch_indirect{(part + 31) / 32, 1, 1};
  }
}

// ------------------------------------------------------ cluster_classify

__device__ __forceinline__ ViewParams view_params(const KArgs& a) {
  ViewParams vp;
  vp.eye = mk(a.build->viewPos[0], a.build->viewPos[1], a.build->viewPos[2]);
  vp.nearPlane = a.view->nearPlane;
  vp.vpy = a.view->viewportf[1];
  vp.rate = a.view->tessRate;
  return vp;
}

__device__ __forceinline__ bool has_disp(const KArgs& a, int32_t idx) {
  return (a.cfg.flags & TC_FLAG_DISPLACEMENT) && idx >= 0 && (uint32_t)idx < a.numTextures;
}

__device__ __forceinline__ float displaced_height(const KArgs& a, const tc_displacement_texture& tex, float scale,
                                                  float offset, v2 uv) {
  float h = sample_height(tex, uv.x, uv.y);
  return (((h * scale) * a.view->displacementScale) + offset) + a.view->displacementOffset;
}

__device__ __forceinline__ uint64_t max_data_bytes(const tc_config& c) {
  return (uint64_t)c.maxGeneratedClusterMegs * 1024ull * 1024ull;
}

// cluster_classify.comp.glsl:150-666, one visible cluster per tile, one
// wave per workgroup.  Counters: 0 split, 1 part, 2 genCluster,
// 3 genVertex, 4 genData.
__global__ __launch_bounds__(64) void k_cluster_classify(KArgs a) {
  __shared__ float s_pos[kMaxClusterVerts * 3];
  __shared__ uint32_t s_fac[kMaxClusterTris * 3];

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id();
  const uint32_t epoch = h->epoch, numTiles = h->numTiles;
  const uint64_t B0 = h->base[0], B1 = h->base[1], B2 = h->base[2], B3 = h->base[3], B4 = h->base[4], B5 = h->base[5];
  const ViewParams vp = view_params(a);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;

  while (true) {
    const uint32_t tile = take_ticket(h);
    if (tile >= numTiles) break;

    tc_cluster_info ci = ptr<const tc_cluster_info>(b->visibleClusters)[tile];
    // malformed entries are reduced to an in-bounds cluster
    if (ci.instanceID >= a.numInstances) ci.instanceID = 0;
    const tc_render_instance& inst = a.instances[ci.instanceID];
    if (ci.clusterID >= inst.numClusters) ci.clusterID = 0;
    const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ci.clusterID];
    const uint32_t state = ptr<const uint32_t>(b->instanceStates)[ci.instanceID];
    const uint32_t nv = min((uint32_t)cl.numVertices, kMaxClusterVerts), nt = min((uint32_t)cl.numTriangles, kMaxClusterTris);
    const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;

    // :177-195 stage object-space vertices
    for (uint32_t v = lane; v < nv; v += kWave) {
      v3 p = ld3(inst.positions, cl.firstLocalVertex + v);
      s_pos[v * 3 + 0] = p.x;
      s_pos[v * 3 + 1] = p.y;
      s_pos[v * 3 + 2] = p.z;
    }
    __syncthreads();

    // :200-246 tessellation factors on world-space undisplaced vertices
    uint32_t simple = 0;
    const bool invisible = (cfg.flags & TC_FLAG_CULLING) && !(state & TC_INSTANCE_VISIBLE_BIT);
    if (invisible) {
      simple = nt;
    } else {
      float m[16];
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = inst.worldMatrix[i];
      for (uint32_t t0 = 0; t0 < nt; t0 += kWave) {
        uint32_t t = t0 + lane;
        bool valid = t < nt;
        uint32_t f[3] = {1, 1, 1};
        if (valid) {
          uint32_t i0 = lt[t * 3 + 0], i1 = lt[t * 3 + 1], i2 = lt[t * 3 + 2];
          v3 w0 = xform_point(m, mk(s_pos[i0 * 3], s_pos[i0 * 3 + 1], s_pos[i0 * 3 + 2]));
          v3 w1 = xform_point(m, mk(s_pos[i1 * 3], s_pos[i1 * 3 + 1], s_pos[i1 * 3 + 2]));
          v3 w2 = xform_point(m, mk(s_pos[i2 * 3], s_pos[i2 * 3 + 1], s_pos[i2 * 3 + 2]));
          tess_factors(w0, w1, w2, vp, f);
          s_fac[t * 3 + 0] = f[0];
          s_fac[t * 3 + 1] = f[1];
          s_fac[t * 3 + 2] = f[2];
        }
        simple += __popcll(ballot(valid && umax3(f[0], f[1], f[2]) == 1));
      }
    }
    const bool full = simple == nt;

    // tile aggregates
    uint64_t agg[5] = {0, 0, 0, 0, 0};
    uint64_t dataSize = 0;
    if (full) {
      dataSize = ptr<const uint32_t>(inst.clusterTemplateInstantiatonSizes)[ci.clusterID];
      agg[2] = 1;
      agg[3] = nv;
      agg[4] = dataSize;
    } else {
      uint32_t ns = 0, np = 0;
      for (uint32_t t0 = 0; t0 < nt; t0 += kWave) {
        uint32_t t = t0 + lane;
        bool valid = t < nt;
        uint32_t mx = valid ? umax3(s_fac[t * 3], s_fac[t * 3 + 1], s_fac[t * 3 + 2]) : 0;
        ns += __popcll(ballot(valid && mx > TC_TESSTABLE_SIZE));
        np += __popcll(ballot(valid && mx <= TC_TESSTABLE_SIZE));
      }
      agg[0] = ns;
      agg[1] = np;
    }
    uint64_t excl[5];
    lookback<5>(a.status, a.maxTiles, epoch, tile, agg, excl);

    if (full) {
      // :300-445 template instantiation of the whole cluster
      const uint32_t genOffset = (uint32_t)(B2 + excl[2]);
      const uint32_t vertexOffset = (uint32_t)(B3 + excl[3]);
      const uint64_t dataOffset = B4 + excl[4];
      const bool ok = !((uint64_t)vertexOffset + nv > cfg.maxGeneratedVertices ||
                        (uint64_t)genOffset + 1 > cfg.maxGeneratedClusters || dataOffset + dataSize > max_data_bytes(cfg));
      if (lane == 0) {
        atomicAdd(&a.readback->numFullClusters, 1u);
        if (ok) {
          // accepted clusters form a prefix of the canonical order, so the
          // template slot is the launch's first slot plus the cluster rank
          const uint32_t tempOffset = (uint32_t)(B5 + (genOffset - B2));
          atomicAdd(&b->tempInstantiateCounter, 1u);
          tc_template_instantiate_info ti;
          ti.clusterIdOffset = 0;
          ti.geometryIndexOffset = 0;
          ti.clusterTemplateAddress = ptr<const uint64_t>(inst.clusterTemplateAdresses)[ci.clusterID];
          ti.vertexBufferAddress = b->genVertices + (uint64_t)vertexOffset * 12ull;
          ti.vertexBufferStride = 12;
          ptr<tc_template_instantiate_info>(b->tempInstantiations)[tempOffset] = ti;
          ptr<uint32_t>(b->tempInstanceIDs)[tempOffset] = ci.instanceID;
          ptr<uint64_t>(b->tempClusterAddresses)[tempOffset] = b->genClusterData + dataOffset;
          atomicAdd(&ptr<tc_blas_build_info>(b->blasBuildInfos)[ci.instanceID].clusterReferencesCount, 1u);
          atomicAdd(&a.readback->numTotalTriangles, simple);
        }
      }
      if (ok) {
        // :459-488 displaced cluster vertices
        const bool disp = has_disp(a, inst.displacementIndex);
        tc_displacement_texture tex;
        if (disp) tex = a.textures[inst.displacementIndex];
        for (uint32_t v = lane; v < nv; v += kWave) {
          v3 p = mk(s_pos[v * 3], s_pos[v * 3 + 1], s_pos[v * 3 + 2]);
          if (disp) {
            uint32_t vi = cl.firstLocalVertex + v;
            float hgt = displaced_height(a, tex, inst.displacementScale, inst.displacementOffset, ld2(inst.texcoords, vi));
            p = add(p, mul(normalize(ld3(inst.normals, vi)), hgt));
          }
          st3(b->genVertices, (uint64_t)vertexOffset + v, p);
        }
      }
    } else {
      // :543-666 split / part emission in triangle order
      uint32_t runSplit = (uint32_t)(B0 + excl[0]);
      uint32_t runPart = (uint32_t)(B1 + excl[1]);
      for (uint32_t t0 = 0; t0 < nt; t0 += kWave) {
        uint32_t t = t0 + lane;
        bool valid = t < nt;
        uint32_t f[3] = {0, 0, 0};
        if (valid) f[0] = s_fac[t * 3], f[1] = s_fac[t * 3 + 1], f[2] = s_fac[t * 3 + 2];
        uint32_t mx = umax3(f[0], f[1], f[2]);
        bool needSplit = valid && mx > TC_TESSTABLE_SIZE;
        bool needPart = valid && mx <= TC_TESSTABLE_SIZE;
        uint64_t vs = ballot(needSplit), vpm = ballot(needPart);
        uint32_t offSplit = runSplit + mask_prefix(vs);
        uint32_t offPart = runPart + mask_prefix(vpm);
        runSplit += __popcll(vs);
        runPart += __popcll(vpm);
        tc_tess_triangle_info ti;
        ti.cluster = ci;
        uint32_t vtx[3] = {0u, 32768u, 32768u << 16};
        if (needSplit && offSplit < cfg.maxSplitTriangles) {
          f[0] = split_factor(f[0], cfg.maxSplitFactor);
          f[1] = split_factor(f[1], cfg.maxSplitFactor);
          f[2] = split_factor(f[2], cfg.maxSplitFactor);
          uint32_t cf = get_config(f, vtx);
          ti.subTriangle.vtxEncoded[0] = vtx[0];
          ti.subTriangle.vtxEncoded[1] = vtx[1];
          ti.subTriangle.vtxEncoded[2] = vtx[2];
          ti.subTriangle.triangleID_config = t | (cf << 16);
          ptr<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
        } else if (needPart && offPart < cfg.maxPartTriangles) {
          uint32_t cf = get_config(f, vtx);
          ti.subTriangle.vtxEncoded[0] = vtx[0];
          ti.subTriangle.vtxEncoded[1] = vtx[1];
          ti.subTriangle.vtxEncoded[2] = vtx[2];
          ti.subTriangle.triangleID_config = t | (cf << 16);
          ptr<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
        }
      }
    }

    if (tile == numTiles - 1 && lane == 0) {
      b->splitTriangleCounter = (int32_t)(uint32_t)(B0 + excl[0] + agg[0]);
      b->partTriangleCounter = (uint32_t)(B1 + excl[1] + agg[1]);
      b->genClusterCounter = (uint32_t)(B2 + excl[2] + agg[2]);
      b->genVertexCounter = (uint32_t)(B3 + excl[3] + agg[3]);
      b->genClusterDataCounter = B4 + excl[4] + agg[4];
    }
    __syncthreads();
  }
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

__device__ __forceinline__ uint32_t find_owner(const uint32_t* s_start, uint32_t t) {
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t step = 32; step > 0; step >>= 1)
    if (lo + step < (uint32_t)kWave && s_start[lo + step] <= t) lo += step;
  return lo;
}

// triangle_split.comp.glsl:598-632 + processAllSubTasks :350-472, one tile
// of 64 split items per wave; children packed densely over the wave.
// Counters: 0 split, 1 part.
__global__ __launch_bounds__(64) void k_triangle_split(KArgs a) {
  __shared__ SplitItem s_items[kWave];
  __shared__ uint32_t s_start[kWave];

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id();
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, itemBegin = h->itemBegin, itemEnd = h->itemEnd;
  const uint64_t B0 = h->base[0], B1 = h->base[1];
  const ViewParams vp = view_params(a);
  const TableView T = load_table(a.tessTable);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;

  while (true) {
    const uint32_t tile = take_ticket(h);
    if (tile >= numTiles) break;

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
    lookback<2>(a.status, a.maxTiles, epoch, tile, agg, excl);

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
        if (needSplit && offSplit < cfg.maxSplitTriangles)
          ptr<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
        else if (needPart && offPart < cfg.maxPartTriangles)
          ptr<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
      }
    }

    if (tile == numTiles - 1 && lane == 0) {
      b->splitWriteCounter = (uint32_t)(B0 + excl[0] + agg[0]);
      b->partTriangleCounter = (uint32_t)(B1 + excl[1] + agg[1]);
    }
    __syncthreads();
  }
}

// ---------------------------------------- triangle_tess_template_instantiate

struct PartBase {
  v3 bary[3];
  v3 pos[3];
  v3 nrm[3];
  v2 uv[3];
  uint32_t cfg;
  uint32_t firstVertex;
  uint32_t vertexOffset;
  uint32_t start;  // first virtual thread of this part
  int32_t dispIndex;
  float dispScale;
  float dispOffset;
  uint32_t _pad;
};

// triangle_tess_template_instantiate.comp.glsl:126-373, 64 part triangles
// per tile.  Counters: 0 genCluster, 1 genVertex, 2 genData.
template <bool PN>
__global__ __launch_bounds__(64) void k_instantiate(KArgs a) {
  __shared__ PartBase s_parts[kWave];
  __shared__ PNBasis s_pn[PN ? kWave : 1];
  __shared__ uint32_t s_start[kWave];

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id();
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numParts = h->itemEnd;
  const uint64_t G0 = h->base[0], V0 = h->base[1], D0 = h->base[2], T0 = h->base[3];
  const TableView T = load_table(a.tessTable);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;

  while (true) {
    const uint32_t tile = take_ticket(h);
    if (tile >= numTiles) break;

    const uint32_t p = tile * kWave + lane;
    const bool inRange = p < numParts;
    tc_tess_triangle_info ti;
    uint32_t cf = 0, nv = 0, nt = 0;
    uint64_t dataSize = 0;
    uint32_t firstVertex = 0;
    if (inRange) ti = ptr<const tc_tess_triangle_info>(b->partTriangles)[p];
    // a malformed record allocates like a (1,1,1) part but generates nothing
    const bool wellFormed = inRange && ti.cluster.instanceID < a.numInstances &&
                            cfg_index(ti.subTriangle.triangleID_config >> 16) < TC_TESSTABLE_NUM_LOOKUP &&
                            ti.cluster.clusterID < a.instances[ti.cluster.instanceID].numClusters;
    const bool valid = inRange;
    if (valid) {
      if (!wellFormed) ti.subTriangle.triangleID_config = 0, ti.cluster.instanceID = 0, ti.cluster.clusterID = 0;
      cf = ti.subTriangle.triangleID_config >> 16;
      tc_tess_table_entry e = cfg_entry(T, cf);
      nv = e.numVertices;
      nt = e.numTriangles;
      firstVertex = e.firstVertex;
      dataSize = T.templateSizes[cfg_index(cf)];
    }
    // :155-160 allocation in part order
    const uint64_t validMask = ballot(valid);
    const uint32_t vIncl = wave_incl_scan(nv);
    const uint64_t dIncl = wave_incl_scan64(dataSize);
    uint64_t agg[3] = {(uint64_t)__popcll(validMask), (uint64_t)shfl(vIncl, kWave - 1), shfl64(dIncl, kWave - 1)};
    uint64_t excl[3];
    lookback<3>(a.status, a.maxTiles, epoch, tile, agg, excl);

    const uint32_t genOffset = (uint32_t)(G0 + excl[0] + mask_prefix(validMask));
    const uint32_t vertexOffset = (uint32_t)(V0 + excl[1] + (vIncl - nv));
    const uint64_t dataOffset = D0 + excl[2] + (dIncl - dataSize);
    const bool ok = valid && !((uint64_t)vertexOffset + nv > cfg.maxGeneratedVertices ||
                               (uint64_t)genOffset + 1 > cfg.maxGeneratedClusters || dataOffset + dataSize > max_data_bytes(cfg));
    const uint64_t okMask = ballot(ok);
    if (ok) {
      // :172-198; accepted parts are a prefix of the canonical order
      const uint32_t tempOffset = (uint32_t)(T0 + (genOffset - G0));
      tc_template_instantiate_info info;
      info.clusterIdOffset = p | (TC_RT_CLUSTER_MODE_SINGLE_TESSELLATED << 30);
      info.geometryIndexOffset = 0;
      info.clusterTemplateAddress = T.templateAddresses[cfg_index(cf)];
      info.vertexBufferAddress = b->genVertices + (uint64_t)vertexOffset * 12ull;
      info.vertexBufferStride = 12;
      ptr<tc_template_instantiate_info>(b->tempInstantiations)[tempOffset] = info;
      ptr<uint32_t>(b->tempInstanceIDs)[tempOffset] = ti.cluster.instanceID;
      ptr<uint64_t>(b->tempClusterAddresses)[tempOffset] = b->genClusterData + dataOffset;
      atomicAdd(&ptr<tc_blas_build_info>(b->blasBuildInfos)[ti.cluster.instanceID].clusterReferencesCount, 1u);
    }
    const uint32_t triSum = wave_sum(ok ? nt : 0u);
    if (lane == 0 && okMask) {
      atomicAdd(&b->tempInstantiateCounter, (uint32_t)__popcll(okMask));
      atomicAdd(&a.readback->numTotalTriangles, triSum);
    }

    // stage the base triangle of every accepted part (:295-337)
    const uint32_t gen = ok ? nv : 0u;
    const uint32_t gIncl = wave_incl_scan(gen);
    const uint32_t totalVerts = shfl(gIncl, kWave - 1);
    s_start[lane] = ok ? gIncl - gen : 0xFFFFFFFFu;
    if (ok) {
      const tc_render_instance& inst = a.instances[ti.cluster.instanceID];
      const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ti.cluster.clusterID];
      const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
      const uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
      PartBase pb;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        uint32_t vi = lt[tri * 3 + v] + cl.firstLocalVertex;
        pb.bary[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
        pb.pos[v] = ld3(inst.positions, vi);
        pb.nrm[v] = normalize(ld3(inst.normals, vi));
        pb.uv[v] = ld2(inst.texcoords, vi);
      }
      pb.cfg = cf;
      pb.firstVertex = firstVertex;
      pb.vertexOffset = vertexOffset;
      pb.start = gIncl - gen;
      pb.dispIndex = has_disp(a, inst.displacementIndex) ? inst.displacementIndex : -1;
      pb.dispScale = inst.displacementScale;
      pb.dispOffset = inst.displacementOffset;
      s_parts[lane] = pb;
      if (PN) setup_pn(s_pn[lane], pb.pos[0], pb.pos[1], pb.pos[2], pb.nrm[0], pb.nrm[1], pb.nrm[2]);
    }
    __syncthreads();

    // :346-372 one generated vertex per virtual thread; output indices of
    // consecutive virtual threads are consecutive
    for (uint32_t t0 = 0; t0 < totalVerts; t0 += kWave) {
      const uint32_t t = t0 + lane;
      if (t >= totalVerts) break;
      const uint32_t owner = find_owner(s_start, t);
      const PartBase& pb = s_parts[owner];
      const uint32_t vert = t - pb.start;
      v3 w = cfg_vertex_bary(T, pb.cfg, pb.firstVertex, vert);
      w = interp(pb.bary[0], pb.bary[1], pb.bary[2], w);
      v3 pos = PN ? get_pn(s_pn[owner], w) : interp(pb.pos[0], pb.pos[1], pb.pos[2], w);
      if (pb.dispIndex >= 0) {
        const tc_displacement_texture tex = a.textures[pb.dispIndex];
        v3 n = interp(pb.nrm[0], pb.nrm[1], pb.nrm[2], w);
        v2 uv = interp2(pb.uv[0], pb.uv[1], pb.uv[2], w);
        float hgt = displaced_height(a, tex, pb.dispScale, pb.dispOffset, uv);
        pos = add(pos, mul(normalize(n), hgt));
      }
      st3(b->genVertices, (uint64_t)pb.vertexOffset + vert, pos);
    }

    if (tile == numTiles - 1 && lane == 0) {
      b->genClusterCounter = (uint32_t)(G0 + excl[0] + agg[0]);
      b->genVertexCounter = (uint32_t)(V0 + excl[1] + agg[1]);
      b->genClusterDataCounter = D0 + excl[2] + agg[2];
    }
    __syncthreads();
  }
}

template __global__ void k_instantiate<false>(KArgs);
template __global__ void k_instantiate<true>(KArgs);

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
  const tc_frame_constants* vw = a.view;
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
  const bool visible = inFrustum && (!clipValid || sizeOk);
  ptr<uint32_t>(a.build->instanceStates)[i] = (inFrustum ? TC_INSTANCE_FRUSTUM_BIT : 0u) | (visible ? TC_INSTANCE_VISIBLE_BIT : 0u);
  ptr<tc_blas_build_info>(a.build->blasBuildInfos)[i].clusterReferencesCount = 0;
}

// clusters_cull.comp.glsl:112-161 in ray-tracing mode (every cluster of
// every instance, instance by instance).  Workgroup i appends instance i at
// the offset of the exclusive scan held in `offsets`.
__global__ void k_clusters_cull_scan(KArgs a, uint32_t* offsets) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t run = a.build->visibleClusterCounter;
  for (uint32_t i = 0; i < a.numInstances; i++) {
    offsets[i] = run;
    run += a.instances[i].numClusters;
  }
  a.build->visibleClusterCounter = run;
}

__global__ void k_clusters_cull_fill(KArgs a, const uint32_t* offsets) {
  const uint32_t i = blockIdx.x;
  const uint32_t n = a.instances[i].numClusters, off = offsets[i];
  tc_cluster_info* out = ptr<tc_cluster_info>(a.build->visibleClusters);
  for (uint32_t c = threadIdx.x; c < n; c += blockDim.x)
    if (off + c < a.cfg.maxVisibleClusters) out[off + c] = tc_cluster_info{i, c};
}

}  // namespace tc
