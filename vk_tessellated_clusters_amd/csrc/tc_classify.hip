// tc_classify.hip -- cluster_classify.comp.glsl:150-666 on gfx950.
//
// A workgroup of four waves owns a tile of 64 visible clusters and works
// across the whole tile at each step, so memory latency overlaps over many
// clusters instead of being paid cluster by cluster:
//   0. one lane per cluster loads the cluster and instance records into LDS;
//   1. the 256 threads run flat over (cluster, triangle) pairs: factors from
//      world-space vertices (read through L1/L2), a 16-bit code per triangle
//      in LDS (split flag plus three 4-bit factors, split factors for
//      triangles going to the split list) and per-cluster counts;
//   2. wave 0 scans the 64 clusters' counts, takes the tile's offsets with
//      one decoupled look-back and makes the allocations of the full
//      clusters (template instantiations) and, with TC_FLAG_1X_TRANSIENT,
//      of the clusters whose simple triangles form a transient CLAS;
//   3. the 256 threads run flat over the displaced vertices of the accepted
//      clusters; each wave then emits the split / part records (and the
//      transient index and mapping bytes) of its share of the clusters.
// Offsets are those of a serial run in (visible cluster, triangle) order.
#include "tc_stage.h"

namespace tc {

struct ClusterSlot {
  tc_cluster_info ci;
  uint32_t nv, nt;
  uint32_t firstLocalVertex;
  uint32_t full, trans;  // template instantiation / 1X transient build
  uint32_t simple;
  uint32_t ns, np;
  uint32_t dataSize, vertexSize, partSize;
  uint32_t offSplit, offPart;
  uint32_t genOffset, vertexOffset;
  uint32_t partOffset;  // transient part slots (top of the part list)
  uint32_t yIncl;       // upper part counter after this cluster
  uint32_t ok;
  uint64_t dataOffset;
};

constexpr uint32_t kSimpleCode = 1u | (1u << 4) | (1u << 8);  // factors (1,1,1), no split

__device__ __forceinline__ uint32_t pack_code(bool split, const uint32_t f[3]) {
  return (split ? 0x8000u : 0u) | f[0] | (f[1] << 4) | (f[2] << 8);
}

struct ClusterHead {
  tc_cluster_info ci;
  uint32_t nv, nt;
  uint32_t firstLocalVertex, firstLocalTriangle;
  uint32_t invisible;
  uint32_t disp;  // displaced (texture index valid)
  // instance fields the flat phases read, so their loads do not chase the
  // instance record
  uint64_t positions, normals, texcoords, localTriangles;
  float dispScale, dispOffset;
  float geoSize;  // geoHi.w, the ripple scale
  tc_displacement_texture tex;
};

__global__ __launch_bounds__(256) void k_cluster_classify(KArgs a) {
  __shared__ uint16_t s_code[kClassifyTile][kMaxClusterTris];
  __shared__ ClusterSlot s_cl[kClassifyTile];
  __shared__ ClusterHead s_h[kClassifyTile];
  __shared__ float s_m[kClassifyTile][16];  // instance world matrices
  __shared__ uint32_t s_simple[kClassifyTile], s_ns[kClassifyTile], s_np[kClassifyTile];
  __shared__ uint32_t s_vstart[kClassifyTile], s_vcluster[kClassifyTile];
  __shared__ uint32_t s_vtotal;
  __shared__ uint32_t s_tile;

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numItems = h->itemEnd;
  const uint64_t B0 = h->base[0], B1 = h->base[1], B2 = h->base[2], B3 = h->base[3], B4 = h->base[4], B5 = h->base[5];
  const uint64_t B6 = h->base[6], B7 = h->base[7];
  const ViewParams vp = view_params(a);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;

  while (true) {
    if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= numTiles) break;
    const uint32_t tileBase = tile * kClassifyTile;
    const uint32_t cnt = min(kClassifyTile, numItems - tileBase);

    // ---- 0. cluster and instance records (:150-176)
    if (threadIdx.x < cnt) {
      const uint32_t k = threadIdx.x;
      tc_cluster_info ci = ptr<const tc_cluster_info>(b->visibleClusters)[tileBase + k];
      if (ci.instanceID >= a.numInstances) ci.instanceID = 0;
      const tc_render_instance& inst = a.instances[ci.instanceID];
      if (ci.clusterID >= inst.numClusters) ci.clusterID = 0;
      const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ci.clusterID];
      const uint32_t state = ptr<const uint32_t>(b->instanceStates)[ci.instanceID];
      ClusterHead hd;
      hd.ci = ci;
      hd.nv = min((uint32_t)cl.numVertices, kMaxClusterVerts);
      hd.nt = min((uint32_t)cl.numTriangles, kMaxClusterTris);
      hd.firstLocalVertex = cl.firstLocalVertex;
      hd.firstLocalTriangle = cl.firstLocalTriangle;
      hd.invisible = (cfg.flags & TC_FLAG_CULLING) && !(state & TC_INSTANCE_VISIBLE_BIT);
      hd.disp = has_disp(a, inst.displacementIndex);
      hd.positions = inst.positions;
      hd.normals = inst.normals;
      hd.texcoords = inst.texcoords;
      hd.localTriangles = inst.clusterLocalTriangles;
      hd.dispScale = inst.displacementScale;
      hd.dispOffset = inst.displacementOffset;
      hd.geoSize = inst.geoHi[3];
      if (hd.disp) hd.tex = a.textures[inst.displacementIndex];
      s_h[k] = hd;
#pragma unroll
      for (int i = 0; i < 16; i++) s_m[k][i] = inst.worldMatrix[i];
      s_simple[k] = hd.invisible ? hd.nt : 0u;  // :203-211 invisible: all 1x1x1
      s_ns[k] = 0;
      s_np[k] = 0;
    }
    __syncthreads();

    // ---- 1. factors and counts, flat over (cluster, triangle) (:177-258).
    // A wave covers 64 triangles of one cluster.
    for (uint32_t idx = threadIdx.x; idx < cnt * kMaxClusterTris; idx += kClassifyWaves * kWave) {
      const uint32_t k = idx / kMaxClusterTris, t = idx % kMaxClusterTris;
      const ClusterHead& hd = s_h[k];
      const bool valid = t < hd.nt;
      uint32_t f[3] = {1, 1, 1};
      if (valid && !hd.invisible) {
        const uint8_t* lt = ptr<const uint8_t>(hd.localTriangles) + hd.firstLocalTriangle + 3 * t;
        const uint32_t i0 = lt[0], i1 = lt[1], i2 = lt[2];
        const v3 w0 = xform_point(s_m[k], ld3(hd.positions, hd.firstLocalVertex + i0));
        const v3 w1 = xform_point(s_m[k], ld3(hd.positions, hd.firstLocalVertex + i1));
        const v3 w2 = xform_point(s_m[k], ld3(hd.positions, hd.firstLocalVertex + i2));
        tess_factors(w0, w1, w2, vp, f);
      }
      const uint32_t mx = umax3(f[0], f[1], f[2]);
      const bool split = valid && mx > TC_TESSTABLE_SIZE;
      if (split) {
        f[0] = split_factor(f[0], cfg.maxSplitFactor);
        f[1] = split_factor(f[1], cfg.maxSplitFactor);
        f[2] = split_factor(f[2], cfg.maxSplitFactor);
      }
      if (valid) s_code[k][t] = (uint16_t)pack_code(split, f);
      if (!hd.invisible) {
        const uint32_t simple = __popcll(ballot(valid && mx == 1));
        const uint32_t ns = __popcll(ballot(split));
        const uint32_t np = __popcll(ballot(valid && !split));
        if (lane == 0) {
          if (simple) atomicAdd(&s_simple[k], simple);
          if (ns) atomicAdd(&s_ns[k], ns);
          if (np) atomicAdd(&s_np[k], np);
        }
      }
    }
    __syncthreads();

    // ---- 1b. per-cluster classification (wave 0, one lane per cluster)
    if (wave == 0 && lane < cnt) {
      const uint32_t k = lane;
      const ClusterHead hd = s_h[k];
      const uint32_t simple = s_simple[k], ns = s_ns[k], np = s_np[k];
      ClusterSlot s;
      s.ci = hd.ci;
      s.nv = hd.nv;
      s.nt = hd.nt;
      s.firstLocalVertex = hd.firstLocalVertex;
      s.simple = simple;
      s.full = simple == hd.nt;
      // :261-276 transientSimpleThreshold = 1
      s.trans = (cfg.flags & TC_FLAG_1X_TRANSIENT) && !s.full && simple > 1;
      s.ns = s.full ? 0 : ns;
      s.np = s.full ? 0 : (s.trans ? np - simple : np);
      s.dataSize = s.full ? ptr<const uint32_t>(a.instances[hd.ci.instanceID].clusterTemplateInstantiatonSizes)[hd.ci.clusterID]
                          : (s.trans ? clas_block_size(simple, hd.nv) : 0u);
      s.vertexSize = s.full ? hd.nv : (s.trans ? hd.nv + (simple * 3 + 11) / 12 : 0u);
      s.partSize = s.trans ? (8 + simple + 24 - 1) / 24 : 0u;
      s_cl[k] = s;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");

    // ---- 2. offsets: scan over the tile's clusters + look-back (wave 0)
    if (wave == 0) {
      const uint32_t k = lane;
      const bool valid = k < cnt;
      ClusterSlot s;
      if (valid) s = s_cl[k];
      const bool alloc = valid && (s.full || s.trans);
      const uint32_t cs = valid ? s.ns : 0u, cp = valid ? s.np : 0u;
      const uint32_t cg = alloc ? 1u : 0u, cv = alloc ? s.vertexSize : 0u;
      const uint64_t cd = alloc ? (uint64_t)s.dataSize : 0ull;
      const uint32_t cy = valid ? s.partSize : 0u;
      const uint32_t cf = valid && s.full ? 1u : 0u, ct = valid && s.trans ? 1u : 0u;
      const uint32_t is = wave_incl_scan(cs), ip = wave_incl_scan(cp), ig = wave_incl_scan(cg), iv = wave_incl_scan(cv);
      const uint64_t id = wave_incl_scan64(cd);
      const uint32_t iy = wave_incl_scan(cy), iff = wave_incl_scan(cf), it = wave_incl_scan(ct);
      uint64_t agg[8] = {shfl(is, kWave - 1), shfl(ip, kWave - 1), shfl(ig, kWave - 1), shfl(iv, kWave - 1),
                         shfl64(id, kWave - 1), shfl(iy, kWave - 1), shfl(iff, kWave - 1), shfl(it, kWave - 1)};
      uint64_t excl[8];
      lookback<8>(a.status, a.maxTiles, epoch, tile, agg, excl);
      bool ok = false;
      if (valid) {
        s.offSplit = (uint32_t)(B0 + excl[0] + (is - cs));
        s.offPart = (uint32_t)(B1 + excl[1] + (ip - cp));
        s.genOffset = (uint32_t)(B2 + excl[2] + (ig - cg));
        s.vertexOffset = (uint32_t)(B3 + excl[3] + (iv - cv));
        s.dataOffset = B4 + excl[4] + (id - cd);
        const uint64_t yBefore = B6 + excl[5] + (iy - cy);
        s.yIncl = (uint32_t)(yBefore + cy);
        // build.glsl:54-66 transient part slots from the top of the list
        s.partOffset = s.trans && (uint64_t)s.offPart + yBefore + s.partSize + 1 <= cfg.maxPartTriangles
                           ? (uint32_t)(cfg.maxPartTriangles - yBefore - s.partSize)
                           : cfg.maxPartTriangles;
        ok = alloc && !((uint64_t)s.vertexOffset + s.vertexSize > cfg.maxGeneratedVertices ||
                        (uint64_t)s.genOffset + 1 > cfg.maxGeneratedClusters ||
                        s.dataOffset + s.dataSize > max_data_bytes(cfg) ||
                        (s.trans && (uint64_t)s.partOffset + s.partSize > cfg.maxPartTriangles));
        s.ok = ok;
        s_cl[k] = s;
        if (ok) {
          // accepted template / transient clusters are prefixes of their
          // kind in the canonical order, so their slots are ranks
          const tc_render_instance& inst = a.instances[s.ci.instanceID];
          const uint64_t vbuf = b->genVertices + (uint64_t)s.vertexOffset * 12ull;
          if (s.full) {
            // :372-392
            const uint32_t tempOffset = (uint32_t)(B5 + excl[6] + (iff - cf));
            tc_template_instantiate_info ti;
            ti.clusterIdOffset = 0;
            ti.geometryIndexOffset = 0;
            ti.clusterTemplateAddress = ptr<const uint64_t>(inst.clusterTemplateAdresses)[s.ci.clusterID];
            ti.vertexBufferAddress = vbuf;
            ti.vertexBufferStride = 12;
            ptr<tc_template_instantiate_info>(b->tempInstantiations)[tempOffset] = ti;
            ptr<uint32_t>(b->tempInstanceIDs)[tempOffset] = s.ci.instanceID;
            ptr<uint64_t>(b->tempClusterAddresses)[tempOffset] = b->genClusterData + s.dataOffset;
          } else {
            // :395-436 transient build record
            const uint32_t transOffset = (uint32_t)(B7 + excl[7] + (it - ct));
            tc_clas_build_info bi;
            bi.clusterID = (TC_RT_CLUSTER_MODE_1X_SUBSET_CLUSTER << 30) | s.partOffset;
            bi.clusterFlags = 0;
            bi.packed = s.simple | (s.nv << 9) | ((b->positionTruncateBitCount & 63u) << 18) | (1u << 24);
            bi.baseGeometryIndexAndFlags = TC_CLAS_GEOMETRY_OPAQUE;
            bi.indexBufferStride = 1;
            bi.vertexBufferStride = 12;
            bi.geometryIndexAndFlagsBufferStride = 0;
            bi.opacityMicromapIndexBufferStride = 0;
            bi.indexBuffer = vbuf + (uint64_t)s.nv * 12ull;
            bi.vertexBuffer = vbuf;
            bi.geometryIndexAndFlagsBuffer = 0;
            bi.opacityMicromapArray = 0;
            bi.opacityMicromapIndexBuffer = 0;
            ptr<tc_clas_build_info>(b->transBuilds)[transOffset] = bi;
            ptr<uint32_t>(b->transInstanceIDs)[transOffset] = s.ci.instanceID;
            ptr<uint64_t>(b->transClusterAddresses)[transOffset] = b->genClusterData + s.dataOffset;
            ptr<tc_tess_triangle_info>(b->partTriangles)[s.partOffset].cluster = s.ci;
          }
        }
      }
      // accepted clusters, compacted, with their first vertex in the tile's
      // flat vertex range (phase 3)
      {
        const uint64_t vm = ballot(ok);
        const uint32_t m = mask_prefix(vm);
        const uint32_t nvOk = ok ? s.nv : 0u;
        const uint32_t vincl = wave_incl_scan(nvOk);
        if (ok) {
          s_vstart[m] = vincl - nvOk;
          s_vcluster[m] = k;
        }
        const uint32_t nOk = __popcll(vm);
        if (lane >= nOk) s_vstart[lane] = 0xFFFFFFFFu;
        if (lane == kWave - 1) s_vtotal = vincl;
      }
      // :302, :442-444 statistics and BLAS reference counts, aggregated
      add_blas_reference_counts(ptr<tc_blas_build_info>(b->blasBuildInfos), valid ? s.ci.instanceID : 0u, ok ? 1u : 0u, ok);
      const uint32_t nFull = __popcll(ballot(cf != 0));
      const uint32_t nOkFull = __popcll(ballot(ok && cf)), nOkTrans = __popcll(ballot(ok && ct));
      const uint32_t okTris = wave_sum(ok ? s.simple : 0u);
      if (lane == 0) {
        if (nFull) atomicAdd(&a.readback->numFullClusters, nFull);
        if (nOkFull) atomicAdd(&b->tempInstantiateCounter, nOkFull);
        if (nOkTrans) atomicAdd(&b->transBuildCounter, nOkTrans);
        if (okTris) atomicAdd(&a.readback->numTotalTriangles, okTris);
        if (tile == numTiles - 1) {
          b->splitTriangleCounter = (int32_t)(uint32_t)(B0 + excl[0] + agg[0]);
          b->partTriangleCounter = (uint32_t)(B1 + excl[1] + agg[1]);
          b->genClusterCounter = (uint32_t)(B2 + excl[2] + agg[2]);
          b->genVertexCounter = (uint32_t)(B3 + excl[3] + agg[3]);
          b->genClusterDataCounter = B4 + excl[4] + agg[4];
          b->dualPartTriangleCounter = ((B6 + excl[5] + agg[5]) << 32) | (uint32_t)(B1 + excl[1] + agg[1]);
        }
      }
    }
    __syncthreads();

    // ---- 3a. displaced vertices of the accepted clusters, flat (:459-488)
    {
      const uint32_t total = s_vtotal;
      for (uint32_t v = threadIdx.x; v < total; v += kClassifyWaves * kWave) {
        const uint32_t m = find_owner(s_vstart, v);
        const uint32_t k = s_vcluster[m];
        const ClusterHead& hd = s_h[k];
        const uint32_t local = v - s_vstart[m];
        const uint64_t out = (uint64_t)s_cl[k].vertexOffset + local;
        const uint32_t vi = hd.firstLocalVertex + local;
        v3 p = ld3(hd.positions, vi);
        if (hd.disp) {
          const v3 n = normalize(ld3(hd.normals, vi));
          const v2 uv = ld2(hd.texcoords, vi);
          if (a.tape) write_tape(a, out, p, n, uv, hd.dispScale, hd.dispOffset, (uint32_t)a.instances[hd.ci.instanceID].displacementIndex);
          const float hgt = displaced_height(a, hd.tex, hd.dispScale, hd.dispOffset, uv);
          p = add(p, mul(n, hgt));
        } else if (a.tape) {
          write_tape(a, out, p, mk(0.f, 0.f, 0.f), v2{0.f, 0.f}, hd.dispScale, hd.dispOffset, ~0u);
        }
        if (cfg.flags & TC_FLAG_ANIMATION) p = ripple_deform(p, hd.ci.instanceID, hd.geoSize, a.view);  // :483-485
        st3(b->genVertices, out, p);
      }
    }

    // ---- 3b. per cluster: transient bytes, split / part records
    for (uint32_t k = wave; k < cnt; k += kClassifyWaves) {
      const ClusterSlot s = s_cl[k];
      if (s.full) continue;
      if (s.ok && s.trans) {
        // :497-534 triangle mapping bytes after the ClusterInfo of the part
        // slot, u8 indices after the vertices, in triangle order
        const uint8_t* lt = ptr<const uint8_t>(s_h[k].localTriangles) + s_h[k].firstLocalTriangle;
        uint8_t* map = ptr<uint8_t>(b->partTriangles) + (uint64_t)s.partOffset * 24ull + 8ull;
        uint8_t* idx = ptr<uint8_t>(b->genVertices) + ((uint64_t)s.vertexOffset + s.nv) * 12ull;
        uint32_t run = 0;
        for (uint32_t t0 = 0; t0 < s.nt; t0 += kWave) {
          const uint32_t t = t0 + lane;
          const bool simple = t < s.nt && s_code[k][t] == kSimpleCode;
          const uint64_t vm = ballot(simple);
          const uint32_t o = run + mask_prefix(vm);
          run += __popcll(vm);
          if (simple) {
            map[o] = (uint8_t)t;
            idx[o * 3 + 0] = lt[t * 3 + 0];
            idx[o * 3 + 1] = lt[t * 3 + 1];
            idx[o * 3 + 2] = lt[t * 3 + 2];
          }
        }
        // zero the rest of the index slots so the vertex stream is deterministic
        for (uint32_t j = s.simple * 3 + lane; j < (s.vertexSize - s.nv) * 12; j += kWave) idx[j] = 0;
      }
      // :543-666 split / part records in triangle order (the simple
      // triangles of a transient cluster are in its CLAS)
      uint32_t runSplit = s.offSplit, runPart = s.offPart;
      for (uint32_t t0 = 0; t0 < s.nt; t0 += kWave) {
        const uint32_t t = t0 + lane;
        const bool valid = t < s.nt && !(s.trans && s_code[k][t] == kSimpleCode);
        const uint32_t code = valid ? (uint32_t)s_code[k][t] : 0u;
        const bool split = valid && (code & 0x8000u);
        const bool part = valid && !split;
        const uint64_t vs = ballot(split), vpm = ballot(part);
        const uint32_t offSplit = runSplit + mask_prefix(vs);
        const uint32_t offPart = runPart + mask_prefix(vpm);
        runSplit += __popcll(vs);
        runPart += __popcll(vpm);
        const bool writeSplit = split && offSplit < cfg.maxSplitTriangles;
        // build.glsl:68-83 with transient builds: x + y + 2 <= MAX_PART_TRIANGLES
        const bool writePart = part && ((cfg.flags & TC_FLAG_1X_TRANSIENT)
                                            ? (uint64_t)offPart + s.yIncl + 2 <= cfg.maxPartTriangles
                                            : offPart < cfg.maxPartTriangles);
        if (writeSplit || writePart) {
          uint32_t f[3] = {code & 15u, (code >> 4) & 15u, (code >> 8) & 15u};
          uint32_t vtx[3] = {0u, 32768u, 32768u << 16};
          const uint32_t cf = get_config(f, vtx);
          tc_tess_triangle_info ti;
          ti.cluster = s.ci;
          ti.subTriangle.vtxEncoded[0] = vtx[0];
          ti.subTriangle.vtxEncoded[1] = vtx[1];
          ti.subTriangle.vtxEncoded[2] = vtx[2];
          ti.subTriangle.triangleID_config = t | (cf << 16);
          if (writeSplit)
            ptr<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
          else
            ptr<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace tc
