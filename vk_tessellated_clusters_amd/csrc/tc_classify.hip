// tc_classify.hip -- cluster_classify.comp.glsl:150-666 on gfx950.
//
// A workgroup of four waves owns a tile of 64 visible clusters:
//   1. each wave classifies clusters of the tile one at a time (vertices
//      staged in LDS, one triangle per lane) and leaves a 16-bit code per
//      triangle in LDS: split flag plus three 4-bit factors (split factors
//      for triangles going to the split list);
//   2. wave 0 scans the 64 clusters' counts, takes the tile's offsets with
//      one decoupled look-back and makes the allocations of the full
//      clusters (template instantiations) and, with TC_FLAG_1X_TRANSIENT,
//      of the clusters whose simple triangles form a transient CLAS;
//   3. the waves emit split / part records and displaced full-cluster
//      vertices at those offsets.
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

__global__ __launch_bounds__(256) void k_cluster_classify(KArgs a) {
  __shared__ float s_pos[kClassifyWaves][kMaxClusterVerts * 3];
  __shared__ uint16_t s_code[kClassifyTile][kMaxClusterTris];
  __shared__ ClusterSlot s_cl[kClassifyTile];
  __shared__ uint32_t s_tile;

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numItems = h->itemEnd;
  const uint64_t B0 = h->base[0], B1 = h->base[1], B2 = h->base[2], B3 = h->base[3], B4 = h->base[4], B5 = h->base[5];
  const uint64_t B6 = h->base[6], B7 = h->base[7];
  const ViewParams vp = view_params(a);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;
  float* pos = s_pos[wave];

  while (true) {
    if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= numTiles) break;
    const uint32_t tileBase = tile * kClassifyTile;
    const uint32_t cnt = min(kClassifyTile, numItems - tileBase);

    // ---- 1. factors and counts (:177-258)
    for (uint32_t k = wave; k < cnt; k += kClassifyWaves) {
      tc_cluster_info ci = ptr<const tc_cluster_info>(b->visibleClusters)[tileBase + k];
      if (ci.instanceID >= a.numInstances) ci.instanceID = 0;
      const tc_render_instance& inst = a.instances[ci.instanceID];
      if (ci.clusterID >= inst.numClusters) ci.clusterID = 0;
      const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ci.clusterID];
      const uint32_t state = ptr<const uint32_t>(b->instanceStates)[ci.instanceID];
      const uint32_t nv = min((uint32_t)cl.numVertices, kMaxClusterVerts);
      const uint32_t nt = min((uint32_t)cl.numTriangles, kMaxClusterTris);
      const bool invisible = (cfg.flags & TC_FLAG_CULLING) && !(state & TC_INSTANCE_VISIBLE_BIT);
      uint32_t simple = 0, ns = 0, np = 0;
      if (invisible) {
        simple = nt;  // :203-211
      } else {
        const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
        for (uint32_t v = lane; v < nv; v += kWave) {
          v3 p = ld3(inst.positions, cl.firstLocalVertex + v);
          pos[v * 3 + 0] = p.x;
          pos[v * 3 + 1] = p.y;
          pos[v * 3 + 2] = p.z;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        float m[16];
#pragma unroll
        for (int i = 0; i < 16; i++) m[i] = inst.worldMatrix[i];
        for (uint32_t t0 = 0; t0 < nt; t0 += kWave) {
          const uint32_t t = t0 + lane;
          const bool valid = t < nt;
          uint32_t f[3] = {1, 1, 1};
          if (valid) {
            const uint32_t i0 = lt[t * 3 + 0], i1 = lt[t * 3 + 1], i2 = lt[t * 3 + 2];
            v3 w0 = xform_point(m, mk(pos[i0 * 3], pos[i0 * 3 + 1], pos[i0 * 3 + 2]));
            v3 w1 = xform_point(m, mk(pos[i1 * 3], pos[i1 * 3 + 1], pos[i1 * 3 + 2]));
            v3 w2 = xform_point(m, mk(pos[i2 * 3], pos[i2 * 3 + 1], pos[i2 * 3 + 2]));
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
          simple += __popcll(ballot(valid && mx == 1));
          ns += __popcll(ballot(split));
          np += __popcll(ballot(valid && !split));
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (invisible)
        for (uint32_t t = lane; t < nt; t += kWave) s_code[k][t] = (uint16_t)kSimpleCode;
      if (lane == 0) {
        ClusterSlot s;
        s.ci = ci;
        s.nv = nv;
        s.nt = nt;
        s.firstLocalVertex = cl.firstLocalVertex;
        s.simple = simple;
        s.full = simple == nt;
        // :261-276 transientSimpleThreshold = 1
        s.trans = (cfg.flags & TC_FLAG_1X_TRANSIENT) && !s.full && simple > 1;
        s.ns = s.full ? 0 : ns;
        s.np = s.full ? 0 : (s.trans ? np - simple : np);
        s.dataSize = s.full ? ptr<const uint32_t>(inst.clusterTemplateInstantiatonSizes)[ci.clusterID]
                            : (s.trans ? clas_block_size(simple, nv) : 0u);
        s.vertexSize = s.full ? nv : (s.trans ? nv + (simple * 3 + 11) / 12 : 0u);
        s.partSize = s.trans ? (8 + simple + 24 - 1) / 24 : 0u;
        s_cl[k] = s;
      }
    }
    __syncthreads();

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

    // ---- 3. emission
    for (uint32_t k = wave; k < cnt; k += kClassifyWaves) {
      const ClusterSlot s = s_cl[k];
      if (s.ok) {
        // :459-488 displaced cluster vertices
        const tc_render_instance& inst = a.instances[s.ci.instanceID];
        const bool disp = has_disp(a, inst.displacementIndex);
        tc_displacement_texture tex;
        if (disp) tex = a.textures[inst.displacementIndex];
        for (uint32_t v = lane; v < s.nv; v += kWave) {
          const uint32_t vi = s.firstLocalVertex + v;
          v3 p = ld3(inst.positions, vi);
          if (a.tape) write_tape(a, (uint64_t)s.vertexOffset + v, p, disp ? normalize(ld3(inst.normals, vi)) : mk(0.f, 0.f, 0.f),
                                 disp ? ld2(inst.texcoords, vi) : v2{0.f, 0.f}, inst.displacementScale,
                                 inst.displacementOffset, disp ? (uint32_t)inst.displacementIndex : ~0u);
          if (disp) {
            const float hgt = displaced_height(a, tex, inst.displacementScale, inst.displacementOffset, ld2(inst.texcoords, vi));
            p = add(p, mul(normalize(ld3(inst.normals, vi)), hgt));
          }
          st3(b->genVertices, (uint64_t)s.vertexOffset + v, p);
        }
        if (s.trans) {
          // :497-534 triangle mapping bytes after the ClusterInfo of the part
          // slot, u8 indices after the vertices, in triangle order
          const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[s.ci.clusterID];
          const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
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
      }
      if (!s.full) {
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
    }
    __syncthreads();
  }
}

}  // namespace tc
