// tc_instantiate.hip -- triangle_tess_template_instantiate.comp.glsl:126-373
// on gfx950.
//
// A workgroup of four waves owns a tile of 256 part triangles.  The
// allocation of every part (generated cluster, vertices, CLAS bytes) is a
// prefix sum over parts in index order: wave-level scans, a workgroup
// combine and one decoupled look-back per tile.  Each wave then stages the
// base triangles of its 64 parts in LDS and generates their displaced
// vertices, one vertex per lane; consecutive lanes write consecutive
// vertices.
#include "tc_stage.h"

namespace tc {

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
  uint32_t instanceID;  // ripple seed
  float geoSize;        // geoHi.w, the ripple scale
};

// Counters: 0 genCluster, 1 genVertex, 2 genData.
template <bool PN>
__global__ __launch_bounds__(256) void k_instantiate(KArgs a) {
  __shared__ PartBase s_parts[kInstWaves][kWave];
  __shared__ PNBasis s_pn[PN ? kInstWaves : 1][PN ? kWave : 1];
  __shared__ uint32_t s_start[kInstWaves][kWave];
  __shared__ uint64_t s_wave[kInstWaves][3];
  __shared__ uint64_t s_excl[3];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_okCount, s_okTris;  // the tile's accepted parts and triangles

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numParts = h->itemEnd;
  const uint64_t G0 = h->base[0], V0 = h->base[1], D0 = h->base[2], T0 = h->base[3];
  const TableView T = load_table(a.tessTable);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;
  PartBase* parts = s_parts[wave];
  uint32_t* starts = s_start[wave];

  while (true) {
    if (threadIdx.x == 0) {
      s_tile = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_okCount = s_okTris = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= numTiles) break;

    const uint32_t p = tile * kInstTile + wave * kWave + lane;
    const bool valid = p < numParts;
    tc_tess_triangle_info ti;
    uint32_t cf = 0, nv = 0, nt = 0, firstVertex = 0;
    uint64_t dataSize = 0;
    if (valid) {
      ti = ptr<const tc_tess_triangle_info>(b->partTriangles)[p];
      // a malformed record allocates like a (1,1,1) part of instance 0
      if (!(ti.cluster.instanceID < a.numInstances && cfg_index(ti.subTriangle.triangleID_config >> 16) < TC_TESSTABLE_NUM_LOOKUP &&
            ti.cluster.clusterID < a.instances[ti.cluster.instanceID].numClusters))
        ti.subTriangle.triangleID_config = 0, ti.cluster.instanceID = 0, ti.cluster.clusterID = 0;
      cf = ti.subTriangle.triangleID_config >> 16;
      const tc_tess_table_entry e = cfg_entry(T, cf);
      nv = e.numVertices;
      nt = e.numTriangles;
      firstVertex = e.firstVertex;
      dataSize = T.templateSizes[cfg_index(cf)];
    }
    // :155-160 allocation in part order
    const uint64_t validMask = ballot(valid);
    const uint32_t gIncl = mask_prefix(validMask) + (valid ? 1u : 0u);
    const uint32_t vIncl = wave_incl_scan(nv);
    const uint64_t dIncl = wave_incl_scan64(dataSize);
    if (lane == kWave - 1) {
      s_wave[wave][0] = gIncl;
      s_wave[wave][1] = vIncl;
      s_wave[wave][2] = dIncl;
    }
    __syncthreads();
    if (wave == 0) {
      uint64_t agg[3] = {0, 0, 0}, excl[3];
#pragma unroll
      for (uint32_t w = 0; w < kInstWaves; w++) agg[0] += s_wave[w][0], agg[1] += s_wave[w][1], agg[2] += s_wave[w][2];
      lookback<3>(a.status, a.maxTiles, epoch, tile, agg, excl);
      if (lane == 0) {
        s_excl[0] = excl[0], s_excl[1] = excl[1], s_excl[2] = excl[2];
        if (tile == numTiles - 1) {
          b->genClusterCounter = (uint32_t)(G0 + excl[0] + agg[0]);
          b->genVertexCounter = (uint32_t)(V0 + excl[1] + agg[1]);
          b->genClusterDataCounter = D0 + excl[2] + agg[2];
        }
      }
    }
    __syncthreads();
    uint64_t wb[3] = {s_excl[0], s_excl[1], s_excl[2]};
    for (uint32_t w = 0; w < wave; w++) wb[0] += s_wave[w][0], wb[1] += s_wave[w][1], wb[2] += s_wave[w][2];

    const uint32_t genOffset = (uint32_t)(G0 + wb[0] + (gIncl - (valid ? 1u : 0u)));
    const uint32_t vertexOffset = (uint32_t)(V0 + wb[1] + (vIncl - nv));
    const uint64_t dataOffset = D0 + wb[2] + (dIncl - dataSize);
    const bool ok = valid && !((uint64_t)vertexOffset + nv > cfg.maxGeneratedVertices ||
                               (uint64_t)genOffset + 1 > cfg.maxGeneratedClusters || dataOffset + dataSize > max_data_bytes(cfg));
    const uint64_t okMask = ballot(ok);
    if (ok) {
      // :172-190; accepted parts are a prefix of the canonical order
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
    }
    // :196-198 BLAS reference counts and statistics, aggregated
    add_blas_reference_counts(ptr<tc_blas_build_info>(b->blasBuildInfos), ok ? ti.cluster.instanceID : 0u, 1u, ok);
    const uint32_t triSum = wave_sum(ok ? nt : 0u);
    if (lane == 0 && okMask) {
      atomicAdd(&s_okCount, (uint32_t)__popcll(okMask));
      atomicAdd(&s_okTris, triSum);
    }

    // :295-337 stage the base triangle of every accepted part
    const uint32_t gen = ok ? nv : 0u;
    const uint32_t genIncl = wave_incl_scan(gen);
    const uint32_t totalVerts = shfl(genIncl, kWave - 1);
    starts[lane] = ok ? genIncl - gen : 0xFFFFFFFFu;
    if (ok) {
      const tc_render_instance& inst = a.instances[ti.cluster.instanceID];
      const tc_cluster cl = ptr<const tc_cluster>(inst.clusters)[ti.cluster.clusterID];
      const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
      const uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
      PartBase pb;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const uint32_t vi = lt[tri * 3 + v] + cl.firstLocalVertex;
        pb.bary[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
        pb.pos[v] = ld3(inst.positions, vi);
        pb.nrm[v] = normalize(ld3(inst.normals, vi));
        pb.uv[v] = ld2(inst.texcoords, vi);
      }
      pb.cfg = cf;
      pb.firstVertex = firstVertex;
      pb.vertexOffset = vertexOffset;
      pb.start = genIncl - gen;
      pb.dispIndex = has_disp(a, inst.displacementIndex) ? inst.displacementIndex : -1;
      pb.dispScale = inst.displacementScale;
      pb.dispOffset = inst.displacementOffset;
      pb.instanceID = ti.cluster.instanceID;
      pb.geoSize = inst.geoHi[3];
      parts[lane] = pb;
      if (PN) setup_pn(s_pn[PN ? wave : 0][lane], pb.pos[0], pb.pos[1], pb.pos[2], pb.nrm[0], pb.nrm[1], pb.nrm[2]);
    }
    __builtin_amdgcn_wave_barrier();

    // :346-372 one generated vertex per lane
    for (uint32_t t0 = 0; t0 < totalVerts; t0 += kWave) {
      const uint32_t t = t0 + lane;
      if (t >= totalVerts) break;
      const uint32_t owner = find_owner(starts, t);
      const PartBase& pb = parts[owner];
      const uint32_t vert = t - pb.start;
      v3 w = cfg_vertex_bary(T, pb.cfg, pb.firstVertex, vert);
      w = interp(pb.bary[0], pb.bary[1], pb.bary[2], w);
      v3 pos = PN ? get_pn(s_pn[PN ? wave : 0][owner], w) : interp(pb.pos[0], pb.pos[1], pb.pos[2], w);
      if (pb.dispIndex >= 0) {
        const tc_displacement_texture tex = a.textures[pb.dispIndex];
        const v3 n = interp(pb.nrm[0], pb.nrm[1], pb.nrm[2], w);
        const v2 uv = interp2(pb.uv[0], pb.uv[1], pb.uv[2], w);
        const float hgt = displaced_height(a, tex, pb.dispScale, pb.dispOffset, uv);
        if (a.tape)
          write_tape(a, (uint64_t)pb.vertexOffset + vert, pos, normalize(n), uv, pb.dispScale, pb.dispOffset,
                     (uint32_t)pb.dispIndex);
        pos = add(pos, mul(normalize(n), hgt));
      } else if (a.tape) {
        write_tape(a, (uint64_t)pb.vertexOffset + vert, pos, mk(0.f, 0.f, 0.f), v2{0.f, 0.f}, 0.f, 0.f, ~0u);
      }
      if (a.cfg.flags & TC_FLAG_ANIMATION) pos = ripple_deform(pos, pb.instanceID, pb.geoSize, a.view);  // :366-368
      st3(b->genVertices, (uint64_t)pb.vertexOffset + vert, pos);
    }
    __syncthreads();
    // one device-scope atomic per counter and tile, not per wave
    if (threadIdx.x == 0 && s_okCount) {
      atomicAdd(&b->tempInstantiateCounter, s_okCount);
      atomicAdd(&a.readback->numTotalTriangles, s_okTris);
    }
  }
}

template __global__ void k_instantiate<false>(KArgs);
template __global__ void k_instantiate<true>(KArgs);

}  // namespace tc
