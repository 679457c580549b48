// tc_instantiate.hip -- triangle_tess_template_instantiate.comp.glsl:126-373
// on gfx950, as two launches.
//
// k_instantiate allocates: a workgroup of four waves owns a tile of 256 part
// triangles, and the allocation of every part (generated cluster, vertices,
// CLAS bytes) is a prefix sum over parts in index order: wave-level scans, a
// workgroup combine and one decoupled look-back per tile.  It writes the
// instantiation records and the BLAS reference counts.  Its tiles carry no
// vertex work, so a tile's dependent chain (part record, table entry,
// look-back, records) is short and many tiles are in flight.
//
// k_instantiate_vertices generates the displaced vertices of the accepted
// parts.  Acceptance is monotone in part order, so the accepted parts are
// the prefix [0, accepted) and part p's record is instantiation T0 + p,
// whose vertex buffer address names its vertices.  A wave stages the base
// triangles of 64 parts in LDS and generates their vertices one per lane;
// consecutive lanes write consecutive vertices.  No wave waits on another,
// so the launch runs at full occupancy and hides the gathers' latency.
#include "tc_stage.h"

namespace tc {

// Parts per lane of the allocation kernel: a wave covers kPartsPerLane x 64
// consecutive parts, a tile kInstWaves such runs (kInstTile parts).
constexpr uint32_t kPartsPerLane = kInstTile / (kInstWaves * kWave);
// Parts a wave stages at a time in the vertex kernel.
constexpr uint32_t kVertChunk = 32;

// Base triangle of one part, staged in LDS for its vertices.
struct PartBase {
  v3 pos[3];
  v3 nrm[3];
  v2 uv[3];
  uint32_t vtx[3];  // encoded barycentrics of the sub-triangle corners
  uint32_t cfg;
  uint32_t firstVertex;
  uint32_t vertexOffset;
  uint32_t start;  // first virtual thread of this part
  int32_t dispIndex;
  float dispScale;
  float dispOffset;
  uint32_t instanceID;  // ripple seed
  float geoSize;        // geoHi.w, the ripple scale
  tc_displacement_texture tex;  // the height map (dispIndex >= 0)
};

// part record as the stages read it: a malformed one allocates like the
// (1,1,1) part of instance 0
__device__ __forceinline__ tc_tess_triangle_info load_part(const KArgs& a, uint32_t p) {
  tc_tess_triangle_info ti = gld_record(ptr<const tc_tess_triangle_info>(a.build->partTriangles) + p);
  if (!(ti.cluster.instanceID < a.numInstances && cfg_index(ti.subTriangle.triangleID_config >> 16) < TC_TESSTABLE_NUM_LOOKUP &&
        ti.cluster.clusterID < a.instances[ti.cluster.instanceID].numClusters))
    ti.subTriangle.triangleID_config = 0, ti.cluster.instanceID = 0, ti.cluster.clusterID = 0;
  return ti;
}

// Counters: 0 genCluster, 1 genVertex, 2 genData, 3 triangles.  Accepted
// parts are a prefix [0, A) of the part list (every allocation ends at or
// after the one before it), so exactly one lane sees the end of that
// prefix: the owner of the first rejected part, or of the last part when
// none is rejected.  It alone writes tempInstantiateCounter = T0 + A and
// adds the prefix's triangles to the readback.
__global__ __launch_bounds__(256) void k_instantiate(KArgs a) {
  __shared__ uint64_t s_wave[kInstWaves][4];
  __shared__ uint64_t s_excl[4];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_inst[kInstWaves], s_count[kInstWaves];  // a wave's single instance and its accepted parts

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numParts = h->itemEnd;
  const uint64_t G0 = h->base[0], V0 = h->base[1], D0 = h->base[2], T0 = h->base[3];
  const TableView T = load_table(a.tessTable);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;
  const uint64_t maxData = max_data_bytes(cfg);
  tc_blas_build_info* infos = ptr<tc_blas_build_info>(b->blasBuildInfos);

  while (true) {
    if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= numTiles) break;

    // this lane's parts: base + k * 64, k < kPartsPerLane (coalesced rounds)
    const uint32_t base = tile * kInstTile + wave * (kPartsPerLane * kWave) + lane;
    tc_tess_triangle_info ti[kPartsPerLane];
    uint32_t nv[kPartsPerLane], nt[kPartsPerLane];
    uint64_t ds[kPartsPerLane];
#pragma unroll
    for (uint32_t k = 0; k < kPartsPerLane; k++) {
      const uint32_t p = base + k * kWave;
      nv[k] = nt[k] = 0;
      ds[k] = 0;
      if (p < numParts) {
        ti[k] = load_part(a, p);
        const uint32_t cf = ti[k].subTriangle.triangleID_config >> 16;
        const tc_tess_table_entry e = cfg_entry(T, cf);
        nv[k] = e.numVertices;
        nt[k] = e.numTriangles;
        ds[k] = gld(T.templateSizes + cfg_index(cf));
      }
    }
    // :155-160 allocation in part order: scans over the wave's rounds
    uint64_t carry[4] = {0, 0, 0, 0};
    uint64_t excl[kPartsPerLane][4];
#pragma unroll
    for (uint32_t k = 0; k < kPartsPerLane; k++) {
      const uint32_t g = base + k * kWave < numParts ? 1u : 0u;
      const uint32_t gi = wave_incl_scan(g), vi = wave_incl_scan(nv[k]), ti3 = wave_incl_scan(nt[k]);
      const uint64_t di = wave_incl_scan64(ds[k]);
      excl[k][0] = carry[0] + gi - g;
      excl[k][1] = carry[1] + vi - nv[k];
      excl[k][2] = carry[2] + di - ds[k];
      excl[k][3] = carry[3] + ti3 - nt[k];
      carry[0] += shfl(gi, kWave - 1);
      carry[1] += shfl(vi, kWave - 1);
      carry[2] += shfl64(di, kWave - 1);
      carry[3] += shfl(ti3, kWave - 1);
    }
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 4; c++) s_wave[wave][c] = carry[c];
    }
    __syncthreads();
    if (wave == 0) {
      uint64_t agg[4] = {0, 0, 0, 0}, ex[4];
#pragma unroll
      for (uint32_t w = 0; w < kInstWaves; w++)
#pragma unroll
        for (int c = 0; c < 4; c++) agg[c] += s_wave[w][c];
      lookback<4>(a.status, a.maxTiles, epoch, tile, agg, ex);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 4; c++) s_excl[c] = ex[c];
        if (tile == numTiles - 1) {
          b->genClusterCounter = (uint32_t)(G0 + ex[0] + agg[0]);
          b->genVertexCounter = (uint32_t)(V0 + ex[1] + agg[1]);
          b->genClusterDataCounter = D0 + ex[2] + agg[2];
        }
      }
    }
    __syncthreads();
    uint64_t wb[4] = {s_excl[0], s_excl[1], s_excl[2], s_excl[3]};
    for (uint32_t w = 0; w < wave; w++)
#pragma unroll
      for (int c = 0; c < 4; c++) wb[c] += s_wave[w][c];

    uint32_t myInst = 0xFFFFFFFFu, myCount = 0;
    bool single = true;
#pragma unroll
    for (uint32_t k = 0; k < kPartsPerLane; k++) {
      const uint32_t p = base + k * kWave;
      const bool valid = p < numParts;
      const uint64_t genOffset = G0 + wb[0] + excl[k][0];
      const uint64_t vertexOffset = V0 + wb[1] + excl[k][1];
      const uint64_t dataOffset = D0 + wb[2] + excl[k][2];
      const bool ok = valid && vertexOffset + nv[k] <= cfg.maxGeneratedVertices && genOffset + 1 <= cfg.maxGeneratedClusters &&
                      dataOffset + ds[k] <= maxData;
      // the end of the accepted prefix: the first rejected part (its
      // predecessor's allocation, ending where this one starts, fits), or
      // the last part
      const bool predOk = p == 0 || (vertexOffset <= cfg.maxGeneratedVertices && genOffset <= cfg.maxGeneratedClusters &&
                                     dataOffset <= maxData);
      const bool endHere = valid && ((!ok && predOk) || (ok && p + 1 == numParts));
      if (endHere) {
        const uint32_t A = ok ? p + 1 : p;
        const uint64_t tris = wb[3] + excl[k][3] + (ok ? nt[k] : 0u);
        b->tempInstantiateCounter = (uint32_t)(T0 + A);
        atomicAdd(&a.readback->numTotalTriangles, (uint32_t)tris);
      }
      if (ok) {
        // :172-190; part p's record is instantiation T0 + p
        const uint32_t tempOffset = (uint32_t)(T0 + (genOffset - G0));
        const uint32_t cf = ti[k].subTriangle.triangleID_config >> 16;
        tc_template_instantiate_info info;
        info.clusterIdOffset = p | (TC_RT_CLUSTER_MODE_SINGLE_TESSELLATED << 30);
        info.geometryIndexOffset = 0;
        info.clusterTemplateAddress = gld(T.templateAddresses + cfg_index(cf));
        info.vertexBufferAddress = b->genVertices + vertexOffset * 12ull;
        info.vertexBufferStride = 12;
        gst_record(ptr<tc_template_instantiate_info>(b->tempInstantiations) + tempOffset, info);
        gst(ptr<uint32_t>(b->tempInstanceIDs) + tempOffset, ti[k].cluster.instanceID);
        gst(ptr<uint64_t>(b->tempClusterAddresses) + tempOffset, b->genClusterData + dataOffset);
      }
      // :196-198 BLAS reference counts: a wave whose accepted parts all
      // belong to one instance reports one count; otherwise per round
      const uint64_t okMask = ballot(ok);
      const uint32_t first = okMask ? shfl(ti[k].cluster.instanceID, (uint32_t)__builtin_ctzll(okMask)) : 0u;
      const bool same = !okMask || ballot(ok && ti[k].cluster.instanceID != first) == 0;
      if (same && single && okMask) {
        if (myInst == 0xFFFFFFFFu || myInst == first) {
          myInst = first;
          myCount += (uint32_t)__popcll(okMask);
        } else {
          single = false;
        }
      } else if (okMask) {
        single = false;
      }
      if (!single) add_blas_reference_counts(infos, ok ? ti[k].cluster.instanceID : 0u, 1u, ok);
    }
    // a wave that left the single-instance form part way flushes the count
    // it had gathered
    if (!single && myCount && lane == 0) atomicAdd(&infos[myInst].clusterReferencesCount, myCount);
    if (lane == 0) s_inst[wave] = single ? myInst : 0xFFFFFFFFu, s_count[wave] = single ? myCount : 0u;
    __syncthreads();
    // one atomic per distinct instance of the single-instance waves
    if (threadIdx.x == 0) {
      for (uint32_t w = 0; w < kInstWaves; w++) {
        if (s_inst[w] == 0xFFFFFFFFu || !s_count[w]) continue;
        uint32_t c = s_count[w];
        for (uint32_t u = w + 1; u < kInstWaves; u++)
          if (s_inst[u] == s_inst[w]) c += s_count[u], s_count[u] = 0;
        atomicAdd(&infos[s_inst[w]].clusterReferencesCount, c);
      }
    }
  }
}

// :295-372 for the accepted parts [0, accepted): accepted is what
// k_instantiate wrote to tempInstantiateCounter above its start value.
// Waves take kVertChunk parts at a time, grid-stride.
template <bool PN>
__global__ __launch_bounds__(256) void k_instantiate_vertices(KArgs a) {
  __shared__ PartBase s_parts[kInstWaves][kVertChunk];
  __shared__ PNBasis s_pn[PN ? kInstWaves : 1][PN ? kVertChunk : 1];
  __shared__ uint32_t s_start[kInstWaves][kVertChunk];

  const WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t T0 = (uint32_t)h->base[3];
  const TableView T = load_table(a.tessTable);
  const tc_scene_building* b = a.build;
  const uint32_t accepted = b->tempInstantiateCounter - T0;
  const tc_template_instantiate_info* infos = ptr<const tc_template_instantiate_info>(b->tempInstantiations) + T0;
  // loop invariants the vertex loop would otherwise reload after each store
  const uint64_t genVertices = b->genVertices;
  const DispGlobal dg = disp_global(a);
  PartBase* parts = s_parts[wave];
  uint32_t* starts = s_start[wave];
  const uint32_t numChunks = (accepted + kVertChunk - 1) / kVertChunk;

  for (uint32_t c = blockIdx.x * kInstWaves + wave; c < numChunks; c += gridDim.x * kInstWaves) {
    const uint32_t p = c * kVertChunk + lane;
    const bool ok = lane < kVertChunk && p < accepted;
    uint32_t nv = 0;
    PartBase pb;
    if (ok) {
      const tc_tess_triangle_info ti = load_part(a, p);
      const uint32_t cf = ti.subTriangle.triangleID_config >> 16;
      const tc_tess_table_entry e = cfg_entry(T, cf);
      nv = e.numVertices;
      const tc_render_instance& inst = a.instances[ti.cluster.instanceID];
      const tc_cluster cl = gld_record(ptr<const tc_cluster>(inst.clusters) + ti.cluster.clusterID);
      const uint8_t* lt = ptr<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
      const uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const uint32_t vi = gld(lt + tri * 3 + v) + cl.firstLocalVertex;
        pb.vtx[v] = ti.subTriangle.vtxEncoded[v];
        pb.pos[v] = ld3(inst.positions, vi);
        pb.nrm[v] = normalize(ld3(inst.normals, vi));
        pb.uv[v] = ld2(inst.texcoords, vi);
      }
      pb.cfg = cf;
      pb.firstVertex = e.firstVertex;
      pb.vertexOffset = (uint32_t)((gld(&infos[p].vertexBufferAddress) - genVertices) / 12ull);
      pb.dispIndex = has_disp(a, inst.displacementIndex) ? inst.displacementIndex : -1;
      if (pb.dispIndex >= 0) pb.tex = a.textures[pb.dispIndex];
      pb.dispScale = inst.displacementScale;
      pb.dispOffset = inst.displacementOffset;
      pb.instanceID = ti.cluster.instanceID;
      pb.geoSize = inst.geoHi[3];
    }
    const uint32_t genIncl = wave_incl_scan(nv);
    const uint32_t totalVerts = shfl(genIncl, kWave - 1);
    if (lane < kVertChunk) starts[lane] = ok ? genIncl - nv : 0xFFFFFFFFu;
    if (ok) {
      pb.start = genIncl - nv;
      parts[lane] = pb;
      if (PN) setup_pn(s_pn[PN ? wave : 0][lane], pb.pos[0], pb.pos[1], pb.pos[2], pb.nrm[0], pb.nrm[1], pb.nrm[2]);
    }
    __builtin_amdgcn_wave_barrier();

    // :346-372 one generated vertex per lane
    for (uint32_t t0 = 0; t0 < totalVerts; t0 += kWave) {
      const uint32_t t = t0 + lane;
      if (t >= totalVerts) break;
      const uint32_t owner = find_owner_n<kVertChunk>(starts, t);
      const PartBase& q = parts[owner];
      const uint32_t vert = t - q.start;
      v3 w = cfg_vertex_bary(T, q.cfg, q.firstVertex, vert);
      w = interp(decode_bary(q.vtx[0]), decode_bary(q.vtx[1]), decode_bary(q.vtx[2]), w);
      v3 pos = PN ? get_pn(s_pn[PN ? wave : 0][owner], w) : interp(q.pos[0], q.pos[1], q.pos[2], w);
      if (q.dispIndex >= 0) {
        const v3 n = interp(q.nrm[0], q.nrm[1], q.nrm[2], w);
        const v2 uv = interp2(q.uv[0], q.uv[1], q.uv[2], w);
        const float hgt = displaced_height(dg, q.tex, q.dispScale, q.dispOffset, uv);
        if (a.tape)
          write_tape(a, (uint64_t)q.vertexOffset + vert, pos, normalize(n), uv, q.dispScale, q.dispOffset,
                     (uint32_t)q.dispIndex);
        pos = add(pos, mul(normalize(n), hgt));
      } else if (a.tape) {
        write_tape(a, (uint64_t)q.vertexOffset + vert, pos, mk(0.f, 0.f, 0.f), v2{0.f, 0.f}, 0.f, 0.f, ~0u);
      }
      if (a.cfg.flags & TC_FLAG_ANIMATION) pos = ripple_deform(pos, q.instanceID, q.geoSize, a.view);  // :366-368
      st3(genVertices, (uint64_t)q.vertexOffset + vert, pos);
    }
    // the next chunk reuses this wave's LDS
    __builtin_amdgcn_wave_barrier();
  }
}

template __global__ void k_instantiate_vertices<false>(KArgs);
template __global__ void k_instantiate_vertices<true>(KArgs);

}  // namespace tc
