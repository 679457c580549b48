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
// With TC_FLAG_2X_TRANSIENT (the TWO_X instance of the kernel) the
// triangles whose factors are all <= 2 form mini batches of up to 8 base
// triangles, each one transient CLAS (:667-905): phase 1 counts the
// batches, phase 2 allocates them after the cluster's own allocation and
// phase 3 writes their vertices, indices, mappings and build records.
// Offsets are those of a serial run in (visible cluster, triangle) order;
// within a cluster the transient allocations (1X, then the 2X batches in
// triangle order) come before the split / part allocations.
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
  // 2X mini batches: count, accepted count (a prefix), and the vertex slot,
  // CLAS byte offset, upper part counter and transient rank of batch 0
  uint32_t nb, nbOk;
  uint32_t batchV0, batchY0, batchRank0;
  uint64_t batchD0;
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
  tc_displacement_texture tex;
  float geoSize;  // geoHi.w, the ripple scale
};

__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// :794-903 the mini triangles of one cluster, one lane per base triangle:
// each writes its config's vertices into its 6 slots (unused slots zero),
// its triangles' u8 indices after the batch's 48 vertex slots and its u16
// mappings (triangle | sub-triangle << 8 | (factor - 1) bits << 12) after
// the ClusterInfo of the batch's part slot; the first lane of a batch writes
// the build record and zeroes the slots and index bytes no member uses.
// Returns the wave's sum of batch triangles (lane 0).
__device__ __forceinline__ uint32_t emit_mini_batches(const KArgs& a, const TableView& T, const ClusterSlot& s,
                                                      const ClusterHead& hd, const uint16_t* code) {
  tc_scene_building* b = a.build;
  const tc_config& cfg = a.cfg;
  const uint32_t lane = lane_id();
  const uint32_t half = lane >> 5;
  const uint64_t halfMask = half ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
  uint32_t batchTotal = 0;
  uint32_t batchBase = 0;
  for (uint32_t t0 = 0; t0 < s.nt && batchBase < s.nbOk; t0 += kWave) {
    const uint32_t t = t0 + lane;
    const uint32_t c = t < s.nt ? (uint32_t)code[t] : 0x8000u;
    uint32_t f[3] = {c & 15u, (c >> 4) & 15u, (c >> 8) & 15u};
    const uint32_t mx = umax3(f[0], f[1], f[2]);
    const bool mini = !(c & 0x8000u) && mx <= 2 && !(s.trans && mx == 1);
    const uint64_t vm = ballot(mini);
    const uint32_t nLo = __popcll(vm & 0xFFFFFFFFull), nHi = __popcll(vm >> 32);
    const uint32_t bLo = (nLo + kMiniBatch - 1) / kMiniBatch, bHi = (nHi + kMiniBatch - 1) / kMiniBatch;
    const uint32_t r = __popcll(vm & halfMask & lanes_below());  // rank in the 32-triangle group
    const uint32_t j = batchBase + (half ? bLo : 0u) + r / kMiniBatch;
    const uint32_t rel = r % kMiniBatch;
    batchBase += bLo + bHi;
    uint32_t vtx[3] = {0u, 32768u, 32768u << 16};
    uint32_t cf = 0;
    tc_tess_table_entry e = {0, 0, 0, 0};
    if (mini) {
      uint32_t fr[3] = {f[0], f[1], f[2]};  // get_config rotates its factors
      cf = get_config(fr, vtx);
      e = cfg_entry(T, cf);
    }
    const uint32_t ntri = mini ? (uint32_t)e.numTriangles : 0u;
    const uint32_t incl = wave_incl_scan(ntri);
    const bool start = mini && rel == 0;
    const uint64_t starts = ballot(start);
    const uint64_t upTo = starts & (lanes_below() | (1ull << lane));
    const uint32_t startLane = upTo ? 63u - (uint32_t)__builtin_clzll(upTo) : lane;
    const uint32_t firstTris = shfl(incl - ntri, startLane);
    // the batch ends before the next start of the same group
    const uint64_t later = starts & ~(lanes_below() | (1ull << lane)) & halfMask;
    const uint32_t endLane = later ? (uint32_t)__builtin_ctzll(later) - 1u : (half ? 63u : 31u);
    const uint32_t endIncl = shfl(incl, endLane);
    const bool ok = mini && j < s.nbOk;
    if (ok) {
      const uint32_t vOff = s.batchV0 + kMiniVertexSize * j;
      const uint32_t pOff = cfg.maxPartTriangles - (s.batchY0 + kMiniPartSize * j) - kMiniPartSize;
      const uint32_t baseTris = incl - ntri - firstTris;
      // base triangle
      const uint8_t* lt = ptr<const uint8_t>(hd.localTriangles) + hd.firstLocalTriangle + 3 * t;
      v3 bp[3], bn[3];
      v2 bt[3];
#pragma unroll
      for (int v = 0; v < 3; v++) {
        const uint32_t vi = hd.firstLocalVertex + lt[v];
        bp[v] = ld3(hd.positions, vi);
        bn[v] = normalize(ld3(hd.normals, vi));
        bt[v] = ld2(hd.texcoords, vi);
      }
      const v3 b0 = decode_bary(vtx[0]), b1 = decode_bary(vtx[1]), b2 = decode_bary(vtx[2]);
      const bool pn = (cfg.flags & TC_FLAG_PN) != 0;
      PNBasis pnb;
      if (pn) setup_pn(pnb, bp[0], bp[1], bp[2], bn[0], bn[1], bn[2]);
      for (uint32_t vert = 0; vert < kMiniVerts; vert++) {
        v3 pos = mk(0.f, 0.f, 0.f);
        if (vert < e.numVertices) {
          const v3 w = interp(b0, b1, b2, cfg_vertex_bary(T, cf, e.firstVertex, vert));
          pos = pn ? get_pn(pnb, w) : interp(bp[0], bp[1], bp[2], w);
          if (hd.disp) {
            const v3 n = interp(bn[0], bn[1], bn[2], w);
            const v2 uv = interp2(bt[0], bt[1], bt[2], w);
            const float hgt = displaced_height(a, hd.tex, hd.dispScale, hd.dispOffset, uv);
            pos = add(pos, mul(normalize(n), hgt));
          }
          if (cfg.flags & TC_FLAG_ANIMATION) pos = ripple_deform(pos, hd.ci.instanceID, hd.geoSize, a.view);
        }
        st3(b->genVertices, (uint64_t)vOff + rel * kMiniVerts + vert, pos);
      }
      const uint32_t packedFactors = (f[0] - 1) | ((f[1] - 1) << 1) | ((f[2] - 1) << 2);
      uint16_t* map = reinterpret_cast<uint16_t*>(ptr<uint8_t>(b->partTriangles) + (uint64_t)pOff * 24ull + 8ull);
      uint8_t* idx = ptr<uint8_t>(b->genVertices) + ((uint64_t)vOff + kMiniBatchVerts) * 12ull;
      for (uint32_t i = 0; i < ntri; i++) {
        map[baseTris + i] = (uint16_t)(t | (i << 8) | (packedFactors << 12));
        uint32_t tv[3];
        cfg_tri_vertices(T, cf, e.firstTriangle, i, tv);
#pragma unroll
        for (int q = 0; q < 3; q++) idx[(baseTris + i) * 3 + q] = (uint8_t)(tv[q] + rel * kMiniVerts);
      }
      if (start) {
        // :750-790 the batch's build record
        const uint32_t batchTris = endIncl - (incl - ntri);
        const uint32_t members = min(kMiniBatch, (half ? nHi : nLo) - r);
        const uint64_t vbuf = b->genVertices + (uint64_t)vOff * 12ull;
        tc_clas_build_info bi;
        bi.clusterID = (TC_RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED << 30) | pOff;
        bi.clusterFlags = 0;
        bi.packed = batchTris | (kMiniBatchVerts << 9) | ((b->positionTruncateBitCount & 63u) << 18) | (1u << 24);
        bi.baseGeometryIndexAndFlags = TC_CLAS_GEOMETRY_OPAQUE;
        bi.indexBufferStride = 1;
        bi.vertexBufferStride = 12;
        bi.geometryIndexAndFlagsBufferStride = 0;
        bi.opacityMicromapIndexBufferStride = 0;
        bi.indexBuffer = vbuf + (uint64_t)kMiniBatchVerts * 12ull;
        bi.vertexBuffer = vbuf;
        bi.geometryIndexAndFlagsBuffer = 0;
        bi.opacityMicromapArray = 0;
        bi.opacityMicromapIndexBuffer = 0;
        const uint32_t transOffset = s.batchRank0 + j;
        ptr<tc_clas_build_info>(b->transBuilds)[transOffset] = bi;
        ptr<uint32_t>(b->transInstanceIDs)[transOffset] = s.ci.instanceID;
        ptr<uint64_t>(b->transClusterAddresses)[transOffset] = b->genClusterData + s.batchD0 + (uint64_t)kMiniDataSize * j;
        ptr<tc_tess_triangle_info>(b->partTriangles)[pOff].cluster = s.ci;
        for (uint32_t q = members * kMiniVerts; q < kMiniBatchVerts; q++) st3(b->genVertices, (uint64_t)vOff + q, mk(0.f, 0.f, 0.f));
        for (uint32_t q = batchTris * 3; q < (kMiniVertexSize - kMiniBatchVerts) * 12; q++) idx[q] = 0;
        batchTotal += batchTris;
      }
    }
  }
  return wave_sum(batchTotal);
}

template <bool TWO_X>
__global__ __launch_bounds__(256) void k_cluster_classify(KArgs a) {
  __shared__ uint16_t s_code[kClassifyTile][kMaxClusterTris];
  __shared__ ClusterSlot s_cl[kClassifyTile];
  __shared__ ClusterHead s_h[kClassifyTile];
  __shared__ float s_m[kClassifyTile][16];  // instance world matrices
  __shared__ uint32_t s_simple[kClassifyTile], s_ns[kClassifyTile], s_np[kClassifyTile];
  __shared__ uint32_t s_vstart[kClassifyTile], s_vcluster[kClassifyTile];
  __shared__ uint32_t s_vtotal;
  __shared__ uint32_t s_tile;
  // 2X: batches / mini triangles counting all triangles with factors <= 2,
  // and counting only those with a factor 2 (the rest of a 1X cluster)
  __shared__ uint32_t s_nb12[TWO_X ? kClassifyTile : 1], s_nb2[TWO_X ? kClassifyTile : 1];
  __shared__ uint32_t s_nm12[TWO_X ? kClassifyTile : 1], s_nm2[TWO_X ? kClassifyTile : 1];

  WsHeader* h = a.ws;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t epoch = h->epoch, numTiles = h->numTiles, numItems = h->itemEnd;
  const uint64_t B0 = h->base[0], B1 = h->base[1], B2 = h->base[2], B3 = h->base[3], B4 = h->base[4], B5 = h->base[5];
  const uint64_t B6 = h->base[6], B7 = h->base[7];
  const ViewParams vp = view_params(a);
  tc_scene_building* b = a.build;
  const tc_config cfg = a.cfg;
  const TableView T = load_table(a.tessTable);

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
      // field by field, so the record stays in registers
      const tc_displacement_texture* tp = a.textures + (hd.disp ? inst.displacementIndex : 0);
      hd.tex.texels = hd.disp ? tp->texels : 0ull;
      hd.tex.width = hd.disp ? tp->width : 0u;
      hd.tex.height = hd.disp ? tp->height : 0u;
      hd.tex.format = hd.disp ? tp->format : 0u;
      hd.tex._pad = 0;
      s_h[k] = hd;
#pragma unroll
      for (int i = 0; i < 16; i++) s_m[k][i] = inst.worldMatrix[i];
      s_simple[k] = hd.invisible ? hd.nt : 0u;  // :203-211 invisible: all 1x1x1
      s_ns[k] = 0;
      s_np[k] = 0;
      if (TWO_X) s_nb12[k] = s_nb2[k] = s_nm12[k] = s_nm2[k] = 0;
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
        if (TWO_X) {
          // a wave covers two whole 32-triangle groups of one cluster
          const uint64_t m12 = ballot(valid && mx <= 2), m2 = ballot(valid && mx == 2);
          const uint32_t l12 = __popcll(m12 & 0xFFFFFFFFull), h12 = __popcll(m12 >> 32);
          const uint32_t l2 = __popcll(m2 & 0xFFFFFFFFull), h2 = __popcll(m2 >> 32);
          if (lane == 0 && (l12 | h12)) {
            atomicAdd(&s_nb12[k], (l12 + kMiniBatch - 1) / kMiniBatch + (h12 + kMiniBatch - 1) / kMiniBatch);
            atomicAdd(&s_nm12[k], l12 + h12);
            if (l2 | h2) {
              atomicAdd(&s_nb2[k], (l2 + kMiniBatch - 1) / kMiniBatch + (h2 + kMiniBatch - 1) / kMiniBatch);
              atomicAdd(&s_nm2[k], l2 + h2);
            }
          }
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
      // :586-592 mini triangles go to 2X batches, not to the part list
      const uint32_t mini = TWO_X && !s.full ? (s.trans ? s_nm2[k] : s_nm12[k]) : 0u;
      s.nb = TWO_X && !s.full ? (s.trans ? s_nb2[k] : s_nb12[k]) : 0u;
      s.np = s.full ? 0 : (s.trans ? np - simple : np) - mini;
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
      const uint32_t nb = TWO_X && valid ? s.nb : 0u;
      const uint32_t cs = valid ? s.ns : 0u, cp = valid ? s.np : 0u;
      const uint32_t cg = (alloc ? 1u : 0u) + nb, cv = (alloc ? s.vertexSize : 0u) + nb * kMiniVertexSize;
      const uint64_t cd = (alloc ? (uint64_t)s.dataSize : 0ull) + (uint64_t)nb * kMiniDataSize;
      const uint32_t cy = (valid ? s.partSize : 0u) + nb * kMiniPartSize;
      const uint32_t cf = valid && s.full ? 1u : 0u, ct1 = valid && s.trans ? 1u : 0u, ct = ct1 + nb;
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
        if (TWO_X) {
          // :711-748 the cluster's mini batches, after its own allocation;
          // accepted batches are a prefix
          s.batchV0 = s.vertexOffset + (alloc ? s.vertexSize : 0u);
          s.batchD0 = s.dataOffset + (alloc ? s.dataSize : 0u);
          s.batchY0 = (uint32_t)(yBefore + s.partSize);
          s.batchRank0 = (uint32_t)(B7 + excl[7] + (it - ct)) + ct1;
          const uint32_t gen0 = s.genOffset + (alloc ? 1u : 0u);
          uint32_t nbOk = 0;
          for (uint32_t j = 0; j < nb; j++) {
            const uint64_t y = (uint64_t)s.batchY0 + (uint64_t)kMiniPartSize * j;
            if ((uint64_t)s.offPart + y + kMiniPartSize + 1 > cfg.maxPartTriangles ||
                (uint64_t)s.batchV0 + (uint64_t)kMiniVertexSize * (j + 1) > cfg.maxGeneratedVertices ||
                (uint64_t)gen0 + j + 1 > cfg.maxGeneratedClusters ||
                s.batchD0 + (uint64_t)kMiniDataSize * (j + 1) > max_data_bytes(cfg))
              break;
            nbOk++;
          }
          s.nbOk = nbOk;
        } else {
          s.nb = s.nbOk = 0;
        }
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
      const uint32_t nbOk = TWO_X && valid ? s.nbOk : 0u;
      const uint32_t refs = (ok ? 1u : 0u) + nbOk;
      add_blas_reference_counts(ptr<tc_blas_build_info>(b->blasBuildInfos), valid ? s.ci.instanceID : 0u, refs, refs != 0);
      const uint32_t nFull = __popcll(ballot(cf != 0));
      const uint32_t nOkFull = __popcll(ballot(ok && cf));
      const uint32_t nOkTrans = __popcll(ballot(ok && ct1)) + (TWO_X ? wave_sum(nbOk) : 0u);
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

    // ---- 3b. per cluster: transient bytes, 2X batches, split / part records
    uint32_t miniTris = 0;
    for (uint32_t k = wave; k < cnt; k += kClassifyWaves) {
      const ClusterSlot s = s_cl[k];
      if (s.full) continue;
      if (TWO_X && s.nbOk) miniTris += emit_mini_batches(a, T, s, s_h[k], s_code[k]);
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
        bool valid = t < s.nt && !(s.trans && s_code[k][t] == kSimpleCode);
        if (TWO_X && valid) {
          // :586-592 in a 2X batch
          const uint32_t c = s_code[k][t];
          const uint32_t mx = umax3(c & 15u, (c >> 4) & 15u, (c >> 8) & 15u);
          if (!(c & 0x8000u) && mx <= 2) valid = false;
        }
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
        const bool writePart = part && (transient_builds(cfg.flags)
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
    if (TWO_X && lane == 0 && miniTris) atomicAdd(&a.readback->numTotalTriangles, miniTris);
    __syncthreads();
  }
}

template __global__ void k_cluster_classify<false>(KArgs);
template __global__ void k_cluster_classify<true>(KArgs);

}  // namespace tc
