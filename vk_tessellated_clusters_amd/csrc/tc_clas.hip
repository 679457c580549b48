// tc_clas.hip -- cluster acceleration structure (CLAS) instantiation.
//
// Replaces vkCmdBuildClusterAccelerationStructureIndirectNV in template
// instantiation mode (renderer_raytrace_clusters_tess.cpp:595-623): every
// TemplateInstantiateInfo names a template (triangle order of a pattern or
// a base cluster) and a vertex range; the result at tempClusterAddresses[i]
// is a self-contained block
//   tc_clas_header | group boxes | triangle words | vertices
// and tempClusterSizes[i] receives its size.  Transient build records
// (ClasBuildInfo, 1X clusters) produce the same block format.
#include <algorithm>

#include "tc_kernels.h"

namespace tc {

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v = fminf(v, __shfl_xor(v, d, kWave));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
  return v;
}

__host__ __device__ __forceinline__ uint32_t r16(uint32_t v) { return (v + 15u) & ~15u; }

// Group boxes and header of a block whose triangle words (s_t) and
// vertices (s_v) are in LDS and already copied to the block; returns the
// block size.  One wave.
__device__ __forceinline__ uint32_t finish_block(char* dst, const float* s_v, const uint32_t* s_t, uint32_t nt, uint32_t nv,
                                                 uint32_t clusterID, uint32_t vertexBase) {
  const uint32_t lane = lane_id();
  const uint32_t ng = (nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
  tc_aabb* groups = reinterpret_cast<tc_aabb*>(dst + sizeof(tc_clas_header));
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t g = lane; g < ng; g += kWave) {
    float glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint32_t t1 = min(nt, (g + 1) * TC_CLAS_GROUP_TRIANGLES);
    for (uint32_t t = g * TC_CLAS_GROUP_TRIANGLES; t < t1; t++) {
      const uint32_t w = s_t[t];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const uint32_t vi = (w >> (8 * k)) & 0xFF;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          glo[c] = fminf(glo[c], s_v[vi * 3 + c]);
          ghi[c] = fmaxf(ghi[c], s_v[vi * 3 + c]);
        }
      }
    }
    tc_aabb box;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      box.lo[c] = glo[c];
      box.hi[c] = ghi[c];
      lo[c] = fminf(lo[c], glo[c]);
      hi[c] = fmaxf(hi[c], ghi[c]);
    }
    gst_record(groups + g, box);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    lo[c] = wave_min(lo[c]);
    hi[c] = wave_max(hi[c]);
  }
  // zero the 16-byte padding after each part so blocks are deterministic
  {
    const uint32_t g0 = (uint32_t)sizeof(tc_clas_header) + ng * 24u, g1 = (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u);
    const uint32_t t0 = g1 + nt * 4u, t1 = g1 + r16(nt * 4u);
    const uint32_t v0 = t1 + nv * 12u, v1 = t1 + r16(nv * 12u);
    uint32_t* w = reinterpret_cast<uint32_t*>(dst);
    const uint32_t npad = (g1 - g0) / 4 + (t1 - t0) / 4 + (v1 - v0) / 4;
    for (uint32_t k = lane; k < npad; k += kWave) {
      const uint32_t a = (g1 - g0) / 4, b = a + (t1 - t0) / 4;
      const uint32_t byte = k < a ? g0 + 4 * k : (k < b ? t0 + 4 * (k - a) : v0 + 4 * (k - b));
      gst(w + byte / 4, 0u);
    }
  }
  if (lane == 0) {
    tc_clas_header h;
    h.lo[0] = lo[0], h.lo[1] = lo[1], h.lo[2] = lo[2];
    h.hi[0] = hi[0], h.hi[1] = hi[1], h.hi[2] = hi[2];
    h.clusterID = clusterID;
    h.counts = nt | (nv << 8) | (ng << 16);
    h.vertexBase = vertexBase;
    h._pad[0] = h._pad[1] = h._pad[2] = 0;
    gst_record(reinterpret_cast<tc_clas_header*>(dst), h);
  }
  return (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u) + r16(nv * 12u);
}

__device__ __forceinline__ uint32_t vertex_base(const tc_scene_building* b, uint64_t vbuf) {
  return vbuf >= b->genVertices ? (uint32_t)((vbuf - b->genVertices) / 12u) : 0xFFFFFFFFu;
}

// Templates of at most kSubTris triangles and kSubVerts vertices (the
// tessellated parts and the whole clusters of up to 128 triangles) are
// built by 16-lane subgroups, four per wave, each staging its block in a
// quarter of the wave's LDS; larger ones by the whole wave.  Both paths
// write the same bytes (box min / max are exact in any order).
constexpr uint32_t kSub = 16;
constexpr uint32_t kSubTris = 128;
constexpr uint32_t kSubVerts = 128;
constexpr uint32_t kSubsPerWave = kWave / kSub;
constexpr uint32_t kWaveFloatsPerLane = (kMaxClusterVerts * 3 + kWave - 1) / kWave;
constexpr uint32_t kWaveTrisPerLane = (kMaxClusterTris + kWave - 1) / kWave;
constexpr uint32_t kSubFloatsPerLane = (kSubVerts * 3 + kSub - 1) / kSub;
constexpr uint32_t kSubTrisPerLane = (kSubTris + kSub - 1) / kSub;
constexpr uint32_t kClasWaves = 4;

__device__ __forceinline__ float sub_min(float v) {
#pragma unroll
  for (int d = 1; d < (int)kSub; d <<= 1) v = fminf(v, __shfl_xor(v, d, kWave));
  return v;
}
__device__ __forceinline__ float sub_max(float v) {
#pragma unroll
  for (int d = 1; d < (int)kSub; d <<= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
  return v;
}

// Each wave takes 4 consecutive records per step (one per subgroup),
// grid-stride over tempInstantiateCounter.
__global__ __launch_bounds__(kClasWaves * kWave) void k_clas_instantiate(KArgs a) {
  // per wave: four subgroup quarters, or one whole-wave block (at most
  // kMaxClusterVerts vertices and kMaxClusterTris triangles)
  static_assert(kSubsPerWave * kSubVerts >= kMaxClusterVerts && kSubsPerWave * kSubTris >= kMaxClusterTris, "LDS per wave");
  __shared__ float s_vall[kClasWaves][kSubsPerWave * kSubVerts * 3];
  __shared__ uint32_t s_tall[kClasWaves][kSubsPerWave * kSubTris];
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->tempInstantiateCounter, a.cfg.maxGeneratedClusters);
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint32_t sub = lane / kSub, sl = lane % kSub;
  float* s_v = s_vall[wave];
  uint32_t* s_t = s_tall[wave];
  // a subgroup stages its small block in its quarter of the wave's LDS
  float* s_sv = s_v + sub * (kSubVerts * 3);
  uint32_t* s_st = s_t + sub * kSubTris;
  const tc_template_instantiate_info* infos = ptr<const tc_template_instantiate_info>(b->tempInstantiations);
  const uint64_t* dsts = ptr<const uint64_t>(b->tempClusterAddresses);
  uint32_t* sizes = ptr<uint32_t>(b->tempClusterSizes);
  const uint32_t perStep = kWave / kSub, stride = gridDim.x * kClasWaves * perStep;
  // two-stage prefetch along the grid-stride loop: while a record is built,
  // the next record's template (its info is already here) and the info of
  // the record after it load, so a step waits on its vertex loads alone
  // instead of the info -> template -> vertices chain
  uint32_t base = (blockIdx.x * kClasWaves + wave) * perStep;
  tc_template_instantiate_info info1{}, info2{};
  tc_clas_template T1{};
  if (base + sub < n) {
    info1 = gld_record(infos + base + sub);
    T1 = gld_record(ptr<const tc_clas_template>(info1.clusterTemplateAddress));
  }
  if (base + stride + sub < n) info2 = gld_record(infos + base + stride + sub);
  for (; base < n; base += stride) {
    const uint32_t i = base + sub;
    const bool valid = i < n;
    const tc_template_instantiate_info info = info1;
    const tc_clas_template T = T1;
    info1 = info2;
    if (i + stride < n) T1 = gld_record(ptr<const tc_clas_template>(info1.clusterTemplateAddress));
    if (i + 2 * stride < n) info2 = gld_record(infos + i + 2 * stride);
    const bool small = valid && T.numTriangles <= kSubTris && T.numVertices <= kSubVerts;
    const uint32_t nt = small ? T.numTriangles : 0u, nv = small ? T.numVertices : 0u;
    const uint32_t ng = (nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
    char* dst = small ? ptr<char>(gld(dsts + i)) : nullptr;
    const char* src = small ? ptr<const char>(info.vertexBufferAddress) : nullptr;
    if (small) {
      uint32_t* tris = reinterpret_cast<uint32_t*>(dst + sizeof(tc_clas_header) + r16(ng * 24u));
      float* verts = reinterpret_cast<float*>(dst + sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u));
      // every load of the block is issued before the first is used, so the
      // copy waits for one round trip rather than one per element; packed
      // vertices (stride 12, what the instantiation records carry) are read
      // from one base, other strides element by element
      const uint32_t nf = nv * 3;
      if (info.vertexBufferStride == 12u) {
        const float* vs = reinterpret_cast<const float*>(src) + sl;
        float vf[kSubFloatsPerLane];
#pragma unroll
        for (uint32_t j = 0; j < kSubFloatsPerLane; j++) vf[j] = sl + j * kSub < nf ? gld(vs + j * kSub) : 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < kSubFloatsPerLane; j++) {
          const uint32_t k = sl + j * kSub;
          if (k < nf) s_sv[k] = vf[j], gst(verts + k, vf[j]);
        }
      } else {
        for (uint32_t k = sl; k < nf; k += kSub) {
          const uint32_t v = k / 3, c = k - v * 3;
          const float f = gld(reinterpret_cast<const float*>(src + (uint64_t)v * info.vertexBufferStride) + c);
          s_sv[k] = f;
          gst(verts + k, f);
        }
      }
      const uint32_t* ts = ptr<const uint32_t>(T.triangles) + sl;
      uint32_t tw[kSubTrisPerLane];
#pragma unroll
      for (uint32_t j = 0; j < kSubTrisPerLane; j++) tw[j] = sl + j * kSub < nt ? gld(ts + j * kSub) : 0u;
#pragma unroll
      for (uint32_t j = 0; j < kSubTrisPerLane; j++) {
        const uint32_t t = sl + j * kSub;
        if (t < nt) s_st[t] = tw[j], gst(tris + t, tw[j]);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t g = sl; small && g < ng; g += kSub) {
      float glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
      const uint32_t t1 = min(nt, (g + 1) * TC_CLAS_GROUP_TRIANGLES);
      for (uint32_t t = g * TC_CLAS_GROUP_TRIANGLES; t < t1; t++) {
        const uint32_t w = s_st[t];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const uint32_t vi = (w >> (8 * k)) & 0xFF;
#pragma unroll
          for (int c = 0; c < 3; c++) glo[c] = fminf(glo[c], s_sv[vi * 3 + c]), ghi[c] = fmaxf(ghi[c], s_sv[vi * 3 + c]);
        }
      }
      tc_aabb box;
#pragma unroll
      for (int c = 0; c < 3; c++) box.lo[c] = glo[c], box.hi[c] = ghi[c], lo[c] = fminf(lo[c], glo[c]), hi[c] = fmaxf(hi[c], ghi[c]);
      gst_record(reinterpret_cast<tc_aabb*>(dst + sizeof(tc_clas_header)) + g, box);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) lo[c] = sub_min(lo[c]), hi[c] = sub_max(hi[c]);
    if (small) {
      // zero the 16-byte padding after each part
      uint32_t* w32 = reinterpret_cast<uint32_t*>(dst);
      const uint32_t g0 = (uint32_t)sizeof(tc_clas_header) + ng * 24u, g1 = (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u);
      const uint32_t t0 = g1 + nt * 4u, t1 = g1 + r16(nt * 4u), v0 = t1 + nv * 12u, v1 = t1 + r16(nv * 12u);
      const uint32_t na = (g1 - g0) / 4, nb = na + (t1 - t0) / 4, nc = nb + (v1 - v0) / 4;
      for (uint32_t q = sl; q < nc; q += kSub)
        gst(w32 + (q < na ? g0 + 4 * q : (q < nb ? t0 + 4 * (q - na) : v0 + 4 * (q - nb))) / 4, 0u);
      if (sl == 0) {
        tc_clas_header h;
        h.lo[0] = lo[0], h.lo[1] = lo[1], h.lo[2] = lo[2];
        h.hi[0] = hi[0], h.hi[1] = hi[1], h.hi[2] = hi[2];
        h.clusterID = T.clusterID + info.clusterIdOffset;
        h.counts = nt | (nv << 8) | (ng << 16);
        h.vertexBase = vertex_base(b, info.vertexBufferAddress);
        h._pad[0] = h._pad[1] = h._pad[2] = 0;
        gst_record(reinterpret_cast<tc_clas_header*>(dst), h);
        gst(sizes + i, (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u) + r16(nv * 12u));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // large templates: the whole wave, one at a time
    uint64_t big = ballot(valid && !small && sl == 0);
    while (big) {
      const uint32_t ij = base + (uint32_t)__builtin_ctzll(big) / kSub;
      big &= big - 1;
      const tc_template_instantiate_info bi = gld_record(infos + ij);
      const tc_clas_template BT = gld_record(ptr<const tc_clas_template>(bi.clusterTemplateAddress));
      const uint32_t bnt = BT.numTriangles, bnv = BT.numVertices, bng = BT.numGroups;
      char* bdst = ptr<char>(gld(dsts + ij));
      uint32_t* tris = reinterpret_cast<uint32_t*>(bdst + sizeof(tc_clas_header) + r16(bng * 24u));
      float* verts = reinterpret_cast<float*>(bdst + sizeof(tc_clas_header) + r16(bng * 24u) + r16(bnt * 4u));
      const char* bsrc = ptr<const char>(bi.vertexBufferAddress);
      for (uint32_t k = lane; k < bnv * 3; k += kWave) {
        const uint32_t v = k / 3, c = k - v * 3;
        const float f = gld(reinterpret_cast<const float*>(bsrc + (uint64_t)v * bi.vertexBufferStride) + c);
        s_v[k] = f;
        gst(verts + k, f);
      }
      for (uint32_t t = lane; t < bnt; t += kWave) {
        const uint32_t w = gld(ptr<const uint32_t>(BT.triangles) + t);
        s_t[t] = w;
        gst(tris + t, w);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint32_t size = finish_block(bdst, s_v, s_t, bnt, bnv, BT.clusterID + bi.clusterIdOffset,
                                         vertex_base(b, bi.vertexBufferAddress));
      if (lane == 0) gst(sizes + ij, size);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// Transient builds (stands for the NV transient CLAS build,
// renderer_raytrace_clusters_tess.cpp:625-655): triangles in record order
// with the transient triangle index as primitive, all the record's
// vertices.  One wave per CLAS, grid-stride over transBuildCounter.
__global__ __launch_bounds__(64) void k_clas_build_transient(KArgs a) {
  __shared__ float s_v[kMaxClusterVerts * 3];
  __shared__ uint32_t s_t[kMaxClusterTris];
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->transBuildCounter, a.cfg.maxGeneratedClusters);
  const uint32_t lane = lane_id();
  const tc_clas_build_info* infos = ptr<const tc_clas_build_info>(b->transBuilds);
  const uint64_t* dsts = ptr<const uint64_t>(b->transClusterAddresses);
  uint32_t* sizes = ptr<uint32_t>(b->transClusterSizes);
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const tc_clas_build_info info = gld_record(infos + i);
    const uint32_t nt = min(info.packed & 511u, kMaxClusterTris), nv = min((info.packed >> 9) & 511u, kMaxClusterVerts);
    const uint32_t ng = (nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
    char* dst = ptr<char>(gld(dsts + i));
    uint32_t* tris = reinterpret_cast<uint32_t*>(dst + sizeof(tc_clas_header) + r16(ng * 24u));
    float* verts = reinterpret_cast<float*>(dst + sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u));
    const char* src = ptr<const char>(info.vertexBuffer);
    const uint8_t* ix = ptr<const uint8_t>(info.indexBuffer);
    const uint32_t is = info.indexBufferStride;
    // the loads of a record are issued together (packed vertices from one
    // base), so the copy waits for one round trip rather than one per element
    const uint32_t nf = nv * 3;
    if (info.vertexBufferStride == 12u) {
      const float* vs = reinterpret_cast<const float*>(src) + lane;
      float vf[kWaveFloatsPerLane];
#pragma unroll
      for (uint32_t j = 0; j < kWaveFloatsPerLane; j++) vf[j] = lane + j * kWave < nf ? gld(vs + j * kWave) : 0.0f;
#pragma unroll
      for (uint32_t j = 0; j < kWaveFloatsPerLane; j++) {
        const uint32_t k = lane + j * kWave;
        if (k < nf) s_v[k] = vf[j], gst(verts + k, vf[j]);
      }
    } else {
      for (uint32_t k = lane; k < nf; k += kWave) {
        const uint32_t v = k / 3, c = k - v * 3;
        const float f = gld(reinterpret_cast<const float*>(src + (uint64_t)v * info.vertexBufferStride) + c);
        s_v[k] = f;
        gst(verts + k, f);
      }
    }
    uint32_t tb[kWaveTrisPerLane][3];
#pragma unroll
    for (uint32_t j = 0; j < kWaveTrisPerLane; j++) {
      const uint32_t t = lane + j * kWave;
      const uint8_t* q = ix + (uint64_t)(t < nt ? t : 0u) * 3u * is;
      tb[j][0] = t < nt ? (uint32_t)gld(q) : 0u;
      tb[j][1] = t < nt ? (uint32_t)gld(q + is) : 0u;
      tb[j][2] = t < nt ? (uint32_t)gld(q + 2u * is) : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kWaveTrisPerLane; j++) {
      const uint32_t t = lane + j * kWave;
      if (t < nt) {
        const uint32_t w = tb[j][0] | (tb[j][1] << 8) | (tb[j][2] << 16) | (t << 24);
        s_t[t] = w;
        gst(tris + t, w);
      }
    }
    __syncthreads();
    const uint32_t size = finish_block(dst, s_v, s_t, nt, nv, info.clusterID, vertex_base(b, info.vertexBuffer));
    if (lane == 0) gst(sizes + i, size);
    __syncthreads();
  }
}

}  // namespace tc

using namespace tc;

extern "C" tc_status tc_clas_instantiate(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args_public(b, config, a);
  if (s != TC_OK) return s;
  const uint32_t waves = std::min<uint32_t>((config.maxGeneratedClusters + 3) / 4,
                                            (uint32_t)num_cu_public() * tuning_snapshot().clasWavesPerCU);
  const uint32_t grid = (waves + kClasWaves - 1) / kClasWaves;
  hipLaunchKernelGGL(k_clas_instantiate, dim3(grid), dim3(kClasWaves * kWave), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

extern "C" tc_status tc_clas_build_transient(const tc_bindings* b, tc_config config, tc_stream stream) {
  KArgs a;
  tc_status s = make_args_public(b, config, a);
  if (s != TC_OK) return s;
  if (!transient_builds(config.flags)) return TC_OK;
  uint32_t grid = std::min<uint32_t>(config.maxGeneratedClusters, (uint32_t)num_cu_public() * tuning_snapshot().clasWavesPerCU);
  hipLaunchKernelGGL(k_clas_build_transient, dim3(grid), dim3(64), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}
