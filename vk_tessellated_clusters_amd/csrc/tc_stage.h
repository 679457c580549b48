// tc_stage.h -- helpers shared by the stage kernels.
#pragma once

#include "tc_kernels.h"

namespace tc {

enum Stage : uint32_t { kStageClassify = 0, kStageSplit = 1, kStageInstantiate = 2 };

constexpr uint32_t kClassifyWaves = 4;   // waves per classify workgroup
constexpr uint32_t kClassifyTile = 64;   // visible clusters per classify tile
constexpr uint32_t kVertCache = 128;     // vertices per wave a classify cluster caches
constexpr uint32_t kInstWaves = 4;       // waves per instantiate workgroup
constexpr uint32_t kInstTile = kInstWaves * 4 * 64;  // part triangles per instantiate tile (4 per lane)

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

// The frame's displacement scale and offset, read once per kernel: vertex
// loops store through pointers the compiler cannot tell apart from the
// frame constants, so a read inside the loop is repeated every iteration.
struct DispGlobal {
  float scale, offset;
};
__device__ __forceinline__ DispGlobal disp_global(const KArgs& a) {
  return DispGlobal{a.view->displacementScale, a.view->displacementOffset};
}

// cluster_classify.comp.glsl:477-479
__device__ __forceinline__ float displaced_height(const DispGlobal& g, const tc_displacement_texture& tex, float scale,
                                                  float offset, v2 uv) {
  float h = sample_height(tex, uv.x, uv.y);
  return (((h * scale) * g.scale) + offset) + g.offset;
}
__device__ __forceinline__ float displaced_height(const KArgs& a, const tc_displacement_texture& tex, float scale,
                                                  float offset, v2 uv) {
  return displaced_height(disp_global(a), tex, scale, offset, uv);
}

// Records how generated vertex `index` was displaced (tc_vertex_tape).
__device__ __forceinline__ void write_tape(const KArgs& a, uint64_t index, v3 base, v3 normal, v2 uv, float scale,
                                           float offset, uint32_t texture) {
  float4* t = reinterpret_cast<float4*>(a.tape + index);
  t[0] = make_float4(base.x, base.y, base.z, normal.x);
  t[1] = make_float4(normal.y, normal.z, uv.x, uv.y);
  t[2] = make_float4(scale * a.view->displacementScale, offset + a.view->displacementOffset, __uint_as_float(texture), 0.0f);
}

// Bytes of our CLAS block for nt triangles and nv vertices (tc_clas_size).
__host__ __device__ constexpr uint32_t clas_block_size(uint32_t nt, uint32_t nv) {
  return (uint32_t)sizeof(tc_clas_header) + ((((nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES) * 24u + 15u) & ~15u) +
         ((nt * 4u + 15u) & ~15u) + ((nv * 12u + 15u) & ~15u);
}

// CLAS bytes reserved per 2X mini batch: the worst case of the batch
// (:732-733 reserves basicClusterSizes[miniBatchTriangles])
constexpr uint32_t kMiniDataSize = clas_block_size(kMiniBatchTris, kMiniBatchVerts);

__device__ __forceinline__ uint64_t max_data_bytes(const tc_config& c) {
  return (uint64_t)c.maxGeneratedClusterMegs * 1024ull * 1024ull;
}

// Last index j of the non-decreasing s_start[0..N) with s_start[j] <= t (0
// when there is none), as a count in two rounds of independent LDS reads:
// the group of eight holding j from every eighth entry, then j within it.
template <uint32_t N>
__device__ __forceinline__ uint32_t find_owner_n(const uint32_t* s_start, uint32_t t) {
  static_assert(N % 8 == 0 && N >= 16, "groups of eight");
  uint32_t g = 0;
#pragma unroll
  for (uint32_t j = 1; j < N / 8; j++) g += s_start[8 * j] <= t ? 1u : 0u;
  const uint32_t* s = s_start + 8 * g;
  uint32_t k = 0;
#pragma unroll
  for (uint32_t i = 1; i < 8; i++) k += s[i] <= t ? 1u : 0u;
  return 8 * g + k;
}
__device__ __forceinline__ uint32_t find_owner(const uint32_t* s_start, uint32_t t) { return find_owner_n<64>(s_start, t); }

// Adds `count` of every active lane to infos[instance].clusterReferencesCount
// with one atomic per distinct instance in the wave (neighbouring lanes
// usually share the instance).
__device__ __forceinline__ void add_blas_reference_counts(tc_blas_build_info* infos, uint32_t key, uint32_t count,
                                                          bool active) {
  uint64_t pending = ballot(active && count != 0);
  while (pending) {
    const uint32_t leader = (uint32_t)__builtin_ctzll(pending);
    const uint32_t lkey = shfl(key, leader);
    const bool same = ((pending >> lane_id()) & 1ull) && key == lkey;
    const uint64_t sameMask = ballot(same);
    const uint32_t sum = wave_sum(same ? count : 0u);
    if (lane_id() == leader) atomicAdd(&infos[lkey].clusterReferencesCount, sum);
    pending &= ~sameMask;
  }
}

}  // namespace tc
