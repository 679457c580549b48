// tc_sort.hip -- stable LSD radix sort of (u64 key, u32 value) pairs and the
// deterministic BLAS insertion built on it.
//
// One pass per 8-bit digit: (1) per-tile digit histograms, (2) per-digit
// exclusive scan over tiles, (3) stable scatter.  A tile is handled by one
// wave in rounds of 64 keys; inside a round, lanes holding the same digit
// find each other with 8 ballots, so ranks follow key order and the sort is
// stable.  The element count is read on the device, and so is the tile
// size: kSortTileSmall keys (4 rounds) for sorts of up to kSortSmallMax
// keys, so that a frame's few hundred thousand keys spread over four times
// as many waves, else kSortTile keys (16 rounds), which keeps the per-digit
// scan over tiles short for large sorts.
#include "tc_sort.h"

namespace tc {

__device__ __forceinline__ uint32_t sort_count(const uint32_t* count, uint32_t capacity) {
  return min(*count, capacity);
}
__device__ __forceinline__ uint32_t sort_tile(uint32_t n) { return n <= kSortSmallMax ? kSortTileSmall : kSortTile; }

// lanes of this wave whose digit equals mine
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t peers = ballot(valid);
#pragma unroll
  for (int bit = 0; bit < 8; bit++) {
    uint64_t m = ballot(valid && ((d >> bit) & 1u));
    peers &= ((d >> bit) & 1u) ? m : ~m;
  }
  return valid ? peers : 0ull;
}

__global__ __launch_bounds__(64) void k_sort_hist(const uint64_t* keys, const uint32_t* count, uint32_t capacity,
                                                  uint32_t* hist, uint32_t maxTiles, uint32_t shift) {
  __shared__ uint32_t s_cnt[kSortRadix];
  const uint32_t n = sort_count(count, capacity), ts = sort_tile(n);
  const uint32_t tile = blockIdx.x;
  const uint32_t begin = tile * ts;
  if (begin >= n) return;
  const uint32_t lane = lane_id();
  for (uint32_t d = lane; d < kSortRadix; d += kWave) s_cnt[d] = 0;
  __syncthreads();
  for (uint32_t r = 0; r < ts; r += kWave) {
    const uint32_t i = begin + r + lane;
    if (i < n) atomicAdd(&s_cnt[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  for (uint32_t d = lane; d < kSortRadix; d += kWave) hist[(uint64_t)d * maxTiles + tile] = s_cnt[d];
}

// one workgroup per digit: exclusive scan of its per-tile counts
__global__ __launch_bounds__(256) void k_sort_scan(uint32_t* hist, uint32_t* digitTotals, const uint32_t* count,
                                                   uint32_t capacity, uint32_t maxTiles) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_carry;
  const uint32_t n = sort_count(count, capacity), ts = sort_tile(n);
  const uint32_t numTiles = (n + ts - 1) / ts;
  const uint32_t d = blockIdx.x;
  uint32_t* h = hist + (uint64_t)d * maxTiles;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < numTiles; t0 += blockDim.x) {
    const uint32_t t = t0 + threadIdx.x;
    const uint32_t v = t < numTiles ? h[t] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t waveBase = 0, total = 0;
    for (uint32_t w = 0; w < 4; w++) {
      if (w < wave) waveBase += s_wave[w];
      total += s_wave[w];
    }
    if (t < numTiles) h[t] = s_carry + waveBase + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) digitTotals[d] = s_carry;
}

__global__ __launch_bounds__(64) void k_sort_scatter(const uint64_t* keysIn, const uint32_t* valsIn, uint64_t* keysOut,
                                                     uint32_t* valsOut, const uint32_t* hist, const uint32_t* digitTotals,
                                                     const uint32_t* count, uint32_t capacity, uint32_t maxTiles,
                                                     uint32_t shift) {
  __shared__ uint32_t s_base[kSortRadix];
  const uint32_t n = sort_count(count, capacity), ts = sort_tile(n);
  const uint32_t tile = blockIdx.x;
  const uint32_t begin = tile * ts;
  if (begin >= n) return;
  const uint32_t lane = lane_id();
  // digit bases: exclusive scan of digit totals (4 digits per lane)
  {
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = digitTotals[lane * 4 + k], s += v[k];
    uint32_t run = wave_incl_scan(s) - s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint32_t d = lane * 4 + k;
      s_base[d] = run + hist[(uint64_t)d * maxTiles + tile];
      run += v[k];
    }
  }
  __syncthreads();
  for (uint32_t r = 0; r < ts; r += kWave) {
    const uint32_t i = begin + r + lane;
    const bool valid = i < n;
    uint64_t key = valid ? keysIn[i] : 0ull;
    const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t rank = mask_prefix(peers);
    uint32_t dst = 0;
    if (valid) dst = s_base[d] + rank;
    __builtin_amdgcn_wave_barrier();
    // the highest lane of each digit group advances the digit base
    if (valid && (peers >> lane) == 1ull) s_base[d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      keysOut[dst] = key;
      valsOut[dst] = valsIn[i];
    }
  }
}

void radix_sort_pairs(const SortBuffers& sb, const uint32_t* count, uint32_t bit0, uint32_t bits, hipStream_t stream,
                      uint32_t* outIndex) {
  uint32_t cur = 0;
  for (uint32_t shift = bit0; shift < bit0 + bits; shift += 8) {
    hipLaunchKernelGGL(k_sort_hist, dim3(sb.maxTiles), dim3(64), 0, stream, sb.keys[cur], count, sb.capacity, sb.hist,
                       sb.maxTiles, shift);
    hipLaunchKernelGGL(k_sort_scan, dim3(kSortRadix), dim3(256), 0, stream, sb.hist, sb.digitTotals, count, sb.capacity,
                       sb.maxTiles);
    hipLaunchKernelGGL(k_sort_scatter, dim3(sb.maxTiles), dim3(64), 0, stream, sb.keys[cur], sb.vals[cur], sb.keys[cur ^ 1],
                       sb.vals[cur ^ 1], sb.hist, sb.digitTotals, count, sb.capacity, sb.maxTiles, shift);
    cur ^= 1;
  }
  if (outIndex) *outIndex = cur;
}

void radix_sort_digits(const SortBuffers& sb, const uint32_t* count, uint32_t digitMask, hipStream_t stream,
                       uint32_t* outIndex) {
  uint32_t cur = 0;
  for (uint32_t d = 0; d < 8; d++) {
    if (!((digitMask >> d) & 1u)) continue;
    const uint32_t shift = 8 * d;
    hipLaunchKernelGGL(k_sort_hist, dim3(sb.maxTiles), dim3(64), 0, stream, sb.keys[cur], count, sb.capacity, sb.hist,
                       sb.maxTiles, shift);
    hipLaunchKernelGGL(k_sort_scan, dim3(kSortRadix), dim3(256), 0, stream, sb.hist, sb.digitTotals, count, sb.capacity,
                       sb.maxTiles);
    hipLaunchKernelGGL(k_sort_scatter, dim3(sb.maxTiles), dim3(64), 0, stream, sb.keys[cur], sb.vals[cur], sb.keys[cur ^ 1],
                       sb.vals[cur ^ 1], sb.hist, sb.digitTotals, count, sb.capacity, sb.maxTiles, shift);
    cur ^= 1;
  }
  if (outIndex) *outIndex = cur;
}

// ------------------------------------------------- blas_clusters_insert

// The list an insert pass reads: template instantiations, or (transient)
// the 1X transient builds, which follow the templates in every instance's
// reference list (blas_clusters_insert runs twice, :99-128).
struct InsertList {
  uint32_t n;
  const uint32_t* ids;
  const uint64_t* addrs;
  const uint32_t* sizes;
};

__device__ __forceinline__ InsertList insert_list(const tc_scene_building* b, int transient, uint32_t capacity) {
  InsertList l;
  l.n = min(transient ? b->transBuildCounter : b->tempInstantiateCounter, capacity);
  l.ids = ptr<const uint32_t>(transient ? b->transInstanceIDs : b->tempInstanceIDs);
  l.addrs = ptr<const uint64_t>(transient ? b->transClusterAddresses : b->tempClusterAddresses);
  l.sizes = ptr<const uint32_t>(transient ? b->transClusterSizes : b->tempClusterSizes);
  return l;
}

__global__ void k_insert_keys(KArgs a, uint64_t* keys, uint32_t* vals, uint32_t capacity, int transient) {
  const InsertList l = insert_list(a.build, transient, capacity);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < l.n; i += gridDim.x * blockDim.x) {
    keys[i] = l.ids[i];
    vals[i] = i;
  }
}

// segStart[instance] = first sorted position of the instance minus the
// references it already holds (templates, for the transient pass), so that
// k - segStart is the slot in its list (modulo 2^32).
__global__ void k_insert_seg_start(KArgs a, const uint64_t* keys, uint32_t* segStart, uint32_t capacity, int transient) {
  const InsertList l = insert_list(a.build, transient, capacity);
  const tc_blas_build_info* infos = ptr<const tc_blas_build_info>(a.build->blasBuildInfos);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < l.n; k += gridDim.x * blockDim.x)
    if (k == 0 || keys[k - 1] != keys[k]) {
      const uint32_t inst = (uint32_t)keys[k];
      segStart[inst] = k - (transient ? infos[inst].clusterReferencesCount : 0u);
    }
}

// blas_clusters_insert.comp.glsl:97-135: every CLAS lands at its rank among
// the CLAS of its instance (CLAS order), in that instance's reference list.
__global__ void k_insert_scatter(KArgs a, const uint64_t* keys, const uint32_t* vals, const uint32_t* segStart,
                                 uint32_t capacity, int transient) {
  __shared__ unsigned long long s_sizeSum;
  if (threadIdx.x == 0) s_sizeSum = 0;
  __syncthreads();
  tc_scene_building* b = a.build;
  const InsertList l = insert_list(b, transient, capacity);
  tc_blas_build_info* infos = ptr<tc_blas_build_info>(b->blasBuildInfos);
  uint64_t sizeSum = 0;
  for (uint32_t k0 = blockIdx.x * blockDim.x; k0 < l.n; k0 += gridDim.x * blockDim.x) {
    const uint32_t k = k0 + threadIdx.x;
    if (k < l.n) {
      const uint32_t inst = (uint32_t)keys[k];
      const uint32_t idx = vals[k];
      const uint32_t rank = k - segStart[inst];
      ptr<uint64_t>(infos[inst].clusterReferences)[rank] = l.addrs[idx];
      if (k + 1 == l.n || keys[k + 1] != keys[k]) infos[inst].clusterReferencesCount = rank + 1;
      sizeSum += l.sizes[idx];
    }
  }
  // one device-scope atomic per workgroup: every wave adding to the one
  // counter serialised thousands of atomics at one L2 address
  sizeSum = wave_sum64(sizeSum);
  if (lane_id() == 0 && sizeSum) atomicAdd(&s_sizeSum, (unsigned long long)sizeSum);
  __syncthreads();
  if (threadIdx.x == 0 && s_sizeSum) atomicAdd((unsigned long long*)&a.readback->numGenActualDatas, s_sizeSum);
}

}  // namespace tc
