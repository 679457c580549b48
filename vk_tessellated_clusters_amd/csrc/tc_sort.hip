// tc_sort.hip -- stable LSD radix sort of (u64 key, u32 value) pairs and the
// deterministic BLAS insertion built on it.
//
// A sort of P 8-bit digits is P + 1 launches:
//   upsweep  every pass's digit counts in one read of the keys; the last
//            workgroup to finish turns them into each pass's exclusive
//            digit bases, resets the pass tickets and clears the counts,
//            so the next sort finds the state as it found it
//   pass     (one per digit) a workgroup takes the next tile ticket, ranks
//            its kSortTileKeys keys by digit in LDS (lanes with the same
//            digit find each other with 8 ballots, so ranks follow key
//            order), publishes the tile's 256 digit counts and walks back
//            over earlier tiles' counts (decoupled look-back, one digit per
//            thread) for the keys of each digit before the tile, then
//            scatters.  Stable, so LSD passes compose.
// Status words carry the pass's epoch (a sort takes P consecutive epochs),
// so words of earlier passes read as unpublished and are never cleared,
// except when the epoch counter wraps.
#include "tc_sort.h"

namespace tc {

constexpr uint32_t kUpsweepGroups = 64;
constexpr uint32_t kUpsweepThreads = 1024;
constexpr uint32_t kSortWaves = kSortThreads / kWave;
constexpr uint32_t kSortLookback = 8;  // predecessors a look-back step reads per digit

struct PassShifts {
  uint32_t shift[kSortMaxPasses];
  uint32_t numPasses;
};

__device__ __forceinline__ uint32_t sort_count(const uint32_t* count, uint32_t capacity) {
  return min(*count, capacity);
}

// The u32 key list a sort may start from instead of a key slot (LIST
// instances): `field` is the device address of a field holding the list's
// address.  Its keys take their positions as values.
__device__ __forceinline__ const uint32_t* sort_list(const uint64_t* field) {
  return ptr<const uint32_t>(gld(field));
}

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

// First epoch of the sort that follows state epoch `last` with P passes.
__device__ __forceinline__ uint32_t sort_epoch0(uint32_t last, uint32_t P) {
  const uint32_t e0 = last + 1;
  return e0 + P > (1u << kEpochBits) ? 1u : e0;
}

template <bool LIST>
__global__ __launch_bounds__(kUpsweepThreads) void k_sort_upsweep(const uint64_t* keys, const uint64_t* listField,
                                                                   const uint32_t* count, uint32_t capacity, PassShifts ps,
                                                                   SortState* st, uint64_t* status, uint32_t maxTiles) {
  __shared__ uint32_t s_cnt[kSortMaxPasses][kSortRadix];
  __shared__ uint32_t s_last;
  const uint32_t n = sort_count(count, capacity), P = ps.numPasses;
  const uint32_t* list = LIST ? sort_list(listField) : nullptr;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  // read before any workgroup counts in; only the last one rewrites it
  const uint32_t last = st->epoch, e0 = sort_epoch0(last, P);
  for (uint32_t i = threadIdx.x; i < P * kSortRadix; i += blockDim.x) (&s_cnt[0][0])[i] = 0;
  if (e0 != last + 1) {
    // the epoch wraps: clear every status word, so none matches again
    const uint64_t words = (uint64_t)maxTiles * kSortRadix;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
      status[i] = 0;
  }
  __syncthreads();
  for (uint32_t i0 = blockIdx.x * blockDim.x + wave * kWave; i0 < n; i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + lane;
    const bool valid = i < n;
    const uint64_t k = valid ? (LIST ? (uint64_t)gld(list + i) : gld(keys + i)) : 0ull;
    const uint64_t live = ballot(valid);
#pragma unroll
    for (uint32_t p = 0; p < kSortMaxPasses; p++) {
      if (p >= P) break;
      const uint32_t d = (uint32_t)(k >> ps.shift[p]) & 0xFFu;
      // a wave whose keys share the digit (sorted high digits) adds once
      if (ballot(valid && d == (uint32_t)__builtin_amdgcn_readfirstlane(d)) == live) {
        if (lane == 0) atomicAdd(&s_cnt[p][d], (uint32_t)__popcll(live));
      } else if (valid) {
        atomicAdd(&s_cnt[p][d], 1u);
      }
    }
  }
  __syncthreads();
  // The counts are device-scope atomics whose results are waited for, so
  // each is performed before this workgroup counts in; an agent-scope fence
  // would instead write back the whole L2 from every wave.
  uint32_t seen = 0;
  for (uint32_t i = threadIdx.x; i < P * kSortRadix; i += blockDim.x) {
    const uint32_t v = (&s_cnt[0][0])[i];
    if (v) seen |= __hip_atomic_fetch_add(&st->totals[0][0] + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (seen == 0xFFFFFFFFu) s_last = 2;  // never: makes the wait on the results explicit
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(&st->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1 ? 1u : 0u;
  __syncthreads();
  if (s_last != 1u) return;
  // last workgroup: every other one has counted in; the totals are read
  // with device-scope loads
  for (uint32_t i = threadIdx.x; i < P * kSortRadix; i += blockDim.x) {
    uint32_t* t = &st->totals[0][0] + i;
    (&s_cnt[0][0])[i] = __hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(t, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  if (wave < P) {  // wave p scans pass p's 256 totals, four per lane
    uint32_t v[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = s_cnt[wave][4 * lane + j], sum += v[j];
    uint32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      st->bases[wave][4 * lane + j] = run;
      run += v[j];
    }
  }
  if (threadIdx.x == 0) {
    st->done = 0;
    for (uint32_t p = 0; p < kSortMaxPasses; p++) st->ticket[p] = 0;
    st->epoch0 = e0;
    st->epoch = e0 + P - 1;
  }
}

// Publishes the count `agg` of digit d in `tile` and walks back over the
// earlier tiles' words of digit d, kSortLookback a step, until an
// inclusive prefix; publishes the tile's own and returns the keys of digit
// d in tiles 0..tile-1.  Predecessors are held by running workgroups
// (tickets are taken in order) that publish their counts before they
// wait, so the walk ends.
__device__ __forceinline__ uint32_t digit_lookback(uint64_t* status, uint32_t epoch, uint32_t tile, uint32_t d,
                                                   uint32_t agg) {
  uint64_t* row = status + (uint64_t)tile * kSortRadix;
  atomic_store_agent(&row[d], pack_status(epoch, tile == 0 ? 2 : 1, agg));
  if (tile == 0) return 0;
  uint64_t excl = 0;
  int64_t pred = (int64_t)tile - 1;
  while (true) {
    uint64_t w[kSortLookback];
#pragma unroll
    for (uint32_t j = 0; j < kSortLookback; j++) {
      const int64_t q = pred - (int64_t)j;
      w[j] = q >= 0 ? atomic_load_agent(&status[(uint64_t)q * kSortRadix + d]) : pack_status(epoch, 2, 0);
    }
    bool stop = false, blocked = false;
    uint32_t used = 0;
#pragma unroll
    for (uint32_t j = 0; j < kSortLookback; j++) {
      if (stop || blocked) continue;
      const uint32_t flag = (uint32_t)(w[j] >> 42) & 3u;
      if ((uint32_t)(w[j] >> 44) != epoch || flag == 0) {
        blocked = true;
      } else {
        excl += w[j] & kValueMask;
        used++;
        stop = flag == 2;
      }
    }
    if (stop) break;
    pred -= used;
    if (used == 0) __builtin_amdgcn_s_sleep(1);
  }
  atomic_store_agent(&row[d], pack_status(epoch, 2, excl + agg));
  return (uint32_t)excl;
}

template <bool LIST>
__global__ __launch_bounds__(kSortThreads) void k_sort_pass(const uint64_t* keysIn, const uint32_t* valsIn,
                                                            const uint64_t* listField, uint64_t* keysOut, uint32_t* valsOut,
                                                            const uint32_t* count, uint32_t capacity, SortState* st,
                                                            uint64_t* status, uint32_t pass, uint32_t shift) {
  constexpr uint32_t K = kSortKeysPerThread;
  __shared__ uint32_t s_wave[kSortWaves][kSortRadix];  // per-wave digit counts, then per-wave offsets
  __shared__ uint32_t s_base[kSortRadix];
  __shared__ uint32_t s_tile;
  // the count, the pass's epoch and its digit bases load together
  const uint32_t n = sort_count(count, capacity);
  const uint32_t* list = LIST ? sort_list(listField) : nullptr;  // a first pass over a u32 list
  const uint32_t epoch = st->epoch0 + pass;
  const uint32_t digitBase = threadIdx.x < kSortRadix ? st->bases[pass][threadIdx.x] : 0u;
  const uint32_t numTiles = (n + kSortTileKeys - 1) / kSortTileKeys;
  if (blockIdx.x >= numTiles) return;  // exactly numTiles workgroups take tickets
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_tile = __hip_atomic_fetch_add(&st->ticket[pass], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t i = threadIdx.x; i < kSortWaves * kSortRadix; i += kSortThreads) (&s_wave[0][0])[i] = 0;
  // wave w ranks the keys [first, first + 64 K) of the tile, 64 per round.
  // Tickets almost always follow dispatch order, so the keys of tile
  // blockIdx.x load while the ticket is taken, and load again otherwise.
  uint64_t key[K];
  uint32_t val[K], rank[K];
  uint32_t tile = blockIdx.x;
  uint32_t first = tile * kSortTileKeys + wave * (kWave * K);
#pragma unroll
  for (uint32_t r = 0; r < K; r++) {
    const uint32_t i = first + r * kWave + lane;
    key[r] = i < n ? (LIST ? (uint64_t)gld(list + i) : gld(keysIn + i)) : 0ull;
    val[r] = i < n ? (LIST ? i : gld(valsIn + i)) : 0u;
  }
  __syncthreads();
  if (s_tile >= numTiles) return;  // uniform; only a state left inconsistent gets here
  if (s_tile != tile) {
    tile = s_tile;
    first = tile * kSortTileKeys + wave * (kWave * K);
#pragma unroll
    for (uint32_t r = 0; r < K; r++) {
      const uint32_t i = first + r * kWave + lane;
      key[r] = i < n ? (LIST ? (uint64_t)gld(list + i) : gld(keysIn + i)) : 0ull;
      val[r] = i < n ? (LIST ? i : gld(valsIn + i)) : 0u;
    }
  }
#pragma unroll
  for (uint32_t r = 0; r < K; r++) {
    const bool valid = first + r * kWave + lane < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & 0xFFu;
    const uint64_t peers = match_digit(d, valid);
    rank[r] = s_wave[wave][d] + mask_prefix(peers);
    __builtin_amdgcn_wave_barrier();
    // the highest lane of each digit group advances the wave's count
    if (valid && (peers >> lane) == 1ull) s_wave[wave][d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // one digit per thread of the first kSortRadix: the tile's count, the
  // waves' offsets, the keys of the digit in earlier tiles
  if (threadIdx.x < kSortRadix) {
    const uint32_t d = threadIdx.x;
    uint32_t agg = 0;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; w++) {
      const uint32_t c = s_wave[w][d];
      s_wave[w][d] = agg;
      agg += c;
    }
    s_base[d] = digitBase + digit_lookback(status, epoch, tile, d, agg);
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < K; r++) {
    if (first + r * kWave + lane >= n) continue;
    const uint32_t dd = (uint32_t)(key[r] >> shift) & 0xFFu;
    const uint32_t dst = s_base[dd] + s_wave[wave][dd] + rank[r];
    gst(keysOut + dst, key[r]);
    gst(valsOut + dst, val[r]);
  }
}

// The first pass reads slot 0, or the list `listField` names and writes
// slot 0; each later pass reads the slot the one before wrote.
static void run_passes(const SortBuffers& sb, const uint32_t* count, const PassShifts& ps, hipStream_t stream,
                       uint32_t* outIndex, const uint64_t* listField) {
  if (listField)
    hipLaunchKernelGGL(k_sort_upsweep<true>, dim3(kUpsweepGroups), dim3(kUpsweepThreads), 0, stream, sb.keys[0], listField,
                       count, sb.capacity, ps, sb.state, sb.status, sb.maxTiles);
  else
    hipLaunchKernelGGL(k_sort_upsweep<false>, dim3(kUpsweepGroups), dim3(kUpsweepThreads), 0, stream, sb.keys[0], listField,
                       count, sb.capacity, ps, sb.state, sb.status, sb.maxTiles);
  const uint32_t grid = std::max(1u, (sb.capacity + kSortTileKeys - 1) / kSortTileKeys);
  int cur = listField ? -1 : 0;  // -1: the list
  for (uint32_t p = 0; p < ps.numPasses; p++) {
    const int out = cur < 0 ? 0 : cur ^ 1;
    if (cur < 0)
      hipLaunchKernelGGL(k_sort_pass<true>, dim3(grid), dim3(kSortThreads), 0, stream, nullptr, nullptr, listField,
                         sb.keys[out], sb.vals[out], count, sb.capacity, sb.state, sb.status, p, ps.shift[p]);
    else
      hipLaunchKernelGGL(k_sort_pass<false>, dim3(grid), dim3(kSortThreads), 0, stream, sb.keys[cur], sb.vals[cur], nullptr,
                         sb.keys[out], sb.vals[out], count, sb.capacity, sb.state, sb.status, p, ps.shift[p]);
    cur = out;
  }
  if (outIndex) *outIndex = cur < 0 ? 0u : (uint32_t)cur;
}

void radix_sort_pairs(const SortBuffers& sb, const uint32_t* count, uint32_t bit0, uint32_t bits, hipStream_t stream,
                      uint32_t* outIndex, const uint64_t* listField) {
  PassShifts ps{};
  for (uint32_t shift = bit0; shift < bit0 + bits && ps.numPasses < kSortMaxPasses; shift += 8)
    ps.shift[ps.numPasses++] = shift;
  run_passes(sb, count, ps, stream, outIndex, listField);
}

void radix_sort_digits(const SortBuffers& sb, const uint32_t* count, uint32_t digitMask, hipStream_t stream,
                       uint32_t* outIndex) {
  PassShifts ps{};
  for (uint32_t d = 0; d < 8; d++)
    if ((digitMask >> d) & 1u) ps.shift[ps.numPasses++] = 8 * d;
  run_passes(sb, count, ps, stream, outIndex, nullptr);
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
  // segStart: as k_insert_seg_start leaves it; for a template pass sorted
  // in one digit pass it may be that pass's digit bases instead (the first
  // sorted position of each instance, as every instance is its own digit)
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
