// tc_sort.h -- device radix sort and BLAS insertion kernels.
#pragma once

#include "tc_kernels.h"

namespace tc {

constexpr uint32_t kSortMaxPasses = 8;  // 8-bit digits of a u64 key

// Sort state in the workspace (zeroed once by tc_workspace_init): the
// upsweep leaves the digit bases of every pass and the pass tickets ready,
// and clears what it accumulated, so a sort needs no clearing launch.
struct SortState {
  uint32_t epoch;                    // look-back epoch of the latest sort's last pass
  uint32_t epoch0;                   // ... and of its first pass
  uint32_t done;                     // upsweep workgroups finished (the last resets it)
  uint32_t ticket[kSortMaxPasses];   // tile tickets of each pass
  uint32_t _pad[5];
  uint32_t totals[kSortMaxPasses][kSortRadix];  // digit counts being accumulated
  uint32_t bases[kSortMaxPasses][kSortRadix];   // exclusive digit bases of each pass
};

struct SortBuffers {
  uint64_t* keys[2];
  uint32_t* vals[2];
  uint64_t* status;   // maxTiles x kSortRadix look-back words
  SortState* state;
  uint32_t capacity;  // max element count
  uint32_t maxTiles;  // ceil(capacity / kSortTileKeys)
};

// Sorts keys[0]/vals[0] by key bits [bit0, bit0 + bits) (whole passes of 8
// bits); *outIndex receives the ping-pong slot holding the result.  With
// listField (the device address of a field holding a u32 list's address)
// the sort starts from that list instead, each key's value its position.
void radix_sort_pairs(const SortBuffers& sb, const uint32_t* count, uint32_t bit0, uint32_t bits, hipStream_t stream,
                      uint32_t* outIndex, const uint64_t* listField = nullptr);
// The same over the 8-bit digits named by digitMask (bit d: key bits
// [8d, 8d + 8)), lowest digit first; digits where every key agrees may be
// left out.
void radix_sort_digits(const SortBuffers& sb, const uint32_t* count, uint32_t digitMask, hipStream_t stream,
                       uint32_t* outIndex);

__global__ void k_insert_seg_start(KArgs a, const uint64_t* keys, uint32_t* segStart, uint32_t capacity, int transient);
__global__ void k_insert_scatter(KArgs a, const uint64_t* keys, const uint32_t* vals, const uint32_t* segStart,
                                 uint32_t capacity, int transient);

}  // namespace tc
