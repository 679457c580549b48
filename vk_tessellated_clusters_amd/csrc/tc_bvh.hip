// tc_bvh.hip -- per-instance BLAS and the TLAS over instances.
//
// Replaces the NV cluster BLAS build and the TLAS update
// (renderer_raytrace_clusters_tess.cpp:695-745, renderer.cpp:318-337):
//   keys    24-bit Morton code (8 bits per axis) of every CLAS centre inside
//           its instance's geometry bounds, the instance index above it:
//           (instance << 32) | morton24, one 64-bit key;
//   sort    the radix sort of tc_sort.hip (one upsweep, then one
//           single-pass launch per digit) over the three Morton digits and
//           the one or two instance digits (stable, CLAS order breaks ties);
//   build   Karras' binary radix tree over each instance's key segment
//           (internal nodes of instance i live at [s_i, s_i + n_i - 1));
//   refit   bottom-up with one arrival counter per internal node, in LDS
//           for nodes whose key range lies inside one workgroup's chunk of
//           sorted positions, in global memory for the few that straddle;
//           a node's four-wide record is written with it;
//   TLAS    the same builder over the instances' world-space boxes, with a
//           30-bit Morton code (morton_of) per instance.
#include <algorithm>

#include "tc_bvh.h"

namespace tc {

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of the box centre, normalised to [lo, hi]
__host__ __device__ __forceinline__ uint32_t morton_of(const float* blo, const float* bhi, const float* lo,
                                                       const float* hi) {
  uint32_t code = 0;
  for (int k = 0; k < 3; k++) {
    const float c = (blo[k] + bhi[k]) * 0.5f;
    const float ext = hi[k] - lo[k];
    float n = ext > 0.0f ? (c - lo[k]) / ext : 0.5f;
    n = fminf(fmaxf(n, 0.0f), 1.0f);
    uint32_t q = (uint32_t)(n * 1023.0f + 0.5f);
    q = (q | (q << 16)) & 0x030000FFu;
    q = (q | (q << 8)) & 0x0300F00Fu;
    q = (q | (q << 4)) & 0x030C30C3u;
    q = (q | (q << 2)) & 0x09249249u;
    code |= q << k;
  }
  return code;
}

// 24-bit Morton code (8 bits per axis) of the box centre, normalised to
// [lo, hi]: the BLAS key below the instance, three radix digits
__device__ __forceinline__ uint32_t morton24_of(const float* blo, const float* bhi, const float* lo, const float* hi) {
  uint32_t code = 0;
  for (int k = 0; k < 3; k++) {
    const float c = (blo[k] + bhi[k]) * 0.5f;
    const float ext = hi[k] - lo[k];
    float n = ext > 0.0f ? (c - lo[k]) / ext : 0.5f;
    n = fminf(fmaxf(n, 0.0f), 1.0f);
    uint32_t q = (uint32_t)(n * 255.0f + 0.5f);
    q = (q | (q << 8)) & 0x0300F00Fu;
    q = (q | (q << 4)) & 0x030C30C3u;
    q = (q | (q << 2)) & 0x09249249u;
    code |= q << k;
  }
  return code;
}

// instance owning global reference position k (lists are contiguous in
// instance order after blas_setup_insertion)
__device__ __forceinline__ uint32_t owner_instance(const tc_blas_build_info* infos, uint32_t numInstances, uint64_t base,
                                                   uint32_t k) {
  uint32_t lo = 0, hi = numInstances;  // last i with start(i) <= k
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    const uint64_t s = (infos[mid].clusterReferences - base) / 8;
    if (s <= k)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Instances whose list starts a keys workgroup stages in LDS for the owner
// search (more instances search the lists in global memory).
constexpr uint32_t kKeysLdsInstances = 4096;

// BLAS keys over every inserted CLAS reference.  A reference's header loads
// while its owner is searched among the list starts.
__global__ __launch_bounds__(256) void k_blas_keys(BvhArgs a) {
  __shared__ uint32_t s_start[kKeysLdsInstances];
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->blasClusterCounter, a.capacity);
  const tc_blas_build_info* infos = ptr<const tc_blas_build_info>(b->blasBuildInfos);
  const uint64_t* refs = ptr<const uint64_t>(b->blasClusterAddresses);
  const uint64_t base = b->blasClusterAddresses;
  const bool lds = a.numInstances <= kKeysLdsInstances;
  if (lds)
    for (uint32_t i = threadIdx.x; i < a.numInstances; i += blockDim.x)
      s_start[i] = (uint32_t)((gld(&infos[i].clusterReferences) - base) / 8);
  __syncthreads();
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const tc_clas_header h = gld_record(ptr<const tc_clas_header>(gld(refs + k)));
    uint32_t inst;
    if (lds) {
      uint32_t lo = 0, hi = a.numInstances;  // last i with start(i) <= k
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_start[mid] <= k)
          lo = mid;
        else
          hi = mid;
      }
      inst = lo;
    } else {
      inst = owner_instance(infos, a.numInstances, base, k);
    }
    const tc_render_instance& ri = a.instances[inst];
    const uint32_t m = morton24_of(h.lo, h.hi, ri.geoLo, ri.geoHi);
    a.keys[k] = ((uint64_t)inst << 32) | m;
    a.vals[k] = k;
  }
}

// length of the common prefix of keys i and j (ties broken by index);
// -1 outside [first, last]
__device__ __forceinline__ int delta(const uint64_t* keys, int64_t i, int64_t j, int64_t first, int64_t last) {
  if (j < first || j > last) return -1;
  const uint64_t ki = keys[i], kj = keys[j];
  if (ki == kj) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(ki ^ kj);
}

// Karras (2012) radix tree for the segment [first, last] of sorted keys;
// thread handles internal node `g` (global index first + local j).
__device__ __forceinline__ void karras_node(const uint64_t* keys, int64_t first, int64_t last, int64_t i, uint32_t* child,
                                            uint32_t* parent, uint32_t leafParentBase, uint32_t* rangeEnd = nullptr) {
  const int d = (delta(keys, i, i + 1, first, last) - delta(keys, i, i - 1, first, last)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, i, i - d, first, last);
  int64_t lmax = 2;
  while (delta(keys, i, i + lmax * d, first, last) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, i, i + (l + t) * d, first, last) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(keys, i, j, first, last);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) / 2;
    if (delta(keys, i, i + (s + t) * d, first, last) > dnode) s += t;
  } while (t > 1);
  const int64_t gamma = i + s * d + min(d, 0);
  const int64_t lo = min(i, j), hi = max(i, j);
  if (rangeEnd) rangeEnd[i] = (uint32_t)j;
  // children: leaf flag + position relative to `first`, or internal node
  const uint32_t c0 = (lo == gamma) ? (TC_BVH_LEAF | (uint32_t)(gamma - first)) : (uint32_t)(gamma - first);
  const uint32_t c1 = (hi == gamma + 1) ? (TC_BVH_LEAF | (uint32_t)(gamma + 1 - first)) : (uint32_t)(gamma + 1 - first);
  const int64_t node = i - first;
  child[2 * (first + node) + 0] = c0;
  child[2 * (first + node) + 1] = c1;
  // parents (global indices), leaves offset by leafParentBase
  if (c0 & TC_BVH_LEAF)
    parent[leafParentBase + first + (c0 & ~TC_BVH_LEAF)] = (uint32_t)(first + node);
  else
    parent[first + c0] = (uint32_t)(first + node);
  if (c1 & TC_BVH_LEAF)
    parent[leafParentBase + first + (c1 & ~TC_BVH_LEAF)] = (uint32_t)(first + node);
  else
    parent[first + c1] = (uint32_t)(first + node);
}

// segment [first, last] of instance inst in the sorted reference list
__device__ __forceinline__ void blas_segment(const BvhArgs& a, uint32_t inst, uint32_t n, int64_t& first, int64_t& last) {
  const tc_blas_build_info& bi = ptr<const tc_blas_build_info>(a.build->blasBuildInfos)[inst];
  first = (int64_t)((bi.clusterReferences - a.build->blasClusterAddresses) / 8);
  last = first + (int64_t)bi.clusterReferencesCount - 1;
}

// One thread per sorted position k: internal node k of its instance (when
// k < last) and the sorted CLAS address.
__global__ void k_blas_hierarchy(BvhArgs a) {
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->blasClusterCounter, a.capacity);
  const uint64_t* refs = ptr<const uint64_t>(b->blasClusterAddresses);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t inst = (uint32_t)(a.sortedKeys[k] >> 32);
    int64_t first, last;
    blas_segment(a, inst, n, first, last);
    a.leafAddr[k] = refs[a.sortedVals[k]];
    a.counter[k] = 0;
    if ((int64_t)k < last) karras_node(a.sortedKeys, first, last, k, a.child, a.parent, a.capacity, a.rangeEnd);
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.numInstances; i += gridDim.x * blockDim.x)
    a.pendingCount[i] = 0;
}

// Box of sorted leaf k from its CLAS header; also writes the leaf's info
// word (counts | clusterID << 32) from the same header.  The refit stages
// every leaf once; the leaves of straddling nodes are read again.
__device__ __forceinline__ void leaf_box(const BvhArgs& a, uint32_t k, float lo[3], float hi[3]) {
  const tc_clas_header h = gld_record(ptr<const tc_clas_header>(a.leafAddr[k]));
  for (int c = 0; c < 3; c++) lo[c] = h.lo[c], hi[c] = h.hi[c];
  a.leafInfo[k] = (uint64_t)h.counts | (uint64_t)h.clusterID << 32;
}

// One slot of a four-wide record: a box and its child word.
struct Slot4 {
  float lo[3], hi[3];
  uint32_t c;
};
__device__ __forceinline__ Slot4 slot4(const float* lo, const float* hi, uint32_t c) {
  Slot4 r;
  for (int q = 0; q < 3; q++) r.lo[q] = lo[q], r.hi[q] = hi[q];
  r.c = c;
  return r;
}

// Four-wide record of a binary node: its children in order, each internal
// child replaced by that child's two children, empty slots last.  Child 0
// gives a0 (and a1 when internal), child 1 gives b0 (and b1).  The slots
// are chosen by selects, so the record stays in registers.
__device__ __forceinline__ Slot4 sel4(bool c, const Slot4& x, const Slot4& y) {
  Slot4 r;
  for (int q = 0; q < 3; q++) r.lo[q] = c ? x.lo[q] : y.lo[q], r.hi[q] = c ? x.hi[q] : y.hi[q];
  r.c = c ? x.c : y.c;
  return r;
}
__device__ __forceinline__ tc_bvh4_node make4(bool leaf0, const Slot4& a0, const Slot4& a1, bool leaf1, const Slot4& b0,
                                              const Slot4& b1) {
  Slot4 e;
  for (int q = 0; q < 3; q++) e.lo[q] = INFINITY, e.hi[q] = -INFINITY;
  e.c = TC_BVH4_EMPTY;
  const Slot4 s[4] = {a0, sel4(leaf0, b0, a1), sel4(leaf0, sel4(leaf1, e, b1), b0), sel4(leaf0 || leaf1, e, b1)};
  tc_bvh4_node o;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    for (int q = 0; q < 3; q++) o.lo[q][k] = s[k].lo[q], o.hi[q][k] = s[k].hi[q];
    o.child[k] = s[k].c;
  }
  for (int q = 0; q < 4; q++) o._pad[q] = 0;
  return o;
}

// Four-wide record of binary node n of a tree whose node array starts at
// `nodes` (child indices relative to it; its internal children are final).
__device__ __forceinline__ tc_bvh4_node collapse4_of(const tc_bvh_node& n, const tc_bvh_node* nodes) {
  Slot4 g[2][2];
  bool leaf[2];
  for (int c = 0; c < 2; c++) {
    const uint32_t ch = c ? n.child1 : n.child0;
    leaf[c] = (ch & TC_BVH_LEAF) != 0;
    if (leaf[c]) {
      g[c][0] = g[c][1] = slot4(c ? n.lo1 : n.lo0, c ? n.hi1 : n.hi0, ch);
    } else {
      const tc_bvh_node m = gld_record(nodes + ch);
      g[c][0] = slot4(m.lo0, m.hi0, m.child0);
      g[c][1] = slot4(m.lo1, m.hi1, m.child1);
    }
  }
  return make4(leaf[0], g[0][0], g[0][1], leaf[1], g[1][0], g[1][1]);
}
__device__ __forceinline__ void collapse4(const tc_bvh_node* nodes, uint32_t i, tc_bvh4_node* out) {
  gst_record(out + i, collapse4_of(gld_record(nodes + i), nodes));
}

// Bottom-up refit in two launches, with no device-scope fence.
//
// k_blas_refit: a workgroup owns a chunk of kRefitChunk sorted positions,
// one thread per position.  It first stages the chunk's topology (child
// words, parents, range ends of the internal nodes, the leaves' parents)
// and its leaves' boxes (from their CLAS headers, which also give the leaf
// info words) in LDS, all loads in parallel.  Then every leaf climbs; the
// second arrival at a node writes it.  Nodes whose key range lies inside
// the chunk are read and counted in LDS (workgroup-scope ordering), so a
// climb waits on no global load.  A climb stops at the first node whose
// range straddles chunks; the second arrival there lists the node as
// pending for its instance (both children are then final).  The chunk's
// four-wide records are written from LDS once every climb is done.
//
// k_blas_refit_top: one workgroup per instance climbs from its pending
// nodes through the straddling nodes with global counters (continuing the
// counts of the first launch) and workgroup-scope ordering, since every
// climber of an instance is in that workgroup; each node's four-wide
// record is written where the node is refit.  A pending node's range
// holds exactly one chunk boundary (its split), so an instance has fewer
// pending nodes than chunks it spans; its list starts at
// (first >> kRefitChunkBits) + instance.
constexpr uint32_t kRefitChunkBits = 10;
constexpr uint32_t kRefitChunk = 1u << kRefitChunkBits;
constexpr uint32_t kRefitThreads = kRefitChunk;
constexpr uint32_t kRefitTopThreads = 1024;

// Writes internal node `node` from its children's boxes (both final) and
// returns its union box in `self`.
__device__ __forceinline__ tc_bvh_node write_node(const BvhArgs& a, uint32_t node, uint32_t c0w, uint32_t c1w,
                                                  const float lo[2][3], const float hi[2][3], tc_aabb& self) {
  tc_bvh_node out;
  for (int q = 0; q < 3; q++) {
    out.lo0[q] = lo[0][q], out.hi0[q] = hi[0][q], out.lo1[q] = lo[1][q], out.hi1[q] = hi[1][q];
  }
  out.child0 = c0w;
  out.child1 = c1w;
  out._pad0 = out._pad1 = 0;
  gst_record(a.nodes + node, out);
  for (int q = 0; q < 3; q++) self.lo[q] = fminf(lo[0][q], lo[1][q]), self.hi[q] = fmaxf(hi[0][q], hi[1][q]);
  gst_record(a.nodeBox + node, self);  // read by straddling parents and the TLAS
  return out;
}

__global__ __launch_bounds__(kRefitThreads) void k_blas_refit(BvhArgs a) {
  __shared__ uint32_t s_cnt[kRefitChunk];
  __shared__ tc_aabb s_box[kRefitChunk];   // internal nodes' union boxes
  __shared__ tc_aabb s_leaf[kRefitChunk];  // leaves' boxes
  __shared__ uint32_t s_child[kRefitChunk][2];
  __shared__ uint32_t s_parent[kRefitChunk], s_leafParent[kRefitChunk], s_range[kRefitChunk];
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->blasClusterCounter, a.capacity);
  const uint32_t j = threadIdx.x;
  for (uint32_t c0 = blockIdx.x * kRefitChunk; c0 < n; c0 += gridDim.x * kRefitChunk) {
    const uint32_t c1 = min(c0 + kRefitChunk, n);
    const uint32_t k = c0 + j;
    int64_t first = 0, last = 0;
    s_cnt[j] = 0;
    if (k < c1) {
      blas_segment(a, (uint32_t)(a.sortedKeys[k] >> 32), n, first, last);
      s_leafParent[j] = a.parent[a.capacity + k];
      // internal node k exists when k < last; the other entries are unread
      s_parent[j] = a.parent[k];
      s_range[j] = a.rangeEnd[k];
      s_child[j][0] = a.child[2 * k + 0];
      s_child[j][1] = a.child[2 * k + 1];
      float lo[3], hi[3];
      leaf_box(a, k, lo, hi);
      for (int q = 0; q < 3; q++) s_leaf[j].lo[q] = lo[q], s_leaf[j].hi[q] = hi[q];
    }
    __syncthreads();
    if (k < c1 && first != last) {  // a single CLAS is its instance's root
      uint32_t node = s_leafParent[j];
      while (true) {
        const bool local = node >= c0 && node < c1 && s_range[node - c0] >= c0 && s_range[node - c0] < c1;
        if (!local) {
          // straddling: the second arrival lists it for the second launch
          if (atomicAdd(&a.counter[node], 1u) == 1u) {
            const uint32_t inst = (uint32_t)(a.sortedKeys[k] >> 32);
            const uint32_t slot = atomicAdd(&a.pendingCount[inst], 1u);
            a.pendingList[((uint32_t)first >> kRefitChunkBits) + inst + slot] = node;
          }
          break;
        }
        const uint32_t ln = node - c0;
        // the arrival orders LDS only (s_box, s_cnt): no wait for the node
        // stores, which no climber of this chunk reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        if (atomicAdd(&s_cnt[ln], 1u) == 0) break;  // first arrival stops
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        // a local node's children (leaves or internal nodes) lie in the chunk
        float lo[2][3], hi[2][3];
        const uint32_t cw[2] = {s_child[ln][0], s_child[ln][1]};
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const uint32_t g = (uint32_t)first + (cw[c] & ~TC_BVH_LEAF) - c0;
          const tc_aabb& bx = (cw[c] & TC_BVH_LEAF) ? s_leaf[g] : s_box[g];
          for (int q = 0; q < 3; q++) lo[c][q] = bx.lo[q], hi[c][q] = bx.hi[q];
        }
        tc_aabb self;
        write_node(a, node, cw[0], cw[1], lo, hi, self);
        s_box[ln] = self;
        if ((int64_t)node == first) break;  // instance root
        node = s_parent[ln];
      }
    }
    __syncthreads();
    // four-wide records of the chunk's local internal nodes, one per
    // thread, from LDS (the grandchildren of a local node lie in the chunk)
    if (k < c1 && (int64_t)k < last && s_range[j] >= c0 && s_range[j] < c1) {
      const uint32_t cw[2] = {s_child[j][0], s_child[j][1]};
      Slot4 gs[2][2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        if (cw[c] & TC_BVH_LEAF) {
          const tc_aabb& bx = s_leaf[(uint32_t)first + (cw[c] & ~TC_BVH_LEAF) - c0];
          gs[c][0] = gs[c][1] = slot4(bx.lo, bx.hi, cw[c]);
        } else {
          const uint32_t g = (uint32_t)first + cw[c] - c0;
#pragma unroll
          for (int d = 0; d < 2; d++) {
            const uint32_t gw = s_child[g][d];
            const uint32_t h = (uint32_t)first + (gw & ~TC_BVH_LEAF) - c0;
            const tc_aabb& bx = (gw & TC_BVH_LEAF) ? s_leaf[h] : s_box[h];
            gs[c][d] = slot4(bx.lo, bx.hi, gw);
          }
        }
      }
      gst_record(a.nodes4 + k, make4((cw[0] & TC_BVH_LEAF) != 0, gs[0][0], gs[0][1], (cw[1] & TC_BVH_LEAF) != 0,
                                     gs[1][0], gs[1][1]));
    }
    __syncthreads();
  }
}

// Arrivals of this launch are counted in LDS, one byte per internal node
// (four per word), for instances of up to kRefitTopLdsNodes internal nodes;
// the count the chunk launch left in the global counter is only read.
// Larger instances count with global atomics.
constexpr uint32_t kRefitTopLdsNodes = 131072;

// Refit of straddling node `node` from its child words cw, with its
// four-wide record: a leaf child's box comes from its CLAS, an internal
// child's from its binary record, whose two slots are also the grandchildren
// the four-wide record takes (the record's union is the child's node box,
// to the bit).
__device__ __forceinline__ void refit_top_node(const BvhArgs& a, uint32_t node, const uint32_t cw[2], int64_t first) {
  float lo[2][3], hi[2][3];
  Slot4 g[2][2];
  bool leaf[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const uint32_t ch = cw[c];
    leaf[c] = (ch & TC_BVH_LEAF) != 0;
    if (leaf[c]) {
      leaf_box(a, (uint32_t)first + (ch & ~TC_BVH_LEAF), lo[c], hi[c]);
      g[c][0] = g[c][1] = slot4(lo[c], hi[c], ch);
    } else {
      const tc_bvh_node m = gld_record(a.nodes + (uint32_t)first + ch);
      for (int q = 0; q < 3; q++) lo[c][q] = fminf(m.lo0[q], m.lo1[q]), hi[c][q] = fmaxf(m.hi0[q], m.hi1[q]);
      g[c][0] = slot4(m.lo0, m.hi0, m.child0);
      g[c][1] = slot4(m.lo1, m.hi1, m.child1);
    }
  }
  tc_aabb self;
  write_node(a, node, cw[0], cw[1], lo, hi, self);
  gst_record(a.nodes4 + node, make4(leaf[0], g[0][0], g[0][1], leaf[1], g[1][0], g[1][1]));
}

__global__ __launch_bounds__(kRefitTopThreads) void k_blas_refit_top(BvhArgs a) {
  __shared__ uint32_t s_arrive[kRefitTopLdsNodes / 4];
  const uint32_t inst = blockIdx.x;
  if (inst >= a.numInstances) return;
  const tc_scene_building* b = a.build;
  const uint32_t n = min(b->blasClusterCounter, a.capacity);
  const tc_blas_build_info& bi = ptr<const tc_blas_build_info>(b->blasBuildInfos)[inst];
  if (bi.clusterReferencesCount < 2) return;
  const int64_t first = (int64_t)((bi.clusterReferences - b->blasClusterAddresses) / 8);
  if (first >= (int64_t)n) return;
  const uint32_t count = a.pendingCount[inst];
  const uint32_t base = ((uint32_t)first >> kRefitChunkBits) + inst;
  const bool ldsCount = bi.clusterReferencesCount - 1 <= kRefitTopLdsNodes;
  if (ldsCount)
    for (uint32_t w = threadIdx.x; w < (bi.clusterReferencesCount + 2) / 4; w += kRefitTopThreads) s_arrive[w] = 0;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < count; t += kRefitTopThreads) {
    uint32_t node = a.pendingList[base + t];
    // A node's child words and parent (written by the hierarchy launch) and
    // the chunk launch's arrivals there are final, so the parent's load
    // together with the node's child records: one round trip a level.
    uint32_t cw[2] = {gld(a.child + 2 * node), gld(a.child + 2 * node + 1)};
    uint32_t up = gld(a.parent + node);
    while (true) {
      const bool root = (int64_t)node == first;
      uint32_t ncw[2] = {0u, 0u}, nup = 0u, prior = 0u;
      if (!root) {
        ncw[0] = gld(a.child + 2 * up), ncw[1] = gld(a.child + 2 * up + 1);
        nup = gld(a.parent + up);
        prior = ldsCount ? gld(a.counter + up) : 0u;
      }
      refit_top_node(a, node, cw, first);
      if (root) break;
      node = up;
      cw[0] = ncw[0], cw[1] = ncw[1];
      up = nup;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      uint32_t before;
      if (ldsCount) {
        const uint32_t rel = node - (uint32_t)first, sh = 8 * (rel & 3);
        before = ((atomicAdd(&s_arrive[rel >> 2], 1u << sh) >> sh) & 0xFFu) + prior;
      } else {
        before = atomicAdd(&a.counter[node], 1u);
      }
      if (before == 0) break;  // first arrival stops
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
}

// TLAS record and world box of instance i.
__device__ void tlas_prep_one(const BvhArgs& a, const float* xforms /* 24 floats per instance */, uint32_t i) {
  const tc_blas_build_info bi = ptr<const tc_blas_build_info>(a.build->blasBuildInfos)[i];
  const uint32_t first = (uint32_t)((bi.clusterReferences - a.build->blasClusterAddresses) / 8);
  const uint32_t cnt = bi.clusterReferencesCount;
  tc_tlas_instance t;
  for (int q = 0; q < 12; q++) t.objectToWorld[q] = xforms[24 * i + q], t.worldToObject[q] = xforms[24 * i + 12 + q];
  t.blasNodes = reinterpret_cast<uint64_t>(a.nodes + first);
  t.blasLeaves = reinterpret_cast<uint64_t>(a.leafAddr + first);
  t.blasRoot = cnt == 1 ? TC_BVH_LEAF : 0u;
  t.numClusters = cnt;
  t.flags = 1;
  t._pad = 0;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (cnt == 1) {
    // a single CLAS is the root and has no parent to refit: its box and info here
    const tc_clas_header h = *ptr<const tc_clas_header>(a.leafAddr[first]);
    for (int q = 0; q < 3; q++) lo[q] = h.lo[q], hi[q] = h.hi[q];
    a.leafInfo[first] = (uint64_t)h.counts | (uint64_t)h.clusterID << 32;
  } else if (cnt > 1) {
    const tc_aabb bx = a.nodeBox[first];
    for (int q = 0; q < 3; q++) lo[q] = bx.lo[q], hi[q] = bx.hi[q];
  }
  a.tlasInst[i] = t;
  // world box of the object box corners
  float wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (cnt > 0)
    for (int c = 0; c < 8; c++) {
      const float p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
      for (int r = 0; r < 3; r++) {
        const float* m = t.objectToWorld + 4 * r;
        const float w = ((m[0] * p[0] + m[1] * p[1]) + m[2] * p[2]) + m[3];
        wlo[r] = fminf(wlo[r], w);
        whi[r] = fmaxf(whi[r], w);
      }
    }
  // widen by a few ULP-scale steps so rounding of the transformed corners
  // never cuts off a hit found in object space
  tc_aabb wb;
  for (int q = 0; q < 3; q++) {
    const float e = cnt > 0 ? 1e-6f * fmaxf(fmaxf(fabsf(wlo[q]), fabsf(whi[q])), 1e-3f) : 0.0f;
    wb.lo[q] = wlo[q] - e, wb.hi[q] = whi[q] + e;
  }
  a.tlasLeafBox[i] = wb;
}

// Per-instance TLAS records and world boxes (one thread per instance).
__global__ void k_tlas_prep(BvhArgs a, const float* xforms) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.numInstances) tlas_prep_one(a, xforms, i);
}

// Scene bounds and TLAS keys over instances (one workgroup).
__global__ __launch_bounds__(256) void k_tlas_keys(BvhArgs a) {
  __shared__ float s_lo[3][256], s_hi[3][256];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = threadIdx.x; i < a.numInstances; i += blockDim.x) {
    const tc_aabb bx = a.tlasLeafBox[i];
    if (!(bx.lo[0] <= bx.hi[0])) continue;
    for (int q = 0; q < 3; q++) lo[q] = fminf(lo[q], (bx.lo[q] + bx.hi[q]) * 0.5f), hi[q] = fmaxf(hi[q], (bx.lo[q] + bx.hi[q]) * 0.5f);
  }
  for (int q = 0; q < 3; q++) s_lo[q][threadIdx.x] = lo[q], s_hi[q][threadIdx.x] = hi[q];
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 3; q++) {
        s_lo[q][threadIdx.x] = fminf(s_lo[q][threadIdx.x], s_lo[q][threadIdx.x + s]);
        s_hi[q][threadIdx.x] = fmaxf(s_hi[q][threadIdx.x], s_hi[q][threadIdx.x + s]);
      }
    __syncthreads();
  }
  const float slo[3] = {s_lo[0][0], s_lo[1][0], s_lo[2][0]}, shi[3] = {s_hi[0][0], s_hi[1][0], s_hi[2][0]};
  for (uint32_t i = threadIdx.x; i < a.numInstances; i += blockDim.x) {
    const tc_aabb bx = a.tlasLeafBox[i];
    const bool empty = !(bx.lo[0] <= bx.hi[0]);
    // empty instances sort last and are never entered
    a.tlasKeys[i] = empty ? (0xFFFFFFFFull << 32) | i : ((uint64_t)morton_of(bx.lo, bx.hi, slo, shi) << 32) | i;
    a.tlasVals[i] = i;
  }
}

// TLAS over the (sorted) instances: one thread per sorted position.
__global__ void k_tlas_hierarchy(BvhArgs a) {
  const uint32_t n = a.numInstances;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    a.tlasCounter[k] = 0;
    if (k + 1 < n) karras_node(a.tlasSortedKeys, 0, (int64_t)n - 1, k, a.tlasChild, a.tlasParent, n);
  }
}

// Bottom-up refit from sorted TLAS leaf k: the second arrival at a node
// writes it.  With WORKGROUP every climber is in one workgroup, so
// workgroup-scope ordering suffices (a device-scope release writes back the
// XCD's L2 at every level).
template <bool WORKGROUP>
__device__ void tlas_refit_leaf(const BvhArgs& a, uint32_t k) {
  const uint32_t n = a.numInstances;
  {
    uint32_t node = a.tlasParent[n + k];
    while (true) {
      if (WORKGROUP)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      else
        __threadfence();
      if (atomicAdd(&a.tlasCounter[node], 1u) == 0) break;
      if (WORKGROUP)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      else
        __threadfence();
      tc_bvh_node out;
      float lo[2][3], hi[2][3];
      for (int c = 0; c < 2; c++) {
        const uint32_t ch = a.tlasChild[2 * node + c];
        tc_aabb bx = (ch & TC_BVH_LEAF) ? a.tlasLeafBox[a.tlasSortedVals[ch & ~TC_BVH_LEAF]] : a.tlasNodeBox[ch];
        for (int q = 0; q < 3; q++) lo[c][q] = bx.lo[q], hi[c][q] = bx.hi[q];
      }
      for (int q = 0; q < 3; q++) out.lo0[q] = lo[0][q], out.hi0[q] = hi[0][q], out.lo1[q] = lo[1][q], out.hi1[q] = hi[1][q];
      const uint32_t c0 = a.tlasChild[2 * node], c1 = a.tlasChild[2 * node + 1];
      // leaves of the TLAS name instances directly
      out.child0 = (c0 & TC_BVH_LEAF) ? (TC_BVH_LEAF | a.tlasSortedVals[c0 & ~TC_BVH_LEAF]) : c0;
      out.child1 = (c1 & TC_BVH_LEAF) ? (TC_BVH_LEAF | a.tlasSortedVals[c1 & ~TC_BVH_LEAF]) : c1;
      out._pad0 = out._pad1 = 0;
      a.tlasNodes[node] = out;
      tc_aabb self;
      for (int q = 0; q < 3; q++) self.lo[q] = fminf(lo[0][q], lo[1][q]), self.hi[q] = fmaxf(hi[0][q], hi[1][q]);
      a.tlasNodeBox[node] = self;
      if (node == 0) break;
      node = a.tlasParent[node];
    }
  }
}

__global__ void k_tlas_refit(BvhArgs a) {
  const uint32_t n = a.numInstances;
  if (n == 1) return;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) tlas_refit_leaf<false>(a, k);
}

__global__ void k_tlas_collapse4(BvhArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 < a.numInstances) collapse4(a.tlasNodes, i, a.tlasNodes4);
}

// instances the one-workgroup TLAS build (k_tlas_build_small) takes
constexpr uint32_t kTlasRankSort = 1024;

// TLAS root word (internal node 0, or the single instance) and the TLAS
// element count read by the sort
__global__ void k_tlas_root(BvhArgs a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    a.tlasRoot[0] = a.numInstances == 1 ? (TC_BVH_LEAF | 0u) : 0u;
    a.tlasRoot[1] = a.numInstances;
  }
}

// The whole TLAS in one workgroup for up to kTlasRankSort instances, one
// thread per instance: records and world boxes, scene bounds and keys, the
// rank sort, hierarchy, refit and four-wide view, the steps the launches
// above make, separated by workgroup barriers instead of kernel boundaries.
__global__ __launch_bounds__(kTlasRankSort) void k_tlas_build_small(BvhArgs a, const float* xforms) {
  __shared__ float s_b[6][kTlasRankSort / kWave];
  __shared__ uint64_t s_key[kTlasRankSort], s_sorted[kTlasRankSort];
  __shared__ uint32_t s_val[kTlasRankSort], s_cnt[kTlasRankSort];
  __shared__ uint32_t s_child[2 * kTlasRankSort], s_parent[2 * kTlasRankSort];
  __shared__ tc_aabb s_leaf[kTlasRankSort], s_box[kTlasRankSort];  // by instance, by internal node
  const uint32_t n = a.numInstances, i = threadIdx.x, lane = lane_id(), wave = threadIdx.x / kWave;
  if (i < n) tlas_prep_one(a, xforms, i);
  if (i == 0) {
    a.tlasRoot[0] = n == 1 ? (TC_BVH_LEAF | 0u) : 0u;
    a.tlasRoot[1] = n;
  }
  __syncthreads();
  // scene bounds of the instance box centres (k_tlas_keys)
  tc_aabb bx;
  bool empty = true;
  float c[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    bx = a.tlasLeafBox[i];
    s_leaf[i] = bx;
    empty = !(bx.lo[0] <= bx.hi[0]);
    if (!empty)
      for (int q = 0; q < 3; q++) c[q] = c[3 + q] = (bx.lo[q] + bx.hi[q]) * 0.5f;
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {
    for (int d = 1; d < kWave; d <<= 1) {
      c[q] = fminf(c[q], __shfl_xor(c[q], d, kWave));
      c[3 + q] = fmaxf(c[3 + q], __shfl_xor(c[3 + q], d, kWave));
    }
  }
  if (lane == 0)
    for (int q = 0; q < 6; q++) s_b[q][wave] = c[q];
  __syncthreads();
  float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t w = 0; w < kTlasRankSort / kWave; w++)
    for (int q = 0; q < 3; q++) slo[q] = fminf(slo[q], s_b[q][w]), shi[q] = fmaxf(shi[q], s_b[3 + q][w]);
  // keys (empty instances sort last and are never entered), then ranks
  if (i < n) s_key[i] = empty ? (0xFFFFFFFFull << 32) | i : ((uint64_t)morton_of(bx.lo, bx.hi, slo, shi) << 32) | i;
  __syncthreads();
  if (i < n) {
    const uint64_t k = s_key[i];
    uint32_t rank = 0;
    for (uint32_t j = 0; j < n; j++) rank += s_key[j] < k ? 1u : 0u;
    a.tlasKeys[rank] = k;
    a.tlasVals[rank] = i;
    s_sorted[rank] = k;
    s_val[rank] = i;
  }
  __syncthreads();
  // hierarchy over the sorted keys (as k_tlas_hierarchy), refit (as
  // tlas_refit_leaf) and four-wide view, with the topology, counters and
  // boxes in LDS: a climb waits on no global load
  if (i < n) {
    s_cnt[i] = 0;
    if (i + 1 < n) karras_node(s_sorted, 0, (int64_t)n - 1, i, s_child, s_parent, n);
  }
  __syncthreads();
  // a TLAS leaf names its instance: the sorted position's value
  const auto leaf_word = [&](uint32_t w) { return (w & TC_BVH_LEAF) ? (TC_BVH_LEAF | s_val[w & ~TC_BVH_LEAF]) : w; };
  const auto box_of = [&](uint32_t w) -> const tc_aabb& {
    return (w & TC_BVH_LEAF) ? s_leaf[s_val[w & ~TC_BVH_LEAF]] : s_box[w];
  };
  if (i < n && n > 1) {
    uint32_t node = s_parent[n + i];
    while (true) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      if (atomicAdd(&s_cnt[node], 1u) == 0) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const uint32_t cw[2] = {s_child[2 * node], s_child[2 * node + 1]};
      float lo[2][3], hi[2][3];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const tc_aabb& bx = box_of(cw[c]);
        for (int q = 0; q < 3; q++) lo[c][q] = bx.lo[q], hi[c][q] = bx.hi[q];
      }
      tc_bvh_node out;
      for (int q = 0; q < 3; q++) out.lo0[q] = lo[0][q], out.hi0[q] = hi[0][q], out.lo1[q] = lo[1][q], out.hi1[q] = hi[1][q];
      out.child0 = leaf_word(cw[0]);
      out.child1 = leaf_word(cw[1]);
      out._pad0 = out._pad1 = 0;
      gst_record(a.tlasNodes + node, out);
      tc_aabb self;
      for (int q = 0; q < 3; q++) self.lo[q] = fminf(lo[0][q], lo[1][q]), self.hi[q] = fmaxf(hi[0][q], hi[1][q]);
      s_box[node] = self;
      if (node == 0) break;
      node = s_parent[node];
    }
  }
  __syncthreads();
  if (i + 1 < n) {
    const uint32_t cw[2] = {s_child[2 * i], s_child[2 * i + 1]};
    Slot4 gs[2][2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (cw[c] & TC_BVH_LEAF) {
        const tc_aabb& bx = box_of(cw[c]);
        gs[c][0] = gs[c][1] = slot4(bx.lo, bx.hi, leaf_word(cw[c]));
      } else {
#pragma unroll
        for (int d = 0; d < 2; d++) {
          const uint32_t gw = s_child[2 * cw[c] + d];
          const tc_aabb& bx = box_of(gw);
          gs[c][d] = slot4(bx.lo, bx.hi, leaf_word(gw));
        }
      }
    }
    gst_record(a.tlasNodes4 + i, make4((cw[0] & TC_BVH_LEAF) != 0, gs[0][0], gs[0][1], (cw[1] & TC_BVH_LEAF) != 0,
                                       gs[1][0], gs[1][1]));
  }
}

}  // namespace tc

// ------------------------------------------------------------ host entry

namespace tc {

static uint64_t a256(uint64_t v) { return (v + 255) & ~255ull; }

AccelLayout accel_layout(uint32_t capacity, uint32_t numInstances) {
  AccelLayout L{};
  uint64_t o = 0;
  auto put = [&](uint64_t& f, uint64_t bytes) {
    f = o;
    o = a256(o + bytes);
  };
  const uint64_t c = capacity, n = std::max(numInstances, 1u);
  put(L.blasNodes, c * sizeof(tc_bvh_node));
  put(L.blasNodes4, c * sizeof(tc_bvh4_node));
  put(L.blasLeaves, c * 8);
  put(L.blasLeafInfo, c * 8);
  put(L.nodeBox, c * sizeof(tc_aabb));
  put(L.child, 2 * c * 4);
  put(L.parent, 2 * c * 4);
  put(L.counter, c * 4);
  put(L.rangeEnd, c * 4);
  put(L.pendingCount, n * 4);
  put(L.pendingList, (c / kRefitChunk + n + 2) * 4);
  put(L.tlasInstances, n * sizeof(tc_tlas_instance));
  put(L.tlasNodes, n * sizeof(tc_bvh_node));
  put(L.tlasNodes4, n * sizeof(tc_bvh4_node));
  put(L.tlasNodeBox, n * sizeof(tc_aabb));
  put(L.tlasLeafBox, n * sizeof(tc_aabb));
  put(L.tlasChild, 2 * n * 4);
  put(L.tlasParent, 2 * n * 4);
  put(L.tlasCounter, n * 4);
  put(L.tlasKeys, n * 8);
  put(L.tlasVals, n * 4);
  put(L.tlasRoot, 256);
  L.total = o;
  return L;
}

}  // namespace tc

using namespace tc;

extern "C" {

uint64_t tc_accel_size(uint32_t capacity, uint32_t numInstances) { return accel_layout(capacity, numInstances).total; }

tc_status tc_accel_bind(void* memory, uint64_t bytes, uint32_t capacity, uint32_t numInstances, tc_accel* out) {
  if (!memory || !out || !capacity || !numInstances) return TC_ERROR_INVALID_ARGUMENT;
  AccelLayout L = accel_layout(capacity, numInstances);
  if (bytes < L.total) return TC_ERROR_INVALID_ARGUMENT;
  char* m = static_cast<char*>(memory);
  out->blasNodes = reinterpret_cast<tc_bvh_node*>(m + L.blasNodes);
  out->blasLeaves = reinterpret_cast<uint64_t*>(m + L.blasLeaves);
  out->tlasInstances = reinterpret_cast<tc_tlas_instance*>(m + L.tlasInstances);
  out->tlasNodes = reinterpret_cast<tc_bvh_node*>(m + L.tlasNodes);
  out->tlasRoot = reinterpret_cast<uint32_t*>(m + L.tlasRoot);
  out->blasNodes4 = reinterpret_cast<tc_bvh4_node*>(m + L.blasNodes4);
  out->tlasNodes4 = reinterpret_cast<tc_bvh4_node*>(m + L.tlasNodes4);
  out->blasLeafInfo = reinterpret_cast<uint64_t*>(m + L.blasLeafInfo);
  out->numInstances = numInstances;
  out->capacity = capacity;
  out->memory = memory;
  out->memoryBytes = bytes;
  return TC_OK;
}

tc_status tc_accel_build(const tc_bindings* b, tc_config config, const tc_accel* accel, const float* instanceXforms,
                         tc_stream stream) {
  KArgs ka;
  tc_status s = make_args_public(b, config, ka);
  if (s != TC_OK) return s;
  if (!accel || !accel->memory || !instanceXforms || accel->numInstances != b->numRenderInstances) return TC_ERROR_INVALID_ARGUMENT;
  const uint32_t capacity = accel->capacity;
  if (capacity < config.maxGeneratedClusters) return TC_ERROR_INVALID_ARGUMENT;
  AccelLayout L = accel_layout(capacity, accel->numInstances);
  if (accel->memoryBytes < L.total) return TC_ERROR_INVALID_ARGUMENT;
  WsLayout W = ws_layout(config, b->numRenderInstances);
  if (W.sortCapacity < capacity) return TC_ERROR_INVALID_ARGUMENT;
  char* m = static_cast<char*>(accel->memory);
  char* w = static_cast<char*>(b->workspace);
  hipStream_t st = (hipStream_t)stream;

  BvhArgs a;
  a.build = b->build;
  a.instances = b->instances;
  a.numInstances = b->numRenderInstances;
  a.capacity = capacity;
  SortBuffers sb;
  for (int i = 0; i < 2; i++) {
    sb.keys[i] = reinterpret_cast<uint64_t*>(w + W.sortKeys[i]);
    sb.vals[i] = reinterpret_cast<uint32_t*>(w + W.sortVals[i]);
  }
  sb.status = reinterpret_cast<uint64_t*>(w + W.sortStatus);
  sb.state = reinterpret_cast<SortState*>(w + W.sortState);
  sb.capacity = W.sortCapacity;
  sb.maxTiles = W.sortMaxTiles;
  a.keys = sb.keys[0];
  a.vals = sb.vals[0];
  a.leafAddr = reinterpret_cast<uint64_t*>(m + L.blasLeaves);
  a.leafInfo = reinterpret_cast<uint64_t*>(m + L.blasLeafInfo);
  a.child = reinterpret_cast<uint32_t*>(m + L.child);
  a.parent = reinterpret_cast<uint32_t*>(m + L.parent);
  a.counter = reinterpret_cast<uint32_t*>(m + L.counter);
  a.rangeEnd = reinterpret_cast<uint32_t*>(m + L.rangeEnd);
  a.pendingCount = reinterpret_cast<uint32_t*>(m + L.pendingCount);
  a.pendingList = reinterpret_cast<uint32_t*>(m + L.pendingList);
  a.nodes = reinterpret_cast<tc_bvh_node*>(m + L.blasNodes);
  a.nodes4 = reinterpret_cast<tc_bvh4_node*>(m + L.blasNodes4);
  a.nodeBox = reinterpret_cast<tc_aabb*>(m + L.nodeBox);
  a.tlasInst = reinterpret_cast<tc_tlas_instance*>(m + L.tlasInstances);
  a.tlasNodes = reinterpret_cast<tc_bvh_node*>(m + L.tlasNodes);
  a.tlasNodes4 = reinterpret_cast<tc_bvh4_node*>(m + L.tlasNodes4);
  a.tlasNodeBox = reinterpret_cast<tc_aabb*>(m + L.tlasNodeBox);
  a.tlasLeafBox = reinterpret_cast<tc_aabb*>(m + L.tlasLeafBox);
  a.tlasChild = reinterpret_cast<uint32_t*>(m + L.tlasChild);
  a.tlasParent = reinterpret_cast<uint32_t*>(m + L.tlasParent);
  a.tlasCounter = reinterpret_cast<uint32_t*>(m + L.tlasCounter);
  a.tlasRoot = reinterpret_cast<uint32_t*>(m + L.tlasRoot);

  const uint32_t grid = std::min<uint32_t>((capacity + 255) / 256, (uint32_t)num_cu_public() * tuning_snapshot().bvhBlocksPerCU);
  // BLAS: keys, sort by (instance, Morton), hierarchy, refit
  hipLaunchKernelGGL(k_blas_keys, dim3(grid), dim3(256), 0, st, a);
  uint32_t instBits = 1;
  while ((1u << instBits) < a.numInstances) instBits++;
  uint32_t res = 0;
  // digits 0-2: the 24-bit Morton code; digits 4-5: the instance
  radix_sort_digits(sb, &b->build->blasClusterCounter, 0x7u | (instBits > 8 ? 0x30u : 0x10u), st, &res);
  a.sortedKeys = sb.keys[res];
  a.sortedVals = sb.vals[res];
  hipLaunchKernelGGL(k_blas_hierarchy, dim3(grid), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_blas_refit, dim3(std::min<uint32_t>((capacity + kRefitChunk - 1) / kRefitChunk, (uint32_t)num_cu_public() * 4)),
                     dim3(kRefitThreads), 0, st, a);
  hipLaunchKernelGGL(k_blas_refit_top, dim3(a.numInstances), dim3(kRefitTopThreads), 0, st, a);
  // TLAS over instances, with key and value arrays of its own
  a.tlasKeys = reinterpret_cast<uint64_t*>(m + L.tlasKeys);
  a.tlasVals = reinterpret_cast<uint32_t*>(m + L.tlasVals);
  if (a.numInstances <= kTlasRankSort) {
    hipLaunchKernelGGL(k_tlas_build_small, dim3(1), dim3(kTlasRankSort), 0, st, a, instanceXforms);
    return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
  }
  hipLaunchKernelGGL(k_tlas_prep, dim3((a.numInstances + 63) / 64), dim3(64), 0, st, a, instanceXforms);
  hipLaunchKernelGGL(k_tlas_keys, dim3(1), dim3(256), 0, st, a);
  // the TLAS sort's scratch slot is the TLAS child / parent arrays, written
  // only afterwards (32 key bits = 4 passes, so the result lands back in
  // slot 0); it shares the workspace sort state and status words, which
  // hold kMaxInstances keys' tiles
  SortBuffers tb = sb;
  tb.keys[0] = a.tlasKeys;
  tb.vals[0] = a.tlasVals;
  tb.keys[1] = reinterpret_cast<uint64_t*>(m + L.tlasChild);
  tb.vals[1] = reinterpret_cast<uint32_t*>(m + L.tlasParent);
  tb.capacity = a.numInstances;
  hipLaunchKernelGGL(k_tlas_root, dim3(1), dim3(64), 0, st, a);
  uint32_t tres = 0;
  radix_sort_pairs(tb, a.tlasRoot + 1, 32, 32, st, &tres);
  a.tlasSortedKeys = tb.keys[tres];
  a.tlasSortedVals = tb.vals[tres];
  const uint32_t tgrid = (a.numInstances + 255) / 256;
  hipLaunchKernelGGL(k_tlas_hierarchy, dim3(tgrid), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_tlas_refit, dim3(tgrid), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_tlas_collapse4, dim3(tgrid), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

}  // extern "C"
