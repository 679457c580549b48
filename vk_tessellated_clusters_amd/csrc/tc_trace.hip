// tc_trace.hip -- primary / shadow ray traversal and the closest-hit decode.
//
// Stands for vkCmdTraceRaysKHR with render_raytrace.rgen.glsl:52-86 (rays),
// the hardware traversal of the cluster BVH, render_raytrace_clusters.rchit.
// glsl:128-236, 338-339 (decode, depth) and render_shading.glsl:263-274
// (shadow rays).  One ray per lane over the four-wide views of the TLAS and
// the BLAS; the lane's stack (LDS, overflowing into private memory) holds
// TLAS and BLAS entries (the BLAS entries of the instance being walked lie
// above the depth it was entered at).  Triangles use the watertight test of
// Woop, Benthin and Wald (JCGT 2013) with an FP64 recomputation of zero edge
// functions; closest-hit ties are broken by (instance, cluster, primitive)
// so the result does not depend on traversal order.
#include <algorithm>

#include "tc_trace.h"

namespace tc {


struct WoopRay {
  float o[3], d[3], inv[3];
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

// (x, y, z)[k] for a lane-varying k in {0, 1, 2} as two selects of
// registers (a dynamic index into an array would go through private memory)
__device__ __forceinline__ float pick3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ void setup_ray(WoopRay& r, const float o[3], const float dir[3]) {
  float d[3];
  for (int k = 0; k < 3; k++) {
    d[k] = dir[k];
    asm volatile("" : "+v"(d[k]));  // register values: the picks below stay selects
    r.o[k] = o[k], r.d[k] = d[k], r.inv[k] = 1.0f / d[k];
  }
  const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  if (d[r.kz] < 0.0f) {
    const int t = r.kx;
    r.kx = r.ky;
    r.ky = t;
  }
  const float dz = pick3(d[0], d[1], d[2], r.kz);
  r.Sx = pick3(d[0], d[1], d[2], r.kx) / dz;
  r.Sy = pick3(d[0], d[1], d[2], r.ky) / dz;
  r.Sz = 1.0f / dz;
}

// conservative slab test; returns entry distance or +inf
__device__ __forceinline__ float box_hit(const WoopRay& r, const float lo[3], const float hi[3], float tmin, float tmax) {
  float tn = tmin, tf = tmax;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float t0 = (lo[k] - r.o[k]) * r.inv[k];
    const float t1 = (hi[k] - r.o[k]) * r.inv[k];
    tn = fmaxf(tn, fminf(t0, t1));
    tf = fminf(tf, fmaxf(t0, t1));
  }
  return tn <= tf * 1.0000005f ? tn : INFINITY;
}

struct TriHit {
  float t, u, v;
};

// Woop et al. 2013 with barycentrics (u: vertex 1, v: vertex 2)
__device__ __forceinline__ bool tri_hit(const WoopRay& r, const float* p0, const float* p1, const float* p2, bool cullBack,
                                        float tmin, float tmax, TriHit& h) {
  const float A0 = p0[0] - r.o[0], A1 = p0[1] - r.o[1], A2 = p0[2] - r.o[2];
  const float B0 = p1[0] - r.o[0], B1 = p1[1] - r.o[1], B2 = p1[2] - r.o[2];
  const float C0 = p2[0] - r.o[0], C1 = p2[1] - r.o[1], C2 = p2[2] - r.o[2];
  const float Akz = pick3(A0, A1, A2, r.kz), Bkz = pick3(B0, B1, B2, r.kz), Ckz = pick3(C0, C1, C2, r.kz);
  const float Ax = pick3(A0, A1, A2, r.kx) - r.Sx * Akz, Ay = pick3(A0, A1, A2, r.ky) - r.Sy * Akz;
  const float Bx = pick3(B0, B1, B2, r.kx) - r.Sx * Bkz, By = pick3(B0, B1, B2, r.ky) - r.Sy * Bkz;
  const float Cx = pick3(C0, C1, C2, r.kx) - r.Sx * Ckz, Cy = pick3(C0, C1, C2, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = (U + V) + W;
  if (det == 0.0f) return false;
  if (cullBack && det < 0.0f) return false;
  const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  const float T = (U * Az + V * Bz) + W * Cz;
  const float t = T / det;
  if (!(t >= tmin && t <= tmax)) return false;
  h.t = t;
  h.u = V / det;
  h.v = W / det;
  return true;
}

struct HitState {
  tc_hit h;
  const char* clas;  // CLAS block of the best hit
  uint32_t word;     // its triangle word
};

__device__ __forceinline__ bool key_less(float t, uint32_t i, uint32_t c, uint32_t p, const tc_hit& best) {
  if (t != best.t) return t < best.t;
  if (i != best.instanceID) return i < best.instanceID;
  if (c != best.clusterID) return c < best.clusterID;
  return p < best.primitiveID;
}

__device__ __forceinline__ void xform34(const float* m, const float v[3], float w, float out[3]) {
#pragma unroll
  for (int r = 0; r < 3; r++) out[r] = ((m[4 * r] * v[0] + m[4 * r + 1] * v[1]) + m[4 * r + 2] * v[2]) + m[4 * r + 3] * w;
}

__device__ __forceinline__ const float* clas_vertices(const char* clas, uint32_t nt, uint32_t ng) {
  return reinterpret_cast<const float*>(clas + sizeof(tc_clas_header) + ((ng * 24u + 15u) & ~15u) + ((nt * 4u + 15u) & ~15u));
}

// Loads through the global address space: pointers rebuilt from 64-bit
// addresses would otherwise become flat accesses, which also count against
// the LDS counter the stack's ds_* accesses wait on.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// (the HIP vector structs copy through a generic reference: load them as
// native vectors)
__device__ __forceinline__ float4 gload(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const __attribute__((address_space(1))) v4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gload(const uint4* p) {
  typedef uint32_t v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const __attribute__((address_space(1))) v4*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Slab test of a group box given as six floats; 1 if entered.
__device__ __forceinline__ uint32_t group_hit(const WoopRay& r, float l0, float l1, float l2, float h0, float h1, float h2,
                                              float tmin, float tmax) {
  const float lo[3] = {l0, l1, l2}, hi[3] = {h0, h1, h2};
  return box_hit(r, lo, hi, tmin, tmax) != INFINITY ? 1u : 0u;
}

// Group-box hit mask of the pairs [0, npairs) of a block, P pairs (2P
// boxes) per step with the step's loads issued together; pair p holds
// groups 2p and 2p + 1 in three float4.
template <uint32_t P>
__device__ __forceinline__ uint32_t group_mask(const WoopRay& r, const float4* gq, uint32_t npairs, float tmin, float tmax) {
  uint32_t mask = 0;
  for (uint32_t p0 = 0; p0 < npairs; p0 += P) {
    float4 v[P][3];
    uint32_t pp[P];
#pragma unroll
    for (uint32_t j = 0; j < P; j++) {
      pp[j] = min(p0 + j, npairs - 1);
      v[j][0] = gload(gq + 3 * pp[j]), v[j][1] = gload(gq + 3 * pp[j] + 1), v[j][2] = gload(gq + 3 * pp[j] + 2);
    }
#pragma unroll
    for (uint32_t j = 0; j < P; j++) {
      const float4 a0 = v[j][0], a1 = v[j][1], a2 = v[j][2];
      mask |= group_hit(r, a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, tmin, tmax) << (2 * pp[j]);
      mask |= group_hit(r, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w, tmin, tmax) << (2 * pp[j] + 1);
    }
  }
  return mask;
}

// Intersects one CLAS block, whose header's counts and cluster ID come in
// `info` (tc_accel.blasLeafInfo); updates `best` (closest) or returns true
// on the first hit (any-hit).  The loads of a step are issued together so a
// lane waits for memory once per step rather than once per box or
// triangle: group boxes eight at a time (four 48-byte pairs), then the four
// triangles of each group entered (one 16-byte load of their words, then
// their twelve vertices, issued with the next group's words).  Reads past
// the last group or triangle stay inside the block (its sections are
// rounded up to 16 bytes) or are clamped, and their results are masked off.
__device__ __forceinline__ bool clas_hit(const WoopRay& r, const char* clas, uint64_t info, uint32_t instanceID, bool cullBack,
                                         bool anyHit, float tmin, HitState& best) {
  const uint32_t counts = (uint32_t)info, clusterID = (uint32_t)(info >> 32);
  const uint32_t nt = counts & 0xFF, ng = (counts >> 16) & 0xFF;
  const float4* gq = reinterpret_cast<const float4*>(clas + sizeof(tc_clas_header));
  const uint32_t* tris = reinterpret_cast<const uint32_t*>(clas + sizeof(tc_clas_header) + ((ng * 24u + 15u) & ~15u));
  const float* verts = clas_vertices(clas, nt, ng);
  // group boxes (ng <= 32): eight per step, or two per step when every
  // lane of the wave holds a block of at most two groups (most tessellated
  // parts), which would otherwise test six clamped repeats
  const uint32_t npairs = (ng + 1) >> 1;
  uint32_t mask = __all(npairs <= 1) ? group_mask<1>(r, gq, npairs, tmin, best.h.t)
                                     : group_mask<4>(r, gq, npairs, tmin, best.h.t);
  if (ng < 32) mask &= (1u << ng) - 1u;
  // then the triangles of the groups entered, so the lanes of a wave run
  // the triangle loop together
  if (!mask) return false;
  uint32_t g = __builtin_ctz(mask);
  mask &= mask - 1;
  uint4 w4 = gload(reinterpret_cast<const uint4*>(tris + g * TC_CLAS_GROUP_TRIANGLES));
  while (true) {
    const uint32_t t0 = g * TC_CLAS_GROUP_TRIANGLES;
    uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
    // the next group's words load while this group's vertices do
    const uint32_t gn = mask ? __builtin_ctz(mask) : g;
    w4 = gload(reinterpret_cast<const uint4*>(tris + gn * TC_CLAS_GROUP_TRIANGLES));
    float v[4][9];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (t0 + j >= nt) w[j] = 0;  // vertex 0: an in-bounds read, result unused
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float* pk = verts + 3 * ((w[j] >> (8 * k)) & 0xFF);
        v[j][3 * k] = gload(pk), v[j][3 * k + 1] = gload(pk + 1), v[j][3 * k + 2] = gload(pk + 2);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (t0 + j >= nt) break;
      TriHit th;
      if (!tri_hit(r, v[j], v[j] + 3, v[j] + 6, cullBack, tmin, best.h.t, th)) continue;
      const uint32_t prim = w[j] >> 24;
      if (anyHit) {
        best.h.t = th.t;
        return true;
      }
      if (key_less(th.t, instanceID, clusterID, prim, best.h)) {
        best.h.t = th.t;
        best.h.u = th.u;
        best.h.v = th.v;
        best.h.instanceID = instanceID;
        best.h.clusterID = clusterID;
        best.h.primitiveID = prim;
        best.clas = clas;
        best.word = w[j];
      }
    }
    if (!mask) return false;
    g = gn;
    mask &= mask - 1;
  }
}

// Traversal stack of one lane: the first kLdsStack entries in LDS (stride
// T, the workgroup size), the rest in a private array.  TLAS and BLAS
// entries share it: the BLAS entries of the instance being walked sit above
// the stack depth at which it was entered.  Bound: every visited node
// pushes at most 3 of its 4 slots, and the four-wide view visits every
// other level, so a ray holds at most 3 * (ceil(dT / 2) + ceil(dB / 2))
// entries for binary depths dT (TLAS) and dB (BLAS).  A Karras tree's depth
// is at most its distinct key bits plus the index bits that break ties.
// TLAS keys are a 30-bit Morton code above the instance index, so dT <=
// 30 + 16 = 46 (2^16 instances); a BLAS covers one instance's segment,
// whose keys differ only in the 24-bit Morton code, so dB <= 24 + 23 = 47
// (2^23 references).  That is at most 3 * (23 + 24) = 141 < kStackDepth.
constexpr uint32_t kLdsStack = 24;
constexpr uint32_t kStackDepth = 160;
constexpr uint32_t kDone = 0xFFFFFFFFu;

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(5))) uint32_t private_u32;

// Holds no array itself, so the stack pointer lives in a register; the
// typed pointers make every access a ds_* or scratch_* instruction.
template <uint32_t T>
struct RayStack {
  lds_u32* lds;         // this lane's first entry, stride T
  private_u32* spill;   // kStackDepth private entries
  uint32_t ldsDepth;    // kLdsStack, or 0 with TC_TRACE_PRIVATE_STACK
  uint32_t sp;
  __device__ __forceinline__ void push(uint32_t e) {
    if (sp < ldsDepth)
      lds[sp * T] = e;
    else if (sp < kStackDepth)
      spill[sp - ldsDepth] = e;
    sp++;
  }
  __device__ __forceinline__ uint32_t pop() {
    --sp;
    return sp < ldsDepth ? lds[sp * T] : spill[sp - ldsDepth];
  }
  // next entry above stack depth `base`, or kDone
  __device__ __forceinline__ uint32_t pop_above(uint32_t base) { return sp > base ? pop() : kDone; }
  // pushes c[3], c[2], c[1] where t[k] < inf, with t ascending (the misses
  // are a suffix).  With room in LDS every entry is written and only the
  // pointer advance depends on the hit, so the wave runs no branch.
  __device__ __forceinline__ void push3(const uint32_t c[4], const float t[4]) {
    if (sp + 3 <= ldsDepth) {
      lds[sp * T] = c[3];
      sp += t[3] != INFINITY;
      lds[sp * T] = c[2];
      sp += t[2] != INFINITY;
      lds[sp * T] = c[1];
      sp += t[1] != INFINITY;
    } else {
      if (t[3] != INFINITY) push(c[3]);
      if (t[2] != INFINITY) push(c[2]);
      if (t[1] != INFINITY) push(c[1]);
    }
  }
};

__device__ __forceinline__ void load_node4(const tc_bvh4_node* p, float lo[3][4], float hi[3][4], uint32_t c[4]) {
  const float4* q = reinterpret_cast<const float4*>(p);
  float4 v[7];
#pragma unroll
  for (int i = 0; i < 7; i++) v[i] = q[i];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lo[a][0] = v[a].x, lo[a][1] = v[a].y, lo[a][2] = v[a].z, lo[a][3] = v[a].w;
    hi[a][0] = v[3 + a].x, hi[a][1] = v[3 + a].y, hi[a][2] = v[3 + a].z, hi[a][3] = v[3 + a].w;
  }
  c[0] = __float_as_uint(v[6].x), c[1] = __float_as_uint(v[6].y), c[2] = __float_as_uint(v[6].z), c[3] = __float_as_uint(v[6].w);
}

__device__ __forceinline__ void cswap(float t[4], uint32_t c[4], int i, int j) {
  if (t[j] < t[i]) {
    const float tt = t[i];
    t[i] = t[j], t[j] = tt;
    const uint32_t cc = c[i];
    c[i] = c[j], c[j] = cc;
  }
}

// Entry distances of the four slots of a node (+inf: missed or empty).
// All four slab tests run unconditionally and the empty slots are selected
// away afterwards, so the wave runs them as straight-line code.
__device__ __forceinline__ void node_slots(const tc_bvh4_node* node, const WoopRay& r, float tmin, float tmax, float t[4],
                                           uint32_t c[4]) {
  float lo[3][4], hi[3][4];
  load_node4(node, lo, hi, c);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float bl[3] = {lo[0][k], lo[1][k], lo[2][k]}, bh[3] = {hi[0][k], hi[1][k], hi[2][k]};
    float bt = box_hit(r, bl, bh, tmin, tmax);
    asm volatile("" : "+v"(bt));  // keep the test out of a branch on the slot
    t[k] = c[k] == TC_BVH4_EMPTY ? INFINITY : bt;
  }
}

// Visits one node: pushes the hit slots but the nearest (farthest first)
// and returns the nearest one's child, or the next entry above `base`
// when no slot is hit.
template <uint32_t T>
__device__ __forceinline__ uint32_t visit_node(const tc_bvh4_node* node, const WoopRay& r, float tmin, float tmax,
                                               RayStack<T>& stack, uint32_t base) {
  float t[4];
  uint32_t c[4];
  node_slots(node, r, tmin, tmax, t, c);
  cswap(t, c, 0, 1);
  cswap(t, c, 2, 3);
  cswap(t, c, 0, 2);
  cswap(t, c, 1, 3);
  cswap(t, c, 1, 2);
  if (t[0] == INFINITY) return stack.pop_above(base);
  stack.push3(c, t);
  return c[0];
}

// Full two-level traversal of one ray over the four-wide views.  Returns
// true if anything was hit.  The TLAS walk stops at each instance whose
// BLAS has clusters; that BLAS is walked with the object-space ray until
// the stack is back at the depth it was entered at.
template <uint32_t T>
__device__ bool trace_ray(const TraceArgs& a, const float o[3], const float d[3], float tmin, float tmax, bool cullBack,
                          bool anyHit, RayStack<T>& stack, HitState& best) {
  best.h.t = tmax;
  best.h.u = best.h.v = 0.0f;
  best.h.instanceID = best.h.clusterID = best.h.primitiveID = 0xFFFFFFFFu;
  best.clas = nullptr;
  best.word = 0;
  WoopRay world;
  setup_ray(world, o, d);
  stack.sp = 0;
  uint32_t e = *a.tlasRoot;
  while (true) {
    // TLAS: find the next instance leaf with clusters
    uint32_t inst = kDone;
    while (e != kDone) {
      if (e & TC_BVH_LEAF) {
        const uint32_t i = e & 0x3FFFFFFFu;
        if (a.tlasInstances[i].numClusters != 0) {
          inst = i;
          break;
        }
        e = stack.pop_above(0);
      } else {
        e = visit_node(a.tlasNodes4 + e, world, tmin, best.h.t, stack, 0);
      }
    }
    if (inst == kDone) break;
    const tc_tlas_instance& ti = a.tlasInstances[inst];
    WoopRay ray;
    {
      float ol[3], dl[3];
      xform34(ti.worldToObject, o, 1.0f, ol);
      xform34(ti.worldToObject, d, 0.0f, dl);
      setup_ray(ray, ol, dl);
    }
    const tc_bvh4_node* blas4 = a.blasNodes4 + (ti.blasNodes - a.blasNodesBase) / sizeof(tc_bvh_node);
    const uint64_t* blasLeaves = reinterpret_cast<const uint64_t*>(ti.blasLeaves);
    const uint64_t* leafInfo = a.blasLeafInfo + (ti.blasLeaves - a.blasLeavesBase) / 8;
    const uint32_t base = stack.sp;
    uint32_t b = ti.blasRoot;
    // while-while: every lane walks nodes until it holds a CLAS leaf (or has
    // finished), then the lanes with leaves intersect them together
    while (b != kDone) {
      while (b != kDone && !(b & TC_BVH_LEAF)) b = visit_node(blas4 + b, ray, tmin, best.h.t, stack, base);
      if (b == kDone) break;
      const char* clas = reinterpret_cast<const char*>(gload(blasLeaves + (b & ~TC_BVH_LEAF)));
      if (clas_hit(ray, clas, gload(leafInfo + (b & ~TC_BVH_LEAF)), inst, cullBack, anyHit, tmin, best)) return true;
      b = stack.pop_above(base);
    }
    e = stack.pop_above(0);
  }
  return best.h.instanceID != 0xFFFFFFFFu;
}

// Wave-wide packet traversal: the 64 lanes of a wave walk one shared node
// stack (in LDS, uniform entries).  A node is entered when any live lane's
// ray enters it before that lane's current closest hit, so every lane sees
// a superset of the nodes its own traversal would visit and finds the same
// closest hit (ties are broken by key).  Node records are fetched once per
// wave.  Suited to coherent rays: camera rays in 8 x 8 pixel tiles and their
// shadow rays.
constexpr uint32_t kPacketStack = 160;
constexpr uint32_t kBlasEntry = 1u << 30;  // packet stack: marks BLAS entries

__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v = fminf(v, __shfl_xor(v, d, kWave));
  return v;
}

__device__ bool trace_packet(const TraceArgs& a, const float o[3], const float d[3], float tmin, float tmax, bool live,
                             bool cullBack, bool anyHit, uint32_t* stk, HitState& best) {
  best.h.t = tmax;
  best.h.u = best.h.v = 0.0f;
  best.h.instanceID = best.h.clusterID = best.h.primitiveID = 0xFFFFFFFFu;
  best.clas = nullptr;
  best.word = 0;
  WoopRay world, ray;
  setup_ray(world, o, d);
  ray = world;
  uint32_t sp = 0;
  uint32_t e = __builtin_amdgcn_readfirstlane(*a.tlasRoot);
  uint32_t inst = 0xFFFFFFFFu;
  const tc_bvh4_node* blas4 = nullptr;
  const uint64_t* blasLeaves = nullptr;
  const uint64_t* leafInfo = nullptr;
  bool found = false;
  while (true) {
    if (!ballot(live)) break;
    bool next = false;  // pop the next entry
    if (!(e & kBlasEntry)) {
      if (inst != 0xFFFFFFFFu) {
        ray = world;
        inst = 0xFFFFFFFFu;
      }
      if (e & TC_BVH_LEAF) {
        const uint32_t i = e & 0x3FFFFFFFu;
        const tc_tlas_instance& ti = a.tlasInstances[i];
        if (ti.numClusters != 0) {
          float ol[3], dl[3];
          xform34(ti.worldToObject, o, 1.0f, ol);
          xform34(ti.worldToObject, d, 0.0f, dl);
          setup_ray(ray, ol, dl);
          inst = i;
          blas4 = a.blasNodes4 + (ti.blasNodes - a.blasNodesBase) / sizeof(tc_bvh_node);
          blasLeaves = reinterpret_cast<const uint64_t*>(ti.blasLeaves);
          leafInfo = a.blasLeafInfo + (ti.blasLeaves - a.blasLeavesBase) / 8;
          e = __builtin_amdgcn_readfirstlane(kBlasEntry | ti.blasRoot);
          continue;
        }
        next = true;
      }
    } else if (e & TC_BVH_LEAF) {
      const uint32_t leaf = e & ~(kBlasEntry | TC_BVH_LEAF);
      const char* clas = reinterpret_cast<const char*>(blasLeaves[leaf]);
      if (live && clas_hit(ray, clas, leafInfo[leaf], inst, cullBack, anyHit, tmin, best)) {
        found = true;
        live = false;  // any-hit: this lane is done
      }
      next = true;
    }
    if (!next) {
      const tc_bvh4_node* node = (e & kBlasEntry) ? blas4 + (e & ~kBlasEntry) : a.tlasNodes4 + e;
      const uint32_t tag = e & kBlasEntry;
      float lo[3][4], hi[3][4], t[4], tw[4];
      uint32_t c[4];
      load_node4(node, lo, hi, c);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float bl[3] = {lo[0][k], lo[1][k], lo[2][k]}, bh[3] = {hi[0][k], hi[1][k], hi[2][k]};
        float bt = box_hit(ray, bl, bh, tmin, best.h.t);
        asm volatile("" : "+v"(bt));
        t[k] = (c[k] == TC_BVH4_EMPTY || !live) ? INFINITY : bt;
        tw[k] = wave_min_f(t[k]);  // nearest entry over the packet (uniform)
        c[k] |= tag;
      }
      cswap(tw, c, 0, 1);
      cswap(tw, c, 2, 3);
      cswap(tw, c, 0, 2);
      cswap(tw, c, 1, 3);
      cswap(tw, c, 1, 2);
      if (tw[0] == INFINITY) {
        next = true;
      } else {
        if (tw[3] != INFINITY && sp < kPacketStack) stk[sp++] = c[3];
        if (tw[2] != INFINITY && sp < kPacketStack) stk[sp++] = c[2];
        if (tw[1] != INFINITY && sp < kPacketStack) stk[sp++] = c[1];
        e = __builtin_amdgcn_readfirstlane(c[0]);
      }
    }
    if (next) {
      if (sp == 0) break;
      e = __builtin_amdgcn_readfirstlane(stk[--sp]);
    }
  }
  return anyHit ? found : best.h.instanceID != 0xFFFFFFFFu;
}

__device__ __forceinline__ void mul44(const float* m, const float v[4], float out[4]) {
#pragma unroll
  for (int r = 0; r < 4; r++) out[r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}

__device__ __forceinline__ void normalize4(float v[4]) {
  const float l = sqrtf(((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
  const float s = 1.0f / l;
  for (int k = 0; k < 4; k++) v[k] = v[k] * s;
}

// Camera rays come in 8 x 8 pixel tiles, one per wave.  Consecutive
// workgroups go to the eight XCDs in turn, so wave w of a launch runs on XCD
// (w / wavesPerGroup) % 8.  The image is cut into eight vertical strips of
// tiles and the waves of XCD x walk strip x in row order, so each XCD's L2
// holds the geometry seen through one strip.  Strips differ in width by at
// most one tile; a wave past the end of its strip has no tile (false).  With
// `rowOrder` the strip's rows are walked in that order instead (a
// permutation of the tile rows, TraceOrder).
constexpr uint32_t kXcds = 8;

// A wave's tile time into its row's maximum (no return value: the wave does
// not wait for it).
__device__ __forceinline__ void time_tile_row(TraceOrder* ord, uint32_t row, uint64_t t0) {
  const uint64_t dt = __builtin_amdgcn_s_memrealtime() - t0;
  __hip_atomic_fetch_max(&ord->rowTime[row], (uint32_t)min(dt, (uint64_t)0xFFFFFFFFu), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

__host__ __device__ __forceinline__ uint32_t strip_waves(uint32_t tilesX, uint32_t tilesY, uint32_t wavesPerGroup) {
  const uint32_t perXcd = (tilesX + kXcds - 1) / kXcds * tilesY;
  return kXcds * ((perXcd + wavesPerGroup - 1) / wavesPerGroup) * wavesPerGroup;
}

__device__ __forceinline__ bool strip_tile(uint32_t w, uint32_t wavesPerGroup, uint32_t tilesX, uint32_t tilesY,
                                           const uint32_t* rowOrder, uint32_t& tx, uint32_t& ty) {
  const uint32_t b = w / wavesPerGroup, x = b % kXcds;
  const uint32_t k = (b / kXcds) * wavesPerGroup + w % wavesPerGroup;
  const uint32_t c0 = x * tilesX / kXcds, sw = (x + 1) * tilesX / kXcds - c0;
  if (sw == 0 || k / sw >= tilesY) return false;
  tx = c0 + k % sw;
  ty = rowOrder ? rowOrder[uniform(k / sw)] : k / sw;
  return true;
}


// render_raytrace.rgen.glsl:52-86
__device__ __forceinline__ void primary_ray(const tc_frame_constants* view, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                                            float o[3], float d[3]) {
  const float uvx = ((float)x + 0.5f) / (float)W, uvy = ((float)y + 0.5f) / (float)H;
  const float dx = uvx * 2.0f - 1.0f, dy = uvy * 2.0f - 1.0f;
  const float org[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  float origin[4];
  mul44(view->viewMatrixI, org, origin);
  const float ndc[4] = {dx, dy, 1.0f, 1.0f};
  float target[4];
  mul44(view->projMatrixI, ndc, target);
  normalize4(target);
  const float tv[4] = {target[0], target[1], target[2], 0.0f};
  float dir[4];
  mul44(view->viewMatrixI, tv, dir);
  normalize4(dir);
  for (int k = 0; k < 3; k++) o[k] = origin[k], d[k] = dir[k];
}

// render_raytrace_clusters.rchit.glsl:128-236 and 338-339: base cluster,
// triangle and barycentrics, world position and depth of a hit
__device__ void decode_hit(const TraceArgs& a, const HitState& hs, tc_hit_decoded& out) {
  const tc_hit& h = hs.h;
  out.instanceID = h.instanceID;
  out.shadowed = 0;
  if (h.instanceID == 0xFFFFFFFFu) {
    out.clusterID = out.triangleID = out.subTriangleID = 0xFFFFFFFFu;
    out.baryBase[0] = out.baryBase[1] = out.baryBase[2] = 0.0f;
    out.depth = 1.0f;
    out.wPos[0] = out.wPos[1] = out.wPos[2] = 0.0f;
    return;
  }
  const uint32_t mode = h.clusterID >> 30;
  uint32_t clusterID = h.clusterID & 0x3FFFFFFFu;
  uint32_t triangleID = h.primitiveID;
  uint32_t subTriangle = h.primitiveID;
  const v3 bw = mk((1.0f - h.u) - h.v, h.u, h.v);
  v3 base = bw;
  if (mode == TC_RT_CLUSTER_MODE_1X_SUBSET_CLUSTER) {
    // :173-178 the part slot holds the ClusterInfo, then one byte per
    // transient triangle naming the cluster triangle
    const uint8_t* slot = ptr<const uint8_t>(a.build->partTriangles) + (uint64_t)clusterID * 24ull;
    const tc_cluster_info ci = *reinterpret_cast<const tc_cluster_info*>(slot);
    triangleID = slot[8 + h.primitiveID];
    clusterID = ci.clusterID;
  } else if (mode == TC_RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED) {
    // :152-170 a u16 mapping per batch triangle after the ClusterInfo: the
    // cluster triangle, the sub-triangle of its config (masked with 0xF; the
    // reference's `& 4` drops sub-triangles 1-3) and the (factor - 1) bits
    // the config is rebuilt from
    const uint8_t* slot = ptr<const uint8_t>(a.build->partTriangles) + (uint64_t)clusterID * 24ull;
    const tc_cluster_info ci = *reinterpret_cast<const tc_cluster_info*>(slot);
    const uint32_t packed = reinterpret_cast<const uint16_t*>(slot + 8)[h.primitiveID];
    triangleID = packed & 0xFFu;
    subTriangle = (packed >> 8) & 0xFu;
    uint32_t f[3] = {1u + ((packed >> 12) & 1u), 1u + ((packed >> 13) & 1u), 1u + (packed >> 14)};
    uint32_t vtx[3] = {0u, 32768u, 32768u << 16};
    const uint32_t cfg = get_config(f, vtx);
    clusterID = ci.clusterID;
    const TableView T = load_table(a.tessTable);
    const tc_tess_table_entry e = cfg_entry(T, cfg);
    uint32_t idx[3];
    cfg_tri_vertices(T, cfg, e.firstTriangle, subTriangle, idx);
    const v3 nb = interp(cfg_vertex_bary(T, cfg, e.firstVertex, idx[0]), cfg_vertex_bary(T, cfg, e.firstVertex, idx[1]),
                         cfg_vertex_bary(T, cfg, e.firstVertex, idx[2]), bw);
    base = interp(decode_bary(vtx[0]), decode_bary(vtx[1]), decode_bary(vtx[2]), nb);
  } else if (mode != TC_RT_CLUSTER_MODE_FULL_CLUSTER) {
    const tc_tess_triangle_info ti = ptr<const tc_tess_triangle_info>(a.build->partTriangles)[clusterID];
    triangleID = ti.subTriangle.triangleID_config & 0xFFFF;
    const uint32_t cfg = ti.subTriangle.triangleID_config >> 16;
    clusterID = ti.cluster.clusterID;
    const TableView T = load_table(a.tessTable);
    const tc_tess_table_entry e = cfg_entry(T, cfg);
    uint32_t idx[3];
    cfg_tri_vertices(T, cfg, e.firstTriangle, h.primitiveID, idx);
    const v3 nb = interp(cfg_vertex_bary(T, cfg, e.firstVertex, idx[0]), cfg_vertex_bary(T, cfg, e.firstVertex, idx[1]),
                         cfg_vertex_bary(T, cfg, e.firstVertex, idx[2]), bw);
    base = interp(decode_bary(ti.subTriangle.vtxEncoded[0]), decode_bary(ti.subTriangle.vtxEncoded[1]),
                  decode_bary(ti.subTriangle.vtxEncoded[2]), nb);
  }
  out.clusterID = clusterID;
  out.triangleID = triangleID;
  out.subTriangleID = subTriangle;
  out.baryBase[0] = base.x, out.baryBase[1] = base.y, out.baryBase[2] = base.z;
  // :248-249 object-space hit point from the hit triangle, then world
  const tc_clas_header hdr = *reinterpret_cast<const tc_clas_header*>(hs.clas);
  const float* verts = clas_vertices(hs.clas, hdr.counts & 0xFF, (hdr.counts >> 16) & 0xFF);
  const float* p0 = verts + 3 * (hs.word & 0xFF);
  const float* p1 = verts + 3 * ((hs.word >> 8) & 0xFF);
  const float* p2 = verts + 3 * ((hs.word >> 16) & 0xFF);
  const v3 op = interp(mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), mk(p2[0], p2[1], p2[2]), bw);
  const float opv[3] = {op.x, op.y, op.z};
  float wp[3];
  xform34(a.tlasInstances[h.instanceID].objectToWorld, opv, 1.0f, wp);
  out.wPos[0] = wp[0], out.wPos[1] = wp[1], out.wPos[2] = wp[2];
  const float w4[4] = {wp[0], wp[1], wp[2], 1.0f};
  float pp[4];
  mul44(a.view->viewProjMatrix, w4, pp);
  out.depth = pp[2] / pp[3];
}

// Four waves per SIMD (at most 128 VGPRs): the batched loads of the CLAS
// test need the registers, and a frame's time is set by the longest waves
// (rays with many leaves), which gain more from fewer round trips than from
// a fifth wave.
template <uint32_t T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_trace(TraceArgs a) {
  __shared__ uint32_t s_stack[kLdsStack * T];
  uint32_t spill[kStackDepth];
  RayStack<T> stack;
  stack.lds = (lds_u32*)(s_stack + threadIdx.x);
  stack.spill = (private_u32*)spill;
  stack.ldsDepth = (a.flags & TC_TRACE_PRIVATE_STACK) ? 0u : kLdsStack;
  const bool cull = (a.flags & TC_TRACE_CULL_BACK) != 0;
  const bool shadows = (a.flags & TC_TRACE_SHADOW) != 0 && a.decoded;
  const uint32_t tilesX = (a.width + 7) / 8, tilesY = (a.height + 7) / 8;
  // tile-row order: the last ranking when it ranked this many rows.  Each
  // wave's tile time goes to its row, timed from the top of one iteration
  // to the top of the next (where the wave has reconverged), by an atomic
  // nobody waits for.
  TraceOrder* ord = !a.rays && a.order && tilesY <= kMaxOrderRows ? a.order : nullptr;
  const uint32_t* rowOrder = ord && ord->rows == tilesY ? ord->order : nullptr;
  if (ord && blockIdx.x == 0 && threadIdx.x == 0) ord->pendingRows = tilesY;
  constexpr uint32_t kNoRow = 0xFFFFFFFFu;
  uint32_t timedRow = kNoRow;
  uint64_t t0 = 0;
  for (uint32_t item = blockIdx.x * T + threadIdx.x; item < a.numItems; item += gridDim.x * T) {
    if (ord) {
      if (timedRow != kNoRow && lane_id() == 0)
        time_tile_row(ord, timedRow, t0);
      timedRow = kNoRow;
      t0 = __builtin_amdgcn_s_memrealtime();
    }
    float o[3], d[3], tmin, tmax;
    uint32_t r = item;
    if (a.rays) {
      const tc_ray ray = a.rays[r];
      for (int k = 0; k < 3; k++) o[k] = ray.origin[k], d[k] = ray.direction[k];
      tmin = ray.tMin;
      tmax = ray.tMax;
    } else {
      uint32_t tx, ty;
      if (!strip_tile(item >> 6, T / kWave, tilesX, tilesY, rowOrder, tx, ty)) continue;
      timedRow = ty;
      const uint32_t in = item & 63, x = tx * 8 + (in & 7), y = ty * 8 + (in >> 3);
      if (x >= a.width || y >= a.height) continue;
      r = y * a.width + x;
      primary_ray(a.view, x, y, a.width, a.height, o, d);
      tmin = a.view->nearPlane;
      tmax = a.view->farPlane;
    }
    HitState hs;
    trace_ray(a, o, d, tmin, tmax, cull, false, stack, hs);
    tc_hit hit = hs.h;
    if (hit.instanceID == 0xFFFFFFFFu) hit.t = -1.0f;
    if (a.hits) a.hits[r] = hit;
    if (a.hitVertices) {
      uint4 hv = make_uint4(~0u, ~0u, ~0u, ~0u);
      if (hs.clas) {
        const uint32_t base = reinterpret_cast<const tc_clas_header*>(hs.clas)->vertexBase;
        hv = make_uint4(base + (hs.word & 0xFF), base + ((hs.word >> 8) & 0xFF), base + ((hs.word >> 16) & 0xFF), 0u);
      }
      reinterpret_cast<uint4*>(a.hitVertices)[r] = hv;
    }
    if (!a.decoded) continue;
    tc_hit_decoded dec;
    decode_hit(a, hs, dec);
    if (shadows && hit.instanceID != 0xFFFFFFFFu) {
      // render_shading.glsl:263-274: opaque, terminate on first hit
      const float* sun = a.view->skyParams.sunDirection;
      const float sd[3] = {sun[0], sun[1], sun[2]};
      HitState sh;
      dec.shadowed = trace_ray(a, dec.wPos, sd, 0.001f, 10000000.0f, false, true, stack, sh) ? 1u : 0u;
    }
    a.decoded[r] = dec;
  }
  if (ord && timedRow != kNoRow && lane_id() == 0) time_tile_row(ord, timedRow, t0);
}

// Packet form of k_trace: one wave per 64 work items (an 8 x 8 pixel tile
// for camera rays), shared stack per wave.
template <uint32_t T>
__global__ __launch_bounds__(T) void k_trace_packet(TraceArgs a) {
  __shared__ uint32_t s_stk[T / kWave][kPacketStack];
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  uint32_t* stk = s_stk[wave];
  const bool cull = (a.flags & TC_TRACE_CULL_BACK) != 0;
  const bool shadows = (a.flags & TC_TRACE_SHADOW) != 0 && a.decoded;
  const uint32_t tilesX = (a.width + 7) / 8, tilesY = (a.height + 7) / 8;
  const uint32_t wavesTotal = gridDim.x * (T / kWave);
  for (uint32_t base = (blockIdx.x * (T / kWave) + wave) * kWave; base < a.numItems; base += wavesTotal * kWave) {
    const uint32_t item = base + lane;
    float o[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 1.f}, tmin = 0.f, tmax = 0.f;
    uint32_t r = 0xFFFFFFFFu;
    if (item < a.numItems) {
      if (a.rays) {
        r = item;
        const tc_ray ray = a.rays[r];
        for (int k = 0; k < 3; k++) o[k] = ray.origin[k], d[k] = ray.direction[k];
        tmin = ray.tMin;
        tmax = ray.tMax;
      } else {
        uint32_t tx = 0, ty = 0;
        const bool tiled = strip_tile(item >> 6, T / kWave, tilesX, tilesY, nullptr, tx, ty);
        const uint32_t in = item & 63, x = tx * 8 + (in & 7), y = ty * 8 + (in >> 3);
        if (tiled && x < a.width && y < a.height) {
          r = y * a.width + x;
          primary_ray(a.view, x, y, a.width, a.height, o, d);
          tmin = a.view->nearPlane;
          tmax = a.view->farPlane;
        }
      }
    }
    const bool live = r != 0xFFFFFFFFu;
    HitState hs;
    trace_packet(a, o, d, tmin, tmax, live, cull, false, stk, hs);
    tc_hit hit = hs.h;
    if (hit.instanceID == 0xFFFFFFFFu) hit.t = -1.0f;
    tc_hit_decoded dec;
    if (live) {
      if (a.hits) a.hits[r] = hit;
      if (a.hitVertices) {
        uint4 hv = make_uint4(~0u, ~0u, ~0u, ~0u);
        if (hs.clas) {
          const uint32_t vb = reinterpret_cast<const tc_clas_header*>(hs.clas)->vertexBase;
          hv = make_uint4(vb + (hs.word & 0xFF), vb + ((hs.word >> 8) & 0xFF), vb + ((hs.word >> 16) & 0xFF), 0u);
        }
        reinterpret_cast<uint4*>(a.hitVertices)[r] = hv;
      }
      if (a.decoded) decode_hit(a, hs, dec);
    }
    if (shadows) {
      // render_shading.glsl:263-274: opaque, terminate on first hit
      const bool sLive = live && hit.instanceID != 0xFFFFFFFFu;
      const float* sun = a.view->skyParams.sunDirection;
      const float sd[3] = {sun[0], sun[1], sun[2]};
      const float so[3] = {sLive ? dec.wPos[0] : 0.f, sLive ? dec.wPos[1] : 0.f, sLive ? dec.wPos[2] : 0.f};
      HitState sh;
      const bool shadowed = trace_packet(a, so, sd, 0.001f, 10000000.0f, sLive, false, true, stk, sh);
      if (sLive) dec.shadowed = shadowed ? 1u : 0u;
    }
    if (live && a.decoded) a.decoded[r] = dec;
  }
}

}  // namespace tc

using namespace tc;

tc_status tc::trace_rays(const tc_bindings* b, tc_config config, const tc_accel* accel, const tc_trace_params* p,
                         TraceOrder* order, tc_stream stream) {
  KArgs ka;
  tc_status s = make_args_public(b, config, ka);
  if (s != TC_OK) return s;
  if (!accel || !p || !accel->tlasInstances || !accel->blasLeafInfo || (!p->hits && !p->decoded && !p->hitVertices))
    return TC_ERROR_INVALID_ARGUMENT;
  TraceArgs a;
  a.view = b->view;
  a.tessTable = b->tessTable;
  a.build = b->build;
  a.tlasInstances = accel->tlasInstances;
  a.tlasNodes = accel->tlasNodes;
  a.tlasRoot = accel->tlasRoot;
  a.rays = p->rays;
  a.width = p->width;
  a.height = p->height;
  a.numRays = p->rays ? p->numRays : p->width * p->height;
  a.tlasNodes4 = accel->tlasNodes4;
  a.blasNodes4 = accel->blasNodes4;
  a.blasNodesBase = reinterpret_cast<uint64_t>(accel->blasNodes);
  a.blasLeavesBase = reinterpret_cast<uint64_t>(accel->blasLeaves);
  a.blasLeafInfo = accel->blasLeafInfo;
  a.hits = p->hits;
  a.decoded = p->decoded;
  a.hitVertices = p->hitVertices;
  a.order = order;
  a.flags = p->flags;
  if (!p->rays && (!p->width || !p->height)) return TC_ERROR_INVALID_ARGUMENT;
  if (a.numRays == 0) return TC_OK;
  const tc_tuning tn = tuning_snapshot();
  const uint32_t T = tn.traceThreads;
  a.numItems = p->rays ? p->numRays : strip_waves((p->width + 7) / 8, (p->height + 7) / 8, T / kWave) * kWave;
  const uint32_t grid = std::min<uint32_t>((a.numItems + T - 1) / T, (uint32_t)num_cu_public() * tn.traceBlocksPerCU);
  // resident waves: four per SIMD (128 VGPRs), four SIMDs per CU
  if ((uint64_t)grid * (T / kWave) > (uint64_t)kOrderRounds * num_cu_public() * 16u) a.order = nullptr;
  if (a.flags & TC_TRACE_PACKETS) {
    if (T == 64)
      hipLaunchKernelGGL(k_trace_packet<64>, dim3(grid), dim3(64), 0, (hipStream_t)stream, a);
    else if (T == 256)
      hipLaunchKernelGGL(k_trace_packet<256>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
    else
      hipLaunchKernelGGL(k_trace_packet<128>, dim3(grid), dim3(128), 0, (hipStream_t)stream, a);
    return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
  }
  if (T == 64)
    hipLaunchKernelGGL(k_trace<64>, dim3(grid), dim3(64), 0, (hipStream_t)stream, a);
  else if (T == 256)
    hipLaunchKernelGGL(k_trace<256>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(k_trace<128>, dim3(grid), dim3(128), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

extern "C" tc_status tc_trace_rays(const tc_bindings* b, tc_config config, const tc_accel* accel, const tc_trace_params* p,
                                   tc_stream stream) {
  return trace_rays(b, config, accel, p, nullptr, stream);
}
