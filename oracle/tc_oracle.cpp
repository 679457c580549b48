/*
 * tc_oracle.cpp -- serial CPU restatement of the reference's frame stages.
 *
 * TEST INFRASTRUCTURE ONLY (see tc_oracle.h).  Each function cites the GLSL
 * or C++ lines of nvpro-samples/vk_tessellated_clusters it restates.
 */
#include "tc_oracle.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace {

#include "../vk_tessellated_clusters_amd/csrc/tess_table_nv_data.inc"

/* ------------------------------------------------------------ small math */

struct v3 {
  float x, y, z;
};
struct v2 {
  float x, y;
};

inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length(v3 a) { return sqrtf(dot(a, a)); }
inline float distance(v3 a, v3 b) { return length(sub(a, b)); }
inline v3 normalize(v3 a) { return mul(a, 1.0f / length(a)); }

/* glm column-major mat4 * vec4(p, 1), summed column by column */
inline v3 xform_point(const float* m, v3 p) {
  v3 r;
  r.x = ((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12];
  r.y = ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13];
  r.z = ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14];
  return r;
}

/* tessellation.glsl:191-201 (precise) */
inline v3 interp(const v3 b[3], v3 w) {
  return mk((b[0].x * w.x + b[1].x * w.y) + b[2].x * w.z, (b[0].y * w.x + b[1].y * w.y) + b[2].y * w.z,
            (b[0].z * w.x + b[1].z * w.y) + b[2].z * w.z);
}
inline v2 interp2(const v2 b[3], v3 w) {
  return v2{(b[0].x * w.x + b[1].x * w.y) + b[2].x * w.z, (b[0].y * w.x + b[1].y * w.y) + b[2].y * w.z};
}

template <class T>
inline T* P(uint64_t a) {
  return reinterpret_cast<T*>(static_cast<uintptr_t>(a));
}

inline v3 ld3(uint64_t base, uint32_t i) {
  const float* f = P<const float>(base) + 3ull * i;
  return mk(f[0], f[1], f[2]);
}
inline v2 ld2(uint64_t base, uint32_t i) {
  const float* f = P<const float>(base) + 2ull * i;
  return v2{f[0], f[1]};
}

inline uint32_t amax3(uint32_t a, uint32_t b, uint32_t c) { return std::max(std::max(a, b), c); }

template <class T>
inline T fadd(T* p, T v) {
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}

/* --------------------------------------------- table + 16^3 lookup */

struct Table {
  tc_tess_table_entry lookup[TC_TESSTABLE_NUM_LOOKUP];
  Table() {
    memset(lookup, 0, sizeof(lookup));
    /* tessellation_table.cpp:53-80: raw order x >= y >= z, (x,z,y) aliases */
    uint32_t idx = 0;
    for (uint32_t x = 1; x <= TC_TABLE_MAX_EDGE_SEGMENTS; x++)
      for (uint32_t y = 1; y <= x; y++)
        for (uint32_t z = 1; z <= y; z++, idx++) {
          tc_tess_table_entry e;
          e.firstTriangle = tc_table_nv_configs[idx * 4 + 0];
          e.firstVertex = tc_table_nv_configs[idx * 4 + 1];
          e.numTriangles = tc_table_nv_configs[idx * 4 + 2];
          e.numVertices = tc_table_nv_configs[idx * 4 + 3];
          lookup[lookup_index(x, y, z)] = e;
          if (z != y && x > 1) lookup[lookup_index(x, z, y)] = e;
        }
  }
  /* tessellation_table.hpp:58-62 */
  static uint32_t lookup_index(uint32_t x, uint32_t y, uint32_t z) { return x + y * 16 + z * 256 - (1 + 16 + 256); }
};

const Table& table() {
  static Table t;
  return t;
}

/* ------------------------------------------------ tessellation.glsl */

/* tessellation.glsl:48-59 */
uint32_t encode_bary(v3 wuv) {
  uint32_t ix = (uint32_t)(wuv.x * 32768.0f + 0.5f);
  uint32_t iy = (uint32_t)(wuv.y * 32768.0f + 0.5f);
  uint32_t iz = (uint32_t)(wuv.z * 32768.0f + 0.5f);
  if (ix > std::max(iy, iz))
    ix = 32768u - iy - iz;
  else if (iy > iz)
    iy = 32768u - ix - iz;
  else
    iz = 32768u - ix - iy;
  return iy | (iz << 16);
}

/* tessellation.glsl:66-76 */
v3 decode_bary(uint32_t vtx) {
  v3 w;
  w.y = float(vtx & 0xFFFF) / 32768.0f;
  w.z = float(vtx >> 16) / 32768.0f;
  w.x = (1.0f - w.y) - w.z;
  return w;
}

/* tessellation.glsl:78-99 */
v3 tess_factors(v3 a, v3 b, v3 c, v3 eye, float nearPlane, float vpy, float rate) {
  float dA = distance(a, eye), dB = distance(b, eye), dC = distance(c, eye);
  v3 spl;
  spl.x = 1.0f / std::max(nearPlane, std::min(dA, dB));
  spl.y = 1.0f / std::max(nearPlane, std::min(dB, dC));
  spl.z = 1.0f / std::max(nearPlane, std::min(dC, dA));
  v3 len = mk(distance(a, b), distance(b, c), distance(c, a));
  v3 f;
  f.x = roundf(((len.x * spl.x) * vpy) * rate);
  f.y = roundf(((len.y * spl.y) * vpy) * rate);
  f.z = roundf(((len.z * spl.z) * vpy) * rate);
  f.x = std::min(std::max(f.x, 1.0f), 32768.0f);
  f.y = std::min(std::max(f.y, 1.0f), 32768.0f);
  f.z = std::min(std::max(f.z, 1.0f), 32768.0f);
  return f;
}

/* tessellation.glsl:101-104 */
inline uint32_t split_factor(uint32_t f, uint32_t maxSplit) { return std::min((f + 11 - 1) / 11, maxSplit); }

/* tessellation.glsl:119-144 */
uint32_t get_config(uint32_t f[3], uint32_t vtx[3]) {
  uint32_t m = amax3(f[0], f[1], f[2]);
  if (m == f[1]) {
    uint32_t t0 = f[0], t1 = f[1], t2 = f[2];
    f[0] = t1, f[1] = t2, f[2] = t0;
    uint32_t v0 = vtx[0], v1 = vtx[1], v2 = vtx[2];
    vtx[0] = v1, vtx[1] = v2, vtx[2] = v0;
  } else if (m == f[2]) {
    uint32_t t0 = f[0], t1 = f[1], t2 = f[2];
    f[0] = t2, f[1] = t0, f[2] = t1;
    uint32_t v0 = vtx[0], v1 = vtx[1], v2 = vtx[2];
    vtx[0] = v2, vtx[1] = v0, vtx[2] = v1;
  }
  uint32_t idx = f[0] + f[1] * 16 + f[2] * 256 - (1 + 16 + 256);
  if (f[2] > f[1]) idx |= 1u << 15;
  return idx;
}

inline uint32_t cfg_index(uint32_t cfg) { return cfg & ~(1u << 15); }
inline bool cfg_flipped(uint32_t cfg) { return (cfg & (1u << 15)) != 0; }

/* tessellation.glsl:151-163 */
inline const tc_tess_table_entry& cfg_entry(uint32_t cfg) { return table().lookup[cfg_index(cfg)]; }

/* tessellation.glsl:165-176 */
void cfg_tri_vertices(uint32_t cfg, uint32_t tri, uint32_t out[3]) {
  uint32_t packed = tc_table_nv_triangles[cfg_entry(cfg).firstTriangle + tri];
  out[0] = packed & 0xFF;
  out[1] = (packed >> 8) & 0xFF;
  out[2] = (packed >> 16) & 0xFF;
  if (cfg_flipped(cfg)) std::swap(out[1], out[2]);
}

/* tessellation.glsl:178-189 */
v3 cfg_vertex_bary(uint32_t cfg, uint32_t vert) {
  v3 w = decode_bary(tc_table_nv_vertices[cfg_entry(cfg).firstVertex + vert]);
  if (cfg_flipped(cfg)) std::swap(w.x, w.y);
  return w;
}

/* ------------------------------------------------------ displacement */

inline float texel(const tc_displacement_texture& t, uint32_t x, uint32_t y) {
  uint64_t i = (uint64_t)y * t.width + x;
  if (t.format == TC_TEXTURE_R16_UNORM) return float(P<const uint16_t>(t.texels)[i]) / 65535.0f;
  return P<const float>(t.texels)[i];
}

inline int32_t wrap(int32_t i, int32_t n) {
  int32_t r = i % n;
  return r < 0 ? r + n : r;
}

/* Manual FP32 bilinear, REPEAT addressing (stands for texture() with the
 * linear sampler, cluster_classify.comp.glsl:477, resources.cpp:128). */
float sample_height(const tc_displacement_texture& t, float u, float v) {
  float x = u * float(t.width) - 0.5f;
  float y = v * float(t.height) - 0.5f;
  x = std::min(std::max(x, -16777216.0f), 16777216.0f);
  y = std::min(std::max(y, -16777216.0f), 16777216.0f);
  float x0 = floorf(x), y0 = floorf(y);
  float fx = x - x0, fy = y - y0;
  int32_t ix = (int32_t)x0, iy = (int32_t)y0;
  int32_t W = (int32_t)t.width, H = (int32_t)t.height;
  uint32_t xa = (uint32_t)wrap(ix, W), xb = (uint32_t)wrap(ix + 1, W);
  uint32_t ya = (uint32_t)wrap(iy, H), yb = (uint32_t)wrap(iy + 1, H);
  float t00 = texel(t, xa, ya), t10 = texel(t, xb, ya);
  float t01 = texel(t, xa, yb), t11 = texel(t, xb, yb);
  float top = t00 * (1.0f - fx) + t10 * fx;
  float bot = t01 * (1.0f - fx) + t11 * fx;
  return top * (1.0f - fy) + bot * fy;
}

/* displacement.glsl:30-104: PN triangle (Vlachos et al., Curved PN
 * Triangles, I3D 2001) with verts[0] at u = 1, verts[1] at v = 1 and
 * verts[2] at w = 1; pow(u,3) is written as u*u*u, and each term's scalar
 * weight is formed before it scales the control point (the GLSL scales the
 * point by 3.0, then by each factor: the same value up to rounding) */
struct PN {
  v3 b030, b003, b300, b021, b012, b102, b201, b210, b120, b111;
};

inline v3 project_to_plane(v3 p, v3 plane, v3 n) {
  v3 d = sub(p, plane);
  return sub(p, mul(n, dot(d, n)));
}

void setup_pn(PN& b, const v3 vt[3], const v3 n[3]) {
  b.b030 = vt[0];
  b.b003 = vt[1];
  b.b300 = vt[2];
  v3 e300 = sub(b.b003, b.b030);
  v3 e030 = sub(b.b300, b.b003);
  v3 e003 = sub(b.b030, b.b300);
  b.b021 = add(b.b030, mk(e300.x / 3.0f, e300.y / 3.0f, e300.z / 3.0f));
  b.b012 = add(b.b030, mk((e300.x * 2.0f) / 3.0f, (e300.y * 2.0f) / 3.0f, (e300.z * 2.0f) / 3.0f));
  b.b102 = add(b.b003, mk(e030.x / 3.0f, e030.y / 3.0f, e030.z / 3.0f));
  b.b201 = add(b.b003, mk((e030.x * 2.0f) / 3.0f, (e030.y * 2.0f) / 3.0f, (e030.z * 2.0f) / 3.0f));
  b.b210 = add(b.b300, mk(e003.x / 3.0f, e003.y / 3.0f, e003.z / 3.0f));
  b.b120 = add(b.b300, mk((e003.x * 2.0f) / 3.0f, (e003.y * 2.0f) / 3.0f, (e003.z * 2.0f) / 3.0f));
  b.b021 = project_to_plane(b.b021, b.b030, n[0]);
  b.b012 = project_to_plane(b.b012, b.b003, n[1]);
  b.b102 = project_to_plane(b.b102, b.b003, n[1]);
  b.b201 = project_to_plane(b.b201, b.b300, n[2]);
  b.b210 = project_to_plane(b.b210, b.b300, n[2]);
  b.b120 = project_to_plane(b.b120, b.b030, n[0]);
  v3 s3 = add(add(b.b003, b.b030), b.b300);
  v3 center = mk(s3.x / 3.0f, s3.y / 3.0f, s3.z / 3.0f);
  v3 s6 = add(add(add(add(add(b.b021, b.b012), b.b102), b.b201), b.b210), b.b120);
  b.b111 = mk(s6.x / 6.0f, s6.y / 6.0f, s6.z / 6.0f);
  v3 d = sub(b.b111, center);
  b.b111 = add(b.b111, mk(d.x / 2.0f, d.y / 2.0f, d.z / 2.0f));
}

v3 get_pn(const PN& b, v3 bary) {
  float u = bary.x, v = bary.y, w = bary.z;
  float u3 = u * u * u, v3_ = v * v * v, w3 = w * w * w;
  float u2 = u * u, v2_ = v * v, w2 = w * w;
  v3 r = mul(b.b300, w3);
  r = add(r, mul(b.b030, u3));
  r = add(r, mul(b.b003, v3_));
  r = add(r, mul(b.b210, (3.0f * w2) * u));
  r = add(r, mul(b.b120, (3.0f * w) * u2));
  r = add(r, mul(b.b201, (3.0f * w2) * v));
  r = add(r, mul(b.b021, (3.0f * u2) * v));
  r = add(r, mul(b.b102, (3.0f * w) * v2_));
  r = add(r, mul(b.b012, (3.0f * u) * v2_));
  r = add(r, mul(b.b111, ((6.0f * w) * u) * v));
  return r;
}

/* height -> offset along the normal, cluster_classify.comp.glsl:472-480 */
inline v3 displace(const oracle_ctx* c, const tc_render_instance& inst, v3 pos, v3 nrm, v2 uv) {
  const tc_displacement_texture& tex = c->displacementTextures[inst.displacementIndex];
  float h = sample_height(tex, uv.x, uv.y);
  h = (((h * inst.displacementScale) * c->view->displacementScale) + inst.displacementOffset) + c->view->displacementOffset;
  return add(pos, mul(normalize(nrm), h));
}

/* sin / cos for the ripple: the argument is reduced by multiples of pi/2 in
 * FP64, then the Taylor series to r^9 (sin) / r^8 (cos) is evaluated in FP32.
 * The GPU evaluates the same operations in the same order, so both sides
 * agree bit for bit; the error against the exact function is below 1e-6 for
 * |x| < 1e4 (pinned against libm in the tests). */
float ripple_sincos(float x, bool cosine) {
  const double k = rint((double)x * 0.63661977236758134308);
  const float r = (float)((double)x - k * 1.57079632679489661923);
  const long long q = (long long)k + (cosine ? 1 : 0);
  const float r2 = r * r;
  const float s = r * (1.0f + r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f)))));
  const float co = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f))));
  switch ((int)(q & 3)) {
    case 0: return s;
    case 1: return co;
    case 2: return -s;
    default: return -co;
  }
}

/* displacement.glsl:106-120 rippleDeform, applied after displacement at
 * cluster_classify.comp.glsl:483-485 and
 * triangle_tess_template_instantiate.comp.glsl:366-368 */
v3 ripple_deform(const oracle_ctx* c, v3 p, uint32_t seed, float size) {
  const tc_frame_constants* view = c->view;
  float maxCoord = std::max(fabsf(p.x), std::max(fabsf(p.y), fabsf(p.z)));
  if (maxCoord == 0.0f) return p; /* no direction at the origin: kept, not NaN */
  float frequency = view->animationRippleFrequency / size;
  float phase = view->animationState * view->animationRippleSpeed;
  float fs = (float)seed;
  v3 wave = mk(ripple_sincos((maxCoord * frequency + fs) + phase, false),
               ripple_sincos(((maxCoord * frequency) * 3.0f + fs) + phase, true),
               ripple_sincos(((maxCoord * frequency) * 1.2f + fs) + phase, false));
  float amp = view->animationRippleAmplitude * size;
  v3 n = normalize(mk(p.z, p.y, p.x));
  return add(p, mk(n.x * (amp * wave.x), n.y * (amp * wave.y), n.z * (amp * wave.z)));
}

inline bool has_disp(const oracle_ctx* c, tc_config cfg, const tc_render_instance& inst) {
  return (cfg.flags & TC_FLAG_DISPLACEMENT) && inst.displacementIndex >= 0 &&
         (uint32_t)inst.displacementIndex < c->numDisplacementTextures;
}

inline uint64_t max_data_bytes(tc_config cfg) { return uint64_t(cfg.maxGeneratedClusterMegs) * 1024ull * 1024ull; }

inline uint64_t gen_vertices_value(const oracle_ctx* c) {
  return c->genVerticesValue ? c->genVerticesValue : c->build->genVertices;
}

/* --------------------------------------------------- cluster_classify */

/* Allocation counters the stage advances; the serial run uses the live
 * SceneBuilding values, the threaded run starts each unit at its prefix.
 * partY is the upper half of the dual part counter (build.glsl:25-83):
 * part-list slots taken from the top of the list by 1X transient
 * clusters for their triangle mappings. */
struct ClassifyCounters {
  uint32_t split, part, genCluster, genVertex;
  uint64_t genData;
  uint32_t partY;
};

/* build.glsl:54-83 per item, in the canonical order: an allocation of n
 * part slots is accepted iff x + y + n + 1 <= MAX_PART_TRIANGLES with x and
 * y the counters before it (the reference decides per subgroup batch). */
inline bool dual_part_ok(uint64_t x, uint64_t y, uint64_t n, uint32_t maxParts) { return x + y + n + 1 <= maxParts; }

/* TESS_USE_TRANSIENTBUILDS (shaderio.h:140-141): 1X or 2X transient builds */
inline bool transient_builds(tc_config cfg) { return (cfg.flags & (TC_FLAG_1X_TRANSIENT | TC_FLAG_2X_TRANSIENT)) != 0; }

/* 1X transient cluster data: the exact size of our CLAS block for the
 * cluster's simple triangles and all its vertices (the reference reserves
 * basicClusterSizes[simpleCount], a worst case for NV's opaque format). */
inline uint32_t clas_size(uint32_t nt, uint32_t nv) {
  auto r16 = [](uint32_t v) { return (v + 15u) & ~15u; };
  uint32_t ng = (nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
  return (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u) + r16(nv * 12u);
}

/* 2X mini batches (cluster_classify.comp.glsl:667-905; shaderio.h:149-158):
 * up to 8 base triangles with every factor <= 2 share one transient CLAS of
 * 6 vertex slots and up to 4 triangles each. */
constexpr uint32_t kMiniBatch = 8, kMiniTris = 4, kMiniVerts = 6;
constexpr uint32_t kMiniBatchTris = kMiniBatch * kMiniTris, kMiniBatchVerts = kMiniBatch * kMiniVerts;
/* the reference's SUBGROUP_SIZE: batches never span two 32-triangle groups */
constexpr uint32_t kMiniGroup = 32;
/* worst-case CLAS bytes / vertex slots / part slots of a batch (:732-737) */
inline uint32_t mini_data_size() { return clas_size(kMiniBatchTris, kMiniBatchVerts); }
constexpr uint32_t kMiniVertexSize = kMiniBatchVerts + (kMiniBatchTris * 3 + 11) / 12;
constexpr uint32_t kMiniPartSize = (8 + kMiniBatchTris * 2 + 24 - 1) / 24;

/* One generated vertex of a sub-triangle: barycentric `w` of the base
 * triangle -> base surface (flat or PN), displacement, ripple
 * (triangle_tess_template_instantiate.comp.glsl:346-368, and the same steps
 * at cluster_classify.comp.glsl:846-876). */
v3 generate_vertex(const oracle_ctx* c, tc_config cfg, const tc_render_instance& inst, uint32_t instanceID, const v3 bp[3],
                   const v3 bn[3], const v2 bt[3], const PN* pn, v3 w) {
  v3 pos = pn ? get_pn(*pn, w) : interp(bp, w);
  if (has_disp(c, cfg, inst)) pos = displace(c, inst, pos, interp(bn, w), interp2(bt, w));
  if (cfg.flags & TC_FLAG_ANIMATION) pos = ripple_deform(c, pos, instanceID, inst.geoHi[3]);
  return pos;
}

/* cluster_classify.comp.glsl:150-905 for one visible cluster.  With
 * TC_FLAG_1X_TRANSIENT a cluster with more than one but not all triangles
 * simple (:261-276) gets a transient CLAS over its simple triangles
 * (:314-447, :489-535) and only its other triangles are split or
 * instantiated (:576-584).  With TC_FLAG_2X_TRANSIENT the remaining
 * triangles whose factors are all <= 2 go into 2X mini batches instead of
 * the part list (:586-592, :667-905).  Within a cluster the transient
 * allocations (1X, then the 2X batches in triangle order) come before the
 * split / part allocations.  WRITE=false only advances the counters. */
template <bool WRITE>
void classify_cluster(const oracle_ctx* c, tc_config cfg, uint32_t visIndex, ClassifyCounters& k) {
  tc_scene_building* b = c->build;
  tc_cluster_info ci = P<tc_cluster_info>(b->visibleClusters)[visIndex];
  const tc_render_instance& inst = c->instances[ci.instanceID];
  const tc_cluster& cl = P<const tc_cluster>(inst.clusters)[ci.clusterID];
  uint32_t state = P<const uint32_t>(b->instanceStates)[ci.instanceID];
  uint32_t nv = cl.numVertices, nt = cl.numTriangles;
  const uint8_t* lt = P<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
  const bool flag1x = (cfg.flags & TC_FLAG_1X_TRANSIENT) != 0;

  /* :177-195 */
  v3 sv[256];
  for (uint32_t v = 0; v < nv; v++) sv[v] = ld3(inst.positions, cl.firstLocalVertex + v);

  /* :200-246 */
  uint32_t fac[256][3];
  uint32_t simple = 0;
  bool invisible = (cfg.flags & TC_FLAG_CULLING) && !(state & TC_INSTANCE_VISIBLE_BIT);
  if (invisible) {
    simple = nt;
    for (uint32_t t = 0; t < nt; t++) fac[t][0] = fac[t][1] = fac[t][2] = 1;
  } else {
    v3 eye = mk(b->viewPos[0], b->viewPos[1], b->viewPos[2]);
    for (uint32_t t = 0; t < nt; t++) {
      v3 w0 = xform_point(inst.worldMatrix, sv[lt[t * 3 + 0]]);
      v3 w1 = xform_point(inst.worldMatrix, sv[lt[t * 3 + 1]]);
      v3 w2 = xform_point(inst.worldMatrix, sv[lt[t * 3 + 2]]);
      v3 f = tess_factors(w0, w1, w2, eye, c->view->nearPlane, c->view->viewportf[1], c->view->tessRate);
      fac[t][0] = (uint32_t)f.x, fac[t][1] = (uint32_t)f.y, fac[t][2] = (uint32_t)f.z;
      if (amax3(fac[t][0], fac[t][1], fac[t][2]) == 1) simple++;
    }
  }
  const bool full = simple == nt;
  const bool trans = flag1x && !full && simple > 1; /* transientSimpleThreshold = 1 */

  if (full || trans) {
    /* :300-447 one allocation per cluster */
    uint32_t genOffset = k.genCluster++;
    uint64_t dataSize = full ? P<const uint32_t>(inst.clusterTemplateInstantiatonSizes)[ci.clusterID] : clas_size(simple, nv);
    uint32_t vertexSize = full ? nv : nv + (simple * 3 + 11) / 12;
    uint32_t partSize = full ? 0 : (8 + simple + 24 - 1) / 24;
    uint32_t partOffset = 0;
    if (trans) {
      partOffset = dual_part_ok(k.part, k.partY, partSize, cfg.maxPartTriangles) ? cfg.maxPartTriangles - k.partY - partSize
                                                                                 : cfg.maxPartTriangles;
      k.partY += partSize;
    }
    uint64_t dataOffset = k.genData;
    k.genData += dataSize;
    uint32_t vertexOffset = k.genVertex;
    k.genVertex += vertexSize;
    if (WRITE) {
      if (full) fadd(&c->readback->numFullClusters, 1u);
      bool ok = !(uint64_t(vertexOffset) + vertexSize > cfg.maxGeneratedVertices ||
                  uint64_t(genOffset) + 1 > cfg.maxGeneratedClusters || dataOffset + dataSize > max_data_bytes(cfg) ||
                  (trans && uint64_t(partOffset) + partSize > cfg.maxPartTriangles));
      if (ok) {
        const uint64_t vbuf = gen_vertices_value(c) + uint64_t(vertexOffset) * 12ull;
        if (full) {
          uint32_t tempOffset = fadd(&b->tempInstantiateCounter, 1u);
          tc_template_instantiate_info ti;
          ti.clusterIdOffset = 0;
          ti.geometryIndexOffset = 0;
          ti.clusterTemplateAddress = P<const uint64_t>(inst.clusterTemplateAdresses)[ci.clusterID];
          ti.vertexBufferAddress = vbuf;
          ti.vertexBufferStride = 12;
          P<tc_template_instantiate_info>(b->tempInstantiations)[tempOffset] = ti;
          P<uint32_t>(b->tempInstanceIDs)[tempOffset] = ci.instanceID;
          P<uint64_t>(b->tempClusterAddresses)[tempOffset] = b->genClusterData + dataOffset;
        } else {
          /* :395-436 transient build record */
          uint32_t transOffset = fadd(&b->transBuildCounter, 1u);
          tc_clas_build_info bi;
          memset(&bi, 0, sizeof(bi));
          bi.clusterID = (TC_RT_CLUSTER_MODE_1X_SUBSET_CLUSTER << 30) | partOffset;
          bi.packed = simple | (nv << 9) | ((b->positionTruncateBitCount & 63u) << 18) | (1u << 24);
          bi.baseGeometryIndexAndFlags = TC_CLAS_GEOMETRY_OPAQUE;
          bi.indexBufferStride = 1;
          bi.vertexBufferStride = 12;
          bi.vertexBuffer = vbuf;
          bi.indexBuffer = vbuf + uint64_t(nv) * 12ull;
          P<tc_clas_build_info>(b->transBuilds)[transOffset] = bi;
          P<uint32_t>(b->transInstanceIDs)[transOffset] = ci.instanceID;
          P<uint64_t>(b->transClusterAddresses)[transOffset] = b->genClusterData + dataOffset;
          P<tc_tess_triangle_info>(b->partTriangles)[partOffset].cluster = ci;
        }
        fadd(&P<tc_blas_build_info>(b->blasBuildInfos)[ci.instanceID].clusterReferencesCount, 1u);
        fadd(&c->readback->numTotalTriangles, simple);
        /* :459-488 displaced cluster vertices */
        bool disp = has_disp(c, cfg, inst);
        for (uint32_t v = 0; v < nv; v++) {
          v3 p = sv[v];
          if (disp) {
            uint32_t vi = cl.firstLocalVertex + v;
            p = displace(c, inst, p, ld3(inst.normals, vi), ld2(inst.texcoords, vi));
          }
          if (cfg.flags & TC_FLAG_ANIMATION) p = ripple_deform(c, p, ci.instanceID, inst.geoHi[3]);
          float* o = P<float>(b->genVertices) + 3ull * (uint64_t(vertexOffset) + v);
          o[0] = p.x, o[1] = p.y, o[2] = p.z;
        }
        if (trans) {
          /* :497-534 triangle mapping bytes after the ClusterInfo of the
           * part slot, u8 indices after the vertices */
          uint8_t* map = P<uint8_t>(b->partTriangles) + uint64_t(partOffset) * 24ull + 8ull;
          uint8_t* idx = P<uint8_t>(b->genVertices) + (uint64_t(vertexOffset) + nv) * 12ull;
          uint32_t o = 0;
          for (uint32_t t = 0; t < nt; t++) {
            if (amax3(fac[t][0], fac[t][1], fac[t][2]) != 1) continue;
            map[o] = (uint8_t)t;
            for (int j = 0; j < 3; j++) idx[o * 3 + j] = lt[t * 3 + j];
            o++;
          }
          /* zero the rest of the index slots so the vertex stream is
           * deterministic */
          for (uint32_t j = simple * 3; j < (vertexSize - nv) * 12; j++) idx[j] = 0;
        }
      }
    }
    if (full) return;
  }

  const bool flag2x = (cfg.flags & TC_FLAG_2X_TRANSIENT) != 0;
  auto is_mini = [&](uint32_t t) {
    uint32_t m = amax3(fac[t][0], fac[t][1], fac[t][2]);
    return flag2x && m <= 2 && !(trans && m == 1);
  };
  if (flag2x) {
    /* :694-790 batches of 8 consecutive mini triangles of a 32-triangle
     * group, one allocation each */
    for (uint32_t g0 = 0; g0 < nt; g0 += kMiniGroup) {
      uint32_t members[kMiniGroup], nm = 0;
      for (uint32_t t = g0; t < std::min(nt, g0 + kMiniGroup); t++)
        if (is_mini(t)) members[nm++] = t;
      for (uint32_t m0 = 0; m0 < nm; m0 += kMiniBatch) {
        const uint32_t cnt = std::min(kMiniBatch, nm - m0);
        uint32_t mcfg[kMiniBatch], mvtx[kMiniBatch][3], batchTris = 0;
        for (uint32_t r = 0; r < cnt; r++) {
          uint32_t f[3] = {fac[members[m0 + r]][0], fac[members[m0 + r]][1], fac[members[m0 + r]][2]};
          mvtx[r][0] = 0, mvtx[r][1] = 32768u, mvtx[r][2] = 32768u << 16;
          mcfg[r] = get_config(f, mvtx[r]);
          batchTris += cfg_entry(mcfg[r]).numTriangles;
        }
        const uint32_t genOffset = k.genCluster++;
        const uint32_t partOffset = dual_part_ok(k.part, k.partY, kMiniPartSize, cfg.maxPartTriangles)
                                        ? cfg.maxPartTriangles - k.partY - kMiniPartSize
                                        : cfg.maxPartTriangles;
        k.partY += kMiniPartSize;
        const uint64_t dataOffset = k.genData;
        k.genData += mini_data_size();
        const uint32_t vertexOffset = k.genVertex;
        k.genVertex += kMiniVertexSize;
        if (!WRITE) continue;
        const bool ok = !(uint64_t(vertexOffset) + kMiniVertexSize > cfg.maxGeneratedVertices ||
                          uint64_t(genOffset) + 1 > cfg.maxGeneratedClusters ||
                          dataOffset + mini_data_size() > max_data_bytes(cfg) ||
                          uint64_t(partOffset) + kMiniPartSize > cfg.maxPartTriangles);
        if (!ok) continue;
        /* :750-790 the transient build record */
        const uint64_t vbuf = gen_vertices_value(c) + uint64_t(vertexOffset) * 12ull;
        uint32_t transOffset = fadd(&b->transBuildCounter, 1u);
        tc_clas_build_info bi;
        memset(&bi, 0, sizeof(bi));
        bi.clusterID = (TC_RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED << 30) | partOffset;
        bi.packed = batchTris | (kMiniBatchVerts << 9) | ((b->positionTruncateBitCount & 63u) << 18) | (1u << 24);
        bi.baseGeometryIndexAndFlags = TC_CLAS_GEOMETRY_OPAQUE;
        bi.indexBufferStride = 1;
        bi.vertexBufferStride = 12;
        bi.vertexBuffer = vbuf;
        bi.indexBuffer = vbuf + uint64_t(kMiniBatchVerts) * 12ull;
        P<tc_clas_build_info>(b->transBuilds)[transOffset] = bi;
        P<uint32_t>(b->transInstanceIDs)[transOffset] = ci.instanceID;
        P<uint64_t>(b->transClusterAddresses)[transOffset] = b->genClusterData + dataOffset;
        P<tc_tess_triangle_info>(b->partTriangles)[partOffset].cluster = ci;
        fadd(&P<tc_blas_build_info>(b->blasBuildInfos)[ci.instanceID].clusterReferencesCount, 1u);
        fadd(&c->readback->numTotalTriangles, batchTris);
        /* :794-903 per member: its config's vertices in 6 slots, its
         * triangles' u8 indices after the 48 vertex slots and u16 mappings
         * (triangle | sub-triangle << 8 | (factor - 1) bits << 12) after the
         * ClusterInfo of the part slot.  Slots and bytes no member uses are
         * zero (the reference leaves them unwritten). */
        float* vout = P<float>(b->genVertices) + 3ull * vertexOffset;
        uint8_t* idx = P<uint8_t>(b->genVertices) + (uint64_t(vertexOffset) + kMiniBatchVerts) * 12ull;
        uint16_t* map = reinterpret_cast<uint16_t*>(P<uint8_t>(b->partTriangles) + uint64_t(partOffset) * 24ull + 8ull);
        for (uint32_t j = 0; j < kMiniBatchVerts * 3; j++) vout[j] = 0.0f;
        for (uint32_t j = 0; j < (kMiniVertexSize - kMiniBatchVerts) * 12; j++) idx[j] = 0;
        const bool pn = (cfg.flags & TC_FLAG_PN) != 0;
        uint32_t o = 0;
        for (uint32_t r = 0; r < cnt; r++) {
          const uint32_t t = members[m0 + r];
          v3 bb[3], bp[3], bn[3];
          v2 bt[3];
          for (int v = 0; v < 3; v++) {
            bb[v] = decode_bary(mvtx[r][v]);
            uint32_t vi = lt[t * 3 + v] + cl.firstLocalVertex;
            bp[v] = ld3(inst.positions, vi);
            bn[v] = normalize(ld3(inst.normals, vi));
            bt[v] = ld2(inst.texcoords, vi);
          }
          PN basePN;
          if (pn) setup_pn(basePN, bp, bn);
          const tc_tess_table_entry& e = cfg_entry(mcfg[r]);
          for (uint32_t vert = 0; vert < e.numVertices; vert++) {
            v3 w = interp(bb, cfg_vertex_bary(mcfg[r], vert));
            v3 pos = generate_vertex(c, cfg, inst, ci.instanceID, bp, bn, bt, pn ? &basePN : nullptr, w);
            float* q = vout + 3 * (r * kMiniVerts + vert);
            q[0] = pos.x, q[1] = pos.y, q[2] = pos.z;
          }
          const uint32_t packedFactors = (fac[t][0] - 1) | ((fac[t][1] - 1) << 1) | ((fac[t][2] - 1) << 2);
          for (uint32_t i = 0; i < e.numTriangles; i++, o++) {
            map[o] = (uint16_t)(t | (i << 8) | (packedFactors << 12));
            uint32_t tv[3];
            cfg_tri_vertices(mcfg[r], i, tv);
            for (int j = 0; j < 3; j++) idx[o * 3 + j] = (uint8_t)(tv[j] + r * kMiniVerts);
          }
        }
      }
    }
  }

  /* :543-666 per-triangle split / part emission */
  for (uint32_t t = 0; t < nt; t++) {
    uint32_t f[3] = {fac[t][0], fac[t][1], fac[t][2]};
    uint32_t m = amax3(f[0], f[1], f[2]);
    if (trans && m == 1) continue; /* :576-584 in the transient CLAS */
    if (is_mini(t)) continue;      /* :586-592 in a 2X batch */
    bool needSplit = m > TC_TESSTABLE_SIZE;
    bool needPart = m <= TC_TESSTABLE_SIZE;
    uint32_t offSplit = needSplit ? k.split++ : 0;
    uint32_t offPart = needPart ? k.part++ : 0;
    if (!WRITE) continue;
    bool partOk = (flag1x || flag2x) ? dual_part_ok(offPart, k.partY, 1, cfg.maxPartTriangles)
                                     : offPart < cfg.maxPartTriangles;
    tc_tess_triangle_info ti;
    ti.cluster = ci;
    ti.subTriangle.vtxEncoded[0] = 0;
    ti.subTriangle.vtxEncoded[1] = 32768u;
    ti.subTriangle.vtxEncoded[2] = 32768u << 16;
    ti.subTriangle.triangleID_config = t;
    if (needSplit && offSplit < cfg.maxSplitTriangles) {
      for (int i = 0; i < 3; i++) f[i] = split_factor(f[i], cfg.maxSplitFactor);
      uint32_t cf = get_config(f, ti.subTriangle.vtxEncoded);
      ti.subTriangle.triangleID_config |= cf << 16;
      P<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
    } else if (needPart && partOk) {
      uint32_t cf = get_config(f, ti.subTriangle.vtxEncoded);
      ti.subTriangle.triangleID_config |= cf << 16;
      P<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
    }
  }
}

/* ------------------------------------------------------ triangle_split */

struct SplitCounters {
  uint32_t split, part;
};

/* triangle_split.comp.glsl:146-175 */
void fill_base_vertices(const oracle_ctx* c, const tc_tess_triangle_info& ti, v3 pos[3], v3 bary[3]) {
  const tc_render_instance& inst = c->instances[ti.cluster.instanceID];
  const tc_cluster& cl = P<const tc_cluster>(inst.clusters)[ti.cluster.clusterID];
  const uint8_t* lt = P<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
  uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
  for (int v = 0; v < 3; v++) {
    pos[v] = xform_point(inst.worldMatrix, ld3(inst.positions, lt[tri * 3 + v] + cl.firstLocalVertex));
    bary[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
  }
}

/* triangle_split.comp.glsl:201-331 for all children of one item, in
 * child order (the order the bit-search distributor visits them). */
template <bool WRITE>
void split_item(const oracle_ctx* c, tc_config cfg, uint32_t itemIndex, SplitCounters& k) {
  tc_scene_building* b = c->build;
  tc_tess_triangle_info parent = P<tc_tess_triangle_info>(b->splitTriangles)[itemIndex];
  uint32_t cfgP = parent.subTriangle.triangleID_config >> 16;
  uint32_t n = cfg_entry(cfgP).numTriangles; /* setupTask :193-199 */
  v3 pos[3], bary0[3];
  fill_base_vertices(c, parent, pos, bary0);
  v3 eye = mk(b->viewPos[0], b->viewPos[1], b->viewPos[2]);
  for (uint32_t child = 0; child < n; child++) {
    tc_tess_triangle_info ti = parent;
    uint32_t vi[3];
    cfg_tri_vertices(cfgP, child, vi);
    for (int v = 0; v < 3; v++) {
      v3 w = cfg_vertex_bary(cfgP, vi[v]);
      v3 rebased = interp(bary0, w); /* :240-242 (precise) */
      ti.subTriangle.vtxEncoded[v] = encode_bary(rebased);
    }
    v3 bary[3], wp[3];
    for (int v = 0; v < 3; v++) bary[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
    for (int v = 0; v < 3; v++) wp[v] = interp(pos, bary[v]); /* getFactors :177-189 */
    v3 ff = tess_factors(wp[0], wp[1], wp[2], eye, c->view->nearPlane, c->view->viewportf[1], c->view->tessRate);
    uint32_t f[3] = {(uint32_t)ff.x, (uint32_t)ff.y, (uint32_t)ff.z};
    uint32_t m = amax3(f[0], f[1], f[2]);
    bool needSplit = m > TC_TESSTABLE_SIZE;
    bool needPart = m <= TC_TESSTABLE_SIZE;
    uint32_t offSplit = needSplit ? k.split++ : 0;
    uint32_t offPart = needPart ? k.part++ : 0;
    if (!WRITE) continue;
    if (needSplit)
      for (int i = 0; i < 3; i++) f[i] = split_factor(f[i], cfg.maxSplitFactor);
    uint32_t cf = get_config(f, ti.subTriangle.vtxEncoded);
    ti.subTriangle.triangleID_config = (ti.subTriangle.triangleID_config & 0xFFFFu) | (cf << 16);
    /* with transient builds the part list is shared with their slots at
     * its top (build.glsl:68-83) */
    bool partOk = transient_builds(cfg)
                      ? dual_part_ok(offPart, uint32_t(b->dualPartTriangleCounter >> 32), 1, cfg.maxPartTriangles)
                      : offPart < cfg.maxPartTriangles;
    if (needSplit && offSplit < cfg.maxSplitTriangles)
      P<tc_tess_triangle_info>(b->splitTriangles)[offSplit] = ti;
    else if (needPart && partOk)
      P<tc_tess_triangle_info>(b->partTriangles)[offPart] = ti;
  }
}

/* ------------------------------------------- template instantiate */

/* triangle_tess_template_instantiate.comp.glsl:126-373 for one part
 * triangle with its allocation already made. */
void instantiate_part(const oracle_ctx* c, tc_config cfg, uint32_t p, uint32_t genOffset, uint64_t dataOffset,
                      uint32_t vertexOffset) {
  tc_scene_building* b = c->build;
  const tc_tessellation_table* tt = c->tessTable;
  tc_tess_triangle_info ti = P<tc_tess_triangle_info>(b->partTriangles)[p];
  uint32_t cf = ti.subTriangle.triangleID_config >> 16;
  uint32_t tri = ti.subTriangle.triangleID_config & 0xFFFF;
  const tc_tess_table_entry& e = cfg_entry(cf);
  uint32_t nv = e.numVertices;
  uint64_t dataSize = P<const uint32_t>(tt->templateInstantiationSizes)[cfg_index(cf)];
  bool ok = !(uint64_t(vertexOffset) + nv > cfg.maxGeneratedVertices || uint64_t(genOffset) + 1 > cfg.maxGeneratedClusters ||
              dataOffset + dataSize > max_data_bytes(cfg));
  if (!ok) return;
  /* :172-198 */
  tc_template_instantiate_info info;
  info.clusterIdOffset = p | (TC_RT_CLUSTER_MODE_SINGLE_TESSELLATED << 30);
  info.geometryIndexOffset = 0;
  info.clusterTemplateAddress = P<const uint64_t>(tt->templateAddresses)[cfg_index(cf)];
  info.vertexBufferAddress = gen_vertices_value(c) + uint64_t(vertexOffset) * 12ull;
  info.vertexBufferStride = 12;
  uint32_t tempOffset = fadd(&b->tempInstantiateCounter, 1u);
  P<tc_template_instantiate_info>(b->tempInstantiations)[tempOffset] = info;
  P<uint32_t>(b->tempInstanceIDs)[tempOffset] = ti.cluster.instanceID;
  P<uint64_t>(b->tempClusterAddresses)[tempOffset] = b->genClusterData + dataOffset;
  fadd(&P<tc_blas_build_info>(b->blasBuildInfos)[ti.cluster.instanceID].clusterReferencesCount, 1u);
  fadd(&c->readback->numTotalTriangles, (uint32_t)e.numTriangles);

  /* :295-372 vertex generation */
  const tc_render_instance& inst = c->instances[ti.cluster.instanceID];
  const tc_cluster& cl = P<const tc_cluster>(inst.clusters)[ti.cluster.clusterID];
  const uint8_t* lt = P<const uint8_t>(inst.clusterLocalTriangles) + cl.firstLocalTriangle;
  v3 bb[3], bp[3], bn[3];
  v2 bt[3];
  for (int v = 0; v < 3; v++) {
    bb[v] = decode_bary(ti.subTriangle.vtxEncoded[v]);
    uint32_t idx = lt[tri * 3 + v] + cl.firstLocalVertex;
    bp[v] = ld3(inst.positions, idx);
    bn[v] = normalize(ld3(inst.normals, idx));
    bt[v] = ld2(inst.texcoords, idx);
  }
  bool pn = (cfg.flags & TC_FLAG_PN) != 0;
  PN basePN;
  if (pn) setup_pn(basePN, bp, bn);
  bool disp = has_disp(c, cfg, inst);
  for (uint32_t vert = 0; vert < nv; vert++) {
    v3 w = interp(bb, cfg_vertex_bary(cf, vert));
    v3 pos = pn ? get_pn(basePN, w) : interp(bp, w);
    if (disp) pos = displace(c, inst, pos, interp(bn, w), interp2(bt, w));
    if (cfg.flags & TC_FLAG_ANIMATION) pos = ripple_deform(c, pos, ti.cluster.instanceID, inst.geoHi[3]);
    float* o = P<float>(b->genVertices) + 3ull * (uint64_t(vertexOffset) + vert);
    o[0] = pos.x, o[1] = pos.y, o[2] = pos.z;
  }
}

/* -------------------------------------------------- parallel helper */

template <class F>
void parallel_for(uint32_t n, uint32_t nthreads, F&& fn) {
  if (nthreads <= 1 || n < 2 * nthreads) {
    for (uint32_t i = 0; i < n; i++) fn(i);
    return;
  }
  std::vector<std::thread> th;
  uint32_t chunk = (n + nthreads - 1) / nthreads;
  for (uint32_t t = 0; t < nthreads; t++) {
    uint32_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([lo, hi, &fn] {
      for (uint32_t i = lo; i < hi; i++) fn(i);
    });
  }
  for (auto& x : th) x.join();
}

void classify_all(const oracle_ctx* c, tc_config cfg, uint32_t nthreads) {
  tc_scene_building* b = c->build;
  uint32_t n = b->visibleClusterCounter; /* clamped by build_setup CLASSIFY */
  ClassifyCounters base{(uint32_t)b->splitTriangleCounter, b->partTriangleCounter, b->genClusterCounter, b->genVertexCounter,
                        b->genClusterDataCounter, uint32_t(b->dualPartTriangleCounter >> 32)};
  if (nthreads <= 1) {
    ClassifyCounters k = base;
    for (uint32_t i = 0; i < n; i++) classify_cluster<true>(c, cfg, i, k);
    base = k;
  } else {
    std::vector<ClassifyCounters> cnt(n);
    parallel_for(n, nthreads, [&](uint32_t i) {
      ClassifyCounters k{0, 0, 0, 0, 0, 0};
      classify_cluster<false>(c, cfg, i, k);
      cnt[i] = k;
    });
    ClassifyCounters run = base;
    for (uint32_t i = 0; i < n; i++) {
      ClassifyCounters d = cnt[i];
      cnt[i] = run;
      run.split += d.split, run.part += d.part, run.genCluster += d.genCluster, run.genVertex += d.genVertex,
          run.genData += d.genData, run.partY += d.partY;
    }
    /* the template / transient records of clusters with an allocation are
     * appended with counters whose order must follow cluster order: emit
     * those clusters serially, the rest in parallel. */
    std::vector<uint8_t> isFull(n);
    for (uint32_t i = 0; i < n; i++) isFull[i] = (i + 1 < n ? cnt[i + 1].genCluster : run.genCluster) != cnt[i].genCluster;
    parallel_for(n, nthreads, [&](uint32_t i) {
      if (isFull[i]) return;
      ClassifyCounters k = cnt[i];
      classify_cluster<true>(c, cfg, i, k);
    });
    for (uint32_t i = 0; i < n; i++)
      if (isFull[i]) {
        ClassifyCounters k = cnt[i];
        classify_cluster<true>(c, cfg, i, k);
      }
    base = run;
  }
  b->splitTriangleCounter = (int32_t)base.split;
  b->partTriangleCounter = base.part;
  b->genClusterCounter = base.genCluster;
  b->genVertexCounter = base.genVertex;
  b->genClusterDataCounter = base.genData;
  b->dualPartTriangleCounter = (uint64_t(base.partY) << 32) | base.part;
}

void split_all(const oracle_ctx* c, tc_config cfg, uint32_t nthreads) {
  tc_scene_building* b = c->build;
  uint32_t start = b->splitPassStart, end = b->splitPassEnd;
  uint32_t n = end > start ? end - start : 0;
  SplitCounters base{b->splitWriteCounter, b->partTriangleCounter};
  if (nthreads <= 1) {
    SplitCounters k = base;
    for (uint32_t i = 0; i < n; i++) split_item<true>(c, cfg, start + i, k);
    base = k;
  } else {
    std::vector<SplitCounters> cnt(n);
    parallel_for(n, nthreads, [&](uint32_t i) {
      SplitCounters k{0, 0};
      split_item<false>(c, cfg, start + i, k);
      cnt[i] = k;
    });
    SplitCounters run = base;
    for (uint32_t i = 0; i < n; i++) {
      SplitCounters d = cnt[i];
      cnt[i] = run;
      run.split += d.split, run.part += d.part;
    }
    parallel_for(n, nthreads, [&](uint32_t i) {
      SplitCounters k = cnt[i];
      split_item<true>(c, cfg, start + i, k);
    });
    base = run;
  }
  b->splitWriteCounter = base.split;
  b->partTriangleCounter = base.part;
  b->dualPartTriangleCounter = (b->dualPartTriangleCounter & ~0xFFFFFFFFull) | base.part;
}

void instantiate_all(const oracle_ctx* c, tc_config cfg, uint32_t nthreads) {
  tc_scene_building* b = c->build;
  const tc_tessellation_table* tt = c->tessTable;
  uint32_t n = b->partTriangleCounter; /* clamped by build_setup INSTANTIATE_TESS */
  std::vector<uint32_t> gen(n), vtx(n);
  std::vector<uint64_t> data(n);
  uint32_t g = b->genClusterCounter, vv = b->genVertexCounter;
  uint64_t d = b->genClusterDataCounter;
  /* :155-160 allocation in part order */
  for (uint32_t p = 0; p < n; p++) {
    uint32_t cf = P<tc_tess_triangle_info>(b->partTriangles)[p].subTriangle.triangleID_config >> 16;
    gen[p] = g++;
    data[p] = d;
    d += P<const uint32_t>(tt->templateInstantiationSizes)[cfg_index(cf)];
    vtx[p] = vv;
    vv += cfg_entry(cf).numVertices;
  }
  /* accepted parts form a prefix, so the template slot of part p is the
   * running count of accepted parts before it; emit records serially and
   * the vertices in parallel */
  if (nthreads <= 1) {
    for (uint32_t p = 0; p < n; p++) instantiate_part(c, cfg, p, gen[p], data[p], vtx[p]);
  } else {
    uint32_t t0 = b->tempInstantiateCounter;
    uint32_t accepted = 0;
    for (uint32_t p = 0; p < n; p++) {
      uint32_t cf = P<tc_tess_triangle_info>(b->partTriangles)[p].subTriangle.triangleID_config >> 16;
      uint64_t ds = P<const uint32_t>(tt->templateInstantiationSizes)[cfg_index(cf)];
      bool ok = !(uint64_t(vtx[p]) + cfg_entry(cf).numVertices > cfg.maxGeneratedVertices ||
                  uint64_t(gen[p]) + 1 > cfg.maxGeneratedClusters || data[p] + ds > max_data_bytes(cfg));
      if (!ok) break;
      accepted++;
    }
    parallel_for(accepted, nthreads, [&](uint32_t p) {
      tc_scene_building local = *b;
      local.tempInstantiateCounter = t0 + p;
      oracle_ctx lc = *c;
      lc.build = &local;
      if (!lc.genVerticesValue) lc.genVerticesValue = b->genVertices;
      instantiate_part(&lc, cfg, p, gen[p], data[p], vtx[p]);
    });
    b->tempInstantiateCounter = t0 + accepted;
  }
  b->genClusterCounter = g;
  b->genVertexCounter = vv;
  b->genClusterDataCounter = d;
}

void frame_tessellate(const oracle_ctx* c, tc_config cfg, uint32_t passes, uint32_t nthreads) {
  if (passes == 0) passes = oracle_num_split_passes(cfg.maxSplitFactor);
  oracle_build_setup(c, cfg, TC_BUILD_SETUP_CLASSIFY);
  classify_all(c, cfg, nthreads);
  oracle_build_setup(c, cfg, TC_BUILD_SETUP_SPLIT);
  for (uint32_t p = 0; p < passes; p++) {
    split_all(c, cfg, nthreads);
    if (p + 1 < passes) oracle_build_setup(c, cfg, TC_BUILD_SETUP_SPLIT_PASS);
  }
  oracle_build_setup(c, cfg, TC_BUILD_SETUP_INSTANTIATE_TESS);
  instantiate_all(c, cfg, nthreads);
  oracle_build_setup(c, cfg, TC_BUILD_SETUP_BUILD_BLAS);
}

/* ------------------------------------------- culling front-end fixture */

inline void xform4(const float* m, const float p[4], float o[4]) {
  for (int r = 0; r < 4; r++) o[r] = ((m[r] * p[0] + m[4 + r] * p[1]) + m[8 + r] * p[2]) + m[12 + r] * p[3];
}

void mat4_mul(const float* a, const float* b, float* o) {
  for (int col = 0; col < 4; col++)
    for (int r = 0; r < 4; r++)
      o[col * 4 + r] = ((a[r] * b[col * 4 + 0] + a[4 + r] * b[col * 4 + 1]) + a[8 + r] * b[col * 4 + 2]) + a[12 + r] * b[col * 4 + 3];
}

/* ---------------------------------------------------------------- HiZ */

/* ceil(log2(x)) from the binary exponent, 0 for x <= 0 */
int ceil_log2(float x) {
  if (!(x > 0.0f)) return 0;
  if (std::isinf(x)) return 1 << 20;
  int e;
  float m = frexpf(x, &e);
  return m == 0.5f ? e - 1 : e;
}

uint64_t hiz_level_offset(uint32_t size, uint32_t level) {
  uint64_t o = 0;
  for (uint32_t k = 0; k < level; k++) o += uint64_t(size >> k) * (size >> k);
  return o;
}

/* textureLod through the far sampler (nvhiz_vk.cpp:84-114: linear filter
 * with MAX reduction, nearest mip, clamp to edge): the maximum of the 2x2
 * footprint texels with non-zero weight */
float hiz_sample(const tc_hiz& h, float u, float v, int level) {
  uint32_t lv = (uint32_t)std::min(std::max(level, 0), (int)h.mipLevels - 1);
  int dim = (int)(h.size >> lv);
  const float* t = h.texels + hiz_level_offset(h.size, lv);
  float x = u * (float)dim - 0.5f, y = v * (float)dim - 0.5f;
  float x0 = floorf(x), y0 = floorf(y);
  bool wx = x - x0 > 0.0f, wy = y - y0 > 0.0f;
  int ix0 = std::min(std::max((int)x0, 0), dim - 1), ix1 = std::min(std::max((int)x0 + 1, 0), dim - 1);
  int iy0 = std::min(std::max((int)y0, 0), dim - 1), iy1 = std::min(std::max((int)y0 + 1, 0), dim - 1);
  float m = t[uint64_t(iy0) * dim + ix0];
  if (wx) m = std::max(m, t[uint64_t(iy0) * dim + ix1]);
  if (wy) {
    m = std::max(m, t[uint64_t(iy1) * dim + ix0]);
    if (wx) m = std::max(m, t[uint64_t(iy1) * dim + ix1]);
  }
  return m;
}

/* culling.glsl:96-114 intersectHiz */
bool intersect_hiz(const tc_hiz& h, const tc_frame_constants* vl, const float cmin[4], const float cmax[4]) {
  float x0 = cmin[0] * 0.5f + 0.5f, y0 = cmin[1] * 0.5f + 0.5f;
  float x1 = cmax[0] * 0.5f + 0.5f, y1 = cmax[1] * 0.5f + 0.5f;
  x0 *= vl->hizSizeFactors[0], y0 *= vl->hizSizeFactors[1];
  x1 *= vl->hizSizeFactors[0], y1 *= vl->hizSizeFactors[1];
  x0 = std::min(x0, vl->hizSizeFactors[2]), y0 = std::min(y0, vl->hizSizeFactors[3]);
  x1 = std::min(x1, vl->hizSizeFactors[2]), y1 = std::min(y1, vl->hizSizeFactors[3]);
  float maxsize = std::max(x1 - x0, y1 - y0) * vl->hizSizeMax;
  float depth = hiz_sample(h, (x0 + x1) * 0.5f, (y0 + y1) * 0.5f, ceil_log2(maxsize));
  return cmin[2] <= depth + 2.0f / (float)(1 << 24);
}

}  // namespace

/* ======================================================== extern "C" */

extern "C" {

uint32_t oracle_table_num_vertices(void) { return TC_TABLE_NUM_VERTICES; }
uint32_t oracle_table_num_triangles(void) { return TC_TABLE_NUM_TRIANGLES; }
uint32_t oracle_table_num_configs(void) { return TC_TABLE_NUM_CONFIGS; }
const uint32_t* oracle_table_vertices(void) { return tc_table_nv_vertices; }
const uint32_t* oracle_table_triangles(void) { return tc_table_nv_triangles; }
const uint16_t* oracle_table_configs(void) { return tc_table_nv_configs; }
void oracle_table_lookup(tc_tess_table_entry out[4096]) { memcpy(out, table().lookup, sizeof(table().lookup)); }

uint32_t oracle_encode_barycentrics(const float wuv[3]) { return encode_bary(mk(wuv[0], wuv[1], wuv[2])); }
void oracle_decode_barycentrics(uint32_t vtx, float wuv[3]) {
  v3 w = decode_bary(vtx);
  wuv[0] = w.x, wuv[1] = w.y, wuv[2] = w.z;
}
void oracle_tess_factors(const float a[3], const float b[3], const float c[3], const float eye[3], float nearPlane, float vpy,
                         float rate, float out[3]) {
  v3 f = tess_factors(mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]), mk(c[0], c[1], c[2]), mk(eye[0], eye[1], eye[2]), nearPlane, vpy,
                      rate);
  out[0] = f.x, out[1] = f.y, out[2] = f.z;
}
void oracle_split_factor(const uint32_t f[3], uint32_t maxSplit, uint32_t out[3]) {
  for (int i = 0; i < 3; i++) out[i] = split_factor(f[i], maxSplit);
}
uint32_t oracle_get_config(const uint32_t factors[3], uint32_t vtx[3]) {
  uint32_t f[3] = {factors[0], factors[1], factors[2]};
  return get_config(f, vtx);
}
/* tessellation.glsl:111-117 */
void oracle_config_factors(uint32_t cfg, uint32_t out[3]) {
  uint32_t idx = cfg_index(cfg) + (1 + 16 + 256);
  out[0] = idx & 15, out[1] = (idx / 16) & 15, out[2] = (idx / 256) & 15;
}
void oracle_config_triangle_vertices(uint32_t cfg, uint32_t tri, uint32_t out[3]) { cfg_tri_vertices(cfg, tri, out); }
void oracle_config_vertex_barycentrics(uint32_t cfg, uint32_t vert, float wuv[3]) {
  v3 w = cfg_vertex_bary(cfg, vert);
  wuv[0] = w.x, wuv[1] = w.y, wuv[2] = w.z;
}
float oracle_sample_height(const tc_displacement_texture* tex, float u, float v) { return sample_height(*tex, u, v); }
float oracle_ripple_sin(float x) { return ripple_sincos(x, false); }

/* nvhiz-update.comp.glsl: level 0 the max of 2x2 source blocks (reads
 * clamped to the image), each further level the max of 2x2 texels */
void oracle_hiz_build(const tc_hiz* h, const void* depth, uint32_t stride) {
  const char* d = static_cast<const char*>(depth);
  auto at = [&](uint32_t x, uint32_t y) {
    float v;
    memcpy(&v, d + (uint64_t(std::min(y, h->height - 1)) * h->width + std::min(x, h->width - 1)) * stride, 4);
    return v;
  };
  for (uint32_t y = 0; y < h->size; y++)
    for (uint32_t x = 0; x < h->size; x++)
      h->texels[uint64_t(y) * h->size + x] =
          std::max(std::max(std::max(at(2 * x, 2 * y), at(2 * x + 1, 2 * y)), at(2 * x, 2 * y + 1)), at(2 * x + 1, 2 * y + 1));
  for (uint32_t k = 1; k < h->mipLevels; k++) {
    uint32_t dim = h->size >> k;
    const float* s = h->texels + hiz_level_offset(h->size, k - 1);
    float* o = h->texels + hiz_level_offset(h->size, k);
    for (uint32_t y = 0; y < dim; y++)
      for (uint32_t x = 0; x < dim; x++)
        o[uint64_t(y) * dim + x] = std::max(std::max(std::max(s[2ull * y * 2 * dim + 2 * x], s[2ull * y * 2 * dim + 2 * x + 1]),
                                                     s[(2ull * y + 1) * 2 * dim + 2 * x]),
                                            s[(2ull * y + 1) * 2 * dim + 2 * x + 1]);
  }
}

int oracle_intersect_hiz(const tc_hiz* hiz, const tc_frame_constants* viewLast, const float clipMin[4],
                         const float clipMax[4]) {
  return intersect_hiz(*hiz, viewLast, clipMin, clipMax) ? 1 : 0;
}
float oracle_ripple_cos(float x) { return ripple_sincos(x, true); }

void oracle_pn(const float verts[9], const float normals[9], const float bary[3], float out[3]) {
  v3 vt[3], n[3];
  for (int k = 0; k < 3; k++)
    vt[k] = mk(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2]), n[k] = mk(normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]);
  PN b;
  setup_pn(b, vt, n);
  v3 r = get_pn(b, mk(bary[0], bary[1], bary[2]));
  out[0] = r.x, out[1] = r.y, out[2] = r.z;
}

/* instances_classify.comp.glsl:100-127 with culling.glsl:30-90 (frustum and
 * size test; no HiZ pyramid exists, so the occlusion test passes). */
void oracle_instances_classify(const oracle_ctx* c, tc_config cfg) {
  tc_scene_building* b = c->build;
  const tc_frame_constants* vw = c->viewLast ? c->viewLast : c->view; /* culling.glsl reads viewLast */
  const bool hiz = (cfg.flags & TC_FLAG_HIZ) && c->hiz.texels;
  for (uint32_t i = 0; i < c->numRenderInstances; i++) {
    const tc_render_instance& inst = c->instances[i];
    float wvp[16];
    mat4_mul(vw->viewProjMatrix, inst.worldMatrix, wvp);
    uint32_t bits = 0x7F;
    float cmin[4], cmax[4];
    bool clipValid = true;
    for (int n = 0; n < 8; n++) {
      float p[4] = {(n & 1) ? inst.geoHi[0] : inst.geoLo[0], (n & 2) ? inst.geoHi[1] : inst.geoLo[1],
                    (n & 4) ? inst.geoHi[2] : inst.geoLo[2], 1.0f};
      float h[4];
      xform4(wvp, p, h);
      bool valid = !(-1.2e-07f < h[3] && h[3] < 1.2e-07f);
      float aw = fabsf(h[3]);
      float clip[4] = {h[0] / aw, h[1] / aw, h[2] / aw, h[3]};
      uint32_t cb = 0;
      cb |= h[0] < -h[3] ? 1u : 0u;
      cb |= h[0] > h[3] ? 2u : 0u;
      cb |= h[1] < -h[3] ? 4u : 0u;
      cb |= h[1] > h[3] ? 8u : 0u;
      cb |= h[2] < 0 ? 16u : 0u;
      cb |= h[2] > h[3] ? 32u : 0u;
      cb |= h[3] <= 0 ? 64u : 0u;
      bits &= cb;
      for (int k = 0; k < 4; k++) {
        cmin[k] = n == 0 ? clip[k] : std::min(cmin[k], clip[k]);
        cmax[k] = n == 0 ? clip[k] : std::max(cmax[k], clip[k]);
      }
      clipValid = clipValid && valid;
    }
    for (int k = 0; k < 2; k++) {
      cmin[k] = std::min(std::max(cmin[k], -1.0f), 1.0f);
      cmax[k] = std::min(std::max(cmax[k], -1.0f), 1.0f);
    }
    bool inFrustum = bits == 0;
    bool sizeOk = (cmax[0] - cmin[0]) > 2.0f / vw->viewportf[0] || (cmax[1] - cmin[1]) > 2.0f / vw->viewportf[1];
    /* :120 occlusion against the previous frame's far HiZ */
    bool hizOk = !hiz || !(inFrustum && clipValid && sizeOk) || intersect_hiz(c->hiz, vw, cmin, cmax);
    bool visible = inFrustum && (!clipValid || (sizeOk && hizOk));
    P<uint32_t>(b->instanceStates)[i] = (inFrustum ? TC_INSTANCE_FRUSTUM_BIT : 0u) | (visible ? TC_INSTANCE_VISIBLE_BIT : 0u);
    P<tc_blas_build_info>(b->blasBuildInfos)[i].clusterReferencesCount = 0;
  }
}

/* clusters_cull.comp.glsl:112-161 in ray-tracing mode: every cluster of
 * every instance is appended, instance by instance. */
void oracle_clusters_cull(const oracle_ctx* c, tc_config cfg) {
  tc_scene_building* b = c->build;
  for (uint32_t i = 0; i < c->numRenderInstances; i++)
    for (uint32_t cl = 0; cl < c->instances[i].numClusters; cl++) {
      uint32_t off = b->visibleClusterCounter++;
      if (off < cfg.maxVisibleClusters) P<tc_cluster_info>(b->visibleClusters)[off] = tc_cluster_info{i, cl};
    }
}

/* build_setup.comp.glsl:102-264 (no transient builds) */
void oracle_build_setup(const oracle_ctx* c, tc_config cfg, int mode) {
  tc_scene_building* b = c->build;
  tc_readback* rb = c->readback;
  if (mode == TC_BUILD_SETUP_CLASSIFY) {
    uint32_t counter = b->visibleClusterCounter;
    rb->numVisibleClusters = counter;
    counter = std::min(counter, cfg.maxVisibleClusters);
    b->visibleClusterCounter = counter;
    b->dispatchClassify = tc_dispatch_indirect{counter, 1, 1};
  } else if (mode == TC_BUILD_SETUP_DRAW_TESS) {
    uint32_t full = b->fullClusterCounter, part = b->partTriangleCounter;
    rb->numFullClusters = full;
    rb->numPartTriangles = part;
    rb->numSplitTriangles = b->splitWriteCounter;
    full = std::min(full, cfg.maxVisibleClusters);
    part = std::min(part, cfg.maxPartTriangles);
    b->drawFullClusters = tc_draw_mesh_tasks_indirect{full, 0};
    b->drawPartTriangles = tc_draw_mesh_tasks_indirect{(part + 31) / 32, 0};
    rb->numBlasClusters = full;
  } else if (mode == TC_BUILD_SETUP_SPLIT) {
    /* min(int, uint) in GLSL converts the int operand to uint */
    uint32_t count = std::min((uint32_t)b->splitTriangleCounter, cfg.maxSplitTriangles);
    b->splitWriteCounter = count;
    b->splitTriangleCounter = (int32_t)count;
    b->splitPassStart = 0;
    b->splitPassEnd = count;
    b->dispatchTriangleSplit = tc_dispatch_indirect{(count + 63) / 64, 1, 1};
  } else if (mode == TC_BUILD_SETUP_SPLIT_PASS) {
    b->splitPass = b->splitPass + 1;
    uint32_t start = std::min(b->splitPassEnd, cfg.maxSplitTriangles);
    uint32_t end = std::min(b->splitWriteCounter, cfg.maxSplitTriangles);
    b->splitPassStart = start;
    b->splitPassEnd = end;
    uint32_t count = end - start;
    b->dispatchTriangleSplit = tc_dispatch_indirect{(count + 63) / 64, 1, 1};
  } else if (mode == TC_BUILD_SETUP_BUILD_BLAS) {
    const bool trans = transient_builds(cfg);
    uint32_t temp = b->tempInstantiateCounter;
    uint32_t tr = trans ? b->transBuildCounter : 0u;
    rb->numBlasClusters = temp + tr;
    if (trans) rb->numTransBuilds = tr;
    rb->numTempInstantiations = temp;
    rb->numGenDatas = b->genClusterDataCounter;
    rb->numGenVertices = b->genVertexCounter;
    rb->numBlasReservedSizes = b->numBlasReservedSizes;
    temp = std::min(cfg.maxGeneratedClusters, temp);
    b->tempInstantiateCounter = temp;
    rb->numActualTempInstantiations = temp;
    b->dispatchBlasTempInsert = tc_dispatch_indirect{(temp + 63) / 64, 1, 1};
    if (trans) {
      tr = std::min(cfg.maxGeneratedClusters, temp + tr) - temp;
      b->transBuildCounter = tr;
      rb->numActualTransBuilds = tr;
      b->dispatchBlasTransInsert = tc_dispatch_indirect{(tr + 63) / 64, 1, 1};
    }
  } else if (mode == TC_BUILD_SETUP_INSTANTIATE_TESS) {
    uint32_t part = b->partTriangleCounter;
    rb->numSplitTriangles = b->splitWriteCounter;
    if (transient_builds(cfg)) {
      /* :240-253; the instantiated parts are the accepted prefix
       * x + y + 2 <= MAX_PART_TRIANGLES of the lower counter */
      uint32_t y = uint32_t(b->dualPartTriangleCounter >> 32);
      rb->numPartTriangles = part + y;
      rb->numTransPartTriangles = y;
      uint32_t cap = cfg.maxPartTriangles > y + 1 ? cfg.maxPartTriangles - y - 1 : 0u;
      part = std::min(part, cap);
    } else {
      rb->numPartTriangles = part;
      part = std::min(part, cfg.maxPartTriangles);
    }
    b->partTriangleCounter = part;
    b->dispatchTriangleInstantiate = tc_dispatch_indirect{(part + 31) / 32, 1, 1};
  }
}

void oracle_cluster_classify(const oracle_ctx* c, tc_config cfg) { classify_all(c, cfg, 1); }
void oracle_triangle_split(const oracle_ctx* c, tc_config cfg) { split_all(c, cfg, 1); }
void oracle_triangle_tess_template_instantiate(const oracle_ctx* c, tc_config cfg) { instantiate_all(c, cfg, 1); }

/* blas_setup_insertion.comp.glsl:100-116 in instance order */
void oracle_blas_setup_insertion(const oracle_ctx* c, tc_config) {
  tc_scene_building* b = c->build;
  for (uint32_t i = 0; i < c->numRenderInstances; i++) {
    tc_blas_build_info& bi = P<tc_blas_build_info>(b->blasBuildInfos)[i];
    uint32_t off = b->blasClusterCounter;
    b->blasClusterCounter += bi.clusterReferencesCount;
    bi.clusterReferencesCount = 0;
    bi.clusterReferencesStride = 8;
    bi.clusterReferences = b->blasClusterAddresses + uint64_t(off) * 8ull;
    c->readback->numBlasActualSizes += P<const uint32_t>(b->blasBuildSizes)[i];
  }
}

/* blas_clusters_insert.comp.glsl:97-135 in CLAS order.  clusterReferences
 * holds the address of the per-instance list in the build's address space;
 * blasClusterAddresses is the host pointer of the same array. */
void oracle_blas_clusters_insert(const oracle_ctx* c, tc_config cfg, int transient) {
  if (transient && !transient_builds(cfg)) return;
  tc_scene_building* b = c->build;
  uint32_t n = transient ? b->transBuildCounter : b->tempInstantiateCounter;
  const uint32_t* ids = P<const uint32_t>(transient ? b->transInstanceIDs : b->tempInstanceIDs);
  const uint64_t* addrs = P<const uint64_t>(transient ? b->transClusterAddresses : b->tempClusterAddresses);
  const uint32_t* sizes = P<const uint32_t>(transient ? b->transClusterSizes : b->tempClusterSizes);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t inst = ids[i];
    uint64_t addr = addrs[i];
    uint32_t size = sizes[i];
    tc_blas_build_info& bi = P<tc_blas_build_info>(b->blasBuildInfos)[inst];
    uint32_t idx = bi.clusterReferencesCount++;
    P<uint64_t>(bi.clusterReferences)[idx] = addr;
    c->readback->numGenActualDatas += size;
  }
}

/* Zeroes the 16-byte padding after the group boxes, triangle words and
 * vertices of a CLAS block. */
static void zero_block_padding(char* dst, uint32_t nt, uint32_t nv, uint32_t ng) {
  auto r16 = [](uint32_t v) { return (v + 15u) & ~15u; };
  uint32_t g1 = (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u), t1 = g1 + r16(nt * 4u);
  memset(dst + sizeof(tc_clas_header) + ng * 24u, 0, g1 - (sizeof(tc_clas_header) + ng * 24u));
  memset(dst + g1 + nt * 4u, 0, t1 - (g1 + nt * 4u));
  memset(dst + t1 + nv * 12u, 0, r16(nv * 12u) - nv * 12u);
}

/* Our CLAS build (stands for the NV template instantiation,
 * renderer_raytrace_clusters_tess.cpp:595-623): header, group boxes of
 * TC_CLAS_GROUP_TRIANGLES triangles, triangle words and vertex copy. */
void oracle_clas_instantiate(const oracle_ctx* c, tc_config cfg) {
  tc_scene_building* b = c->build;
  uint32_t n = std::min(b->tempInstantiateCounter, cfg.maxGeneratedClusters);
  auto r16 = [](uint32_t v) { return (v + 15u) & ~15u; };
  for (uint32_t i = 0; i < n; i++) {
    const tc_template_instantiate_info& info = P<const tc_template_instantiate_info>(b->tempInstantiations)[i];
    const tc_clas_template& T = *P<const tc_clas_template>(info.clusterTemplateAddress);
    uint32_t nt = T.numTriangles, nv = T.numVertices, ng = T.numGroups;
    char* dst = P<char>(P<const uint64_t>(b->tempClusterAddresses)[i]);
    tc_aabb* groups = reinterpret_cast<tc_aabb*>(dst + sizeof(tc_clas_header));
    uint32_t* tris = reinterpret_cast<uint32_t*>(dst + sizeof(tc_clas_header) + r16(ng * 24u));
    float* verts = reinterpret_cast<float*>(dst + sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u));
    const char* src = P<const char>(info.vertexBufferAddress);
    for (uint32_t v = 0; v < nv; v++)
      for (int k = 0; k < 3; k++) verts[v * 3 + k] = reinterpret_cast<const float*>(src + uint64_t(v) * info.vertexBufferStride)[k];
    for (uint32_t t = 0; t < nt; t++) tris[t] = P<const uint32_t>(T.triangles)[t];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t g = 0; g < ng; g++) {
      tc_aabb box;
      for (int k = 0; k < 3; k++) box.lo[k] = INFINITY, box.hi[k] = -INFINITY;
      for (uint32_t t = g * TC_CLAS_GROUP_TRIANGLES; t < std::min(nt, (g + 1) * TC_CLAS_GROUP_TRIANGLES); t++)
        for (int k = 0; k < 3; k++) {
          uint32_t vi = (tris[t] >> (8 * k)) & 0xFF;
          for (int a = 0; a < 3; a++) {
            box.lo[a] = std::min(box.lo[a], verts[vi * 3 + a]);
            box.hi[a] = std::max(box.hi[a], verts[vi * 3 + a]);
          }
        }
      groups[g] = box;
      for (int a = 0; a < 3; a++) lo[a] = std::min(lo[a], box.lo[a]), hi[a] = std::max(hi[a], box.hi[a]);
    }
    tc_clas_header h;
    for (int a = 0; a < 3; a++) h.lo[a] = lo[a], h.hi[a] = hi[a];
    h.clusterID = T.clusterID + info.clusterIdOffset;
    h.counts = nt | (nv << 8) | (ng << 16);
    h.vertexBase = info.vertexBufferAddress >= b->genVertices ? (uint32_t)((info.vertexBufferAddress - b->genVertices) / 12u)
                                                               : 0xFFFFFFFFu;
    h._pad[0] = h._pad[1] = h._pad[2] = 0;
    memcpy(dst, &h, sizeof(h));
    zero_block_padding(dst, nt, nv, ng);
    P<uint32_t>(b->tempClusterSizes)[i] = (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u) + r16(nv * 12u);
  }
}

/* Our CLAS build from transient build records (stands for the NV
 * transient CLAS build, renderer_raytrace_clusters_tess.cpp:625-655):
 * triangles in record order (primitive = transient triangle index), group
 * boxes of TC_CLAS_GROUP_TRIANGLES triangles, all the record's vertices. */
void oracle_clas_build_transient(const oracle_ctx* c, tc_config cfg) {
  tc_scene_building* b = c->build;
  if (!transient_builds(cfg)) return;
  uint32_t n = std::min(b->transBuildCounter, cfg.maxGeneratedClusters);
  auto r16 = [](uint32_t v) { return (v + 15u) & ~15u; };
  for (uint32_t i = 0; i < n; i++) {
    const tc_clas_build_info& info = P<const tc_clas_build_info>(b->transBuilds)[i];
    uint32_t nt = info.packed & 511u, nv = (info.packed >> 9) & 511u;
    uint32_t ng = (nt + TC_CLAS_GROUP_TRIANGLES - 1) / TC_CLAS_GROUP_TRIANGLES;
    char* dst = P<char>(P<const uint64_t>(b->transClusterAddresses)[i]);
    tc_aabb* groups = reinterpret_cast<tc_aabb*>(dst + sizeof(tc_clas_header));
    uint32_t* tris = reinterpret_cast<uint32_t*>(dst + sizeof(tc_clas_header) + r16(ng * 24u));
    float* verts = reinterpret_cast<float*>(dst + sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u));
    const char* vsrc = P<const char>(info.vertexBuffer);
    const uint8_t* isrc = P<const uint8_t>(info.indexBuffer);
    for (uint32_t v = 0; v < nv; v++)
      for (int k = 0; k < 3; k++) verts[v * 3 + k] = reinterpret_cast<const float*>(vsrc + uint64_t(v) * info.vertexBufferStride)[k];
    for (uint32_t t = 0; t < nt; t++) {
      const uint8_t* ix = isrc + uint64_t(t) * 3u * info.indexBufferStride;
      tris[t] = uint32_t(ix[0]) | (uint32_t(ix[info.indexBufferStride]) << 8) | (uint32_t(ix[2u * info.indexBufferStride]) << 16) |
                (t << 24);
    }
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t g = 0; g < ng; g++) {
      tc_aabb box;
      for (int k = 0; k < 3; k++) box.lo[k] = INFINITY, box.hi[k] = -INFINITY;
      for (uint32_t t = g * TC_CLAS_GROUP_TRIANGLES; t < std::min(nt, (g + 1) * TC_CLAS_GROUP_TRIANGLES); t++)
        for (int k = 0; k < 3; k++) {
          uint32_t vi = (tris[t] >> (8 * k)) & 0xFF;
          for (int a = 0; a < 3; a++) {
            box.lo[a] = std::min(box.lo[a], verts[vi * 3 + a]);
            box.hi[a] = std::max(box.hi[a], verts[vi * 3 + a]);
          }
        }
      groups[g] = box;
      for (int a = 0; a < 3; a++) lo[a] = std::min(lo[a], box.lo[a]), hi[a] = std::max(hi[a], box.hi[a]);
    }
    tc_clas_header h;
    for (int a = 0; a < 3; a++) h.lo[a] = lo[a], h.hi[a] = hi[a];
    h.clusterID = info.clusterID;
    h.counts = nt | (nv << 8) | (ng << 16);
    h.vertexBase = info.vertexBuffer >= b->genVertices ? (uint32_t)((info.vertexBuffer - b->genVertices) / 12u) : 0xFFFFFFFFu;
    h._pad[0] = h._pad[1] = h._pad[2] = 0;
    memcpy(dst, &h, sizeof(h));
    zero_block_padding(dst, nt, nv, ng);
    P<uint32_t>(b->transClusterSizes)[i] = (uint32_t)sizeof(tc_clas_header) + r16(ng * 24u) + r16(nt * 4u) + r16(nv * 12u);
  }
}

/* ------------------------------------------------------------ ray tracing */

namespace {

struct Ray {
  float o[3], d[3], inv[3];
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

/* Woop, Benthin, Wald: Watertight Ray/Triangle Intersection (JCGT 2013) */
void setup_ray(Ray& r, const float o[3], const float d[3]) {
  for (int k = 0; k < 3; k++) r.o[k] = o[k], r.d[k] = d[k], r.inv[k] = 1.0f / d[k];
  float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  if (d[r.kz] < 0.0f) std::swap(r.kx, r.ky);
  r.Sx = d[r.kx] / d[r.kz];
  r.Sy = d[r.ky] / d[r.kz];
  r.Sz = 1.0f / d[r.kz];
}

float box_hit(const Ray& r, const float lo[3], const float hi[3], float tmin, float tmax) {
  float tn = tmin, tf = tmax;
  for (int k = 0; k < 3; k++) {
    float t0 = (lo[k] - r.o[k]) * r.inv[k], t1 = (hi[k] - r.o[k]) * r.inv[k];
    tn = fmaxf(tn, fminf(t0, t1));
    tf = fminf(tf, fmaxf(t0, t1));
  }
  return tn <= tf * 1.0000005f ? tn : INFINITY;
}

bool tri_hit(const Ray& r, const float* p0, const float* p1, const float* p2, bool cull, float tmin, float tmax, float& t,
             float& u, float& v) {
  float A[3] = {p0[0] - r.o[0], p0[1] - r.o[1], p0[2] - r.o[2]};
  float B[3] = {p1[0] - r.o[0], p1[1] - r.o[1], p1[2] - r.o[2]};
  float Cc[3] = {p2[0] - r.o[0], p2[1] - r.o[1], p2[2] - r.o[2]};
  float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
  float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
  float Cx = Cc[r.kx] - r.Sx * Cc[r.kz], Cy = Cc[r.ky] - r.Sy * Cc[r.kz];
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = (U + V) + W;
  if (det == 0.0f || (cull && det < 0.0f)) return false;
  float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * Cc[r.kz];
  float T = (U * Az + V * Bz) + W * Cz;
  float tt = T / det;
  if (!(tt >= tmin && tt <= tmax)) return false;
  t = tt, u = V / det, v = W / det;
  return true;
}

void xform34(const float* m, const float v[3], float w, float out[3]) {
  for (int r = 0; r < 3; r++) out[r] = ((m[4 * r] * v[0] + m[4 * r + 1] * v[1]) + m[4 * r + 2] * v[2]) + m[4 * r + 3] * w;
}

struct Best {
  tc_hit h;
  const char* clas;
  uint32_t word;
};

bool less_key(float t, uint32_t i, uint32_t c, uint32_t p, const tc_hit& b) {
  if (t != b.t) return t < b.t;
  if (i != b.instanceID) return i < b.instanceID;
  if (c != b.clusterID) return c < b.clusterID;
  return p < b.primitiveID;
}

const float* clas_verts(const char* clas, uint32_t nt, uint32_t ng) {
  return reinterpret_cast<const float*>(clas + sizeof(tc_clas_header) + ((ng * 24u + 15u) & ~15u) + ((nt * 4u + 15u) & ~15u));
}

/* Object-space union of each instance's CLAS boxes.  A ray that enters a
 * CLAS box enters this union too (the slab bounds move monotonically with
 * the box bounds under rounding), so skipping instances whose union is
 * missed changes no result; it only saves the per-CLAS tests. */
std::vector<tc_aabb> instance_boxes(const oracle_ctx* c) {
  std::vector<tc_aabb> out(c->numRenderInstances);
  const tc_blas_build_info* infos = P<const tc_blas_build_info>(c->build->blasBuildInfos);
  for (uint32_t i = 0; i < c->numRenderInstances; i++) {
    tc_aabb& u = out[i];
    for (int k = 0; k < 3; k++) u.lo[k] = INFINITY, u.hi[k] = -INFINITY;
    const uint64_t* refs = P<const uint64_t>(infos[i].clusterReferences);
    for (uint32_t r = 0; r < infos[i].clusterReferencesCount; r++) {
      tc_clas_header hdr;
      memcpy(&hdr, P<const char>(refs[r]), sizeof(hdr));
      for (int k = 0; k < 3; k++) u.lo[k] = fminf(u.lo[k], hdr.lo[k]), u.hi[k] = fmaxf(u.hi[k], hdr.hi[k]);
    }
  }
  return out;
}

/* every CLAS of every instance, tested in CLAS-list order */
bool trace_one(const oracle_ctx* c, const float* xf, const float o[3], const float d[3], float tmin, float tmax, bool cull,
               bool anyHit, Best& best, const tc_aabb* instBoxes) {
  best.h.t = tmax;
  best.h.u = best.h.v = 0.0f;
  best.h.instanceID = best.h.clusterID = best.h.primitiveID = 0xFFFFFFFFu;
  best.clas = nullptr;
  best.word = 0;
  const tc_blas_build_info* infos = P<const tc_blas_build_info>(c->build->blasBuildInfos);
  for (uint32_t i = 0; i < c->numRenderInstances; i++) {
    const tc_blas_build_info& bi = infos[i];
    if (!bi.clusterReferencesCount) continue;
    float ol[3], dl[3];
    xform34(xf + 24 * i + 12, o, 1.0f, ol);
    xform34(xf + 24 * i + 12, d, 0.0f, dl);
    Ray r;
    setup_ray(r, ol, dl);
    if (box_hit(r, instBoxes[i].lo, instBoxes[i].hi, tmin, best.h.t) == INFINITY) continue;
    const uint64_t* refs = P<const uint64_t>(bi.clusterReferences);
    for (uint32_t k = 0; k < bi.clusterReferencesCount; k++) {
      const char* clas = P<const char>(refs[k]);
      tc_clas_header hdr;
      memcpy(&hdr, clas, sizeof(hdr));
      if (box_hit(r, hdr.lo, hdr.hi, tmin, best.h.t) == INFINITY) continue;
      uint32_t nt = hdr.counts & 0xFF, ng = (hdr.counts >> 16) & 0xFF;
      const tc_aabb* groups = reinterpret_cast<const tc_aabb*>(clas + sizeof(tc_clas_header));
      const uint32_t* tris = reinterpret_cast<const uint32_t*>(clas + sizeof(tc_clas_header) + ((ng * 24u + 15u) & ~15u));
      const float* verts = clas_verts(clas, nt, ng);
      for (uint32_t g = 0; g < ng; g++) {
        if (box_hit(r, groups[g].lo, groups[g].hi, tmin, best.h.t) == INFINITY) continue;
        for (uint32_t t = g * TC_CLAS_GROUP_TRIANGLES; t < std::min(nt, (g + 1) * TC_CLAS_GROUP_TRIANGLES); t++) {
          uint32_t w = tris[t];
          float th, tu, tv;
          if (!tri_hit(r, verts + 3 * (w & 0xFF), verts + 3 * ((w >> 8) & 0xFF), verts + 3 * ((w >> 16) & 0xFF), cull, tmin,
                       best.h.t, th, tu, tv))
            continue;
          if (anyHit) {
            best.h.t = th;
            return true;
          }
          uint32_t prim = w >> 24;
          if (less_key(th, i, hdr.clusterID, prim, best.h)) {
            best.h = tc_hit{th, tu, tv, i, hdr.clusterID, prim};
            best.clas = clas;
            best.word = w;
          }
        }
      }
    }
  }
  return best.h.instanceID != 0xFFFFFFFFu;
}

void mul44(const float* m, const float v[4], float out[4]) {
  for (int r = 0; r < 4; r++) out[r] = ((m[r] * v[0] + m[4 + r] * v[1]) + m[8 + r] * v[2]) + m[12 + r] * v[3];
}
void normalize4(float v[4]) {
  float l = sqrtf(((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
  float s = 1.0f / l;
  for (int k = 0; k < 4; k++) v[k] = v[k] * s;
}

/* render_raytrace.rgen.glsl:52-86 */
void primary_ray(const tc_frame_constants* vw, uint32_t x, uint32_t y, uint32_t W, uint32_t H, float o[3], float d[3]) {
  float uvx = ((float)x + 0.5f) / (float)W, uvy = ((float)y + 0.5f) / (float)H;
  float dx = uvx * 2.0f - 1.0f, dy = uvy * 2.0f - 1.0f;
  float org[4] = {0, 0, 0, 1}, origin[4], ndc[4] = {dx, dy, 1.0f, 1.0f}, target[4], dir[4];
  mul44(vw->viewMatrixI, org, origin);
  mul44(vw->projMatrixI, ndc, target);
  normalize4(target);
  float tv[4] = {target[0], target[1], target[2], 0.0f};
  mul44(vw->viewMatrixI, tv, dir);
  normalize4(dir);
  for (int k = 0; k < 3; k++) o[k] = origin[k], d[k] = dir[k];
}

/* render_raytrace_clusters.rchit.glsl:128-236, 338-339 */
void decode(const oracle_ctx* c, const float* xf, const Best& b, tc_hit_decoded& out) {
  const tc_hit& h = b.h;
  out.instanceID = h.instanceID;
  out.shadowed = 0;
  if (h.instanceID == 0xFFFFFFFFu) {
    out.clusterID = out.triangleID = out.subTriangleID = 0xFFFFFFFFu;
    out.baryBase[0] = out.baryBase[1] = out.baryBase[2] = 0.0f;
    out.depth = 1.0f;
    out.wPos[0] = out.wPos[1] = out.wPos[2] = 0.0f;
    return;
  }
  uint32_t mode = h.clusterID >> 30, clusterID = h.clusterID & 0x3FFFFFFFu, triangleID = h.primitiveID;
  uint32_t subTriangle = h.primitiveID;
  v3 bw = mk((1.0f - h.u) - h.v, h.u, h.v), base = bw;
  if (mode == TC_RT_CLUSTER_MODE_1X_SUBSET_CLUSTER) {
    /* rchit:173-178: the part slot holds the ClusterInfo, then one byte per
     * transient triangle naming the cluster triangle */
    const uint8_t* slot = P<const uint8_t>(c->build->partTriangles) + uint64_t(clusterID) * 24ull;
    tc_cluster_info ci;
    memcpy(&ci, slot, sizeof(ci));
    triangleID = slot[8 + h.primitiveID];
    clusterID = ci.clusterID;
  } else if (mode == TC_RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED) {
    /* rchit:152-170: a u16 mapping per batch triangle after the ClusterInfo
     * names the cluster triangle, the sub-triangle of its config and the
     * (factor - 1) bits the config is rebuilt from; the sub-triangle field
     * is masked with 0xF (the reference's `& 4` drops sub-triangles 1-3) */
    const uint8_t* slot = P<const uint8_t>(c->build->partTriangles) + uint64_t(clusterID) * 24ull;
    tc_cluster_info ci;
    memcpy(&ci, slot, sizeof(ci));
    uint16_t packed;
    memcpy(&packed, slot + 8 + 2 * h.primitiveID, 2);
    triangleID = packed & 0xFFu;
    subTriangle = (packed >> 8) & 0xFu;
    uint32_t f[3] = {1u + ((packed >> 12) & 1u), 1u + ((packed >> 13) & 1u), 1u + (uint32_t(packed) >> 14)};
    uint32_t vtx[3] = {0u, 32768u, 32768u << 16};
    uint32_t cfg = get_config(f, vtx);
    clusterID = ci.clusterID;
    uint32_t idx[3];
    cfg_tri_vertices(cfg, subTriangle, idx);
    v3 vb[3] = {cfg_vertex_bary(cfg, idx[0]), cfg_vertex_bary(cfg, idx[1]), cfg_vertex_bary(cfg, idx[2])};
    v3 nb = interp(vb, bw);
    v3 bb[3] = {decode_bary(vtx[0]), decode_bary(vtx[1]), decode_bary(vtx[2])};
    base = interp(bb, nb);
  } else if (mode != TC_RT_CLUSTER_MODE_FULL_CLUSTER) {
    const tc_tess_triangle_info& ti = P<const tc_tess_triangle_info>(c->build->partTriangles)[clusterID];
    triangleID = ti.subTriangle.triangleID_config & 0xFFFF;
    uint32_t cfg = ti.subTriangle.triangleID_config >> 16;
    clusterID = ti.cluster.clusterID;
    uint32_t idx[3];
    cfg_tri_vertices(cfg, h.primitiveID, idx);
    v3 vb[3] = {cfg_vertex_bary(cfg, idx[0]), cfg_vertex_bary(cfg, idx[1]), cfg_vertex_bary(cfg, idx[2])};
    v3 nb = interp(vb, bw);
    v3 bb[3] = {decode_bary(ti.subTriangle.vtxEncoded[0]), decode_bary(ti.subTriangle.vtxEncoded[1]),
                decode_bary(ti.subTriangle.vtxEncoded[2])};
    base = interp(bb, nb);
  }
  out.clusterID = clusterID;
  out.triangleID = triangleID;
  out.subTriangleID = subTriangle;
  out.baryBase[0] = base.x, out.baryBase[1] = base.y, out.baryBase[2] = base.z;
  tc_clas_header hdr;
  memcpy(&hdr, b.clas, sizeof(hdr));
  const float* verts = clas_verts(b.clas, hdr.counts & 0xFF, (hdr.counts >> 16) & 0xFF);
  const float* p0 = verts + 3 * (b.word & 0xFF);
  const float* p1 = verts + 3 * ((b.word >> 8) & 0xFF);
  const float* p2 = verts + 3 * ((b.word >> 16) & 0xFF);
  v3 pv[3] = {mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), mk(p2[0], p2[1], p2[2])};
  v3 op = interp(pv, bw);
  float opv[3] = {op.x, op.y, op.z}, wp[3];
  xform34(xf + 24 * h.instanceID, opv, 1.0f, wp);
  out.wPos[0] = wp[0], out.wPos[1] = wp[1], out.wPos[2] = wp[2];
  float w4[4] = {wp[0], wp[1], wp[2], 1.0f}, pp[4];
  mul44(c->view->viewProjMatrix, w4, pp);
  out.depth = pp[2] / pp[3];
}

/* one ray: closest hit, its decode and (flags & 2) the shadow ray from the
 * hit point toward the sun (render_shading.glsl:263-274) */
void trace_and_decode(const oracle_ctx* c, const float* xforms, const tc_aabb* instBoxes, const float o[3], const float d[3],
                      float tmin, float tmax, uint32_t flags, tc_hit* hitOut, tc_hit_decoded* decOut) {
  bool cull = (flags & 1u) != 0, shadows = (flags & 2u) != 0;
  Best b;
  trace_one(c, xforms, o, d, tmin, tmax, cull, false, b, instBoxes);
  tc_hit hit = b.h;
  if (hit.instanceID == 0xFFFFFFFFu) hit.t = -1.0f;
  if (hitOut) *hitOut = hit;
  if (!decOut) return;
  tc_hit_decoded dec;
  decode(c, xforms, b, dec);
  if (shadows && hit.instanceID != 0xFFFFFFFFu) {
    const float* sun = c->view->skyParams.sunDirection;
    float sd[3] = {sun[0], sun[1], sun[2]};
    Best sb;
    dec.shadowed = trace_one(c, xforms, dec.wPos, sd, 0.001f, 10000000.0f, false, true, sb, instBoxes) ? 1u : 0u;
  }
  *decOut = dec;
}

}  // namespace

extern "C" void oracle_trace(const oracle_ctx* c, const float* xforms, const tc_ray* rays, uint32_t numRays, uint32_t width,
                             uint32_t height, uint32_t flags, tc_hit* hits, tc_hit_decoded* decoded, uint32_t nthreads) {
  uint32_t n = rays ? numRays : width * height;
  const std::vector<tc_aabb> boxes = instance_boxes(c);
  parallel_for(n, nthreads, [&](uint32_t r) {
    float o[3], d[3], tmin, tmax;
    if (rays) {
      for (int k = 0; k < 3; k++) o[k] = rays[r].origin[k], d[k] = rays[r].direction[k];
      tmin = rays[r].tMin, tmax = rays[r].tMax;
    } else {
      primary_ray(c->view, r % width, r / width, width, height, o, d);
      tmin = c->view->nearPlane, tmax = c->view->farPlane;
    }
    trace_and_decode(c, xforms, boxes.data(), o, d, tmin, tmax, flags, hits ? hits + r : nullptr,
                     decoded ? decoded + r : nullptr);
  });
}

extern "C" void oracle_trace_pixels(const oracle_ctx* c, const float* xforms, const uint32_t* pixels, uint32_t numPixels,
                                    uint32_t width, uint32_t height, uint32_t flags, tc_hit* hits, tc_hit_decoded* decoded,
                                    uint32_t nthreads) {
  const std::vector<tc_aabb> boxes = instance_boxes(c);
  parallel_for(numPixels, nthreads, [&](uint32_t r) {
    float o[3], d[3];
    primary_ray(c->view, pixels[r] % width, pixels[r] / width, width, height, o, d);
    trace_and_decode(c, xforms, boxes.data(), o, d, c->view->nearPlane, c->view->farPlane, flags, hits ? hits + r : nullptr,
                     decoded ? decoded + r : nullptr);
  });
}

extern "C" void oracle_primary_ray(const tc_frame_constants* view, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
                                   float origin[3], float direction[3]) {
  primary_ray(view, x, y, width, height, origin, direction);
}

uint32_t oracle_num_split_passes(uint32_t splitFactor) {
  /* renderer_raytrace_clusters_tess.cpp:516-520 */
  uint32_t coord = 32768, passes = 0;
  if (splitFactor < 2) splitFactor = 2;
  while (coord > splitFactor) {
    coord /= splitFactor;
    passes++;
  }
  return passes;
}

void oracle_frame_tessellate(const oracle_ctx* c, tc_config cfg, uint32_t passes) { frame_tessellate(c, cfg, passes, 1); }
void oracle_frame_tessellate_mt(const oracle_ctx* c, tc_config cfg, uint32_t passes, uint32_t nthreads) {
  frame_tessellate(c, cfg, passes, nthreads);
}

}  // extern "C"
