// tc_device.h -- device-side helpers shared by the gfx950 kernels.
//
// Floating-point evaluation order is fixed explicitly (files are built with
// -ffp-contract=off and correctly rounded fp32 division / sqrt) so integer
// decisions taken from float math (tessellation factors, barycentric
// encoding) and generated positions are reproducible bit for bit:
//   dot       (a.x*b.x + a.y*b.y) + a.z*b.z
//   length    sqrtf(dot(v, v))
//   normalize v * (1.0f / length(v))
//   mat4*vec4 column-major, summed column by column
//   round     roundf (half away from zero)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tc_abi.h"

#pragma clang fp contract(off)

namespace tc {

constexpr int kWave = 64;

// ------------------------------------------------------------------ math

struct v3 {
  float x, y, z;
};
struct v2 {
  float x, y;
};

__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float length(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float distance(v3 a, v3 b) { return length(sub(a, b)); }
__device__ __forceinline__ v3 normalize(v3 a) { return mul(a, 1.0f / length(a)); }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ v3 xform_point(const float* m, v3 p) {
  v3 r;
  r.x = ((m[0] * p.x + m[4] * p.y) + m[8] * p.z) + m[12];
  r.y = ((m[1] * p.x + m[5] * p.y) + m[9] * p.z) + m[13];
  r.z = ((m[2] * p.x + m[6] * p.y) + m[10] * p.z) + m[14];
  return r;
}

// barycentric interpolation, tessellation.glsl:191-201 (precise)
__device__ __forceinline__ v3 interp(v3 b0, v3 b1, v3 b2, v3 w) {
  return mk((b0.x * w.x + b1.x * w.y) + b2.x * w.z, (b0.y * w.x + b1.y * w.y) + b2.y * w.z,
            (b0.z * w.x + b1.z * w.y) + b2.z * w.z);
}
__device__ __forceinline__ v2 interp2(v2 b0, v2 b1, v2 b2, v3 w) {
  return v2{(b0.x * w.x + b1.x * w.y) + b2.x * w.z, (b0.y * w.x + b1.y * w.y) + b2.y * w.z};
}

__device__ __forceinline__ uint32_t umax3(uint32_t a, uint32_t b, uint32_t c) { return max(max(a, b), c); }

template <class T>
__device__ __forceinline__ T* ptr(uint64_t a) {
  return reinterpret_cast<T*>(a);
}

// Load from device memory through the global address space.  A pointer
// rebuilt from a 64-bit address is generic, and a generic (flat) load also
// counts against the LDS counter, so LDS accesses after it wait for memory.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
// A record (a multiple of four bytes) through the global address space,
// word by word; neighbouring words merge into wide loads.
template <class T>
__device__ __forceinline__ T gld_record(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  T out;
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
  const uint32_t* s = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; i++) o[i] = gld(s + i);
  return out;
}
// The value unchanged, opaque to the optimiser (no arithmetic is derived
// from it), e.g. to keep neighbouring byte loads from being merged.
__device__ __forceinline__ uint64_t opaque(uint64_t x) {
  asm volatile("" : "+v"(x));
  return x;
}
// Store through the global address space, for the same reason: a flat
// store holds the next LDS access until the store has completed.
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// A record through the global address space, word by word.
template <class T>
__device__ __forceinline__ void gst_record(T* p, const T& v) {
  static_assert(sizeof(T) % 4 == 0, "whole words");
  uint32_t* d = reinterpret_cast<uint32_t*>(p);
  const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; i++) gst(d + i, s[i]);
}

__device__ __forceinline__ v3 ld3(uint64_t base, uint32_t i) {
  const float* f = ptr<const float>(base) + 3ull * i;
  return mk(gld(f), gld(f + 1), gld(f + 2));
}
__device__ __forceinline__ v2 ld2(uint64_t base, uint32_t i) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  const f2 f = gld(ptr<const f2>(base) + i);
  return v2{f.x, f.y};
}
__device__ __forceinline__ void st3(uint64_t base, uint64_t i, v3 p) {
  float* f = ptr<float>(base) + 3ull * i;
  gst(f, p.x);
  gst(f + 1, p.y);
  gst(f + 2, p.z);
}

// --------------------------------------------------- tessellation.glsl

// tessellation.glsl:48-59
__device__ __forceinline__ uint32_t encode_bary(v3 w) {
  uint32_t ix = (uint32_t)(w.x * 32768.0f + 0.5f);
  uint32_t iy = (uint32_t)(w.y * 32768.0f + 0.5f);
  uint32_t iz = (uint32_t)(w.z * 32768.0f + 0.5f);
  if (ix > max(iy, iz))
    ix = 32768u - iy - iz;
  else if (iy > iz)
    iy = 32768u - ix - iz;
  else
    iz = 32768u - ix - iy;
  return iy | (iz << 16);
}

// tessellation.glsl:66-76
__device__ __forceinline__ v3 decode_bary(uint32_t vtx) {
  v3 w;
  w.y = float(vtx & 0xFFFF) / 32768.0f;
  w.z = float(vtx >> 16) / 32768.0f;
  w.x = (1.0f - w.y) - w.z;
  return w;
}

struct ViewParams {
  v3 eye;
  float nearPlane;
  float vpy;
  float rate;
};

// tessellation.glsl:78-99, returned as integers; dA, dB, dC are the
// corners' distances to the eye, distance(a, vp.eye) etc.
__device__ __forceinline__ void tess_factors(v3 a, v3 b, v3 c, float dA, float dB, float dC, const ViewParams& vp,
                                             uint32_t f[3]) {
  float sx = 1.0f / fmaxf(vp.nearPlane, fminf(dA, dB));
  float sy = 1.0f / fmaxf(vp.nearPlane, fminf(dB, dC));
  float sz = 1.0f / fmaxf(vp.nearPlane, fminf(dC, dA));
  float lx = distance(a, b), ly = distance(b, c), lz = distance(c, a);
  float fx = roundf(((lx * sx) * vp.vpy) * vp.rate);
  float fy = roundf(((ly * sy) * vp.vpy) * vp.rate);
  float fz = roundf(((lz * sz) * vp.vpy) * vp.rate);
  f[0] = (uint32_t)fminf(fmaxf(fx, 1.0f), 32768.0f);
  f[1] = (uint32_t)fminf(fmaxf(fy, 1.0f), 32768.0f);
  f[2] = (uint32_t)fminf(fmaxf(fz, 1.0f), 32768.0f);
}
__device__ __forceinline__ void tess_factors(v3 a, v3 b, v3 c, const ViewParams& vp, uint32_t f[3]) {
  tess_factors(a, b, c, distance(a, vp.eye), distance(b, vp.eye), distance(c, vp.eye), vp, f);
}

// tessellation.glsl:101-104
__device__ __forceinline__ uint32_t split_factor(uint32_t f, uint32_t maxSplit) { return min((f + 10u) / 11u, maxSplit); }

// tessellation.glsl:119-144
__device__ __forceinline__ uint32_t get_config(uint32_t f[3], uint32_t vtx[3]) {
  uint32_t m = umax3(f[0], f[1], f[2]);
  if (m == f[1]) {
    uint32_t t0 = f[0], v0 = vtx[0];
    f[0] = f[1], f[1] = f[2], f[2] = t0;
    vtx[0] = vtx[1], vtx[1] = vtx[2], vtx[2] = v0;
  } else if (m == f[2]) {
    uint32_t t2 = f[2], v2 = vtx[2];
    f[2] = f[1], f[1] = f[0], f[0] = t2;
    vtx[2] = vtx[1], vtx[1] = vtx[0], vtx[0] = v2;
  }
  uint32_t idx = f[0] + f[1] * 16u + f[2] * 256u - 273u;
  if (f[2] > f[1]) idx |= 1u << 15;
  return idx;
}

__device__ __forceinline__ uint32_t cfg_index(uint32_t cfg) { return cfg & 0x7FFFu; }
__device__ __forceinline__ bool cfg_flipped(uint32_t cfg) { return (cfg & 0x8000u) != 0; }

// Table view held in registers (wave-uniform addresses).
struct TableView {
  const uint32_t* vertices;
  const uint32_t* triangles;
  const tc_tess_table_entry* entries;
  const uint64_t* templateAddresses;
  const uint32_t* templateSizes;
};

__device__ __forceinline__ TableView load_table(const tc_tessellation_table* t) {
  TableView v;
  v.vertices = ptr<const uint32_t>(t->vertices);
  v.triangles = ptr<const uint32_t>(t->triangles);
  v.entries = ptr<const tc_tess_table_entry>(t->entries);
  v.templateAddresses = ptr<const uint64_t>(t->templateAddresses);
  v.templateSizes = ptr<const uint32_t>(t->templateInstantiationSizes);
  return v;
}

__device__ __forceinline__ tc_tess_table_entry cfg_entry(const TableView& T, uint32_t cfg) {
  // one 8-byte load, through the global address space
  const uint32_t* q = reinterpret_cast<const uint32_t*>(T.entries) + 2 * cfg_index(cfg);
  const uint32_t rx = gld(q), ry = gld(q + 1);
  tc_tess_table_entry e;
  e.firstTriangle = (uint16_t)(rx & 0xFFFF);
  e.firstVertex = (uint16_t)(rx >> 16);
  e.numTriangles = (uint16_t)(ry & 0xFFFF);
  e.numVertices = (uint16_t)(ry >> 16);
  return e;
}

// tessellation.glsl:165-176
// vertex indices of a table triangle word of config cfg
__device__ __forceinline__ void tri_word_vertices(uint32_t packed, uint32_t cfg, uint32_t out[3]) {
  out[0] = packed & 0xFF;
  uint32_t i1 = (packed >> 8) & 0xFF, i2 = (packed >> 16) & 0xFF;
  bool fl = cfg_flipped(cfg);
  out[1] = fl ? i2 : i1;
  out[2] = fl ? i1 : i2;
}
__device__ __forceinline__ void cfg_tri_vertices(const TableView& T, uint32_t cfg, uint32_t firstTriangle, uint32_t tri,
                                                 uint32_t out[3]) {
  uint32_t packed = gld(T.triangles + firstTriangle + tri);
  out[0] = packed & 0xFF;
  uint32_t i1 = (packed >> 8) & 0xFF, i2 = (packed >> 16) & 0xFF;
  bool fl = cfg_flipped(cfg);
  out[1] = fl ? i2 : i1;
  out[2] = fl ? i1 : i2;
}

// tessellation.glsl:178-189
__device__ __forceinline__ v3 cfg_vertex_bary(const TableView& T, uint32_t cfg, uint32_t firstVertex, uint32_t vert) {
  v3 w = decode_bary(gld(T.vertices + firstVertex + vert));
  if (cfg_flipped(cfg)) {
    float t = w.x;
    w.x = w.y;
    w.y = t;
  }
  return w;
}

// --------------------------------------------------------- displacement

// x / 65535 for a 16-bit x, correctly rounded: x times the rounded
// reciprocal, then one residual correction with fused multiply-adds.  Equal
// to the IEEE quotient for every one of the 65536 inputs (an exhaustive
// check runs in the CPU tests); the division takes about ten instructions.
__device__ __forceinline__ float unorm16(uint32_t x) {
  constexpr float R = 1.0f / 65535.0f;
  const float xf = (float)x;
  const float q0 = xf * R;
  const float r = __builtin_fmaf(-q0, 65535.0f, xf);
  return __builtin_fmaf(r, R, q0);
}

__device__ __forceinline__ float texel(const tc_displacement_texture& t, uint32_t x, uint32_t y) {
  uint64_t i = (uint64_t)y * t.width + x;
  if (t.format == TC_TEXTURE_R16_UNORM) return unorm16(gld(ptr<const uint16_t>(t.texels) + i));
  return gld(ptr<const float>(t.texels) + i);
}

// i mod n in [0, n); coordinates already inside the texture (nearly all of
// them) skip the integer division.
__device__ __forceinline__ int32_t wrapi(int32_t i, int32_t n) {
  if ((uint32_t)i < (uint32_t)n) return i;
  int32_t r = i % n;
  return r < 0 ? r + n : r;
}

// FP32 bilinear with REPEAT addressing; stands for texture() with a linear
// sampler (cluster_classify.comp.glsl:477).
__device__ __forceinline__ float sample_height(const tc_displacement_texture& t, float u, float v) {
  float x = u * float(t.width) - 0.5f;
  float y = v * float(t.height) - 0.5f;
  x = fminf(fmaxf(x, -16777216.0f), 16777216.0f);
  y = fminf(fmaxf(y, -16777216.0f), 16777216.0f);
  float x0 = floorf(x), y0 = floorf(y);
  float fx = x - x0, fy = y - y0;
  int32_t ix = (int32_t)x0, iy = (int32_t)y0;
  int32_t W = (int32_t)t.width, H = (int32_t)t.height;
  uint32_t xa = (uint32_t)wrapi(ix, W), xb = (uint32_t)wrapi(ix + 1, W);
  uint32_t ya = (uint32_t)wrapi(iy, H), yb = (uint32_t)wrapi(iy + 1, H);
  // one format branch around all four fetches, so that they are in flight
  // together (a branch per texel waits for each fetch in turn)
  const uint64_t ra = (uint64_t)ya * t.width, rb = (uint64_t)yb * t.width;
  float t00, t10, t01, t11;
  if (t.format == TC_TEXTURE_R16_UNORM) {
    const uint16_t* p = ptr<const uint16_t>(t.texels);
    const uint16_t a00 = gld(p + ra + xa), a10 = gld(p + ra + xb), a01 = gld(p + rb + xa), a11 = gld(p + rb + xb);
    t00 = unorm16(a00), t10 = unorm16(a10), t01 = unorm16(a01), t11 = unorm16(a11);
  } else {
    const float* p = ptr<const float>(t.texels);
    t00 = gld(p + ra + xa), t10 = gld(p + ra + xb), t01 = gld(p + rb + xa), t11 = gld(p + rb + xb);
  }
  float top = t00 * (1.0f - fx) + t10 * fx;
  float bot = t01 * (1.0f - fx) + t11 * fx;
  return top * (1.0f - fy) + bot * fy;
}

// The four texels sample_height(t, u, v) reads and their bilinear weights
// (same coordinate arithmetic), order (x0,y0) (x1,y0) (x0,y1) (x1,y1).
__device__ __forceinline__ void bilinear_footprint(uint32_t width, uint32_t height, float u, float v, uint64_t idx[4],
                                                   float w[4]) {
  float x = u * float(width) - 0.5f;
  float y = v * float(height) - 0.5f;
  x = fminf(fmaxf(x, -16777216.0f), 16777216.0f);
  y = fminf(fmaxf(y, -16777216.0f), 16777216.0f);
  const float x0 = floorf(x), y0 = floorf(y);
  const float fx = x - x0, fy = y - y0;
  const int32_t ix = (int32_t)x0, iy = (int32_t)y0;
  const uint64_t xa = (uint32_t)wrapi(ix, (int32_t)width), xb = (uint32_t)wrapi(ix + 1, (int32_t)width);
  const uint64_t ya = (uint32_t)wrapi(iy, (int32_t)height), yb = (uint32_t)wrapi(iy + 1, (int32_t)height);
  idx[0] = ya * width + xa, idx[1] = ya * width + xb, idx[2] = yb * width + xa, idx[3] = yb * width + xb;
  w[0] = (1.0f - fx) * (1.0f - fy), w[1] = fx * (1.0f - fy), w[2] = (1.0f - fx) * fy, w[3] = fx * fy;
}

// sin / cos that evaluate identically on the CPU oracle and the GPU: the
// argument is reduced by multiples of pi/2 in FP64 (k = rint(x * 2/pi),
// r = x - k * pi/2), then the Taylor series to r^9 / r^8 is evaluated in
// FP32 (|r| <= pi/4 keeps the truncation below 3e-9).
__device__ __forceinline__ float tc_sincos_reduced(float x, bool cosine) {
  const double k = rint((double)x * 0.63661977236758134308);
  const float r = (float)((double)x - k * 1.57079632679489661923);
  long long q = (long long)k + (cosine ? 1 : 0);
  const float r2 = r * r;
  const float s = r * (1.0f + r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f)))));
  const float c = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f))));
  const int m = (int)(q & 3);
  return m == 0 ? s : (m == 1 ? c : (m == 2 ? -s : -c));
}
__device__ __forceinline__ float tc_sin(float x) { return tc_sincos_reduced(x, false); }
__device__ __forceinline__ float tc_cos(float x) { return tc_sincos_reduced(x, true); }

// displacement.glsl:106-120 ripple animation (DO_ANIMATION)
__device__ __forceinline__ v3 ripple_deform(v3 p, uint32_t seed, float size, const tc_frame_constants* view) {
  const float maxCoord = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
  // normalize(p.zyx) has no direction at the object-space origin; the point
  // stays put there instead of turning into NaN
  if (maxCoord == 0.0f) return p;
  const float frequency = view->animationRippleFrequency / size;
  const float phase = view->animationState * view->animationRippleSpeed;
  const float fs = (float)seed;
  const v3 wave = mk(tc_sin((maxCoord * frequency + fs) + phase), tc_cos(((maxCoord * frequency) * 3.0f + fs) + phase),
                     tc_sin(((maxCoord * frequency) * 1.2f + fs) + phase));
  const float amp = view->animationRippleAmplitude * size;
  const v3 n = normalize(mk(p.z, p.y, p.x));
  return add(p, mk(n.x * (amp * wave.x), n.y * (amp * wave.y), n.z * (amp * wave.z)));
}

// ------------------------------------------------------------------ HiZ

// ceil(log2(x)) exactly, from the binary exponent (x = m 2^e, m in
// [0.5, 1)); 0 for x <= 0 (a zero-sized footprint samples level 0)
__device__ __forceinline__ int ceil_log2(float x) {
  if (!(x > 0.0f)) return 0;
  if (isinf(x)) return 1 << 20;
  int e;
  const float m = frexpf(x, &e);
  return m == 0.5f ? e - 1 : e;
}

__device__ __forceinline__ uint64_t hiz_level_offset(uint32_t size, uint32_t level) {
  uint64_t o = 0;
  for (uint32_t k = 0; k < level; k++) o += (uint64_t)(size >> k) * (size >> k);
  return o;
}

// textureLod with the far sampler (nvhiz_vk.cpp:84-114: linear filter with
// MAX reduction, nearest mip, clamp to edge): the maximum of the texels of
// the 2 x 2 bilinear footprint that carry a non-zero weight
__device__ __forceinline__ float hiz_sample(const tc_hiz& h, float u, float v, int level) {
  const uint32_t lv = (uint32_t)min(max(level, 0), (int)h.mipLevels - 1);
  const int dim = (int)(h.size >> lv);
  const float* t = h.texels + hiz_level_offset(h.size, lv);
  const float x = u * (float)dim - 0.5f, y = v * (float)dim - 0.5f;
  const float x0 = floorf(x), y0 = floorf(y);
  const bool wx = x - x0 > 0.0f, wy = y - y0 > 0.0f;
  const int ix0 = min(max((int)x0, 0), dim - 1), ix1 = min(max((int)x0 + 1, 0), dim - 1);
  const int iy0 = min(max((int)y0, 0), dim - 1), iy1 = min(max((int)y0 + 1, 0), dim - 1);
  float m = t[(uint64_t)iy0 * dim + ix0];
  if (wx) m = fmaxf(m, t[(uint64_t)iy0 * dim + ix1]);
  if (wy) {
    m = fmaxf(m, t[(uint64_t)iy1 * dim + ix0]);
    if (wx) m = fmaxf(m, t[(uint64_t)iy1 * dim + ix1]);
  }
  return m;
}

// culling.glsl:96-114 intersectHiz (clipMin / clipMax with xy already
// clamped to [-1, 1], as intersectFrustum returns them)
__device__ __forceinline__ bool intersect_hiz(const tc_hiz& h, const tc_frame_constants* vl, const float cmin[4],
                                              const float cmax[4]) {
  float x0 = cmin[0] * 0.5f + 0.5f, y0 = cmin[1] * 0.5f + 0.5f;
  float x1 = cmax[0] * 0.5f + 0.5f, y1 = cmax[1] * 0.5f + 0.5f;
  x0 *= vl->hizSizeFactors[0], y0 *= vl->hizSizeFactors[1];
  x1 *= vl->hizSizeFactors[0], y1 *= vl->hizSizeFactors[1];
  x0 = fminf(x0, vl->hizSizeFactors[2]), y0 = fminf(y0, vl->hizSizeFactors[3]);
  x1 = fminf(x1, vl->hizSizeFactors[2]), y1 = fminf(y1, vl->hizSizeFactors[3]);
  const float maxsize = fmaxf(x1 - x0, y1 - y0) * vl->hizSizeMax;
  const float depth = hiz_sample(h, (x0 + x1) * 0.5f, (y0 + y1) * 0.5f, ceil_log2(maxsize));
  return cmin[2] <= depth + 2.0f / (float)(1 << 24);
}

// displacement.glsl:30-104 (PN triangle), pow written as products
struct PNBasis {
  v3 b030, b003, b300, b021, b012, b102, b201, b210, b120, b111;
};

__device__ __forceinline__ v3 project_to_plane(v3 p, v3 plane, v3 n) {
  v3 d = sub(p, plane);
  return sub(p, mul(n, dot(d, n)));
}
__device__ __forceinline__ v3 div3(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 mul2div3(v3 a) { return mk((a.x * 2.0f) / 3.0f, (a.y * 2.0f) / 3.0f, (a.z * 2.0f) / 3.0f); }

__device__ __forceinline__ void setup_pn(PNBasis& b, v3 p0, v3 p1, v3 p2, v3 n0, v3 n1, v3 n2) {
  b.b030 = p0;
  b.b003 = p1;
  b.b300 = p2;
  v3 e300 = sub(b.b003, b.b030);
  v3 e030 = sub(b.b300, b.b003);
  v3 e003 = sub(b.b030, b.b300);
  b.b021 = add(b.b030, div3(e300, 3.0f));
  b.b012 = add(b.b030, mul2div3(e300));
  b.b102 = add(b.b003, div3(e030, 3.0f));
  b.b201 = add(b.b003, mul2div3(e030));
  b.b210 = add(b.b300, div3(e003, 3.0f));
  b.b120 = add(b.b300, mul2div3(e003));
  b.b021 = project_to_plane(b.b021, b.b030, n0);
  b.b012 = project_to_plane(b.b012, b.b003, n1);
  b.b102 = project_to_plane(b.b102, b.b003, n1);
  b.b201 = project_to_plane(b.b201, b.b300, n2);
  b.b210 = project_to_plane(b.b210, b.b300, n2);
  b.b120 = project_to_plane(b.b120, b.b030, n0);
  v3 center = div3(add(add(b.b003, b.b030), b.b300), 3.0f);
  v3 s6 = add(add(add(add(add(b.b021, b.b012), b.b102), b.b201), b.b210), b.b120);
  b.b111 = div3(s6, 6.0f);
  b.b111 = add(b.b111, div3(sub(b.b111, center), 2.0f));
}

__device__ __forceinline__ v3 get_pn(const PNBasis& b, v3 bary) {
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

// ------------------------------------------------------------ wave ops

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of `mask` below this lane
__device__ __forceinline__ uint32_t mask_prefix(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t shfl(uint32_t v, uint32_t src) { return (uint32_t)__shfl((int)v, (int)src, kWave); }

__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
  uint32_t lo = shfl((uint32_t)v, src), hi = shfl((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// inclusive wave scan (Hillis-Steele over 64 lanes)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t o = (uint32_t)__shfl_up((int)v, d, kWave);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, kWave);
    uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, kWave);
    uint64_t o = ((uint64_t)hi << 32) | lo;
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, d, kWave);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d, kWave);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += (uint32_t)__shfl_xor((int)v, d, kWave);
  return v;
}

// ------------------------------------------------ workspace + look-back

// Status word of one tile and counter:
//   [63:44] epoch  [43:42] flag (1 aggregate, 2 inclusive)  [41:0] value
constexpr uint32_t kEpochBits = 20;
constexpr uint64_t kValueMask = (1ull << 42) - 1;
constexpr uint32_t kMaxCounters = 8;

__device__ __forceinline__ uint64_t pack_status(uint32_t epoch, uint32_t flag, uint64_t value) {
  return ((uint64_t)epoch << 44) | ((uint64_t)flag << 42) | (value & kValueMask);
}

// Workspace header; the status words follow it.
struct WsHeader {
  uint32_t epoch;     // current look-back epoch (1 .. 2^20-1)
  uint32_t ticket;    // next tile ticket
  uint32_t numTiles;  // tiles of the current launch
  uint32_t itemBegin; // first item of the current launch
  uint32_t itemEnd;   // end of the items of the current launch
  uint32_t _pad[3];
  uint64_t base[8];   // counter values at launch start
};
constexpr uint64_t kWsHeaderBytes = 1024;

// Pass table of the persistent split kernel (TC_FLAG_PERSISTENT_SPLIT),
// after the header: pass p covers global tickets [tileBegin, tileBegin +
// numTiles); a pass is published (passReady > p) once the pass before it
// has completed, and `finished` is set when no pass with tiles is left.
constexpr uint32_t kMaxSplitPasses = 16;
struct SplitPass {
  uint32_t tileBegin, numTiles;
  uint32_t itemBegin, itemEnd;
  uint32_t done;   // completed tiles
  uint32_t epoch;  // look-back epoch of the pass
  uint64_t b0, b1; // split / part counters at pass start
};
struct SplitQueue {
  uint32_t passReady;
  uint32_t finished;
  uint32_t numPasses;
  uint32_t error;      // a wait exceeded its spin limit
  uint32_t waitLimit;  // sleeps before a wait gives up (tc_tuning.splitWaitLimit)
  uint32_t _pad[3];
  SplitPass pass[kMaxSplitPasses];
};
static_assert(256 + sizeof(SplitQueue) <= kWsHeaderBytes, "split queue exceeds the workspace header");
__device__ __forceinline__ SplitQueue* split_queue(WsHeader* h) {
  return reinterpret_cast<SplitQueue*>(reinterpret_cast<char*>(h) + 256);
}

__device__ __forceinline__ uint64_t atomic_load_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Takes the next tile ticket (whole wave; returns a wave-uniform value).
__device__ __forceinline__ uint32_t take_ticket(WsHeader* h) {
  uint32_t t = 0;
  if (lane_id() == 0) t = __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return uniform(t);
}

// The next ticket of a launch with numTiles tiles, or numTiles once they
// are all taken.  Tickets only grow, so a read at or past numTiles means no
// work is left, and the workgroups of an empty launch (a split pass with
// nothing to split: thousands of them) leave without an atomic on the one
// ticket word.  The read costs a round trip per tile, so launches that
// always have work take tickets directly.  One thread.
__device__ __forceinline__ uint32_t take_ticket_below(WsHeader* h, uint32_t numTiles) {
  if (__hip_atomic_load(&h->ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= numTiles) return numTiles;
  return __hip_atomic_fetch_add(&h->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back over N counters for `tile`, executed by one full
// wave.  agg[] is this tile's wave-uniform aggregate; returns the exclusive
// prefix of every counter over tiles 0..tile-1 of this launch.  Tiles are
// handed out in ticket order, so every predecessor is held by a running
// wave that publishes its aggregate before it waits on anyone; the wait
// therefore ends.  Status words carry the launch epoch, so words from
// earlier launches read as "not yet published".
// `abort` (optional): a word another wave sets to end every wait early; the
// results are then invalid (the persistent split flags the launch).
//
// In two steps: lookback_publish writes the tile's aggregates of counters
// [0, N) (its inclusive prefixes for tile 0); lookback_walk then walks back
// over counters [C0, C0 + M) and publishes their inclusive prefixes.  A
// tile may walk its counters in several groups: every predecessor published
// all its aggregates before walking any group, so a later group never waits
// on an unpublished word, and each group holds only its own M counters in
// registers.
template <int N>
__device__ __forceinline__ void lookback_publish(uint64_t* status, uint32_t maxTiles, uint32_t epoch, uint32_t tile,
                                                 const uint64_t (&agg)[N]) {
  const uint32_t lane = lane_id();
  if (lane < (uint32_t)N) {
    uint64_t a = 0;
#pragma unroll
    for (int c = 0; c < N; c++)
      if ((uint32_t)c == lane) a = agg[c];
    atomic_store_agent(&status[lane * (uint64_t)maxTiles + tile], pack_status(epoch, tile == 0 ? 2 : 1, a));
  }
}

template <int C0, int M>
__device__ __forceinline__ void lookback_walk(uint64_t* status, uint32_t maxTiles, uint32_t epoch, uint32_t tile,
                                              const uint64_t (&agg)[M], uint64_t (&excl)[M],
                                              const uint32_t* abort = nullptr) {
  const uint32_t lane = lane_id();
  if (tile == 0) {
#pragma unroll
    for (int c = 0; c < M; c++) excl[c] = 0;
    return;
  }
  uint64_t sum[M];
  bool done[M];
#pragma unroll
  for (int c = 0; c < M; c++) sum[c] = 0, done[c] = false;

  int64_t pred = (int64_t)tile - 1 - (int64_t)lane;
  while (true) {
    uint64_t w[M];
    bool ok = true;
#pragma unroll
    for (int c = 0; c < M; c++) {
      if (pred >= 0) {
        w[c] = atomic_load_agent(&status[(C0 + c) * (uint64_t)maxTiles + (uint64_t)pred]);
        ok = ok && ((uint32_t)(w[c] >> 44) == epoch) && ((w[c] >> 42) & 3u) != 0;
      } else {
        w[c] = pack_status(epoch, 2, 0);  // before tile 0: inclusive zero
      }
    }
    if (ballot(!ok) != 0) {
      if (abort && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    bool all = true;
#pragma unroll
    for (int c = 0; c < M; c++) {
      if (done[c]) continue;
      uint64_t incl = ballot(((w[c] >> 42) & 3u) == 2);
      uint64_t val = w[c] & kValueMask;
      if (incl) {
        uint32_t first = (uint32_t)__builtin_ctzll(incl);
        sum[c] += wave_sum64(lane <= first ? val : 0);
        done[c] = true;
      } else {
        sum[c] += wave_sum64(val);
        all = false;
      }
    }
    if (all) break;
    pred -= kWave;
  }
#pragma unroll
  for (int c = 0; c < M; c++) excl[c] = uniform64(sum[c]);
  if (lane < (uint32_t)M) {
    uint64_t e = 0, a = 0;
#pragma unroll
    for (int c = 0; c < M; c++)
      if ((uint32_t)c == lane) e = excl[c], a = agg[c];
    atomic_store_agent(&status[(C0 + lane) * (uint64_t)maxTiles + tile], pack_status(epoch, 2, e + a));
  }
}

template <int N>
__device__ __forceinline__ void lookback(uint64_t* status, uint32_t maxTiles, uint32_t epoch, uint32_t tile,
                                         const uint64_t (&agg)[N], uint64_t (&excl)[N],
                                         const uint32_t* abort = nullptr) {
  lookback_publish<N>(status, maxTiles, epoch, tile, agg);
  lookback_walk<0, N>(status, maxTiles, epoch, tile, agg, excl, abort);
}

}  // namespace tc
