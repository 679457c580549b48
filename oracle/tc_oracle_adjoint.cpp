/*
 * tc_oracle_adjoint.cpp -- FP64 checker for the depth adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (see tc_oracle.h).  No reference counterpart:
 * the reference renderer is not differentiable.  These functions define the
 * frozen forward map the adjoint differentiates -- for every ray, the
 * distance to the plane of its hit triangle, whose three vertices are
 * re-displaced in FP64 from their tc_vertex_tape records and the height
 * maps -- and evaluate its gradient two ways: analytically (barycentrics
 * recomputed in FP64 from the hit point) and by central differences of the
 * L2 depth loss.
 */
#include <math.h>
#include <string.h>

#include <unordered_map>
#include <vector>

#include "tc_oracle.h"

namespace {

struct D3 {
  double x, y, z;
};
D3 d3(double x, double y, double z) { return {x, y, z}; }
D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 scl(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

int32_t wrap(int32_t i, int32_t n) {
  int32_t r = i % n;
  return r < 0 ? r + n : r;
}

/* Texel indices and weights of the FP32 bilinear REPEAT sample at (u, v):
 * the coordinate arithmetic is the sampler's own (FP32), the weights are
 * carried in FP64. */
void footprint(uint32_t W, uint32_t H, float u, float v, uint64_t idx[4], double w[4]) {
  float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
  x = fminf(fmaxf(x, -16777216.0f), 16777216.0f);
  y = fminf(fmaxf(y, -16777216.0f), 16777216.0f);
  const float x0 = floorf(x), y0 = floorf(y);
  const double fx = (double)(x - x0), fy = (double)(y - y0);
  const int32_t ix = (int32_t)x0, iy = (int32_t)y0;
  const uint64_t xa = (uint64_t)wrap(ix, (int32_t)W), xb = (uint64_t)wrap(ix + 1, (int32_t)W);
  const uint64_t ya = (uint64_t)wrap(iy, (int32_t)H), yb = (uint64_t)wrap(iy + 1, (int32_t)H);
  idx[0] = ya * W + xa, idx[1] = ya * W + xb, idx[2] = yb * W + xa, idx[3] = yb * W + xb;
  w[0] = (1 - fx) * (1 - fy), w[1] = fx * (1 - fy), w[2] = (1 - fx) * fy, w[3] = fx * fy;
}

struct Problem {
  const tc_ray* rays;
  const tc_hit* hits;
  const uint32_t* hitVertices;
  const tc_vertex_tape* tape;
  const oracle_height_map* maps;
  uint32_t numMaps;
  const float* xforms; /* 24 per instance: objectToWorld | worldToObject */
};

bool has_map(const Problem& P, const tc_vertex_tape& t) { return t.texture < P.numMaps && P.maps[t.texture].texels; }

/* vertex position with texel (pmap, ptexel) offset by delta */
D3 vertex64(const Problem& P, uint32_t vi, uint32_t pmap, uint64_t ptexel, double delta) {
  const tc_vertex_tape& t = P.tape[vi];
  const D3 base = d3(t.base[0], t.base[1], t.base[2]);
  if (t.texture == 0xFFFFFFFFu) return base;
  double h = 0.0;
  if (has_map(P, t)) {
    const oracle_height_map& m = P.maps[t.texture];
    uint64_t idx[4];
    double w[4];
    footprint(m.width, m.height, t.u, t.v, idx, w);
    for (int j = 0; j < 4; j++) {
      double tv = (double)m.texels[idx[j]];
      if (t.texture == pmap && idx[j] == ptexel) tv += delta;
      h += w[j] * tv;
    }
  }
  return add(base, scl(d3(t.normal[0], t.normal[1], t.normal[2]), h * (double)t.heightScale + (double)t.heightOffset));
}

struct RayObj {
  D3 o, d;
};

RayObj ray_object(const Problem& P, uint32_t r) {
  const float* m = P.xforms + 24ull * P.hits[r].instanceID + 12;
  const tc_ray& ray = P.rays[r];
  RayObj q;
  double o[3], d[3];
  for (int i = 0; i < 3; i++) {
    o[i] = (double)m[4 * i] * ray.origin[0] + (double)m[4 * i + 1] * ray.origin[1] + (double)m[4 * i + 2] * ray.origin[2] +
           (double)m[4 * i + 3];
    d[i] = (double)m[4 * i] * ray.direction[0] + (double)m[4 * i + 1] * ray.direction[1] +
           (double)m[4 * i + 2] * ray.direction[2];
  }
  q.o = d3(o[0], o[1], o[2]);
  q.d = d3(d[0], d[1], d[2]);
  return q;
}

bool hit_valid(const Problem& P, uint32_t r) {
  return P.hits[r].t >= 0.0f && P.hits[r].instanceID != 0xFFFFFFFFu && P.hitVertices[4ull * r] != 0xFFFFFFFFu;
}

double t64(const Problem& P, uint32_t r, uint32_t pmap, uint64_t ptexel, double delta) {
  const uint32_t* hv = P.hitVertices + 4ull * r;
  const D3 p0 = vertex64(P, hv[0], pmap, ptexel, delta), p1 = vertex64(P, hv[1], pmap, ptexel, delta),
           p2 = vertex64(P, hv[2], pmap, ptexel, delta);
  const RayObj q = ray_object(P, r);
  const D3 n = cross(sub(p1, p0), sub(p2, p0));
  return dot(n, sub(p0, q.o)) / dot(n, q.d);
}

Problem make(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, const tc_vertex_tape* tape,
             const oracle_height_map* maps, uint32_t numMaps, const float* xforms) {
  return Problem{rays, hits, hitVertices, tape, maps, numMaps, xforms};
}

}  // namespace

extern "C" {

void oracle_depth_t64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                      const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps, const float* xforms,
                      uint32_t perturbMap, uint64_t perturbTexel, double delta, double* t) {
  const Problem P = make(rays, hits, hitVertices, tape, maps, numMaps, xforms);
  for (uint32_t r = 0; r < numRays; r++) t[r] = hit_valid(P, r) ? t64(P, r, perturbMap, perturbTexel, delta) : -1.0;
}

void oracle_depth_adjoint64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                            const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps,
                            const float* xforms, const double* gradT, double* const* grads) {
  const Problem P = make(rays, hits, hitVertices, tape, maps, numMaps, xforms);
  for (uint32_t r = 0; r < numRays; r++) {
    if (gradT[r] == 0.0 || !hit_valid(P, r)) continue;
    const uint32_t* hv = P.hitVertices + 4ull * r;
    const D3 p[3] = {vertex64(P, hv[0], ~0u, 0, 0), vertex64(P, hv[1], ~0u, 0, 0), vertex64(P, hv[2], ~0u, 0, 0)};
    const RayObj q = ray_object(P, r);
    const D3 e1 = sub(p[1], p[0]), e2 = sub(p[2], p[0]);
    const D3 n = cross(e1, e2);
    const double nd = dot(n, q.d);
    if (nd == 0.0) continue;
    const double t = dot(n, sub(p[0], q.o)) / nd;
    const D3 x = sub(add(q.o, scl(q.d, t)), p[0]);
    const double nn = dot(n, n);
    const double b1 = dot(cross(x, e2), n) / nn, b2 = dot(cross(e1, x), n) / nn;
    const double b[3] = {1.0 - b1 - b2, b1, b2};
    for (int k = 0; k < 3; k++) {
      const tc_vertex_tape& tk = P.tape[hv[k]];
      if (!has_map(P, tk) || !grads[tk.texture]) continue;
      const double coef =
          gradT[r] * b[k] * (double)tk.heightScale * dot(n, d3(tk.normal[0], tk.normal[1], tk.normal[2])) / nd;
      const oracle_height_map& m = P.maps[tk.texture];
      uint64_t idx[4];
      double w[4];
      footprint(m.width, m.height, tk.u, tk.v, idx, w);
      for (int j = 0; j < 4; j++) grads[tk.texture][idx[j]] += coef * w[j];
    }
  }
}

void oracle_depth_loss_fd64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                            const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps,
                            const float* xforms, const float* target, double scale, uint32_t map, const uint64_t* texels,
                            uint32_t count, double eps, double* out) {
  const Problem P = make(rays, hits, hitVertices, tape, maps, numMaps, xforms);
  std::unordered_map<uint64_t, uint32_t> slot;
  for (uint32_t i = 0; i < count; i++) slot.emplace(texels[i], i);
  std::vector<std::vector<uint32_t>> affected(count);
  for (uint32_t r = 0; r < numRays; r++) {
    if (!hit_valid(P, r) || !(target[r] >= 0.0f)) continue;
    const uint32_t* hv = P.hitVertices + 4ull * r;
    for (int k = 0; k < 3; k++) {
      const tc_vertex_tape& tk = P.tape[hv[k]];
      if (tk.texture != map || !has_map(P, tk)) continue;
      uint64_t idx[4];
      double w[4];
      footprint(P.maps[map].width, P.maps[map].height, tk.u, tk.v, idx, w);
      for (int j = 0; j < 4; j++) {
        auto it = slot.find(idx[j]);
        if (it == slot.end()) continue;
        std::vector<uint32_t>& a = affected[it->second];
        if (a.empty() || a.back() != r) a.push_back(r);
      }
    }
  }
  for (uint32_t i = 0; i < count; i++) {
    double acc = 0.0;
    for (uint32_t r : affected[i]) {
      const double tp = t64(P, r, map, texels[i], eps) - (double)target[r];
      const double tm = t64(P, r, map, texels[i], -eps) - (double)target[r];
      acc += tp * tp - tm * tm;
    }
    out[i] = scale * acc / (2.0 * eps);
  }
}

}  /* extern "C" */
