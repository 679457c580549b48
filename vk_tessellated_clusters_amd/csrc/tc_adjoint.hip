// tc_adjoint.hip -- reverse-mode derivative of hit distance with respect to
// height-map texels (no reference counterpart).
//
// Forward chain per ray, with the tessellation pattern and the hit triangle
// frozen: texels -> bilinear height H(u_k, v_k) at each of the hit
// triangle's three generated vertices -> P_k = base_k + normal_k *
// (heightScale_k * H + heightOffset_k) -> t, the distance along the ray to
// the plane through P_0, P_1, P_2.  Differentiating o + t d = sum_k b_k P_k
// (sum b_k = 1) along the plane normal n gives dt/dP_k = b_k n / (n . d).
// The chain is evaluated in object space, where t equals the world-space
// distance because the direction is carried over linearly.
#include <algorithm>

#include "tc_device.h"

namespace tc {

constexpr uint32_t kAdjointThreads = 256;
int num_cu_public();

__global__ __launch_bounds__(kAdjointThreads) void k_depth_adjoint(tc_adjoint_params p) {
  for (uint32_t r = blockIdx.x * kAdjointThreads + threadIdx.x; r < p.numRays; r += gridDim.x * kAdjointThreads) {
    const float g = p.gradT[r];
    if (g == 0.0f) continue;
    const tc_hit h = p.hits[r];
    if (!(h.t >= 0.0f) || h.instanceID == 0xFFFFFFFFu) continue;
    const uint4 hv = reinterpret_cast<const uint4*>(p.hitVertices)[r];
    if (hv.x == 0xFFFFFFFFu) continue;
    const float* m = p.tlasInstances[h.instanceID].worldToObject;
    const tc_ray ray = p.rays[r];
    const v3 d = mk((m[0] * ray.direction[0] + m[1] * ray.direction[1]) + m[2] * ray.direction[2],
                    (m[4] * ray.direction[0] + m[5] * ray.direction[1]) + m[6] * ray.direction[2],
                    (m[8] * ray.direction[0] + m[9] * ray.direction[1]) + m[10] * ray.direction[2]);
    const uint32_t vi[3] = {hv.x, hv.y, hv.z};
    const uint64_t gv = reinterpret_cast<uint64_t>(p.genVertices);
    const v3 p0 = ld3(gv, vi[0]), p1 = ld3(gv, vi[1]), p2 = ld3(gv, vi[2]);
    const v3 n = cross(sub(p1, p0), sub(p2, p0));
    const float nd = dot(n, d);
    if (nd == 0.0f) continue;
    const float b[3] = {(1.0f - h.u) - h.v, h.u, h.v};
    const float gs = g / nd;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float4* tp = reinterpret_cast<const float4*>(p.tape + vi[k]);
      const float4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
      const uint32_t tex = __float_as_uint(q2.z);
      if (tex >= p.numTextures) continue;
      float* grad = p.gradTexels[tex];
      if (!grad) continue;
      const v3 nk = mk(q0.w, q1.x, q1.y);
      const float coef = ((gs * b[k]) * q2.x) * dot(n, nk);
      if (coef == 0.0f) continue;
      const tc_displacement_texture t = p.textures[tex];
      uint64_t idx[4];
      float w[4];
      bilinear_footprint(t.width, t.height, q1.z, q1.w, idx, w);
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (w[j] != 0.0f) unsafeAtomicAdd(grad + idx[j], coef * w[j]);
    }
  }
}

}  // namespace tc

using namespace tc;

extern "C" tc_status tc_depth_adjoint(const tc_adjoint_params* p, tc_stream stream) {
  if (!p) return TC_ERROR_INVALID_ARGUMENT;
  if (p->numRays == 0) return TC_OK;
  if (!p->rays || !p->hits || !p->hitVertices || !p->gradT || !p->genVertices || !p->tape || !p->tlasInstances ||
      (p->numTextures && (!p->textures || !p->gradTexels)))
    return TC_ERROR_INVALID_ARGUMENT;
  const uint32_t grid =
      std::min<uint32_t>((p->numRays + kAdjointThreads - 1) / kAdjointThreads, (uint32_t)num_cu_public() * 8);
  hipLaunchKernelGGL(k_depth_adjoint, dim3(grid), dim3(kAdjointThreads), 0, (hipStream_t)stream, *p);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}
