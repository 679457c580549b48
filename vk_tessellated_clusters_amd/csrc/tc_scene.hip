// tc_scene.hip -- synthetic height maps on the device.
//
// Tileable value-noise fBm with an integer lattice hash, the same
// arithmetic (FP64, evaluation order of the numpy generator in
// vk_tessellated_clusters_amd/scene.py) so both produce identical texels.
#include "tc_kernels.h"

namespace tc {

__device__ __forceinline__ uint32_t lattice_hash(uint32_t x, uint32_t y, uint32_t seed) {
  uint32_t h = (x * 0x8DA6B343u) ^ (y * 0xD8163841u) ^ seed;
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

__global__ void k_fbm(void* out, uint32_t size, uint32_t seed, uint32_t octaves, uint32_t baseFreq, uint32_t format) {
  const uint32_t X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y;
  if (X >= size || Y >= size) return;
  const double cx = ((double)X + 0.5) / (double)size, cy = ((double)Y + 0.5) / (double)size;
  double acc = 0.0, ampSum = 0.0;
  for (uint32_t o = 0; o < octaves; o++) {
    const uint32_t f = baseFreq << o;
    const double amp = ldexp(1.0, -(int)o);  // 0.5^o, exact
    const double x = cx * (double)f, y = cy * (double)f;
    const double xi = floor(x), yi = floor(y);
    const double xf = x - xi, yf = y - yi;
    const double sx = xf * xf * (3.0 - 2.0 * xf), sy = yf * yf * (3.0 - 2.0 * yf);
    const uint32_t x0 = (uint32_t)xi % f, x1 = ((uint32_t)xi + 1) % f;
    const uint32_t y0 = (uint32_t)yi % f, y1 = ((uint32_t)yi + 1) % f;
    const uint32_t s = seed * 0x27D4EB2Fu + o * 0x165667B1u;
    const double v00 = (double)lattice_hash(x0, y0, s) / 4294967296.0;
    const double v10 = (double)lattice_hash(x1, y0, s) / 4294967296.0;
    const double v01 = (double)lattice_hash(x0, y1, s) / 4294967296.0;
    const double v11 = (double)lattice_hash(x1, y1, s) / 4294967296.0;
    const double top = v00 * (1.0 - sx) + v10 * sx;
    const double bot = v01 * (1.0 - sx) + v11 * sx;
    acc = acc + amp * (top * (1.0 - sy) + bot * sy);
    ampSum = ampSum + amp;
  }
  const float h = (float)(acc / ampSum);
  const uint64_t i = (uint64_t)Y * size + X;
  if (format == TC_TEXTURE_R16_UNORM) {
    const float q = rintf(h * 65535.0f);
    reinterpret_cast<uint16_t*>(out)[i] = (uint16_t)fminf(fmaxf(q, 0.0f), 65535.0f);
  } else {
    reinterpret_cast<float*>(out)[i] = h;
  }
}

}  // namespace tc

extern "C" tc_status tc_fbm_height(void* texelsDev, uint32_t size, uint32_t seed, uint32_t octaves, uint32_t baseFreq,
                                   uint32_t format, tc_stream stream) {
  if (!texelsDev || !size || !octaves || !baseFreq || (format != TC_TEXTURE_R16_UNORM && format != TC_TEXTURE_R32_FLOAT))
    return TC_ERROR_INVALID_ARGUMENT;
  hipLaunchKernelGGL(tc::k_fbm, dim3((size + 255) / 256, size), dim3(256), 0, (hipStream_t)stream, texelsDev, size, seed,
                     octaves, baseFreq, format);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}
