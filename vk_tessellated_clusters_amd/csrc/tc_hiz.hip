// tc_hiz.hip -- the far HiZ pyramid of a depth image (nvhiz_vk.cpp:278-320,
// nvhiz-update.comp.glsl with hizFarLevel 0 and standard depth): level 0
// holds the maximum of each 2 x 2 block of the depth image, every further
// level the maximum of 2 x 2 texels of the level before.  One launch per
// level; reads of the source image clamp to its last row / column, so the
// whole power-of-two allocation is defined.
#include <algorithm>

#include "tc_kernels.h"

namespace tc {

__global__ void k_hiz_fill(float* t, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    t[i] = 1.0f;
}

__global__ void k_hiz_level0(tc_hiz h, const char* depth, uint32_t stride) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= h.size || y >= h.size) return;
  const uint32_t x0 = min(2 * x, h.width - 1), x1 = min(2 * x + 1, h.width - 1);
  const uint32_t y0 = min(2 * y, h.height - 1), y1 = min(2 * y + 1, h.height - 1);
  auto at = [&](uint32_t px, uint32_t py) {
    return *reinterpret_cast<const float*>(depth + ((uint64_t)py * h.width + px) * stride);
  };
  const float m = fmaxf(fmaxf(fmaxf(at(x0, y0), at(x1, y0)), at(x0, y1)), at(x1, y1));
  h.texels[(uint64_t)y * h.size + x] = m;
}

__global__ void k_hiz_down(tc_hiz h, uint32_t level) {
  const uint32_t dim = h.size >> level;
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dim || y >= dim) return;
  const float* src = h.texels + hiz_level_offset(h.size, level - 1);
  float* dst = h.texels + hiz_level_offset(h.size, level);
  const uint64_t sd = 2ull * dim;
  const float m = fmaxf(fmaxf(fmaxf(src[2ull * y * sd + 2 * x], src[2ull * y * sd + 2 * x + 1]),
                              src[(2ull * y + 1) * sd + 2 * x]),
                        src[(2ull * y + 1) * sd + 2 * x + 1]);
  dst[(uint64_t)y * dim + x] = m;
}

static bool hiz_ok(const tc_hiz* h) {
  return h && h->texels && h->size && !(h->size & (h->size - 1)) && h->mipLevels >= 1 && h->mipLevels <= 16 &&
         (h->size >> (h->mipLevels - 1)) == 1 && h->width && h->height;
}

}  // namespace tc

using namespace tc;

extern "C" {

tc_status tc_hiz_setup(uint32_t width, uint32_t height, tc_hiz* hiz, uint64_t* bytes, float sizeFactors[4],
                       float* sizeMax) {
  // at least 3 used texels per axis, so the clamp factors (used - 2) stay
  // positive
  if (!hiz || width < 6 || height < 6 || width > (1u << 16) || height > (1u << 16)) return TC_ERROR_INVALID_ARGUMENT;
  // nvhiz_vk.cpp:290-310 with divisor 2 << hizFarLevel = 2
  const uint32_t dim = std::max(width, height) / 2;
  uint32_t size = 1, mips = 1;
  while (size < dim) size *= 2, mips++;
  hiz->size = size;
  hiz->mipLevels = mips;
  hiz->usedWidth = width / 2;
  hiz->usedHeight = height / 2;
  hiz->width = width;
  hiz->height = height;
  uint64_t n = 0;
  for (uint32_t k = 0; k < mips; k++) n += (uint64_t)(size >> k) * (size >> k);
  if (bytes) *bytes = 4 * n;
  // nvhiz_vk.cpp:29-40 getShaderFactors / getSizeMax
  if (sizeFactors) {
    sizeFactors[0] = float(hiz->usedWidth) / float(size);
    sizeFactors[1] = float(hiz->usedHeight) / float(size);
    sizeFactors[2] = float(hiz->usedWidth - 2) / float(size);
    sizeFactors[3] = float(hiz->usedHeight - 2) / float(size);
  }
  if (sizeMax) *sizeMax = float(size);
  return TC_OK;
}

tc_status tc_hiz_clear(const tc_hiz* hiz, tc_stream stream) {
  if (!hiz_ok(hiz)) return TC_ERROR_INVALID_ARGUMENT;
  uint64_t n = 0;
  for (uint32_t k = 0; k < hiz->mipLevels; k++) n += (uint64_t)(hiz->size >> k) * (hiz->size >> k);
  hipLaunchKernelGGL(k_hiz_fill, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 4096)), dim3(256), 0, (hipStream_t)stream,
                     hiz->texels, n);
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

tc_status tc_hiz_build(const tc_hiz* hiz, const void* depth, uint32_t pixelStride, tc_stream stream) {
  if (!hiz_ok(hiz) || !depth || pixelStride < 4 || (pixelStride & 3) || (reinterpret_cast<uintptr_t>(depth) & 3))
    return TC_ERROR_INVALID_ARGUMENT;
  hipStream_t st = (hipStream_t)stream;
  const tc_hiz h = *hiz;
  hipLaunchKernelGGL(k_hiz_level0, dim3((h.size + 255) / 256, h.size), dim3(256), 0, st, h, (const char*)depth, pixelStride);
  for (uint32_t k = 1; k < h.mipLevels; k++) {
    const uint32_t dim = h.size >> k;
    hipLaunchKernelGGL(k_hiz_down, dim3((dim + 255) / 256, dim), dim3(256), 0, st, h, k);
  }
  return hipGetLastError() == hipSuccess ? TC_OK : TC_ERROR_LAUNCH;
}

}  // extern "C"
