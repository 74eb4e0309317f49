// Device helpers shared by the denoiser's training kernels (p3_nn.hip) and
// its inference kernels (p3_denoise.hip), so both compute the network input
// and the bf16 rounding identically.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace p3nn {

// f32 -> bf16 bits, round to nearest even (NaN stays NaN).
__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// The network input of one pixel (8 channels) from its histogram bin (u32
// RGBA, oversample 1) and the DE-off tonemapped f32 RGBA image: base RGB,
// density log1p(16 count k2) / 4 (count k2 is the bin's share of the
// frame's samples per unit area, so the feature does not depend on the
// quality), mean splat colour RGB (0..255 units / count) and occupancy.
__device__ inline uint4 pixel_features(const uint4 h, const float4 base, float k2) {
  const float cnt = (float)h.w;
  const float inv = cnt > 0.0f ? 1.0f / cnt : 0.0f;
  const float f3 = log1pf(16.0f * cnt * k2) * 0.25f;
  uint4 v;
  v.x = to_bf16(base.x) | ((uint32_t)to_bf16(base.y) << 16);
  v.y = to_bf16(base.z) | ((uint32_t)to_bf16(f3) << 16);
  v.z = to_bf16((float)h.x * inv) | ((uint32_t)to_bf16((float)h.y * inv) << 16);
  v.w = to_bf16((float)h.z * inv) | ((uint32_t)to_bf16(cnt > 0.0f ? 1.0f : 0.0f) << 16);
  return v;
}

}  // namespace p3nn
