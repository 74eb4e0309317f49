// Variation catalog for the gfx950 chaos kernel.
//
// Each entry follows the FP32 WGSL function of the same name in
// src/shaders/chaos.wgsl (cited per entry) with the same operation order
// (the library is built with -ffp-contract=off, so a*b+c keeps two
// roundings). Dispatch is by catalog index (src/variations.ts:24 V table).
#pragma once
#include "p3_device.h"

namespace p3 {

struct VarIn {
  float w;          // slot weight
  float p[10];      // param0..param9
  const float* xf;  // owning xform record (affine rows for waves/popcorn/rings/fan)
};

// chaos.wgsl:524-527
__device__ inline float2 var_linear(float2 p, float w) { return make_float2(p.x * w, p.y * w); }

// chaos.wgsl:528-531
__device__ inline float2 var_sinusoidal(float2 p, float w) {
  return make_float2(w * safe_sin(p.x), w * safe_sin(p.y));
}

// chaos.wgsl:532-536
__device__ inline float2 var_spherical(float2 p, float w) {
  const float r2 = (p.x * p.x + p.y * p.y) + kEps;
  const float s = w / r2;
  return make_float2(p.x * s, p.y * s);
}

__device__ inline float2 apply_variation(int idx, float2 p, const VarIn& a, Rng& rng) {
  (void)rng;
  switch (idx) {
    case 0: return var_linear(p, a.w);
    case 1: return var_sinusoidal(p, a.w);
    case 2: return var_spherical(p, a.w);
    default: return make_float2(0.0f, 0.0f);
  }
}

// Direct-colour override (chaos.wgsl:8192-8210). No dc_* entry is in the
// catalog yet, so no slot overrides the palette.
__device__ inline bool dc_color_for(int idx, float2 p, const VarIn& a, float3& rgb) {
  (void)idx; (void)p; (void)a; (void)rgb;
  return false;
}

}  // namespace p3
