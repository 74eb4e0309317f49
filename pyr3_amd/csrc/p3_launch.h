// Launch-side interface between the C ABI (p3_api.cpp) and the kernels:
// argument blocks passed by value to each kernel, and the launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3 {

struct ChaosArgs {
  const float* xforms;
  const uint32_t* distrib;     // rows 0..n-1 then the fallback row
  const float4* palette;       // 256 RGBA
  uint32_t* hist;              // W x H x 4
  const uint32_t* isaac_in;    // walkers x 36 or null (seed on device)
  float* trace;
  uint32_t fallback_row;
  uint32_t width, height, iters, fuse, seed, walkers;
  int32_t final_idx;
  float scale, cx, cy, cos_r, sin_r, jitter, xform_blend;
  uint32_t palette_mode;
  uint32_t trace_walkers, trace_iters;
};

struct DeArgs {
  const uint4* hist;     // W x H RGBA u32
  float4* pre;           // W x H
  uint8_t* rad;          // W x H
  uint32_t* tile_rmax;   // tiles_x x tiles_y
  float4* out;           // W x H filtered
  const float* lut;      // 257 kernel-sum normalisers
  uint32_t width, height, tiles_x, tiles_y;
  int R;                 // ceil(clamp(max_rad, 0, 256))
  float max_rad, min_rad, curve, k1, k2;
};

struct TmArgs {
  const float4* filtered;   // DE on
  const uint4* hist;        // DE off
  const float* taps;
  uint32_t* out8;           // W x H packed RGBA8
  float4* outf;             // optional W x H RGBA f32
  uint32_t width, height, os, fwidth;
  float k1, k2, gamma, vibrancy, highpow, linrange;
  float bg0, bg1, bg2;
  uint32_t transparent;
};

hipError_t launch_chaos(const ChaosArgs& a, bool saturating, hipStream_t stream);
hipError_t launch_density(const DeArgs& a, hipStream_t stream);
hipError_t launch_tonemap(const TmArgs& a, bool de, hipStream_t stream);
hipError_t launch_fill(void* p, size_t bytes, uint32_t v, hipStream_t stream);

}  // namespace p3
