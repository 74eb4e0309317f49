// Launch-side interface between the C ABI (p3_api.cpp) and the kernels:
// argument blocks passed by value to each kernel, and the launchers. The
// chaos kernel is compiled per genome shape at run time (p3_rtc.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "p3_args.h"

namespace p3 {

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
  const float* curves;      // 5 x 256 channel-curve LUT (read when curves_mask != 0)
  uint32_t curves_mask;     // bit 0 composite, 1 R, 2 G, 3 B, 4 luma
  uint32_t hsl_on;
  float hsl_hue, hsl_sat, hsl_light;   // degrees, factor, offset
};

struct LinArgs {
  const uint4* hist;        // W*os x H*os x 4 u32
  float4* out;              // W x H RGBA f32
  uint32_t width, height, os;
  double scale;             // (k1 * k2) / os^2
};
hipError_t launch_linear_export(const LinArgs& a, hipStream_t stream);
hipError_t launch_hist_merge(uint32_t* hist, uint64_t* acc, uint64_t* carry, size_t bins,
                             hipStream_t stream);
hipError_t launch_atomic_probe(uint64_t* cells_base, uint32_t cells, uint32_t blocks,
                               uint32_t per_thread, uint32_t seed, hipStream_t stream);

// One genome shape of a union kernel: a record blob of that shape (host
// memory), its regular xform count and final slot (-1: none).
struct ShapeRef {
  const float* xforms;
  int n_regular;
  int final_slot;
};
// Union kernel of `shapes` (each shape once, in case order): key, source,
// compile (cached like the one-shape kernels) and the batch launch of
// `n_blocks` blocks over the sub-dispatches of `bargs`.
std::string union_key(const std::vector<ShapeRef>& shapes);
std::string union_source(const std::vector<ShapeRef>& shapes);
bool union_code(const std::vector<ShapeRef>& shapes, const std::vector<char>** out,
                std::string& err);
hipError_t launch_union_batch(const BatchArgs& bargs, uint32_t n_blocks,
                              const std::vector<ShapeRef>& shapes, hipStream_t stream,
                              std::string& err);

// Chaos kernel specialised for the shape of `host_xforms` (p3_rtc.cpp).
hipError_t launch_chaos_rtc(const ChaosArgs& a, const float* host_xforms, int n_regular,
                            int final_slot, hipStream_t stream, std::string& err);
// True when the shape's module is loaded on device `dev`.
bool chaos_module_loaded(int dev, const float* host_xforms, int n_regular, int final_slot);
// As launch_chaos_rtc when the shape's module is already loaded on the
// current device; hipErrorNotReady otherwise (never compiles or loads).
hipError_t launch_chaos_rtc_loaded(const ChaosArgs& a, const float* host_xforms, int n_regular,
                                   int final_slot, hipStream_t stream);
// Batched launch of sub-dispatches sharing the shape of `host_xforms`.
hipError_t launch_chaos_rtc_multi(const MultiArgs& margs, const float* host_xforms, int n_regular,
                                  int final_slot, hipStream_t stream, std::string& err);
bool chaos_code(const float* host_xforms, int n_regular, int final_slot,
                const std::vector<char>** out, std::string& err);
// True when this shape's specialised kernel is compiled (in this process or
// in the on-disk code-object cache); never compiles.
bool chaos_code_ready(const float* host_xforms, int n_regular, int final_slot);
std::string chaos_source(const float* host_xforms, int n_regular, int final_slot);
std::string chaos_shape_key(const float* host_xforms, int n_regular, int final_slot);
size_t chaos_compiled_shapes();
int chaos_block();  // walkers per chaos block
hipError_t launch_density(const DeArgs& a, hipStream_t stream);
hipError_t launch_tonemap(const TmArgs& a, bool de, hipStream_t stream);
hipError_t launch_fill(void* p, size_t bytes, uint32_t v, hipStream_t stream);

}  // namespace p3
