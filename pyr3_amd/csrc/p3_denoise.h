// Learned-denoiser launch interface (kernels in p3_denoise.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace p3 {

constexpr int kDenoiseLayers = 10;
// (c_in, c_out, stride) of the ten 3x3 layers, in weight order.
constexpr int kDenoiseShape[kDenoiseLayers][3] = {
    {4, 16, 1}, {16, 16, 1}, {16, 32, 2}, {32, 32, 1}, {32, 64, 2},
    {64, 64, 1}, {64, 32, 1}, {64, 32, 1}, {32, 16, 1}, {32, 16, 1}};
constexpr int kDenoiseHead = 3 * 16 + 3;  // 1x1 head weights + bias

struct DenoiseArgs {
  const uint16_t* w[kDenoiseLayers];  // packed bf16 [9][cin/16][cout][16]
  const float* b[kDenoiseLayers];     // biases
  const float* head;                  // 3x16 weights then 3 biases
  const float4* img;                  // renderer's RGBA f32 image
  const uint32_t* hist;               // renderer's RGBA histogram (oversample 1)
  uint32_t* out8;                     // RGBA8 output
  uint16_t *s0a, *s0, *h1a, *s1, *h1b, *q0, *q1;  // activations
  int width, height;
};

hipError_t launch_denoise(const DenoiseArgs& d, hipStream_t s);

}  // namespace p3
