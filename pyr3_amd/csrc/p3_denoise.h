// Learned-denoiser launch interface (kernels in p3_denoise.hip; network in
// pyr3_amd/denoise/net.py).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace p3 {

constexpr int kDenoiseLayers = 14;
// (c_in, c_out, output level) of the fourteen 3x3 layers e0 e1 d1 e2 d2 e3
// d3 b0 u3 c3 u2 c2 u1 c1, in weight order (level: 0 full, 1 half, 2
// quarter, 3 eighth resolution).
constexpr int kDenoiseShape[kDenoiseLayers][3] = {
    {8, 16, 0},    {16, 16, 0}, {16, 32, 1},  {32, 32, 1}, {32, 64, 2},
    {64, 64, 2},   {64, 128, 3}, {128, 128, 3}, {128, 64, 2}, {128, 64, 2},
    {64, 32, 1},   {64, 32, 1}, {32, 16, 0},  {32, 16, 0}};
constexpr int kDenoiseHead = 3 * 16 + 3;  // 1x1 head weights + bias
// Padded row length of a layer's packed weights: 9 c_in rounded up to 32.
constexpr int denoise_kp(int l) { return (9 * kDenoiseShape[l][0] + 31) / 32 * 32; }

struct DenoiseArgs {
  const uint16_t* w[kDenoiseLayers];  // packed bf16 [c_out][kp], k = tap * c_in + c
  const float* b[kDenoiseLayers];     // biases
  const float* head;                  // 3 x 16 weights then 3 biases
  const float4* img;                  // renderer's DE-off tonemapped RGBA f32 image
  const uint32_t* hist;               // renderer's RGBA histogram (oversample 1)
  uint32_t* out8;                     // RGBA8 output
  uint16_t* act[kDenoiseLayers - 1];  // activations of layers 0..12 (NHWC bf16)
  float k2;                           // the frame's calibration constant
  int width, height;
};

hipError_t launch_denoise(const DenoiseArgs& d, hipStream_t s);

}  // namespace p3
