// Learned denoiser inference on gfx950: BF16 MFMA implicit-GEMM convolutions.
//
// Network (pyr3_amd/denoise/model.py): ten 3x3 convolutions in a 3-level
// U-Net (16/32/64 channels) and a 1x1 RGB head added to the input image.
// Activations are NHWC bf16 (channel-minor, 32/64/128 B per pixel); weights
// are bf16 in [tap][k-chunk][out channel][16 k] order so one lane reads its
// four k-values of one output channel with one 8-byte load.
//
// One wave computes 16 consecutive output pixels of one row for every output
// channel: for each of the 9 taps and each 16-channel chunk of the input it
// issues one v_mfma_f32_16x16x16_bf16 per 16 output channels with
//   A[m][k] = input(pixel m, channel k)   (lane l: m = l%16, k = 4(l/16)..+3)
//   B[k][n] = weight(out n, in k)         (lane l: n = l%16, k = 4(l/16)..+3)
//   D[m][n] accumulates in FP32           (lane l: m = 4(l/16)+i, n = l%16)
// then adds the bias, applies ReLU and stores bf16. Zero padding, stride 2,
// a nearest-neighbour 2x upsample of the input and a channel concat of two
// inputs are folded into the A-operand addressing. The first layer reads
// its 4 input channels (RGB of the renderer's f32 image, log1p(count/255)/4
// of its histogram) directly.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "p3_denoise.h"

namespace p3 {
namespace {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct ConvArgs {
  const uint16_t* in_a;   // NHWC bf16, ca channels (or null for the feature layer)
  const uint16_t* in_b;   // second input of a concat (cb channels), or null
  const uint16_t* w;      // [9][cin/16][cout][16] bf16
  const float* bias;      // cout
  uint16_t* out;          // NHWC bf16, cout channels
  const float4* img;      // feature layer: the f32 RGBA image
  const uint32_t* hist;   // feature layer: the RGBA u32 histogram (os 1)
  int in_w, in_h;         // input tensor size (low-res size when upsampling)
  int out_w, out_h;
};

// kCin = ca + cb (multiples of 16), kCout multiple of 16.
template <int kCa, int kCb, int kCout, int kStride, bool kUp, bool kFeat>
__global__ void __launch_bounds__(256) conv3x3_mfma(ConvArgs A) {
  constexpr int kCin = kCa + kCb;
  constexpr int kChunks = kCin / 16;
  constexpr int kNt = kCout / 16;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int y = blockIdx.y;
  const int x0 = (blockIdx.x * 4 + wave) * 16;
  if (x0 >= A.out_w) return;  // whole wave out of the row
  const int x = x0 + m;       // output pixel of this lane's A row
  v4f32 acc[kNt];
#pragma unroll
  for (int nt = 0; nt < kNt; ++nt) acc[nt] = v4f32{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    int iy, ix;
    bool ok;
    if (kUp) {  // conv over the 2x nearest upsample, cropped to the output size
      const int uy = y + dy, ux = x + dx;
      ok = uy >= 0 && uy < A.out_h && ux >= 0 && ux < A.out_w;
      iy = uy >> 1;
      ix = ux >> 1;
      ok = ok && iy < A.in_h && ix < A.in_w;
    } else {
      iy = y * kStride + dy;
      ix = x * kStride + dx;
      ok = x < A.out_w && iy >= 0 && iy < A.in_h && ix >= 0 && ix < A.in_w;
    }
    const size_t pix = ok ? (size_t)iy * A.in_w + ix : 0;
#pragma unroll
    for (int kc = 0; kc < kChunks; ++kc) {
      v4i16 a = v4i16{0, 0, 0, 0};
      if (kFeat) {
        if (ok && kq == 0) {
          const float4 c = A.img[pix];
          const float cnt = (float)A.hist[pix * 4 + 3];
          const float d = log1pf(cnt * (1.0f / 255.0f)) * 0.25f;
          a = v4i16{(short)to_bf16(c.x), (short)to_bf16(c.y), (short)to_bf16(c.z),
                    (short)to_bf16(d)};
        }
      } else if (ok) {
        const bool first = kc * 16 < kCa;
        const uint16_t* src = first ? A.in_a + pix * kCa + kc * 16
                                    : A.in_b + pix * (kCb > 0 ? kCb : 16) + (kc * 16 - kCa);
        const uint2 v = *(const uint2*)(src + 4 * kq);
        a = v4i16{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu),
                  (short)(v.y >> 16)};
      }
#pragma unroll
      for (int nt = 0; nt < kNt; ++nt) {
        const uint16_t* wp = A.w + (((size_t)(t * kChunks + kc) * kCout + nt * 16 + m) * 16 + 4 * kq);
        const uint2 v = *(const uint2*)wp;
        const v4i16 b = v4i16{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu),
                              (short)(v.y >> 16)};
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
      }
    }
  }
  // D[m'][n]: lane holds m' = 4 kq + i, n = m
  const size_t row = (size_t)y * A.out_w;
#pragma unroll
  for (int nt = 0; nt < kNt; ++nt) {
    const int n = nt * 16 + m;
    const float bn = A.bias[n];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int xo = x0 + 4 * kq + i;
      if (xo < A.out_w) {
        const float v = fmaxf(acc[nt][i] + bn, 0.0f);
        A.out[(row + xo) * kCout + n] = to_bf16(v);
      }
    }
  }
}

// 1x1 head: RGB = image RGB + W (3x16) . v + b, clamped, stored as RGBA8
// (round to nearest, like the tonemap store); alpha from the image.
__global__ void __launch_bounds__(256) head_rgba8(const uint16_t* __restrict__ v,
                                                  const float* __restrict__ w,  // 3x16 + 3
                                                  const float4* __restrict__ img,
                                                  uint32_t* __restrict__ out8, int n_pix) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n_pix) return;
  const uint4 a = *(const uint4*)(v + (size_t)p * 16);
  const uint4 b = *(const uint4*)(v + (size_t)p * 16 + 8);
  const uint32_t words[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  float f[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    f[2 * i] = from_bf16((uint16_t)(words[i] & 0xffffu));
    f[2 * i + 1] = from_bf16((uint16_t)(words[i] >> 16));
  }
  const float4 c = img[p];
  float rgb[3] = {c.x, c.y, c.z};
  uint32_t px = 0u;
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    float s = w[48 + o];
#pragma unroll
    for (int i = 0; i < 16; ++i) s += w[o * 16 + i] * f[i];
    const float r = fminf(fmaxf(rgb[o] + s, 0.0f), 1.0f);
    px |= (uint32_t)(r * 255.0f + 0.5f) << (8 * o);
  }
  const float al = fminf(fmaxf(c.w, 0.0f), 1.0f);
  px |= (uint32_t)(al * 255.0f + 0.5f) << 24;
  out8[p] = px;
}

template <int kCa, int kCb, int kCout, int kStride, bool kUp, bool kFeat>
hipError_t run(const ConvArgs& a, hipStream_t s) {
  const dim3 grid((a.out_w + 63) / 64, a.out_h);
  hipLaunchKernelGGL((conv3x3_mfma<kCa, kCb, kCout, kStride, kUp, kFeat>), grid, dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_denoise(const DenoiseArgs& d, hipStream_t s) {
  const int W = d.width, H = d.height;
  const int W2 = (W + 1) / 2, H2 = (H + 1) / 2, W4 = (W2 + 1) / 2, H4 = (H2 + 1) / 2;
  hipError_t e;
  ConvArgs a{};
  // e0: features -> 16 (full)
  a = ConvArgs{nullptr, nullptr, d.w[0], d.b[0], d.s0a, d.img, d.hist, W, H, W, H};
  if ((e = run<16, 0, 16, 1, false, true>(a, s)) != hipSuccess) return e;
  // e1: 16 -> 16 (full) = skip s0
  a = ConvArgs{d.s0a, nullptr, d.w[1], d.b[1], d.s0, nullptr, nullptr, W, H, W, H};
  if ((e = run<16, 0, 16, 1, false, false>(a, s)) != hipSuccess) return e;
  // d0: 16 -> 32 stride 2 (half)
  a = ConvArgs{d.s0, nullptr, d.w[2], d.b[2], d.h1a, nullptr, nullptr, W, H, W2, H2};
  if ((e = run<16, 0, 32, 2, false, false>(a, s)) != hipSuccess) return e;
  // d1: 32 -> 32 (half) = skip s1
  a = ConvArgs{d.h1a, nullptr, d.w[3], d.b[3], d.s1, nullptr, nullptr, W2, H2, W2, H2};
  if ((e = run<32, 0, 32, 1, false, false>(a, s)) != hipSuccess) return e;
  // b0: 32 -> 64 stride 2 (quarter)
  a = ConvArgs{d.s1, nullptr, d.w[4], d.b[4], d.q0, nullptr, nullptr, W2, H2, W4, H4};
  if ((e = run<32, 0, 64, 2, false, false>(a, s)) != hipSuccess) return e;
  // b1: 64 -> 64 (quarter)
  a = ConvArgs{d.q0, nullptr, d.w[5], d.b[5], d.q1, nullptr, nullptr, W4, H4, W4, H4};
  if ((e = run<64, 0, 64, 1, false, false>(a, s)) != hipSuccess) return e;
  // u0: up2(64) -> 32 (half)
  a = ConvArgs{d.q1, nullptr, d.w[6], d.b[6], d.h1a, nullptr, nullptr, W4, H4, W2, H2};
  if ((e = run<64, 0, 32, 1, true, false>(a, s)) != hipSuccess) return e;
  // u1: cat(u0, s1) 64 -> 32 (half)
  a = ConvArgs{d.h1a, d.s1, d.w[7], d.b[7], d.h1b, nullptr, nullptr, W2, H2, W2, H2};
  if ((e = run<32, 32, 32, 1, false, false>(a, s)) != hipSuccess) return e;
  // v0: up2(32) -> 16 (full)
  a = ConvArgs{d.h1b, nullptr, d.w[8], d.b[8], d.s0a, nullptr, nullptr, W2, H2, W, H};
  if ((e = run<32, 0, 16, 1, true, false>(a, s)) != hipSuccess) return e;
  // v1: cat(v0, s0) 32 -> 16 (full)
  a = ConvArgs{d.s0a, d.s0, d.w[9], d.b[9], d.v1, nullptr, nullptr, W, H, W, H};
  if ((e = run<16, 16, 16, 1, false, false>(a, s)) != hipSuccess) return e;
  const int n = W * H;
  hipLaunchKernelGGL(head_rgba8, dim3((n + 255) / 256), dim3(256), 0, s, d.v1, d.head, d.img,
                     d.out8, n);
  return hipGetLastError();
}

}  // namespace p3
