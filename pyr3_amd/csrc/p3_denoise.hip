// Learned denoiser inference on gfx950: BF16 MFMA implicit-GEMM convolutions.
//
// Network (pyr3_amd/denoise/model.py): ten 3x3 convolutions in a 3-level
// U-Net (16/32/64 channels) and a 1x1 RGB head added to the input image.
// Activations are NHWC bf16 (channel-minor, 32/64/128 B per pixel); weights
// are bf16 in [tap][k-chunk][out channel][16 k] order so one lane reads its
// four k-values of one output channel with one 8-byte load.
//
// A block of 4 waves computes kTX = 64 * kNG consecutive output pixels of one
// row for every output channel. It first stages the three input rows the
// 3x3 window touches (kTX * stride + halo columns, all input channels) in
// LDS with 16-byte coalesced loads, folding in the zero padding, stride 2,
// the nearest-neighbour 2x upsample of a low-resolution input (staged at
// output-pixel resolution) and the channel concat of two inputs; pixels are
// kCin + 8 bf16 apart in LDS so the 16 lanes of a fragment read hit distinct
// banks. Each wave owns one 16-channel tile of the outputs and a share of the
// block's 16-pixel groups; it holds that tile's B fragments for all 9 taps
// and input chunks in registers (loaded once, before the staging), and for
// each tap and chunk reads one A fragment per group from LDS and issues
//   v_mfma_f32_16x16x16_bf16 with
//   A[m][k] = input(pixel m, channel k)   (lane l: m = l%16, k = 4(l/16)..+3)
//   B[k][n] = weight(out n, in k)         (lane l: n = l%16, k = 4(l/16)..+3)
//   D[m][n] accumulates in FP32           (lane l: m = 4(l/16)+i, n = l%16).
// Bias and ReLU are applied, the bf16 results go through LDS (reusing the
// input tile) and leave as 16-byte coalesced stores. The last layer instead
// feeds the 1x1 head from LDS and stores the RGBA8 image. The first layer
// builds its 4 input channels (RGB of the renderer's f32 image,
// log1p(count/255)/4 of its histogram) while staging.
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
  uint16_t* out;          // NHWC bf16, cout channels (null when the head is fused)
  const float4* img;      // feature layer and head: the f32 RGBA image
  const uint32_t* hist;   // feature layer: the RGBA u32 histogram (os 1)
  const float* head;      // fused head: 3x16 weights then 3 biases
  uint32_t* out8;         // fused head: RGBA8 output
  int in_w, in_h;         // input tensor size (low-res size when upsampling)
  int out_w, out_h;
};

// Block tile: kTY output rows x kTX output pixels. Staged input: rows and
// columns the 3x3 window reaches, at input resolution (for the upsampling
// layers: low-resolution rows/columns (y0/2 - 1) .. (y0 + kTY)/2).
template <int kCin, int kCout, int kStride, bool kUp, int kTX, int kTY>
struct ConvTile {
  static constexpr int kSX = kUp ? kTX / 2 + 2 : (kTX - 1) * kStride + 3;  // staged columns
  static constexpr int kSY = kUp ? kTY / 2 + 2 : (kTY - 1) * kStride + 3;  // staged rows
  static constexpr int kPS = kCin + 8;                                     // LDS pixel stride
  static constexpr int kOS = kCout + 8;                                    // output pixel stride
  static constexpr int kIn = kSY * kSX * kPS, kOut = kTY * kTX * kOS;
  static constexpr int kWords = kIn > kOut ? kIn : kOut;                   // bf16 elements
};

// One 16-byte quad (8 channels) of the staged input pixel (row r, column c)
// of the tile at output (y0, x0); zero outside the input.
template <int kCa, int kCb, int kStride, bool kUp, bool kFeat>
__device__ inline uint4 stage_quad(const ConvArgs& A, int y0, int x0, int r, int c, int q) {
  const int iy = kUp ? (y0 >> 1) - 1 + r : y0 * kStride - 1 + r;
  const int ix = kUp ? (x0 >> 1) - 1 + c : x0 * kStride - 1 + c;
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  if (iy < 0 || iy >= A.in_h || ix < 0 || ix >= A.in_w) return v;
  const size_t pix = (size_t)iy * A.in_w + ix;
  if (kFeat) {
    if (q == 0) {
      const float4 c4 = A.img[pix];
      const float cnt = (float)A.hist[pix * 4 + 3];
      const float d = log1pf(cnt * (1.0f / 255.0f)) * 0.25f;
      v.x = (uint32_t)to_bf16(c4.x) | ((uint32_t)to_bf16(c4.y) << 16);
      v.y = (uint32_t)to_bf16(c4.z) | ((uint32_t)to_bf16(d) << 16);
    }
    return v;
  }
  const int ch = q * 8;
  const uint16_t* src = ch < kCa ? A.in_a + pix * kCa + ch
                                 : A.in_b + pix * (kCb > 0 ? kCb : 8) + (ch - kCa);
  return *(const uint4*)src;
}

// kCin = ca + cb (multiples of 16; the feature layer stages 4 real channels
// as 16), kCout 16, 32 or 64.
template <int kCa, int kCb, int kCout, int kStride, bool kUp, bool kFeat, bool kHead, int kTX,
          int kTY>
__global__ void __launch_bounds__(256) conv3x3_lds(ConvArgs A) {
  constexpr int kCin = kCa + kCb;
  constexpr int kChunks = kCin / 16;
  constexpr int kNt = kCout / 16;
  using T = ConvTile<kCin, kCout, kStride, kUp, kTX, kTY>;
  __shared__ __align__(16) uint16_t lds[T::kWords];
  const int y0 = blockIdx.y * kTY;
  const int x0 = blockIdx.x * kTX;

  // Wave w owns output-channel tile nt = w / kWPN and every kWPN-th of the
  // block's kG groups of 16 pixels (group g: row g / kGX, pixels 16 (g % kGX)
  // + 0..15); its B fragments (9 taps x kChunks) are loaded into registers
  // once, before the staging, and reused for all its groups.
  constexpr int kGX = kTX / 16;
  constexpr int kG = kTY * kGX;
  constexpr int kWPN = 4 / kNt;        // waves per output-channel tile
  constexpr int kGW = kG / kWPN;       // pixel groups per wave
  static_assert(kNt == 1 || kNt == 2 || kNt == 4, "16, 32 or 64 output channels");
  static_assert(!kUp || (kTX % 2 == 0 && kTY % 2 == 0), "even tiles for the upsample");
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int nt = wave / kWPN, wi = wave % kWPN;
  v4i16 b[9][kChunks];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int kc = 0; kc < kChunks; ++kc) {
      const uint2 v = *(const uint2*)(A.w + (((size_t)(t * kChunks + kc) * kCout + nt * 16 + m) * 16 +
                                             4 * kq));
      b[t][kc] = v4i16{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu),
                       (short)(v.y >> 16)};
    }

  constexpr int kQ = kCin / 8;  // quads per pixel
  for (int i = threadIdx.x; i < T::kSY * T::kSX * kQ; i += 256) {
    const int q = i % kQ, c = (i / kQ) % T::kSX, r = i / (kQ * T::kSX);
    *(uint4*)&lds[(r * T::kSX + c) * T::kPS + q * 8] =
        stage_quad<kCa, kCb, kStride, kUp, kFeat>(A, y0, x0, r, c, q);
  }
  __syncthreads();

  v4f32 acc[kGW];
#pragma unroll
  for (int g = 0; g < kGW; ++g) acc[g] = v4f32{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
#pragma unroll
    for (int kc = 0; kc < kChunks; ++kc) {
#pragma unroll
      for (int g = 0; g < kGW; ++g) {
        const int gg = wi + kWPN * g;
        const int ty = gg / kGX, px = (gg % kGX) * 16 + m;
        int r, c;
        bool ok = true;
        if (kUp) {  // output-resolution tap, cropped to the output, then halved
          const int uy = y0 + ty + dy, ux = x0 + px + dx;
          ok = uy >= 0 && uy < A.out_h && ux >= 0 && ux < A.out_w;
          r = (uy >> 1) - ((y0 >> 1) - 1);
          c = (ux >> 1) - ((x0 >> 1) - 1);
        } else {
          r = ty * kStride + dy + 1;
          c = px * kStride + dx + 1;
        }
        v4i16 a = v4i16{0, 0, 0, 0};
        if (ok) {
          const uint2 v = *(const uint2*)&lds[(r * T::kSX + c) * T::kPS + kc * 16 + 4 * kq];
          a = v4i16{(short)(v.x & 0xffffu), (short)(v.x >> 16), (short)(v.y & 0xffffu),
                    (short)(v.y >> 16)};
        }
        acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t][kc], acc[g], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // the input tile is dead; its LDS now holds the output tile
  // D[m'][n]: lane holds m' = 4 kq + i, n = m; tile pixel index ty * kTX + px
  {
    const int n = nt * 16 + m;
    const float bn = A.bias[n];
#pragma unroll
    for (int g = 0; g < kGW; ++g) {
      const int gg = wi + kWPN * g;
      const int base = (gg / kGX) * kTX + (gg % kGX) * 16 + 4 * kq;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds[(base + i) * T::kOS + n] = to_bf16(fmaxf(acc[g][i] + bn, 0.0f));
    }
  }
  __syncthreads();
  if (kHead) {
    // 1x1 head: RGB = image RGB + W (3x16) . v + b, clamped, stored as RGBA8
    // (round to nearest, like the tonemap store); alpha from the image.
    static_assert(!kHead || kCout == 16, "the head reads 16 channels");
    for (int tp = threadIdx.x; tp < kTY * kTX; tp += 256) {
      const int yy = y0 + tp / kTX, xx = x0 + tp % kTX;
      if (yy >= A.out_h || xx >= A.out_w) continue;
      const size_t p = (size_t)yy * A.out_w + xx;
      float f[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) f[i] = from_bf16(lds[tp * T::kOS + i]);
      const float4 c = A.img[p];
      const float rgb[3] = {c.x, c.y, c.z};
      uint32_t out = 0u;
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        float s = A.head[48 + o];
#pragma unroll
        for (int i = 0; i < 16; ++i) s += A.head[o * 16 + i] * f[i];
        const float v = fminf(fmaxf(rgb[o] + s, 0.0f), 1.0f);
        out |= (uint32_t)(v * 255.0f + 0.5f) << (8 * o);
      }
      const float al = fminf(fmaxf(c.w, 0.0f), 1.0f);
      out |= (uint32_t)(al * 255.0f + 0.5f) << 24;
      A.out8[p] = out;
    }
  } else {
    constexpr int kOQ = kCout / 8;
    for (int i = threadIdx.x; i < kTY * kTX * kOQ; i += 256) {
      const int tp = i / kOQ, q = i % kOQ;
      const int yy = y0 + tp / kTX, xx = x0 + tp % kTX;
      if (yy < A.out_h && xx < A.out_w)
        *(uint4*)&A.out[((size_t)yy * A.out_w + xx) * kCout + q * 8] =
            *(const uint4*)&lds[tp * T::kOS + q * 8];
    }
  }
}

template <int kCa, int kCb, int kCout, int kStride, bool kUp, bool kFeat, int kTX, int kTY,
          bool kHead = false>
hipError_t run(const ConvArgs& a, hipStream_t s) {
  const dim3 grid((a.out_w + kTX - 1) / kTX, (a.out_h + kTY - 1) / kTY);
  hipLaunchKernelGGL((conv3x3_lds<kCa, kCb, kCout, kStride, kUp, kFeat, kHead, kTX, kTY>), grid,
                     dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_denoise(const DenoiseArgs& d, hipStream_t s) {
  const int W = d.width, H = d.height;
  const int W2 = (W + 1) / 2, H2 = (H + 1) / 2, W4 = (W2 + 1) / 2, H4 = (H2 + 1) / 2;
  hipError_t e;
  auto conv = [&](const uint16_t* a, const uint16_t* b, int layer, uint16_t* out, int iw, int ih,
                  int ow, int oh) {
    ConvArgs c{};
    c.in_a = a; c.in_b = b; c.w = d.w[layer]; c.bias = d.b[layer]; c.out = out;
    c.img = d.img; c.hist = d.hist; c.head = d.head; c.out8 = d.out8;
    c.in_w = iw; c.in_h = ih; c.out_w = ow; c.out_h = oh;
    return c;
  };
  // e0: features -> 16 (full)
  if ((e = run<16, 0, 16, 1, false, true, 64, 4>(conv(nullptr, nullptr, 0, d.s0a, W, H, W, H), s))) return e;
  // e1: 16 -> 16 (full) = skip s0
  if ((e = run<16, 0, 16, 1, false, false, 64, 4>(conv(d.s0a, nullptr, 1, d.s0, W, H, W, H), s))) return e;
  // d0: 16 -> 32 stride 2 (half)
  if ((e = run<16, 0, 32, 2, false, false, 64, 2>(conv(d.s0, nullptr, 2, d.h1a, W, H, W2, H2), s))) return e;
  // d1: 32 -> 32 (half) = skip s1
  if ((e = run<32, 0, 32, 1, false, false, 64, 2>(conv(d.h1a, nullptr, 3, d.s1, W2, H2, W2, H2), s)))
    return e;
  // b0: 32 -> 64 stride 2 (quarter)
  if ((e = run<32, 0, 64, 2, false, false, 64, 1>(conv(d.s1, nullptr, 4, d.q0, W2, H2, W4, H4), s)))
    return e;
  // b1: 64 -> 64 (quarter)
  if ((e = run<64, 0, 64, 1, false, false, 64, 1>(conv(d.q0, nullptr, 5, d.q1, W4, H4, W4, H4), s)))
    return e;
  // u0: up2(64) -> 32 (half)
  if ((e = run<64, 0, 32, 1, true, false, 64, 2>(conv(d.q1, nullptr, 6, d.h1a, W4, H4, W2, H2), s)))
    return e;
  // u1: cat(u0, s1) 64 -> 32 (half)
  if ((e = run<32, 32, 32, 1, false, false, 64, 2>(conv(d.h1a, d.s1, 7, d.h1b, W2, H2, W2, H2), s)))
    return e;
  // v0: up2(32) -> 16 (full)
  if ((e = run<32, 0, 16, 1, true, false, 64, 4>(conv(d.h1b, nullptr, 8, d.s0a, W2, H2, W, H), s)))
    return e;
  // v1: cat(v0, s0) 32 -> 16 (full), then the 1x1 head -> RGBA8
  return run<16, 16, 16, 1, false, false, 64, 4, true>(conv(d.s0a, d.s0, 9, nullptr, W, H, W, H), s);
}

}  // namespace p3
