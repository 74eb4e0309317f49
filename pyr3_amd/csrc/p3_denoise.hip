// Learned denoiser inference on gfx950: BF16 MFMA implicit-GEMM
// convolutions of the 4-level U-Net (pyr3_amd/denoise/net.py).
//
// Activations are NHWC bf16; a layer's weights are bf16 W[o][k] with
// k = tap * Cin + c, rows padded to kp = 9 Cin rounded up to 32 (the
// training kernels' layout, pyr3nn.h). The GEMM's K runs over that flat k in
// chunks of 32, one
//   v_mfma_f32_16x16x32_bf16 per chunk and 16-pixel group:
//   A[m][k] = input(pixel m, flat k)   lane l: m = l % 16, k = 8 (l / 16) + 0..7
//   B[k][n] = W[n][flat k]             lane l: n = l % 16, k = 8 (l / 16) + 0..7
//   D[m][n] (FP32)                     lane l: n = l % 16, m = 4 (l / 16) + 0..3
// A lane's 8 k lie in one tap (Cin is a multiple of 8), so its A operand is
// one 16-byte LDS read and its B operand one 16-byte load of W.
//
// A block of 4 waves computes a kTY x kTX tile of output pixels for every
// output channel. It stages the input rows and columns the 3x3 window
// reaches (at input resolution; for the upsampling layers the low-resolution
// rows and columns) in LDS with 16-byte loads, folding in the zero padding,
// stride 2 and the concat of two inputs; pixels are Cin + 8 bf16 apart in
// LDS, so the 16 lanes of a fragment read hit distinct banks. The first
// layer computes the network input from the histogram and the DE-off
// tonemapped image while staging (p3nn::pixel_features, the function the
// training path uses). Each wave owns one or two 16-channel output tiles and
// a share of the tile's 16-pixel groups; B operands are loaded per K chunk
// (L1/L2-resident weights). Bias and ReLU are applied, the bf16 results go
// through LDS (reusing the input tile) and leave as 16-byte coalesced
// stores. The last layer instead feeds the 1x1 head from LDS, adds the base
// image and stores RGBA8.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <type_traits>

#include "p3_denoise.h"
#include "p3_nn_common.h"

namespace p3 {
namespace {

using p3nn::from_bf16;
using p3nn::to_bf16;

typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

struct ConvArgs {
  const uint16_t* in_a;   // NHWC bf16, ca channels (null for the feature layer)
  const uint16_t* in_b;   // concat's second input (cb channels), or null
  const uint16_t* w;      // [cout][kp] bf16
  const float* bias;      // cout
  uint16_t* out;          // NHWC bf16, cout channels (null when the head is fused)
  const float4* img;      // feature layer and head: the DE-off tonemapped f32 image
  const uint32_t* hist;   // feature layer: the u32 RGBA histogram (oversample 1)
  const float* head;      // fused head: 3 x 16 weights then 3 biases
  uint32_t* out8;         // fused head: RGBA8 output
  float k2;               // feature layer: the frame's calibration constant
  int in_w, in_h;         // input size (low-resolution size when upsampling)
  int out_w, out_h;
};

template <int kCin, int kCout, int kStride, bool kUp, int kTX, int kTY>
struct ConvTile {
  static constexpr int kSX = kUp ? kTX / 2 + 2 : (kTX - 1) * kStride + 3;  // staged columns
  static constexpr int kSY = kUp ? kTY / 2 + 2 : (kTY - 1) * kStride + 3;  // staged rows
  static constexpr int kPS = kCin + 8;                                     // LDS pixel stride
  static constexpr int kOS = kCout + 8;                                    // output pixel stride
  static constexpr int kIn = kSY * kSX * kPS, kOut = kTY * kTX * kOS;
  static constexpr int kWords = kIn > kOut ? kIn : kOut;                   // bf16 elements
};

// One 16-byte quad (8 channels) of staged input pixel (r, c) of the tile at
// output (y0, x0); zero outside the input.
template <int kCa, int kCb, int kStride, bool kUp, bool kFeat>
__device__ inline uint4 stage_quad(const ConvArgs& A, int y0, int x0, int r, int c, int q) {
  const int iy = kUp ? (y0 >> 1) - 1 + r : y0 * kStride - 1 + r;
  const int ix = kUp ? (x0 >> 1) - 1 + c : x0 * kStride - 1 + c;
  if (iy < 0 || iy >= A.in_h || ix < 0 || ix >= A.in_w) return make_uint4(0u, 0u, 0u, 0u);
  const size_t pix = (size_t)iy * A.in_w + ix;
  if (kFeat) return p3nn::pixel_features(((const uint4*)A.hist)[pix], A.img[pix], A.k2);
  const int ch = q * 8;
  return ch < kCa ? *(const uint4*)(A.in_a + pix * kCa + ch)
                  : *(const uint4*)(A.in_b + pix * (kCb > 0 ? kCb : 8) + (ch - kCa));
}

template <int kCa, int kCb, int kCout, int kStride, bool kUp, bool kFeat, bool kHead, int kTX,
          int kTY>
__global__ void __launch_bounds__(256) conv3x3_mfma(ConvArgs A) {
  constexpr int kCin = kCa + kCb;
  constexpr int kKp = (9 * kCin + 31) / 32 * 32;
  constexpr int kChunks = kKp / 32;
  constexpr int kNt = kCout / 16;                    // 16-channel output tiles
  constexpr int kNtW = kNt > 4 ? kNt / 4 : 1;        // tiles per wave
  constexpr int kWPN = kNt >= 4 ? 1 : 4 / kNt;       // waves per tile
  constexpr int kGX = kTX / 16, kG = kTY * kGX;      // 16-pixel groups of the tile
  constexpr int kGW = kG / kWPN;                     // groups per wave
  static_assert(kCin % 8 == 0 && kCout % 16 == 0 && kG % kWPN == 0, "tile shape");
  static_assert(!kUp || (kTX % 2 == 0 && kTY % 2 == 0), "even tiles for the upsample");
  using T = ConvTile<kCin, kCout, kStride, kUp, kTX, kTY>;
  __shared__ __align__(16) uint16_t lds[T::kWords];
  const int y0 = blockIdx.y * kTY, x0 = blockIdx.x * kTX;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int nt0 = kNt >= 4 ? wave * kNtW : wave / kWPN;   // first output tile of the wave
  const int wi = kNt >= 4 ? 0 : wave % kWPN;              // its share of the groups

  constexpr int kQ = kCin / 8;  // quads per pixel
  for (int i = threadIdx.x; i < T::kSY * T::kSX * kQ; i += 256) {
    const int q = i % kQ, c = (i / kQ) % T::kSX, r = i / (kQ * T::kSX);
    *(uint4*)&lds[(r * T::kSX + c) * T::kPS + q * 8] =
        stage_quad<kCa, kCb, kStride, kUp, kFeat>(A, y0, x0, r, c, q);
  }
  __syncthreads();

  v4f32 acc[kNtW][kGW];
#pragma unroll
  for (int j = 0; j < kNtW; ++j)
#pragma unroll
    for (int g = 0; g < kGW; ++g) acc[j][g] = v4f32{0.0f, 0.0f, 0.0f, 0.0f};
  constexpr int kUnroll = kChunks * kNtW <= 18 ? kChunks : 4;
  // The K loop. An upsampling layer reads output-resolution tap (uy, ux) at
  // staged low-resolution pixel (uy >> 1, ux >> 1); with an even tile origin
  // that is row ((ty + dy) >> 1) + 1, column ((px + dx) >> 1) + 1 of the
  // tile. Taps outside the output read zeros from the staged border, except
  // the row or column just past an odd output size, which halves onto the
  // input's last row or column: blocks on such a border test every tap
  // (kChecked).
  auto k_loop = [&](auto checked) {
    constexpr bool kChecked = decltype(checked)::value;
#pragma unroll kUnroll
    for (int q = 0; q < kChunks; ++q) {
      const int k = q * 32 + kg * 8;               // this lane's first flat k
      const int tap = k / kCin < 9 ? k / kCin : 8;  // padding k: any tap (B is zero there)
      const int ch = k / kCin < 9 ? k % kCin : 0;
      const int dy = tap / 3 - 1, dx = tap % 3 - 1;
      v8bf16 b[kNtW];
#pragma unroll
      for (int j = 0; j < kNtW; ++j)
        b[j] = __builtin_bit_cast(
            v8bf16, *(const uint4*)(A.w + (size_t)((nt0 + j) * 16 + m) * kKp + k));
#pragma unroll
      for (int g = 0; g < kGW; ++g) {
        const int gg = wi + kWPN * g;
        const int ty = gg / kGX, px = (gg % kGX) * 16 + m;
        int r, c;
        bool ok = true;
        if (kUp) {
          r = ((ty + dy) >> 1) + 1;
          c = ((px + dx) >> 1) + 1;
          if (kChecked) {
            const int uy = y0 + ty + dy, ux = x0 + px + dx;
            ok = uy < A.out_h && ux < A.out_w;
          }
        } else {
          r = ty * kStride + dy + 1;
          c = px * kStride + dx + 1;
        }
        uint4 av = make_uint4(0u, 0u, 0u, 0u);
        if (ok) av = *(const uint4*)&lds[(r * T::kSX + c) * T::kPS + ch];
        const v8bf16 a = __builtin_bit_cast(v8bf16, av);
#pragma unroll
        for (int j = 0; j < kNtW; ++j)
          acc[j][g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[j][g], 0, 0, 0);
      }
    }
  };
  const bool border = kUp && (((A.out_h & 1) && y0 + kTY >= A.out_h) ||
                              ((A.out_w & 1) && x0 + kTX >= A.out_w));   // block-uniform
  if (border) k_loop(std::true_type{});
  else k_loop(std::false_type{});
  __syncthreads();  // the input tile is dead; its LDS now holds the output tile
#pragma unroll
  for (int j = 0; j < kNtW; ++j) {
    const int n = (nt0 + j) * 16 + m;
    const float bn = A.bias[n];
#pragma unroll
    for (int g = 0; g < kGW; ++g) {
      const int gg = wi + kWPN * g;
      const int base = (gg / kGX) * kTX + (gg % kGX) * 16 + 4 * kg;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        lds[(base + i) * T::kOS + n] = to_bf16(fmaxf(acc[j][g][i] + bn, 0.0f));
    }
  }
  __syncthreads();
  if (kHead) {
    // 1x1 head: RGB = base RGB + W (3 x 16) . v + b, clamped, RGBA8 with
    // round to nearest (the tonemap's store); alpha from the base image.
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
        float s = rgb[o] + A.head[48 + o];
#pragma unroll
        for (int i = 0; i < 16; ++i) s += A.head[o * 16 + i] * f[i];
        const float v = fminf(fmaxf(s, 0.0f), 1.0f);
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
  hipLaunchKernelGGL((conv3x3_mfma<kCa, kCb, kCout, kStride, kUp, kFeat, kHead, kTX, kTY>), grid,
                     dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_denoise(const DenoiseArgs& d, hipStream_t s) {
  const int W0 = d.width, H0 = d.height;
  const int W1 = (W0 + 1) / 2, H1 = (H0 + 1) / 2, W2 = (W1 + 1) / 2, H2 = (H1 + 1) / 2;
  const int W3 = (W2 + 1) / 2, H3 = (H2 + 1) / 2;
  hipError_t e;
  auto conv = [&](const uint16_t* a, const uint16_t* b, int layer, uint16_t* out, int iw, int ih,
                  int ow, int oh) {
    ConvArgs c{};
    c.in_a = a; c.in_b = b; c.w = d.w[layer]; c.bias = d.b[layer]; c.out = out;
    c.img = d.img; c.hist = d.hist; c.head = d.head; c.out8 = d.out8; c.k2 = d.k2;
    c.in_w = iw; c.in_h = ih; c.out_w = ow; c.out_h = oh;
    return c;
  };
  uint16_t* const* t = d.act;
  // layer order and tensors as kDenoiseShape (p3_denoise.h); tile shapes
  // (pixels x rows per block) measured per layer at 1920x1080
  // (16 x 16 x 32 MFMA, random data, each layer alone): narrow tall tiles
  // for the low-resolution layers (d3 14.4 us at 16x8 against 26.1 at 32x1,
  // b0 17.2 at 16x8 against 26.3 at 64x2, c3 37.3 at 32x4 against 49.8 at
  // 64x2, c2 45.6 at 16x16 against 51.1 at 64x4, e3 19.4 at 16x8 against
  // 22.6 at 64x2, d1 24.7 at 32x8 against 26.3 at 64x2), wide ones at full
  // resolution (e1 34.3 at 64x8, e2 23.5 at 64x4, c1 53.0 at 32x8 against
  // 59.0 at 64x4); the upsampling layers with their compile-time tap
  // indexing: u2 32.3 at 32x8 against 50.1 at 32x4, u1 32.7 at 32x16
  // against 41.0 at 64x4, u3 26.3 at 32x4 (36.0 at 64x4)
  if ((e = run<8, 0, 16, 1, false, true, 64, 8>(conv(nullptr, nullptr, 0, t[0], W0, H0, W0, H0), s))) return e;
  if ((e = run<16, 0, 16, 1, false, false, 64, 8>(conv(t[0], nullptr, 1, t[1], W0, H0, W0, H0), s))) return e;
  if ((e = run<16, 0, 32, 2, false, false, 32, 8>(conv(t[1], nullptr, 2, t[2], W0, H0, W1, H1), s))) return e;
  if ((e = run<32, 0, 32, 1, false, false, 64, 4>(conv(t[2], nullptr, 3, t[3], W1, H1, W1, H1), s))) return e;
  if ((e = run<32, 0, 64, 2, false, false, 64, 1>(conv(t[3], nullptr, 4, t[4], W1, H1, W2, H2), s))) return e;
  if ((e = run<64, 0, 64, 1, false, false, 16, 8>(conv(t[4], nullptr, 5, t[5], W2, H2, W2, H2), s))) return e;
  if ((e = run<64, 0, 128, 2, false, false, 16, 8>(conv(t[5], nullptr, 6, t[6], W2, H2, W3, H3), s))) return e;
  if ((e = run<128, 0, 128, 1, false, false, 16, 8>(conv(t[6], nullptr, 7, t[7], W3, H3, W3, H3), s))) return e;
  if ((e = run<128, 0, 64, 1, true, false, 32, 4>(conv(t[7], nullptr, 8, t[8], W3, H3, W2, H2), s))) return e;
  if ((e = run<64, 64, 64, 1, false, false, 32, 4>(conv(t[8], t[5], 9, t[9], W2, H2, W2, H2), s))) return e;
  if ((e = run<64, 0, 32, 1, true, false, 32, 8>(conv(t[9], nullptr, 10, t[10], W2, H2, W1, H1), s))) return e;
  if ((e = run<32, 32, 32, 1, false, false, 16, 16>(conv(t[10], t[3], 11, t[11], W1, H1, W1, H1), s))) return e;
  if ((e = run<32, 0, 16, 1, true, false, 32, 16>(conv(t[11], nullptr, 12, t[12], W1, H1, W0, H0), s))) return e;
  return run<16, 16, 16, 1, false, false, 32, 8, true>(conv(t[12], t[1], 13, nullptr, W0, H0, W0, H0), s);
}

}  // namespace p3
