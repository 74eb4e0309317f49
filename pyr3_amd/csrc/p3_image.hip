// Density estimation and tonemap kernels for gfx950.
//
// DE follows density_main (src/shaders/density.wgsl:51-123): every
// non-empty super-pixel n scatters its log-density-scaled colour over a disc
// of integer radius r_n (half-even rounded adaptive radius), Gaussian
// sigma = r_n / 3, normalised by the per-radius kernel sum so each bucket
// deposits weight 1. The reference evaluates it as a per-output gather over
// the full (2R+1)^2 box with pow/log/round per visited neighbour. Here it is
// split in two passes:
//   de_prep    once per bin: r_n (u8, 255 = empty) and the bin's RGBA times
//              ls_n / knorm(r_n) (float4), plus the max radius of each
//              16 x 16 bin tile;
//   de_gather  per 16 x 16 output tile: the largest radius among bins that
//              could reach the tile bounds the loop box (Rt <= R), the
//              (16 + 2 Rt)^2 neighbourhood is staged in LDS, and every output
//              sums kw * pre_n over its box in the reference's dy-then-dx order.
// kw = exp(-d^2 / (2 sigma^2)) is the same expression as the reference.
//
// The tonemap follows fs (src/shaders/visualize_f32.wgsl:162-271 with DE,
// src/shaders/visualize_u32.wgsl:148-248 without): separable collapse of
// fwidth^2 super-pixels (clamp to edge), flam3 alpha/vibrancy/highlight
// composite, background blend, optional channel curves and HSV adjust
// (:240-266), RGBA8 store with round-to-nearest.
#include <type_traits>

#include "p3_device.h"
#include "p3_launch.h"

namespace p3 {

constexpr int kTile = 16;
constexpr uint8_t kEmpty = 255;

__global__ void __launch_bounds__(256) de_prep(DeArgs A) {
  __shared__ uint32_t smax;
  const uint32_t x = blockIdx.x * kTile + threadIdx.x;
  const uint32_t y = blockIdx.y * kTile + threadIdx.y;
  if (threadIdx.x == 0 && threadIdx.y == 0) smax = 0u;
  __syncthreads();
  if (x < A.width && y < A.height) {
    const size_t i = (size_t)y * A.width + x;
    const uint4 h = A.hist[i];
    const float cnt = (float)h.w;
    uint8_t r8 = kEmpty;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (cnt > 0.0f) {
      const float hits = cnt / 255.0f;
      const float rf = fminf(fmaxf(A.max_rad / powf(hits + 1.0f, A.curve), A.min_rad), A.max_rad);
      const int ri = min(max((int)rintf(rf), 0), 256);
      const float knorm = fmaxf(A.lut[ri], 1e-6f);
      const float ls = (A.k1 * logf(1.0f + cnt * A.k2)) / cnt;
      const float s = ls / knorm;
      p = make_float4((float)h.x * s, (float)h.y * s, (float)h.z * s, cnt * s);
      // Radii beyond the search box never reach a neighbour inside it; the
      // gather clips its box to R anyway, so store min(ri, 254).
      r8 = (uint8_t)min(ri, 254);
      atomicMax(&smax, (uint32_t)min(ri, A.R));
    }
    A.pre[i] = p;
    A.rad[i] = r8;
  }
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0)
    A.tile_rmax[blockIdx.y * A.tiles_x + blockIdx.x] = smax;
}

__device__ inline int tile_bound(const DeArgs& A, int tx0, int ty0) {
  // Tiles whose bins lie within R of the output tile.
  const int rt = (A.R + kTile - 1) / kTile;
  int m = 0;
  for (int ty = max(ty0 - rt, 0); ty <= min(ty0 + rt, (int)A.tiles_y - 1); ++ty)
    for (int tx = max(tx0 - rt, 0); tx <= min(tx0 + rt, (int)A.tiles_x - 1); ++tx)
      m = max(m, (int)A.tile_rmax[ty * A.tiles_x + tx]);
  return min(m, A.R);
}

__device__ inline float de_kw(int d2, int r) {
  const float rad = (float)r;
  const float sigma = fmaxf(rad / 3.0f, 1e-6f);
  return expf(-(float)d2 / (2.0f * sigma * sigma));
}

// Offset of radius r's weights in the per-block table: radius j holds
// j^2 + 1 weights (d^2 = 0 .. j^2), so radii 0 .. r-1 take
// sum (j^2 + 1) = (r - 1) r (2r - 1) / 6 + r.
__host__ __device__ constexpr int kw_offset(int r) { return (r - 1) * r * (2 * r - 1) / 6 + r; }

// A staged bin as one LDS word. Halos up to 10: 16 bits, its weights'
// offset in the table (kw_offset(r), low 9 bits) and its reach r^2 + 1
// above them (0: empty bin, which every test rejects). Larger halos: one
// byte, its radius (255: empty bin; those halos are walked only through
// masks, which leave empty bins out), the offset read from a table of the
// Rt + 1 offsets. A visit at squared distance d2 takes the bin if
// reaches(word, d2) and reads weight [offset + d2].
template <int kR> struct DePack {
  static constexpr bool kReach = kR <= 10;
  using T = typename std::conditional<kReach, uint16_t, uint8_t>::type;
  static constexpr int kShift = 9;
  static constexpr uint32_t kOff = (1u << kShift) - 1u;
  static constexpr uint32_t kEmptyWord = kReach ? 0u : 255u;
  static_assert(!kReach || (kw_offset(kR) <= (int)kOff && kR * kR + 1 < 128), "16-bit word");
  static_assert(kR < 255, "radius byte");
  __device__ static uint32_t encode(int r) {
    return kReach ? ((uint32_t)kw_offset(r) | ((uint32_t)(r * r + 1) << kShift)) : (uint32_t)r;
  }
  __device__ static bool reaches(uint32_t w, uint32_t d2) {
    return kReach ? d2 < (w >> kShift) : d2 <= w * w;
  }
  __device__ static uint32_t offset(uint32_t w, const uint16_t* off) {
    return kReach ? (w & kOff) : (uint32_t)off[w];
  }
  // radius >= t and the bin is occupied
  __device__ static bool at_least(uint32_t w, int t) {
    return kReach ? (w >> kShift) > (uint32_t)(t * t) : (w != 255u && w >= (uint32_t)t);
  }
};

// One row of an output's box through its mask: bit j of m is staged
// column (window start + j), dx = j + dx0; the bins are visited in
// increasing x.
template <class Pk, class M>
__device__ __forceinline__ void de_walk(M m, const typename Pk::T* pk, const float4* pre,
                                        const float* kw, const uint16_t* off, int dx0, int dy2,
                                        float4& sum) {
  while (m) {
    const int j = sizeof(M) == 8 ? (int)__builtin_ctzll((uint64_t)m) : (int)__builtin_ctz((uint32_t)m);
    m &= m - (M)1;
    const uint32_t w = pk[j];
    const int dx = j + dx0;
    const uint32_t d2 = (uint32_t)(dx * dx + dy2);
    if (!Pk::reaches(w, d2)) continue;
    const float k = kw[Pk::offset(w, off) + d2];
    const float4 p = pre[j];
    sum.x = sum.x + p.x * k;
    sum.y = sum.y + p.y * k;
    sum.z = sum.z + p.z * k;
    sum.w = sum.w + p.w * k;
  }
}

// Per 16 x 16 output tile with search radius Rt <= kR: the (16 + 2 Rt)^2
// neighbourhood (pre-scaled RGBA, packed radius) is staged in LDS, and each
// output sums kw * pre over the bins of its (2 Rt + 1)^2 box that reach it
// (d^2 <= r^2), in dy-then-dx order (the reference's order; bins that do
// not reach it add nothing). A bin reaches an output dy rows away only if
// its radius is at least |dy|, so with kMask one wave ballots every staged
// row into 64-bit masks of the columns whose radius is at least t, for
// t = 0 .. Rt, and row dy of an output's box visits the set bits of mask
// |dy| inside its 2 Rt + 1 columns in increasing x; without kMask every
// column of the box is tested. Both visit the reaching bins in the same
// order, so the sums are bit-identical. The Gaussian weights
// exp(-d^2 / 2 sigma^2) depend only on (d^2, r): they are tabulated per
// block for r <= Rt (packed, kw_offset) with the same expression. A staged
// bin's radius is at most Rt (it is at most max_rad < R, and its tile is
// one of those bounding Rt); the staging clamps it there all the same, so
// no visit reads past the table. Tiles with Rt > kR read HBM directly.
template <int kR, bool kMask>
__global__ void __launch_bounds__(256) de_gather(DeArgs A) {
  using Pk = DePack<kR>;
  constexpr int kSpan = kTile + 2 * kR;
  constexpr int kM = kMask ? kSpan * (kR + 1) : 1;   // masks (none without kMask)
  static_assert(kSpan <= 64, "one wave ballots a staged row into a 64-bit mask");
  static_assert(kMask || Pk::kReach, "a box scan needs words whose test rejects empty bins");
  __shared__ float4 s_pre[kSpan * kSpan];
  __shared__ typename Pk::T s_pk[kSpan * kSpan];
  __shared__ uint64_t s_mask[kM];                  // [row][t]: columns with radius >= t
  __shared__ float s_kw[kw_offset(kR + 1)];
  __shared__ uint16_t s_off[Pk::kReach ? 1 : kR + 1];   // kw_offset(r), r <= Rt
  __shared__ int s_rt;
  const int tid = threadIdx.y * kTile + threadIdx.x;
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  if (tid == 0) s_rt = tile_bound(A, blockIdx.x, blockIdx.y);
  __syncthreads();
  const int Rt = s_rt;
  const int xi = x0 + threadIdx.x, yi = y0 + threadIdx.y;
  const int W = (int)A.width, H = (int)A.height;
  float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

  if (Rt <= kR) {
    const int span = kTile + 2 * Rt;
    for (int k = tid; k < kw_offset(Rt + 1); k += 256) {
      int r = 0;
      while (kw_offset(r + 1) <= k) ++r;
      s_kw[k] = de_kw(k - kw_offset(r), r);
    }
    if (!Pk::kReach && tid <= Rt) s_off[tid] = (uint16_t)kw_offset(tid);
    for (int k = tid; k < span * span; k += 256) {
      const int ly = k / span, lx = k - ly * span;
      const int gx = x0 - Rt + lx, gy = y0 - Rt + ly;
      uint32_t pk = Pk::kEmptyWord;
      float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
        const size_t g = (size_t)gy * W + gx;
        const int r8 = A.rad[g];
        if (r8 != kEmpty) {
          p = A.pre[g];
          pk = Pk::encode(min(r8, Rt));
        }
      }
      s_pk[k] = (typename Pk::T)pk;
      s_pre[k] = p;
    }
    __syncthreads();
    if (kMask) {   // one wave per staged row
      const int lane = tid & 63, wave = tid >> 6;
      for (int ly = wave; ly < span; ly += 4) {
        const uint32_t w = lane < span ? (uint32_t)s_pk[ly * span + lane] : Pk::kEmptyWord;
        for (int t = 0; t <= Rt; ++t) {
          const uint64_t m = __ballot(Pk::at_least(w, t));
          if (lane == 0) s_mask[ly * (kR + 1) + t] = m;
        }
      }
      __syncthreads();
    }
    if (xi < W && yi < H) {
      const int cx = threadIdx.x + Rt, cy = threadIdx.y + Rt;
      const int nb = 2 * Rt + 1;   // window: staged columns cx - Rt .. cx + Rt
      for (int dy = -Rt; dy <= Rt; ++dy) {
        const int rowk = (cy + dy) * span;
        if (kMask) {
          const uint64_t m = s_mask[(cy + dy) * (kR + 1) + (dy < 0 ? -dy : dy)] >> threadIdx.x;
          const int w0 = rowk + (int)threadIdx.x;
          if (nb <= 32)   // block-uniform
            de_walk<Pk, uint32_t>((uint32_t)m & (0xffffffffu >> (32 - nb)), s_pk + w0, s_pre + w0,
                                  s_kw, s_off, -Rt, dy * dy, sum);
          else
            de_walk<Pk, uint64_t>(m & ((1ull << nb) - 1ull), s_pk + w0, s_pre + w0, s_kw, s_off,
                                  -Rt, dy * dy, sum);
        } else {
          for (int dx = -Rt; dx <= Rt; ++dx) {
            const int k = rowk + cx + dx;
            const uint32_t pk = s_pk[k];
            const uint32_t d2 = (uint32_t)(dx * dx + dy * dy);
            if (!Pk::reaches(pk, d2)) continue;
            const float kw = s_kw[Pk::offset(pk, s_off) + d2];
            const float4 p = s_pre[k];
            sum.x = sum.x + p.x * kw;
            sum.y = sum.y + p.y * kw;
            sum.z = sum.z + p.z * kw;
            sum.w = sum.w + p.w * kw;
          }
        }
      }
    }
  } else if (xi < W && yi < H) {
    for (int dy = -Rt; dy <= Rt; ++dy) {
      const int ny = yi + dy;
      if (ny < 0 || ny >= H) continue;
      for (int dx = -Rt; dx <= Rt; ++dx) {
        const int nx = xi + dx;
        if (nx < 0 || nx >= W) continue;
        const size_t g = (size_t)ny * W + nx;
        const int r = A.rad[g];
        if (r == kEmpty) continue;
        const int d2 = dx * dx + dy * dy;
        if (d2 > r * r) continue;
        const float kw = de_kw(d2, r);
        const float4 p = A.pre[g];
        sum.x = sum.x + p.x * kw;
        sum.y = sum.y + p.y * kw;
        sum.z = sum.z + p.z * kw;
        sum.w = sum.w + p.w * kw;
      }
    }
  }
  if (xi < W && yi < H) A.out[(size_t)yi * W + xi] = sum;
}

// The gather's LDS halo follows the search radius R (9 os + 1 super-pixels
// for a 9-pixel DE radius). At R <= 10 (oversample 1) a 10 halo and the box
// scan (24.9 KB of LDS, 6 blocks per CU); the mask walk, which visits 27 %
// fewer bins per wave there, costs more per visit and in building its
// masks than it saves (1080p q30: 148 us box against 196 us masks, q2:
// 161 against 181, with de_prep). Larger R (oversample 2): a 16 halo and
// the mask walk (51.8 KB, 3 blocks per CU), whose radius-thresholded rows
// skip the bins that cannot reach across |dy| rows (tiles with Rt > 16 read
// HBM).
hipError_t launch_density(const DeArgs& a, hipStream_t stream) {
  const dim3 block(kTile, kTile);
  const dim3 grid(a.tiles_x, a.tiles_y);
  de_prep<<<grid, block, 0, stream>>>(a);
  if (a.R <= 10) de_gather<10, false><<<grid, block, 0, stream>>>(a);
  else de_gather<16, true><<<grid, block, 0, stream>>>(a);
  return hipGetLastError();
}

// ------------------------------------------------------------- tonemap --

// pow as WGSL defines it, exp2(y log2 x), on the hardware's log2 and exp2
// (v_log_f32, v_exp_f32): the tonemap's powers all have x >= 0 and y > 0
// (or x > 0 and y = 0), where this is exact in form and within ~3e-7
// relative of a correctly rounded pow.
__device__ inline float pow_wgsl(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

__device__ inline float calc_alpha(float d, float g, float lin) {
  if (d <= 0.0f) return 0.0f;
  if (d < lin) {
    const float fv = pow_wgsl(lin, g);
    const float frac = d / lin;
    return (1.0f - frac) * d * (fv / lin) + frac * pow_wgsl(d, g);
  }
  return pow_wgsl(d, g);
}

__device__ inline float3 rgb2hsv(float3 c) {
  const float mx = fmaxf(fmaxf(c.x, c.y), c.z);
  const float mn = fminf(fminf(c.x, c.y), c.z);
  const float d = mx - mn;
  float h = 0.0f;
  if (d > 0.0f) {
    if (mx == c.x) h = (c.y - c.z) / d;
    else if (mx == c.y) h = 2.0f + (c.z - c.x) / d;
    else h = 4.0f + (c.x - c.y) / d;
    h *= 60.0f;
    if (h < 0.0f) h += 360.0f;
  }
  return make_float3(h, mx > 0.0f ? d / mx : 0.0f, mx);
}

__device__ inline float3 hsv2rgb(float3 c) {
  const float h = c.x / 60.0f, s = c.y, v = c.z;
  const float i = floorf(h);
  const float f = h - i;
  const float p = v * (1.0f - s), q = v * (1.0f - s * f), t = v * (1.0f - s * (1.0f - f));
  const int ix = (int)i % 6;
  if (ix == 0) return make_float3(v, t, p);
  if (ix == 1) return make_float3(q, v, p);
  if (ix == 2) return make_float3(p, v, t);
  if (ix == 3) return make_float3(p, q, v);
  if (ix == 4) return make_float3(t, p, v);
  return make_float3(v, p, q);
}

__device__ inline float3 calc_newrgb(float3 c, float ls, float highpow) {
  if (ls == 0.0f || (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f)) return make_float3(0.f, 0.f, 0.f);
  const float3 cw = make_float3(c.x / 255.0f, c.y / 255.0f, c.z / 255.0f);
  const float3 sc = make_float3(ls * cw.x, ls * cw.y, ls * cw.z);
  float maxa = sc.x, maxc = cw.x;
  if (sc.y > maxa) { maxa = sc.y; maxc = cw.y; }
  if (sc.z > maxa) { maxa = sc.z; maxc = cw.z; }
  if (maxa > 255.0f && highpow >= 0.0f) {
    const float newls = 255.0f / maxc;
    const float lsratio = pow_wgsl(newls / ls, highpow);
    float3 n = make_float3(newls * cw.x / 255.0f, newls * cw.y / 255.0f, newls * cw.z / 255.0f);
    float3 hsv = rgb2hsv(n);
    hsv.y *= lsratio;
    n = hsv2rgb(hsv);
    return make_float3(n.x * 255.0f, n.y * 255.0f, n.z * 255.0f);
  }
  const float newls = 255.0f / maxc;
  float adj = -highpow;
  if (adj > 1.0f) adj = 1.0f;
  if (maxa <= 255.0f) adj = 1.0f;
  const float m = (1.0f - adj) * newls + adj * ls;
  return make_float3(m * cw.x, m * cw.y, m * cw.z);
}

// visualize_f32.wgsl:63-68: linear lookup in channel `ch` of the curves LUT.
__device__ inline float curve_lut(const float* L, uint32_t ch, float x) {
  const float idx = fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f;
  const uint32_t i0 = (uint32_t)floorf(idx);
  const uint32_t i1 = min(i0 + 1u, 255u);
  return mixf(L[ch * 256u + i0], L[ch * 256u + i1], idx - (float)i0);
}

// visualize_f32.wgsl:240-266: channel curves (luma, R, G, B, composite),
// then the HSV adjust.
__device__ inline float3 grade(const TmArgs& A, float3 o) {
  if (A.curves_mask != 0u) {
    if (A.curves_mask & 16u) {
      const float l = o.x * 0.299f + o.y * 0.587f + o.z * 0.114f;
      const float ln = curve_lut(A.curves, 4u, l);
      if (l > 0.0f) {
        const float s = ln / l;
        o = make_float3(o.x * s, o.y * s, o.z * s);
      }
    }
    if (A.curves_mask & 2u) o.x = curve_lut(A.curves, 1u, o.x);
    if (A.curves_mask & 4u) o.y = curve_lut(A.curves, 2u, o.y);
    if (A.curves_mask & 8u) o.z = curve_lut(A.curves, 3u, o.z);
    if (A.curves_mask & 1u)
      o = make_float3(curve_lut(A.curves, 0u, o.x), curve_lut(A.curves, 0u, o.y),
                      curve_lut(A.curves, 0u, o.z));
  }
  if (A.hsl_on != 0u) {
    float3 hsv = rgb2hsv(o);
    hsv.x = hsv.x + A.hsl_hue;
    hsv.x = hsv.x - 360.0f * floorf(hsv.x / 360.0f);
    hsv.y = fminf(fmaxf(hsv.y * A.hsl_sat, 0.0f), 1.0f);
    hsv.z = fminf(fmaxf(hsv.z + A.hsl_light, 0.0f), 1.0f);
    o = hsv2rgb(hsv);
  }
  return o;
}

__device__ inline uint32_t to_u8(float v) {
  return (uint32_t)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

// visualize_f32.wgsl's per-pixel composite of the filtered sums (c0..c3)
// into output pixel (xi, yi): RGBA8, and RGBA f32 when A.outf is set.
__device__ inline void tm_store(const TmArgs& A, uint32_t xi, uint32_t yi, float c0, float c1,
                                float c2, float c3) {
  float4 o;
  if (c3 <= 0.0f) {
    o = A.transparent ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(A.bg0, A.bg1, A.bg2, 1.0f);
  } else {
    const float g = 1.0f / A.gamma;
    const float tmp = c3 / 255.0f;
    float alpha = calc_alpha(tmp, g, A.linrange);
    const float ls_a = A.vibrancy * 256.0f * alpha / fmaxf(tmp, 1e-12f);
    alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
    const float3 n = calc_newrgb(make_float3(c0, c1, c2), ls_a, A.highpow);
    const float cc[3] = {c0, c1, c2};
    const float nn[3] = {n.x, n.y, n.z};
    const float bg[3] = {A.bg0, A.bg1, A.bg2};
    float res[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float perch = (1.0f - A.vibrancy) * 256.0f * pow_wgsl(fmaxf(cc[i] / 255.0f, 0.0f), g);
      const float bgt = A.transparent ? 0.0f : (1.0f - alpha) * 256.0f * bg[i];
      const float composed = nn[i] + perch + bgt;
      res[i] = fminf(fmaxf(composed / 256.0f, 0.0f), 1.0f);
    }
    const float3 gr = grade(A, make_float3(res[0], res[1], res[2]));
    o = make_float4(fminf(fmaxf(gr.x, 0.0f), 1.0f), fminf(fmaxf(gr.y, 0.0f), 1.0f),
                    fminf(fmaxf(gr.z, 0.0f), 1.0f), A.transparent ? alpha : 1.0f);
  }
  const size_t oi = (size_t)yi * A.width + xi;
  A.out8[oi] = to_u8(o.x) | (to_u8(o.y) << 8) | (to_u8(o.z) << 16) | (to_u8(o.w) << 24);
  if (A.outf) A.outf[oi] = o;
}

// Log-density scaling of one bin without DE (visualize_u32.wgsl:175-184):
// the bin's RGBA counts times k1 log(1 + cnt k2) / cnt, zero for an empty bin.
__device__ inline float4 log_scaled(const TmArgs& A, uint4 h) {
  const float cnt = (float)h.w;
  if (!(cnt > 0.0f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float ls = (A.k1 * logf(1.0f + cnt * A.k2)) / cnt;
  return make_float4((float)h.x * ls, (float)h.y * ls, (float)h.z * ls, cnt * ls);
}

// Spatial filter (fw x fw taps centred on the output pixel's os x os
// super-pixels, edges clamped) and composite, one thread per output pixel,
// bins read from HBM per tap. Used when the filter's tile does not fit the
// LDS budget of tonemap_tiled.
template <bool kDe>
__global__ void __launch_bounds__(256) tonemap_kernel(TmArgs A) {
  const uint32_t xi = blockIdx.x * 16 + threadIdx.x;
  const uint32_t yi = blockIdx.y * 16 + threadIdx.y;
  if (xi >= A.width || yi >= A.height) return;
  const uint32_t N = A.os, fw = A.fwidth, half = fw / 2u;
  const int SW = (int)(A.width * N), SH = (int)(A.height * N);
  const uint32_t cx = xi * N + N / 2u, cy = yi * N + N / 2u;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
  for (uint32_t dy = 0; dy < fw; ++dy) {
    const float ky = A.taps[dy];
    const int sy = min(max((int)cy + (int)dy - (int)half, 0), SH - 1);
    for (uint32_t dx = 0; dx < fw; ++dx) {
      const float kx = A.taps[dx];
      const int sx = min(max((int)cx + (int)dx - (int)half, 0), SW - 1);
      const size_t sb = (size_t)sy * SW + sx;
      const float w = kx * ky;
      const float4 f = kDe ? A.filtered[sb] : log_scaled(A, A.hist[sb]);
      c0 = c0 + f.x * w; c1 = c1 + f.y * w; c2 = c2 + f.z * w; c3 = c3 + f.w * w;
    }
  }
  tm_store(A, xi, yi, c0, c1, c2, c3);
}

// The same filter and composite over a 32 x 8 tile of output pixels whose
// super-pixels (plus the filter's halo, edges clamped) are staged in LDS
// once: each bin is read from HBM once per tile (and, without DE, log-scaled
// once instead of once per tap), then every tap reads LDS. The per-pixel
// sums take the same terms in the same order as tonemap_kernel (a bin that
// is empty adds +0), so both kernels give identical bits.
constexpr uint32_t kTmTileX = 32, kTmTileY = 8;
template <bool kDe>
__global__ void __launch_bounds__(256) tonemap_tiled(TmArgs A, uint32_t span_x, uint32_t span_y) {
  extern __shared__ float4 tm_tile[];
  const uint32_t N = A.os, fw = A.fwidth, half = fw / 2u;
  const int SW = (int)(A.width * N), SH = (int)(A.height * N);
  const int x0 = (int)(blockIdx.x * kTmTileX * N + N / 2u) - (int)half;
  const int y0 = (int)(blockIdx.y * kTmTileY * N + N / 2u) - (int)half;
  const uint32_t t = threadIdx.x;
  const uint32_t cells = span_x * span_y;
  for (uint32_t k = t; k < cells; k += 256u) {
    const uint32_t ty = k / span_x, tx = k - ty * span_x;
    const int sy = min(max(y0 + (int)ty, 0), SH - 1);
    const int sx = min(max(x0 + (int)tx, 0), SW - 1);
    const size_t sb = (size_t)sy * SW + sx;
    tm_tile[k] = kDe ? A.filtered[sb] : log_scaled(A, A.hist[sb]);
  }
  __syncthreads();
  const uint32_t lx = t % kTmTileX, ly = t / kTmTileX;
  const uint32_t xi = blockIdx.x * kTmTileX + lx, yi = blockIdx.y * kTmTileY + ly;
  if (xi >= A.width || yi >= A.height) return;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
  for (uint32_t dy = 0; dy < fw; ++dy) {
    const float ky = A.taps[dy];
    const float4* row = tm_tile + (ly * N + dy) * span_x + lx * N;
    for (uint32_t dx = 0; dx < fw; ++dx) {
      const float w = A.taps[dx] * ky;
      const float4 f = row[dx];
      c0 = c0 + f.x * w; c1 = c1 + f.y * w; c2 = c2 + f.z * w; c3 = c3 + f.w * w;
    }
  }
  tm_store(A, xi, yi, c0, c1, c2, c3);
}

// ------------------------------------------------------- linear export --

// histogramToLinearRgba (src/export-linear.ts:24-47) on the device: each
// output pixel sums its os x os super-pixels (u64, exact), scales by
// (k1 k2) / os^2 in f64 and stores f32 (round to nearest), alpha clamped to
// [0, 1] -- the reference's f64 arithmetic, so the result is bit-identical.
// HBM-bound: 16 B read per super-pixel, 16 B written per output pixel.
__global__ void __launch_bounds__(256) linear_export(LinArgs A) {
  const uint32_t x = blockIdx.x * 16 + threadIdx.x;
  const uint32_t y = blockIdx.y * 16 + threadIdx.y;
  if (x >= A.width || y >= A.height) return;
  const uint32_t N = A.os;
  const size_t SW = (size_t)A.width * N;
  uint64_t r = 0, g = 0, b = 0, c = 0;
  for (uint32_t sy = 0; sy < N; ++sy)
    for (uint32_t sx = 0; sx < N; ++sx) {
      const uint4 h = A.hist[((size_t)y * N + sy) * SW + (size_t)x * N + sx];
      r += h.x; g += h.y; b += h.z; c += h.w;
    }
  const double s = A.scale;
  A.out[(size_t)y * A.width + x] =
      make_float4((float)((double)r * s), (float)((double)g * s), (float)((double)b * s),
                  (float)fmin(1.0, fmax(0.0, (double)c * s)));
}

hipError_t launch_linear_export(const LinArgs& a, hipStream_t stream) {
  const dim3 block(16, 16);
  const dim3 grid((a.width + 15) / 16, (a.height + 15) / 16);
  linear_export<<<grid, block, 0, stream>>>(a);
  return hipGetLastError();
}

// ----------------------------------------------------- saturation fix-up --

// After chaos dispatches that used packed adds: bin i's accumulator holds
// four 16-bit fields f_k = (S_k + c_{k-1}) mod 2^16 of the true sums S_k,
// with c_k the counted carries out of field k (c_{-1} = 0), so
// S_k = f_k + 2^16 c_k - c_{k-1} exactly; each u32 word becomes
// min(word + S_k, 2^32 - 1), which is the saturating sum since all adds are
// non-negative. The accumulator and counts are cleared.
__device__ inline uint32_t add_sat64(uint32_t w, uint64_t s) {
  const uint64_t t = (uint64_t)w + s;
  return t > 0xffffffffull ? 0xffffffffu : (uint32_t)t;
}

__global__ void __launch_bounds__(256) hist_merge(uint4* hist, uint64_t* acc, ulonglong2* carry,
                                                  size_t bins) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= bins) return;
  const uint64_t a = acc[i];
  const ulonglong2 c01 = carry[2 * i], c23 = carry[2 * i + 1];
  const bool events = (c01.x | c01.y | c23.x | c23.y) != 0ull;
  if (a == 0ull && !events) return;
  const uint64_t s0 = (a & 0xffffull) + (c01.x << 16);
  const uint64_t s1 = ((a >> 16) & 0xffffull) + (c01.y << 16) - c01.x;
  const uint64_t s2 = ((a >> 32) & 0xffffull) + (c23.x << 16) - c01.y;
  const uint64_t s3 = (a >> 48) + (c23.y << 16) - c23.x;
  uint4 h = hist[i];
  h.x = add_sat64(h.x, s0);
  h.y = add_sat64(h.y, s1);
  h.z = add_sat64(h.z, s2);
  h.w = add_sat64(h.w, s3);
  hist[i] = h;
  acc[i] = 0ull;
  if (events) {
    carry[2 * i] = make_ulonglong2(0ull, 0ull);
    carry[2 * i + 1] = make_ulonglong2(0ull, 0ull);
  }
}

hipError_t launch_hist_merge(uint32_t* hist, uint64_t* acc, uint64_t* carry, size_t bins,
                             hipStream_t stream) {
  const size_t grid = (bins + 255) / 256;
  if (grid == 0) return hipSuccess;
  hist_merge<<<(unsigned)grid, 256, 0, stream>>>((uint4*)hist, acc, (ulonglong2*)carry, bins);
  return hipGetLastError();
}

// ------------------------------------------------------- atomic probe --

// Scatter-add probe in the splat's request form: `per_thread` random cells
// per thread over `cells` 8-byte cells, one relaxed device-scope 64-bit add
// per cell that puts 1 in each of its four 16-bit fields (the packed
// accumulator's form).
__global__ void __launch_bounds__(256) atomic_probe(uint64_t* grid, uint32_t cells,
                                                    uint32_t per_thread, uint32_t seed) {
  uint32_t s = (blockIdx.x * 256u + threadIdx.x) * 2654435761u ^ seed;
  for (uint32_t i = 0; i < per_thread; ++i) {
    s = s * 1664525u + 1013904223u;
    const uint32_t c = (uint32_t)(((uint64_t)s * cells) >> 32);
    (void)__hip_atomic_fetch_add(grid + c, 0x0001000100010001ull, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
  }
}

hipError_t launch_atomic_probe(uint64_t* grid, uint32_t cells, uint32_t blocks,
                               uint32_t per_thread, uint32_t seed, hipStream_t stream) {
  atomic_probe<<<blocks, 256, 0, stream>>>(grid, cells, per_thread, seed);
  return hipGetLastError();
}

// LDS budget of one tonemap_tiled block: 48 KiB holds the 32 x 8 tile up to
// os 2 with 16 taps (66 x 34 cells would be 35.9 KiB) and os 1 with up to
// 20 taps; wider filters take the per-tap HBM kernel.
constexpr size_t kTmTileBytes = 48 * 1024;

hipError_t launch_tonemap(const TmArgs& a, bool de, hipStream_t stream) {
  if (a.width == 0 || a.height == 0) return hipSuccess;
  const uint32_t span_x = (kTmTileX - 1) * a.os + a.fwidth;
  const uint32_t span_y = (kTmTileY - 1) * a.os + a.fwidth;
  const size_t bytes = (size_t)span_x * span_y * sizeof(float4);
  if (bytes <= kTmTileBytes) {
    const dim3 grid((a.width + kTmTileX - 1) / kTmTileX, (a.height + kTmTileY - 1) / kTmTileY);
    if (de) tonemap_tiled<true><<<grid, 256, bytes, stream>>>(a, span_x, span_y);
    else tonemap_tiled<false><<<grid, 256, bytes, stream>>>(a, span_x, span_y);
    return hipGetLastError();
  }
  const dim3 block(16, 16);
  const dim3 grid((a.width + 15) / 16, (a.height + 15) / 16);
  if (de) tonemap_kernel<true><<<grid, block, 0, stream>>>(a);
  else tonemap_kernel<false><<<grid, block, 0, stream>>>(a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ fill --

// Fills n4 16-byte units at p, then the n_tail (< 4) u32 words after them.
__global__ void fill_u32(uint4* p, size_t n4, uint32_t n_tail, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n4; k += stride) p[k] = make_uint4(v, v, v, v);
  if (i < n_tail) reinterpret_cast<uint32_t*>(p + n4)[i] = v;
}

// `bytes` must be a multiple of 4; every word is written, including a tail
// shorter than 16 bytes (an odd count of 8-byte accumulators, or of u32
// capture words not divisible by 4).
hipError_t launch_fill(void* p, size_t bytes, uint32_t v, hipStream_t stream) {
  if (bytes % 4 != 0) return hipErrorInvalidValue;
  const size_t n4 = bytes / 16;
  const uint32_t n_tail = (uint32_t)((bytes / 4) % 4);
  if (n4 == 0 && n_tail == 0) return hipSuccess;
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  fill_u32<<<(unsigned)blocks, 256, 0, stream>>>((uint4*)p, n4, n_tail, v);
  return hipGetLastError();
}

}  // namespace p3
