// Density estimation, tonemap and host-side tables — oracle restatement.
// Test infrastructure only: see oracle.hpp.
//
// kernel_norm_lut   follows src/density.ts:78-99   (f64 sums, stored f32)
// density           follows src/shaders/density.wgsl:51-123
// spatial_kernel    follows src/spatial-filter.ts:217-256 (gaussian, f64 -> f32)
// tonemap           follows src/shaders/visualize_f32.wgsl:96-271 (DE on) and
//                   src/shaders/visualize_u32.wgsl:148-248 (DE off)
// calibration       follows src/calibration.ts:37-43
// build_distrib     follows src/genome.ts:495-534 (f64 cumulative scan)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.hpp"

namespace {

inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }
inline float fget(const uint32_t* u, int i) { float v; std::memcpy(&v, &u[i], 4); return v; }

struct DeU { uint32_t W, H; float max_rad, min_rad, curve, k1, k2; };
DeU read_de(const uint32_t* u) {
  return {u[0], u[1], fget(u, 2), fget(u, 3), fget(u, 4), fget(u, 5), fget(u, 6)};
}

void density_rows(const uint32_t* hist, const DeU& u, const float* lut, float* out, uint32_t y0,
                  uint32_t y1) {
  const int R = (int)std::ceil(clampf(u.max_rad, 0.0f, 256.0f));
  const int W = (int)u.W, H = (int)u.H;
  for (int yi = (int)y0; yi < (int)y1; ++yi) {
    for (int xi = 0; xi < W; ++xi) {
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
      for (int dy = -R; dy <= R; ++dy) {
        int ny = yi + dy;
        if (ny < 0 || ny >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
          int nx = xi + dx;
          if (nx < 0 || nx >= W) continue;
          const uint32_t* b = hist + ((size_t)ny * W + nx) * 4;
          float cnt = (float)b[3];
          if (cnt <= 0.0f) continue;
          float hits = cnt / 255.0f;
          float rf = clampf(u.max_rad / std::pow(hits + 1.0f, u.curve), u.min_rad, u.max_rad);
          // WGSL round() is half-to-even: rintf under the default rounding mode.
          int ri = std::min(std::max((int)std::rint(rf), 0), 256);
          float rad = (float)ri;
          float d2 = (float)(dx * dx + dy * dy);
          if (d2 > rad * rad) continue;
          float sigma = std::max(rad / 3.0f, 1e-6f);
          float kw = std::exp(-d2 / (2.0f * sigma * sigma));
          float knorm = std::max(lut[ri], 1e-6f);
          float ls = (u.k1 * std::log(1.0f + cnt * u.k2)) / cnt;
          float w = (kw / knorm) * ls;
          s0 = s0 + (float)b[0] * w;
          s1 = s1 + (float)b[1] * w;
          s2 = s2 + (float)b[2] * w;
          s3 = s3 + cnt * w;
        }
      }
      float* o = out + ((size_t)yi * W + xi) * 4;
      o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
    }
  }
}

struct VizU {
  uint32_t W, H;
  float k1, k2, gamma, vibrancy, highpow, linrange;
  uint32_t os, fwidth;
  float bg[3];
  uint32_t transparent;
  uint32_t curves_mask, hsl_on;   // grading: words 16..20
  float hsl_hue, hsl_sat, hsl_light;
};
VizU read_viz(const uint32_t* u) {
  VizU v;
  v.W = u[0]; v.H = u[1];
  v.k1 = fget(u, 2); v.k2 = fget(u, 3); v.gamma = fget(u, 4); v.vibrancy = fget(u, 5);
  v.highpow = fget(u, 6); v.linrange = fget(u, 7);
  v.os = u[8]; v.fwidth = u[9];
  v.bg[0] = fget(u, 12); v.bg[1] = fget(u, 13); v.bg[2] = fget(u, 14);
  v.transparent = u[21];
  v.curves_mask = u[16]; v.hsl_on = u[17];
  v.hsl_hue = fget(u, 18); v.hsl_sat = fget(u, 19); v.hsl_light = fget(u, 20);
  return v;
}

float calc_alpha(float d, float g, float lin) {
  if (d <= 0.0f) return 0.0f;
  if (d < lin) {
    float fv = std::pow(lin, g);
    float frac = d / lin;
    return (1.0f - frac) * d * (fv / lin) + frac * std::pow(d, g);
  }
  return std::pow(d, g);
}

void rgb2hsv(const float c[3], float o[3]) {
  float mx = std::max(std::max(c[0], c[1]), c[2]);
  float mn = std::min(std::min(c[0], c[1]), c[2]);
  float d = mx - mn;
  float h = 0.0f;
  if (d > 0.0f) {
    if (mx == c[0]) h = (c[1] - c[2]) / d;
    else if (mx == c[1]) h = 2.0f + (c[2] - c[0]) / d;
    else h = 4.0f + (c[0] - c[1]) / d;
    h *= 60.0f;
    if (h < 0.0f) h += 360.0f;
  }
  o[0] = h; o[1] = mx > 0.0f ? d / mx : 0.0f; o[2] = mx;
}

void hsv2rgb(const float c[3], float o[3]) {
  float h = c[0] / 60.0f, s = c[1], v = c[2];
  float i = std::floor(h);
  float f = h - i;
  float p = v * (1.0f - s), q = v * (1.0f - s * f), t = v * (1.0f - s * (1.0f - f));
  int ix = (int)i % 6;  // WGSL i32 % keeps the dividend's sign
  float r[3];
  switch (ix) {
    case 0: r[0] = v; r[1] = t; r[2] = p; break;
    case 1: r[0] = q; r[1] = v; r[2] = p; break;
    case 2: r[0] = p; r[1] = v; r[2] = t; break;
    case 3: r[0] = p; r[1] = q; r[2] = v; break;
    case 4: r[0] = t; r[1] = p; r[2] = v; break;
    default: r[0] = v; r[1] = p; r[2] = q; break;
  }
  o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}

void calc_newrgb(const float c[3], float ls, float highpow, float o[3]) {
  if (ls == 0.0f || (c[0] == 0.0f && c[1] == 0.0f && c[2] == 0.0f)) {
    o[0] = o[1] = o[2] = 0.0f;
    return;
  }
  float sc[3], cw[3];
  for (int i = 0; i < 3; ++i) { cw[i] = c[i] / 255.0f; sc[i] = ls * cw[i]; }
  float maxa = sc[0], maxc = c[0] / 255.0f;
  if (sc[1] > maxa) { maxa = sc[1]; maxc = c[1] / 255.0f; }
  if (sc[2] > maxa) { maxa = sc[2]; maxc = c[2] / 255.0f; }
  if (maxa > 255.0f && highpow >= 0.0f) {
    float newls = 255.0f / maxc;
    float lsratio = std::pow(newls / ls, highpow);
    float nrgb[3], hsv[3];
    for (int i = 0; i < 3; ++i) nrgb[i] = newls * cw[i] / 255.0f;
    rgb2hsv(nrgb, hsv);
    hsv[1] *= lsratio;
    hsv2rgb(hsv, nrgb);
    for (int i = 0; i < 3; ++i) o[i] = nrgb[i] * 255.0f;
  } else {
    float newls = 255.0f / maxc;
    float adj = -highpow;
    if (adj > 1.0f) adj = 1.0f;
    if (maxa <= 255.0f) adj = 1.0f;
    float m = (1.0f - adj) * newls + adj * ls;
    for (int i = 0; i < 3; ++i) o[i] = m * cw[i];
  }
}

// visualize_f32.wgsl:63-68 lut(): linear lookup in channel `ch` of the
// 5 x 256 curves LUT.
float curve_lut(const float* L, uint32_t ch, float x) {
  const float idx = clampf(x, 0.0f, 1.0f) * 255.0f;
  const uint32_t i0 = (uint32_t)std::floor(idx);
  const uint32_t i1 = std::min(i0 + 1u, 255u);
  const float t = idx - (float)i0;
  return L[ch * 256u + i0] * (1.0f - t) + L[ch * 256u + i1] * t;
}

// visualize_f32.wgsl:240-266: channel curves (luma, then R, G, B, then the
// composite curve on all three), then the HSV adjust (hue offset wrapped
// into [0, 360), saturation scaled, value offset, both clamped to [0, 1]).
void grade(const VizU& u, const float* curves, float o[3]) {
  if (u.curves_mask != 0u) {
    if (u.curves_mask & 16u) {
      const float l = o[0] * 0.299f + o[1] * 0.587f + o[2] * 0.114f;
      const float ln = curve_lut(curves, 4u, l);
      if (l > 0.0f) {
        const float s = ln / l;
        o[0] = o[0] * s; o[1] = o[1] * s; o[2] = o[2] * s;
      }
    }
    if (u.curves_mask & 2u) o[0] = curve_lut(curves, 1u, o[0]);
    if (u.curves_mask & 4u) o[1] = curve_lut(curves, 2u, o[1]);
    if (u.curves_mask & 8u) o[2] = curve_lut(curves, 3u, o[2]);
    if (u.curves_mask & 1u)
      for (int i = 0; i < 3; ++i) o[i] = curve_lut(curves, 0u, o[i]);
  }
  if (u.hsl_on != 0u) {
    float hsv[3];
    rgb2hsv(o, hsv);
    hsv[0] = hsv[0] + u.hsl_hue;
    hsv[0] = hsv[0] - 360.0f * std::floor(hsv[0] / 360.0f);
    hsv[1] = clampf(hsv[1] * u.hsl_sat, 0.0f, 1.0f);
    hsv[2] = clampf(hsv[2] + u.hsl_light, 0.0f, 1.0f);
    hsv2rgb(hsv, o);
  }
}

// One output pixel. `src_f` (DE on, filtered f32) or `src_u` (DE off, raw
// u32 histogram with the per-super-pixel log-density inside the loop).
void tonemap_px(const VizU& u, const float* taps, const float* src_f, const uint32_t* src_u,
                const float* curves, uint32_t xi, uint32_t yi, float out[4]) {
  const uint32_t N = u.os, fw = u.fwidth, half = fw / 2u;
  const int SW = (int)(u.W * N), SH = (int)(u.H * N);
  const uint32_t cx = xi * N + N / 2u, cy = yi * N + N / 2u;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
  for (uint32_t dy = 0; dy < fw; ++dy) {
    float ky = taps[dy];
    int sy = std::min(std::max((int)cy + (int)dy - (int)half, 0), SH - 1);
    for (uint32_t dx = 0; dx < fw; ++dx) {
      float kx = taps[dx];
      int sx = std::min(std::max((int)cx + (int)dx - (int)half, 0), SW - 1);
      size_t sb = ((size_t)sy * SW + sx) * 4;
      float w = kx * ky;
      if (src_f) {
        c0 = c0 + src_f[sb + 0] * w; c1 = c1 + src_f[sb + 1] * w;
        c2 = c2 + src_f[sb + 2] * w; c3 = c3 + src_f[sb + 3] * w;
      } else {
        float cnt = (float)src_u[sb + 3];
        if (cnt > 0.0f) {
          float ls = (u.k1 * std::log(1.0f + cnt * u.k2)) / cnt;
          c0 = c0 + ((float)src_u[sb + 0] * ls) * w;
          c1 = c1 + ((float)src_u[sb + 1] * ls) * w;
          c2 = c2 + ((float)src_u[sb + 2] * ls) * w;
          c3 = c3 + (cnt * ls) * w;
        }
      }
    }
  }
  if (c3 <= 0.0f) {
    if (u.transparent != 0u) { out[0] = out[1] = out[2] = out[3] = 0.0f; return; }
    out[0] = u.bg[0]; out[1] = u.bg[1]; out[2] = u.bg[2]; out[3] = 1.0f;
    return;
  }
  float g = 1.0f / u.gamma;
  float tmp = c3 / 255.0f;
  float alpha = calc_alpha(tmp, g, u.linrange);
  float ls_a = u.vibrancy * 256.0f * alpha / std::max(tmp, 1e-12f);
  alpha = clampf(alpha, 0.0f, 1.0f);
  float c[3] = {c0, c1, c2}, nrgb[3];
  calc_newrgb(c, ls_a, u.highpow, nrgb);
  for (int i = 0; i < 3; ++i) {
    float perch = (1.0f - u.vibrancy) * 256.0f * std::pow(std::max(c[i] / 255.0f, 0.0f), g);
    float bgt = u.transparent != 0u ? 0.0f : (1.0f - alpha) * 256.0f * u.bg[i];
    float composed = nrgb[i] + perch + bgt;
    out[i] = clampf(composed / 256.0f, 0.0f, 1.0f);
  }
  grade(u, curves, out);
  for (int i = 0; i < 3; ++i) out[i] = clampf(out[i], 0.0f, 1.0f);
  out[3] = u.transparent != 0u ? alpha : 1.0f;
}

}  // namespace

extern "C" {

void orc_kernel_norm_lut(float* out257) {
  for (int r = 0; r <= 256; ++r) {
    if (r == 0) { out257[0] = 1.0f; continue; }
    double sigma = r / 3.0;
    double inv = 1.0 / (2.0 * sigma * sigma);
    double r2 = (double)r * r, sum = 0.0;
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        double d2 = (double)dx * dx + (double)dy * dy;
        if (d2 > r2) continue;
        sum += std::exp(-d2 * inv);
      }
    out257[r] = (float)sum;
  }
}

// de_uniforms: the 32-byte DEUniforms block (u32 W_s, H_s; f32 maxRad_s,
// minRad_s, curve, k1, k2; pad).
void orc_density(const uint32_t* hist, const uint32_t* de_uniforms, const float* lut, float* out,
                 int threads) {
  DeU u = read_de(de_uniforms);
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    uint32_t y0 = (uint32_t)((uint64_t)u.H * t / threads);
    uint32_t y1 = (uint32_t)((uint64_t)u.H * (t + 1) / threads);
    pool.emplace_back([=] { density_rows(hist, u, lut, out, y0, y1); });
  }
  for (auto& th : pool) th.join();
}

// viz_uniforms: the 96-byte VizUniforms block. Exactly one of filtered/hist
// is non-null; `curves` is the 5 x 256 LUT (read only when word 16 is non-zero).
// Writes f32 RGBA (pre-quantisation) and RGBA8 (round to nearest).
void orc_tonemap(const uint32_t* viz_uniforms, const float* taps, const float* filtered,
                 const uint32_t* hist, const float* curves, float* out_f32, uint8_t* out_u8,
                 int threads) {
  VizU u = read_viz(viz_uniforms);
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    uint32_t y0 = (uint32_t)((uint64_t)u.H * t / threads);
    uint32_t y1 = (uint32_t)((uint64_t)u.H * (t + 1) / threads);
    pool.emplace_back([=] {
      for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = 0; x < u.W; ++x) {
          float px[4];
          tonemap_px(u, taps, filtered, hist, curves, x, y, px);
          size_t o = ((size_t)y * u.W + x) * 4;
          for (int c = 0; c < 4; ++c) {
            if (out_f32) out_f32[o + c] = px[c];
            if (out_u8) out_u8[o + c] = (uint8_t)(clampf(px[c], 0.0f, 1.0f) * 255.0f + 0.5f);
          }
        }
    });
  }
  for (auto& th : pool) th.join();
}

// Gaussian spatial kernel (filters.c-style support 1.5), returns the tap count.
int orc_spatial_kernel(double radius, int oversample, float* out17) {
  const double supp = 1.5;
  double fw = 2.0 * supp * oversample * radius;
  int fwidth = (int)std::floor(fw) + 1;
  if (((fwidth ^ oversample) & 1) != 0) fwidth++;
  fwidth = std::max(3, std::min(17, fwidth));
  double adjust = fw > 0 ? supp * fwidth / fw : 1.0;
  double sum = 0.0;
  for (int i = 0; i < fwidth; ++i) {
    double x = ((2.0 * i + 1.0) / fwidth - 1.0) * adjust;
    double v = std::exp(-2.0 * x * x);
    out17[i] = (float)v;
    sum += v;
  }
  if (sum != 0.0)
    for (int i = 0; i < fwidth; ++i) out17[i] = (float)((double)out17[i] / sum);
  return fwidth;
}

void orc_calibration(double scale, double samples, double brightness, int oversample,
                     double* k1, double* k2) {
  double os2 = (double)oversample * oversample;
  *k1 = brightness * 255.0 * 268.0 / 256.0;
  *k2 = os2 * scale * scale / (255.0 * samples);
}

// Pick table: rows 0..n-1 ("previous xform = i", weights scaled by xaos[i][j])
// followed by the no-previous row. xaos is n*n row-major or null.
// Returns -1 on a negative/non-finite weight (the reference throws).
int orc_build_distrib(int n, const double* weights, const double* xaos, uint32_t* out) {
  for (int row = 0; row <= n; ++row) {
    const int xi = row < n ? row : -1;
    uint32_t* o = out + (size_t)row * orc::GRAIN;
    auto mul = [&](int j) { return (xi >= 0 && xaos) ? xaos[(size_t)xi * n + j] : 1.0; };
    double total = 0.0;
    for (int j = 0; j < n; ++j) {
      double d = weights[j] * mul(j);
      if (!std::isfinite(d) || d < 0) return -1;
      total += d;
    }
    if (total == 0.0) { std::memset(o, 0, sizeof(uint32_t) * orc::GRAIN); continue; }
    double dr = total / orc::GRAIN;
    int j = 0;
    double t = weights[0] * mul(0);
    double r = 0.0;
    for (int i = 0; i < orc::GRAIN; ++i) {
      while (r >= t && j < n - 1) {
        ++j;
        t += weights[j] * mul(j);
      }
      o[i] = (uint32_t)j;
      r += dr;
    }
  }
  return 0;
}

}  // extern "C"
