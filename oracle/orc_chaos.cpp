// Chaos game + histogram scatter — oracle restatement (FP32).
// Test infrastructure only: see oracle.hpp.
//
// Follows src/shaders/chaos.wgsl:
//   apply_pre_blur      :8164-8181   (5 rand01 per pre_blur slot)
//   apply_xform_chain   :8226-8267   (affine -> pre_blur -> sum -> post -> colour)
//   chaos_main          :8269-8732   (init draws, pick, bad/retry, final lens,
//                                     jitter, splat, saturating u32 adds :8120-8135,
//                                     colour modes flow / trap / phase :8625-8697,
//                                     capture of the colour index :8720-8728)
//   hsv_to_rgb          :1662-1666
// Buffer layouts follow src/chaos.ts:381-456 (uniform slots) and
// src/genome.ts:328-406 (116-float xform record).
#include <algorithm>
#include <thread>
#include <vector>

#include "oracle.hpp"
#include "orc_variations.hpp"

namespace orc {

namespace {

struct WalkerRng {
  Isaac s;
  uint64_t draws = 0;
  uint32_t irand() { ++draws; return isaac_irand(s); }
  float rand01() { return rand01_of(irand()); }
  float rand11() { return rand11_of(irand()); }
};

// WGSL mix(a, b, t) = a * (1 - t) + b * t
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

// WGSL u32(f32): saturating, NaN -> 0.
inline uint32_t f2u(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)f;
}

inline void add_sat(uint32_t& slot, uint32_t d) {
  if (d == 0u) return;
  slot = (0xffffffffu - slot) < d ? 0xffffffffu : slot + d;
}

struct Uniforms {
  uint32_t width, height, iters, fuse;
  float scale, cx, cy;
  uint32_t seed;
  int32_t final_idx;
  float rotation;
  uint32_t palette_mode, trace_mode, walker_count;
  float jitter;
  uint32_t capture_index, color_mode;
  float flow_strength, flow_scale;
  uint32_t trap_kind, trap_mode;
  float trap_cx, trap_cy, trap_radius, trap_nx, trap_ny, trap_falloff, trap_freq, trap_strength;
  float phase_strength, phase_freq;
  float xform_blend;
};

constexpr float kPi = 3.14159265358979323846f;

// WGSL fract(x) = x - floor(x)
inline float fractf(float x) { return x - std::floor(x); }

// chaos.wgsl:1662-1666: k = (1, 2/3, 1/3), p = |fract(h + k) * 6 - 3|,
// rgb = v * mix(1, clamp(p - 1, 0, 1), s).
inline void hsv_to_rgb(float h, float s, float v, float out[3]) {
  const float k[3] = {1.0f, 2.0f / 3.0f, 1.0f / 3.0f};
  for (int c = 0; c < 3; ++c) {
    const float p = std::fabs(fractf(h + k[c]) * 6.0f - 3.0f);
    out[c] = v * mixf(1.0f, clampf(p - 1.0f, 0.0f, 1.0f), s);
  }
}

// Colour-mode overrides of the splat colour (chaos.wgsl:8625-8697): flow
// (direction of this turn's displacement -> hue, saturated magnitude -> value),
// trap distance (palette at a glow / ring coordinate of the distance to a
// point, circle or line) and phase (arg z -> hue, log-modulus rings -> value),
// each blended over the palette colour by its strength; alpha is kept.
inline void colour_mode(const struct Uniforms& u, const float* palette, float sx, float sy,
                        float p0x, float p0y, float pal[4]) {
  float rgb[3];
  float t;
  if (u.color_mode == 1u) {
    const float dx = sx - p0x, dy = sy - p0y;
    const float mag = std::sqrt(dx * dx + dy * dy);
    if (!(mag > 1e-20f)) return;
    const float hue = (std::atan2(dy, dx) + kPi) / (2.0f * kPi);
    const float val = clampf(1.0f - std::exp(-std::max(u.flow_scale, 0.0f) * mag), 0.0f, 1.0f);
    hsv_to_rgb(hue, 1.0f, val, rgb);
    t = clampf(u.flow_strength, 0.0f, 1.0f);
  } else if (u.color_mode == 2u) {
    const float qx = sx - u.trap_cx, qy = sy - u.trap_cy;
    float d;
    if (u.trap_kind == 0u) d = std::sqrt(qx * qx + qy * qy);
    else if (u.trap_kind == 1u) d = std::fabs(std::sqrt(qx * qx + qy * qy) - u.trap_radius);
    else d = std::fabs(qx * u.trap_nx + qy * u.trap_ny);
    float tt;
    if (u.trap_mode == 0u) tt = clampf(std::exp(-std::max(u.trap_falloff, 0.0f) * d), 0.0f, 1.0f);
    else tt = fractf(d * std::max(u.trap_freq, 0.0f));
    const uint32_t idx = std::min<uint32_t>(f2u(tt * 255.0f), 255u);
    for (int c = 0; c < 3; ++c) rgb[c] = palette[idx * 4 + c];
    t = clampf(u.trap_strength, 0.0f, 1.0f);
  } else if (u.color_mode == 3u) {
    const float mag = std::sqrt(sx * sx + sy * sy);
    if (!(mag > 1e-20f)) return;
    const float hue = (std::atan2(sy, sx) + kPi) / (2.0f * kPi);
    const float ring = fractf(std::log(mag) * std::max(u.phase_freq, 0.0f));
    hsv_to_rgb(hue, 1.0f, mixf(1.0f, 0.5f, ring), rgb);
    t = clampf(u.phase_strength, 0.0f, 1.0f);
  } else {
    return;
  }
  for (int c = 0; c < 3; ++c) pal[c] = mixf(pal[c], rgb[c], t);
}

Uniforms read_uniforms(const uint32_t* u) {
  Uniforms r;
  auto f = [&](int i) { float v; std::memcpy(&v, &u[i], 4); return v; };
  r.width = u[0]; r.height = u[1]; r.iters = u[2]; r.fuse = u[3];
  r.scale = f(4); r.cx = f(5); r.cy = f(6); r.seed = u[7];
  r.final_idx = (int32_t)u[8]; r.rotation = f(9); r.palette_mode = u[10];
  r.trace_mode = u[11]; r.walker_count = u[12]; r.jitter = f(13);
  r.capture_index = u[14]; r.color_mode = u[15]; r.flow_strength = f(16); r.flow_scale = f(17);
  r.trap_kind = u[18]; r.trap_mode = u[19]; r.trap_cx = f(20); r.trap_cy = f(21);
  r.trap_radius = f(22); r.trap_nx = f(23); r.trap_ny = f(24); r.trap_falloff = f(25);
  r.trap_freq = f(26); r.trap_strength = f(27); r.phase_strength = f(28); r.phase_freq = f(29);
  r.xform_blend = f(30);
  return r;
}

struct ChainOut {
  V2<float> pv, pv_pre, pa;
  float new_z;
  bool dc_active;
  float dc_rgb[3];
};

inline void load_args(const float* xf, int k, VarArgs<float>& a) {
  const float* v = xf + 20 + k * 4;
  const float* ve = xf + 52 + k * 4;
  const float* ve2 = xf + 84 + k * 4;
  a.w = v[1];
  a.p[0] = v[2]; a.p[1] = v[3];
  a.p[2] = ve[0]; a.p[3] = ve[1]; a.p[4] = ve[2]; a.p[5] = ve[3];
  a.p[6] = ve2[0]; a.p[7] = ve2[1]; a.p[8] = ve2[2]; a.p[9] = ve2[3];
  for (int i = 0; i < 4; ++i) { a.a0[i] = xf[i]; a.a1[i] = xf[4 + i]; }
}

inline int nvars_of(const float* xf) {
  uint32_t n = f2u(xf[7]);
  return (int)std::min<uint32_t>(n, MAX_VARS);
}

// pre_blur (index 97): five draws per slot, Gaussian-ish radius from the
// first four, angle from the fifth; the slot weight is folded into the radius.
V2<float> pre_blur(const float* xf, int nv, V2<float> c, WalkerRng& rng) {
  for (int k = 0; k < nv; ++k) {
    const float* v = xf + 20 + k * 4;
    if (f2u(v[0]) == 97u) {
      float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
      float g = v[1] * (r0 + r1 + r2 + r3 - 2.0f);
      float r4 = rng.rand01();
      float ang = r4 * K<float>::TAU;
      c.x = c.x + std::cos(ang) * g;
      c.y = c.y + std::sin(ang) * g;
    }
  }
  return c;
}

// Variation sum + direct-colour override over one xform's slots.
V2<float> var_sum(const float* xf, int nv, V2<float> pa, WalkerRng& rng,
                  bool& dc_active, float* dc_rgb) {
  V2<float> pv{0.0f, 0.0f};
  VarArgs<float> a;
  for (int k = 0; k < nv; ++k) {
    const float* v = xf + 20 + k * 4;
    uint32_t idx = f2u(v[0]);
    if (idx != 97u) {
      load_args(xf, k, a);
      V2<float> o = apply_variation<float>((int)idx, pa, a, rng);
      pv.x = pv.x + o.x;
      pv.y = pv.y + o.y;
    }
    if (xf[11] > 0.5f && v[1] > 0.0f) {
      float rgb[3];
      if (dc_color_for((int)idx, pa, xf, k, rgb)) {
        dc_active = true;
        dc_rgb[0] = rgb[0]; dc_rgb[1] = rgb[1]; dc_rgb[2] = rgb[2];
      }
    }
  }
  return pv;
}

inline V2<float> post_affine(const float* xf, V2<float> pv) {
  if (xf[15] != 0.0f) {
    float x = xf[12] * pv.x + xf[13] * pv.y + xf[14];
    float y = xf[16] * pv.x + xf[17] * pv.y + xf[18];
    return {x, y};
  }
  return pv;
}

ChainOut xform_chain(const float* xf, float px, float py, float pz, WalkerRng& rng) {
  ChainOut o;
  o.pa = {xf[0] * px + xf[1] * py + xf[2], xf[4] * px + xf[5] * py + xf[6]};
  int nv = nvars_of(xf);
  V2<float> pam = pre_blur(xf, nv, o.pa, rng);
  o.dc_active = false;
  o.dc_rgb[0] = o.dc_rgb[1] = o.dc_rgb[2] = 0.0f;
  V2<float> pv = var_sum(xf, nv, pam, rng, o.dc_active, o.dc_rgb);
  o.pv_pre = pv;
  o.pv = post_affine(xf, pv);
  o.new_z = mixf(pz, xf[8], xf[9]);
  return o;
}

inline bool is_bad(V2<float> v) {
  return std::isnan(v.x) || std::isnan(v.y) || std::fabs(v.x) > 1e10f || std::fabs(v.y) > 1e10f;
}

inline void put(float* t, int i, float v) { t[i] = v; }
inline void putu(float* t, int i, uint32_t u) { std::memcpy(&t[i], &u, 4); }

}  // namespace

// Runs walkers [wb, we) of one dispatch and accumulates into `hist`.
// `force` (optional, trace-shaped): for traced walkers, turn i > 0 starts
// from the recorded end point of turn i - 1 (x, y at [12, 13], z at [9]),
// so each turn is a one-step evaluation from the other implementation's
// state (used to check per-step arithmetic without chaotic amplification).
void chaos(const float* xforms, const uint32_t* distrib, int n_rows, const uint32_t* uni,
           const float* palette, uint32_t* hist, uint32_t wb, uint32_t we,
           float* trace, int trace_walkers, int trace_iters, const float* force = nullptr,
           uint32_t* idx_sum = nullptr) {
  const Uniforms u = read_uniforms(uni);
  const float cos_r = std::cos(u.rotation);
  const float sin_r = std::sin(u.rotation);
  const uint32_t total = u.iters + u.fuse;
  auto row_ptr = [&](int row) -> const uint32_t* {
    int r = row == MAX_XFORMS ? n_rows - 1 : row;
    return distrib + (size_t)r * GRAIN;
  };
  for (uint32_t w = wb; w < we && w < u.walker_count; ++w) {
    WalkerRng rng;
    seed_walker(rng.s, u.seed, w);
    float x = rng.rand11();
    float y = rng.rand11();
    float z = rng.rand01();
    int prev = -1;
    uint32_t consec_bad = 0;
    for (uint32_t i = 0; i < total; ++i) {
      if (force && i > 0 && (int)w < trace_walkers && (int)i <= trace_iters) {
        const float* f = force + ((size_t)w * trace_iters + (i - 1)) * 16;
        x = f[12]; y = f[13]; z = f[9];
      }
      const float p0x = x, p0y = y;  // position before this turn's chain
      int row = prev >= 0 ? prev : MAX_XFORMS;
      uint32_t raw = rng.irand();
      uint32_t fn = row_ptr(row)[raw & (GRAIN - 1)];
      const float* xf = xforms + (size_t)fn * XFORM_FLOATS;
      ChainOut A = xform_chain(xf, x, y, z, rng);
      V2<float> pv = A.pv;
      float new_z = A.new_z;
      bool dc_active = A.dc_active;
      float dc_rgb[3] = {A.dc_rgb[0], A.dc_rgb[1], A.dc_rgb[2]};
      if (u.xform_blend > 0.0f) {
        if (rng.rand01() < u.xform_blend) {
          uint32_t rb = rng.irand();
          uint32_t bi = row_ptr(MAX_XFORMS)[rb & (GRAIN - 1)];
          ChainOut B = xform_chain(xforms + (size_t)bi * XFORM_FLOATS, x, y, z, rng);
          float t = rng.rand01();
          pv.x = mixf(pv.x, B.pv.x, t);
          pv.y = mixf(pv.y, B.pv.y, t);
          new_z = mixf(new_z, B.new_z, t);
        }
      }
      bool bad = is_bad(pv);
      if (bad) {
        float rx = rng.rand11();
        float ry = rng.rand11();
        pv = {rx, ry};
        consec_bad += 1;
      } else {
        consec_bad = 0;
      }
      float* tr = nullptr;
      if (trace && (int)w < trace_walkers && (int)i < trace_iters) {
        tr = trace + ((size_t)w * trace_iters + i) * 16;
        put(tr, 0, (float)i); put(tr, 1, (float)fn);
        put(tr, 2, A.pa.x); put(tr, 3, A.pa.y);
        put(tr, 4, A.pv_pre.x); put(tr, 5, A.pv_pre.y);
        put(tr, 6, pv.x); put(tr, 7, pv.y);
        put(tr, 8, bad ? 1.0f : 0.0f); put(tr, 9, new_z);
        putu(tr, 11, raw);
        put(tr, 12, pv.x); put(tr, 13, pv.y); put(tr, 14, 0.0f); put(tr, 15, 0.0f);
      }
      const float ppx = pv.x, ppy = pv.y, ppz = new_z;  // p_pre_final
      if (consec_bad > 0 && consec_bad < 5) {
        x = ppx; y = ppy; z = ppz;
        if (tr) putu(tr, 10, (uint32_t)rng.draws);
        if (i > 0) i -= 1;
        continue;
      }
      if (consec_bad >= 5) consec_bad = 0;
      prev = (int)fn;
      float sx = ppx, sy = ppy, sz = ppz;
      if (u.final_idx >= 0) {
        const float* fx = xforms + (size_t)u.final_idx * XFORM_FLOATS;
        bool apply = fx[10] == 1.0f ? true : (rng.rand01() < fx[10]);
        if (apply) {
          V2<float> fpa{fx[0] * pv.x + fx[1] * pv.y + fx[2], fx[4] * pv.x + fx[5] * pv.y + fx[6]};
          int fnv = nvars_of(fx);
          V2<float> fpam = pre_blur(fx, fnv, fpa, rng);
          V2<float> fpv = var_sum(fx, fnv, fpam, rng, dc_active, dc_rgb);
          fpv = post_affine(fx, fpv);
          float fz = mixf(new_z, fx[8], fx[9]);
          if (!is_bad(fpv)) { sx = fpv.x; sy = fpv.y; sz = fz; }
        }
      }
      float mag = std::max(std::max(std::fabs(ppx), std::fabs(ppy)), 1e-30f);
      float amp = mag * u.jitter;
      float jx = (rng.rand01() - 0.5f) * amp;
      float jy = (rng.rand01() - 0.5f) * amp;
      x = ppx + jx; y = ppy + jy; z = ppz;
      if (tr) {
        put(tr, 12, x); put(tr, 13, y); put(tr, 14, sx); put(tr, 15, sy);
        putu(tr, 10, (uint32_t)rng.draws);
      }
      if (i >= u.fuse) {
        float opacity = xf[10];
        float dx = sx - u.cx, dy = sy - u.cy;
        float rx = dx * cos_r - dy * sin_r;
        float ry = dx * sin_r + dy * cos_r;
        float px = rx * u.scale + (float)u.width * 0.5f;
        float py = ry * u.scale + (float)u.height * 0.5f;
        // floor(px) in [0, W) <=> px in [0, W) for finite px; NaN fails both.
        if (px >= 0.0f && px < (float)u.width && py >= 0.0f && py < (float)u.height) {
          uint32_t xi = (uint32_t)std::floor(px), yi = (uint32_t)std::floor(py);
          float cxf = clampf(sz, 0.0f, 1.0f) * 255.0f;
          float pal[4];
          if (u.palette_mode == 1u) {
            uint32_t i0 = f2u(cxf);
            float frac = cxf - (float)i0;
            if (i0 >= 255u) { i0 = 254u; frac = 1.0f; }
            for (int c = 0; c < 4; ++c)
              pal[c] = mixf(palette[i0 * 4 + c], palette[(i0 + 1) * 4 + c], frac);
          } else {
            uint32_t pi = std::min<uint32_t>(f2u(cxf), 255u);
            for (int c = 0; c < 4; ++c) pal[c] = palette[pi * 4 + c];
          }
          if (dc_active) { pal[0] = dc_rgb[0]; pal[1] = dc_rgb[1]; pal[2] = dc_rgb[2]; }
          colour_mode(u, palette, sx, sy, p0x, p0y, pal);
          float wgt = opacity * 255.0f;
          uint32_t* b = hist + ((size_t)yi * u.width + xi) * 4;
          add_sat(b[0], f2u(pal[0] * wgt));
          add_sat(b[1], f2u(pal[1] * wgt));
          add_sat(b[2], f2u(pal[2] * wgt));
          add_sat(b[3], f2u(wgt));
          if (u.capture_index == 1u && idx_sum)
            add_sat(idx_sum[(size_t)yi * u.width + xi], f2u(clampf(sz, 0.0f, 1.0f) * wgt));
        }
      }
    }
  }
}

}  // namespace orc

extern "C" {

// Single-threaded run of walkers [wb, we).
void orc_chaos(const float* xforms, const uint32_t* distrib, int n_rows, const uint32_t* uni,
               const float* palette, uint32_t* hist, uint32_t wb, uint32_t we,
               float* trace, int trace_walkers, int trace_iters) {
  orc::chaos(xforms, distrib, n_rows, uni, palette, hist, wb, we, trace, trace_walkers,
             trace_iters);
}

// As orc_chaos, with the capture buffer (one u32 per bin) for capture_index.
void orc_chaos_capture(const float* xforms, const uint32_t* distrib, int n_rows,
                       const uint32_t* uni, const float* palette, uint32_t* hist,
                       uint32_t* idx_sum, uint32_t wb, uint32_t we) {
  orc::chaos(xforms, distrib, n_rows, uni, palette, hist, wb, we, nullptr, 0, 0, nullptr,
             idx_sum);
}

// Whole dispatch on `threads` std::threads: walkers split in contiguous
// ranges, private histograms merged with saturating adds (order-independent).
void orc_chaos_mt(const float* xforms, const uint32_t* distrib, int n_rows, const uint32_t* uni,
                  const float* palette, uint32_t* hist, int threads) {
  const uint32_t W = uni[0], H = uni[1], walkers = uni[12];
  const size_t n = (size_t)W * H * 4;
  if (threads < 1) threads = 1;
  std::vector<std::vector<uint32_t>> priv(threads);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    uint32_t wb = (uint32_t)((uint64_t)walkers * t / threads);
    uint32_t we = (uint32_t)((uint64_t)walkers * (t + 1) / threads);
    pool.emplace_back([&, t, wb, we] {
      priv[t].assign(n, 0u);
      orc::chaos(xforms, distrib, n_rows, uni, palette, priv[t].data(), wb, we, nullptr, 0, 0);
    });
  }
  for (auto& th : pool) th.join();
  for (int t = 0; t < threads; ++t)
    for (size_t i = 0; i < n; ++i) orc::add_sat(hist[i], priv[t][i]);
}

// Single-threaded, with each traced turn restarted from `force` (see chaos()).
void orc_chaos_follow(const float* xforms, const uint32_t* distrib, int n_rows,
                      const uint32_t* uni, const float* palette, uint32_t* hist, float* trace,
                      int trace_walkers, int trace_iters, const float* force) {
  orc::chaos(xforms, distrib, n_rows, uni, palette, hist, 0, (uint32_t)trace_walkers, trace,
             trace_walkers, trace_iters, force);
}

}  // extern "C"
