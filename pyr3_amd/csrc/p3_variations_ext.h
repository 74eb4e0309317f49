// Variation catalog entries V99..V325 for the gfx950 chaos kernel.
//
// FP32 expressions of the WGSL functions of the same names in
// src/shaders/chaos.wgsl (V99..V325 at :1682-7729, dispatch :7848-8106), with
// the reference's operation order and RNG draw order (every draw the WGSL
// makes is made, in order, whether or not its value is used). WGSL
// semantics kept: f32 `%` is x - y*trunc(x/y), i32(f)/u32(f) saturate, i32
// arithmetic wraps (done in uint32_t), select() evaluates both sides,
// round() is half-to-even, clamp(x, lo, hi) = min(max(x, lo), hi),
// mix(a, b, t) = a(1-t) + bt, and vector +/- of a literal (x, 0) keeps the
// +0 / -0 on the zero lane (it can change the sign of a zero).
#pragma once
#include "p3_variations.h"

namespace p3 {

// ---- shared helpers ----------------------------------------------------------

P3_INLINE float fract(float x) { return x - floorf(x); }
P3_INLINE float clampv(float x, float lo, float hi) { return smin(smax(x, lo), hi); }
P3_INLINE float mixv(float a, float b, float t) { return a * ((float)1 - t) + b * t; }
P3_INLINE float rint_even(float x) { return rintf(x); }
P3_INLINE float signv(float x) { return x > (float)0 ? (float)1 : (x < (float)0 ? (float)-1 : (float)0); }
P3_INLINE uint32_t f2u(float f) {  // WGSL u32(f): saturating, NaN -> 0
  if (!(f == f) || f <= (float)0) return 0u;
  if (f >= (float)4294967295.0) return 4294967295u;
  return (uint32_t)f;
}
P3_INLINE int32_t imin(int32_t a, int32_t b) { return b < a ? b : a; }
P3_INLINE int32_t imax(int32_t a, int32_t b) { return a < b ? b : a; }
P3_INLINE uint32_t umin(uint32_t a, uint32_t b) { return b < a ? b : a; }
P3_INLINE int32_t imod(int32_t a, int32_t b) { return b == 0 ? a : a % b; }  // WGSL i32 %
P3_INLINE int32_t idiv(int32_t a, int32_t b) { return b == 0 ? a : a / b; }  // WGSL i32 /
P3_INLINE int32_t iwrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
P3_INLINE int32_t iwrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// Complex helpers, z = (re, im) (chaos.wgsl:389-460).
P3_INLINE float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
P3_INLINE float2 c_sqr(float2 z) { return make_float2(z.x * z.x - z.y * z.y, (float)2 * z.x * z.y); }
// |b|^2 floored at the f32 value of 1e-100 (zero: the floor is a no-op in f32).
P3_INLINE float2 c_div(float2 a, float2 b) {
  float m2 = smax(b.x * b.x + b.y * b.y, (float)(float)1e-100);
  return make_float2((a.x * b.x + a.y * b.y) / m2, (a.y * b.x - a.x * b.y) / m2);
}
P3_INLINE float2 c_recip(float2 z) {
  float m2 = smax(z.x * z.x + z.y * z.y, (float)(float)1e-100);
  return make_float2(z.x / m2, -z.y / m2);
}
P3_INLINE float2 c_sqrt(float2 z) {
  float rad = len2(z);
  float sb = sel((float)1, (float)-1, z.y < (float)0);
  return make_float2(sqrtf(smax((float)0.5 * (rad + z.x), (float)0)),
          sb * sqrtf(smax((float)0.5 * (rad - z.x), (float)0)));
}
P3_INLINE float2 c_log(float2 z) {
  float m2 = (z.x * z.x + z.y * z.y) + (float)1e-20;
  return make_float2((float)0.5 * logf(m2), atan2f(z.y, z.x));
}
P3_INLINE float2 c_exp(float2 z) {
  float e = expf(clampv(z.x, (float)-20, (float)20));
  return make_float2(e * safe_cos(z.y), e * safe_sin(z.y));
}
P3_INLINE float2 c_pow(float2 t, float2 p) { return c_exp(c_mul(p, c_log(t))); }
P3_INLINE float2 c_sin(float2 z) {
  float y = clampv(z.y, (float)-20, (float)20);
  float ep = expf(y), en = expf(-y);
  float ch = (float)0.5 * (ep + en), sh = (float)0.5 * (ep - en);
  return make_float2(safe_sin(z.x) * ch, safe_cos(z.x) * sh);
}

// Improved-noise gradient lattice (src/shaders/noise_perlin.wgsl): Ken
// Perlin's published 256-entry permutation, indexed modulo 256.
__constant__ uint8_t kPerlinPerm[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225,
    140, 36,  103, 30,  69,  142, 8,   99,  37,  240, 21,  10,  23,  190, 6,   148,
    247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203, 117, 35,  11,  32,
    57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175,
    74,  165, 71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122,
    60,  211, 133, 230, 220, 105, 92,  41,  55,  46,  245, 40,  244, 102, 143, 54,
    65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,  18,  169,
    200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,
    52,  217, 226, 250, 124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212,
    207, 206, 59,  227, 47,  16,  58,  17,  182, 189, 28,  42,  223, 183, 170, 213,
    119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104,
    218, 246, 97,  228, 251, 34,  242, 193, 238, 210, 144, 12,  191, 179, 162, 241,
    81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,  181, 199, 106, 157,
    184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,
    222, 114, 67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

P3_INLINE uint32_t perm(uint32_t i) { return kPerlinPerm[i & 255u]; }

P3_INLINE float perlin_fade(float t) {
  return t * t * t * (t * (t * (float)6 - (float)15) + (float)10);
}
// Gradient pick: low 3 bits of the hash choose axis order, signs, and a
// doubled second component.
P3_INLINE float perlin_grad2(uint32_t h, float x, float y) {
  h &= 7u;
  float u = h < 4u ? x : y;
  float v = h < 4u ? y : x;
  float a = (h & 1u) ? -u : u;
  float b = (h & 2u) ? (float)-2 * v : (float)2 * v;
  return a + b;
}
P3_INLINE float perlin2d(float px, float py) {
  float fx = floorf(px), fy = floorf(py);
  uint32_t X = (uint32_t)(f2i(fx) & 255), Y = (uint32_t)(f2i(fy) & 255);
  float xf = px - fx, yf = py - fy;
  float u = perlin_fade(xf), v = perlin_fade(yf);
  uint32_t A = (perm(X) + Y) & 255u, B = (perm(X + 1u) + Y) & 255u;
  float g00 = perlin_grad2(perm(A), xf, yf);
  float g10 = perlin_grad2(perm(B), xf - (float)1, yf);
  float g01 = perlin_grad2(perm(A + 1u), xf, yf - (float)1);
  float g11 = perlin_grad2(perm(B + 1u), xf - (float)1, yf - (float)1);
  return mixv(mixv(g00, g10, u), mixv(g01, g11, u), v);
}
P3_INLINE float perlin_fbm(float px, float py, float octaves, float scale) {
  float total = 0, amp = 1, freq = scale, norm = 0;
  const uint32_t n = f2u(clampv(octaves, (float)1, (float)8));
  for (uint32_t i = 0; i < n; ++i) {
    total = total + perlin2d(px * freq, py * freq) * amp;
    norm = norm + amp;
    amp = amp * (float)0.5;
    freq = freq * (float)2;
  }
  return total / norm;
}

// ---- direct-colour entries, position side (chaos.wgsl:1752) -------------------

P3_INLINE float2 var_dc_cylinder_pos(float2 p, float w) { return make_float2(w * safe_sin(p.x), w * p.y); }

// ---- post-flam3 plugin pack (chaos.wgsl:1784-1931) ---------------------------

// chaos.wgsl:1784 — three draws, always taken.
P3_INLINE float2 var_cpow2(float2 p, float w, float pr, float pa, float divisor, float range, Rng& rng) {
  float div = sel(divisor, (float)1, divisor == (float)0);
  float rmax = smax(range, (float)1);
  float step = kTau / div;
  float c = pr * safe_cos(kPi * (float)0.5 * pa) / div;
  float d = pr * safe_sin(kPi * (float)0.5 * pa) / div;
  float hc = c * (float)0.5;
  float inv_range = (float)0.5 / rmax;
  float full = kTau * rmax;
  float u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01();
  float a = atan2f(p.y, p.x);
  float n = floorf(u0 * rmax) + sel((float)0, (float)1, a < (float)0);
  a = a + kTau * n;
  if (safe_cos(a * inv_range) < (u1 * (float)2 - (float)1)) a = a - full;
  float lnr2 = logf(dot2(p));
  float r = w * expf(hc * lnr2 - d * a);
  float th = c * a + (d * (float)0.5) * lnr2 + step * floorf(div * u2);
  return make_float2(r * safe_cos(th), r * safe_sin(th));
}

// chaos.wgsl:1814 — four draws, always taken.
P3_INLINE float2 var_cpow3(float2 p, float w, float pr, float pd, float divisor, float spread, Rng& rng) {
  float div = sel(divisor, (float)1, divisor == (float)0);
  float step = kTau / div;
  float dabs = smax(fabsf(pd), (float)1e-30);
  float slog = sel(logf(dabs), -logf(dabs), pd < (float)0);
  float ang = atan2f(slog * pr, kTau);
  float ca = safe_cos(ang), sa = safe_sin(ang);
  float tc = ca * pr * ca / div;
  float td = ca * pr * sa / div;
  float hc = tc * (float)0.5, hd = td * (float)0.5;
  float coeff = sel((float)-0.095 * spread / td, (float)0, td == (float)0);
  float u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01(), u3 = rng.rand01();
  float a = atan2f(p.y, p.x);
  if (a < (float)0) a = a + kTau;
  if (safe_cos(a * (float)0.5) < (u0 * (float)2 - (float)1)) a = a - kTau;
  float bsign = sel(-kTau, kTau, u1 < (float)0.5);
  a = a + bsign * rint_even(logf(smax(u2, (float)1e-30)) * coeff);
  float lnr2 = logf(dot2(p));
  float r = w * expf(hc * lnr2 - td * a);
  float th = tc * a + hd * lnr2 + step * floorf(div * u3);
  return make_float2(r * safe_cos(th), r * safe_sin(th));
}

// chaos.wgsl:1856 — n-gon "loonie" with star and circle blends.
P3_INLINE float2 var_loonie2(float2 p, float w, float sides_f, float star, float circle) {
  const int32_t sides = imin(imax(f2i(sides_f), 1), 16);
  float a = kTau / (float)sides;
  float sa = safe_sin(a), ca = safe_cos(a);
  float sins = safe_sin(-star * kPi * (float)0.5), coss = safe_cos(star * kPi * (float)0.5);
  float sinc = safe_sin(circle * kPi * (float)0.5), cosc = safe_cos(circle * kPi * (float)0.5);
  float ww = w * w;
  float xr = p.x, yr = p.y;
  float r2 = xr * coss + fabsf(yr) * sins;
  float cr = sqrtf(xr * xr + yr * yr);
  for (int32_t i = 0; i < sides - 1; ++i) {
    float nx = xr * ca - yr * sa;
    yr = xr * sa + yr * ca;
    xr = nx;
    r2 = smax(r2, xr * coss + fabsf(yr) * sins);
  }
  r2 = r2 * cosc + cr * sinc;
  if (sides > 2) r2 = r2 * r2;
  else r2 = fabsf(r2) * r2;
  if (r2 > (float)0 && r2 < ww) {
    float r = w * sqrtf(fabsf(ww / r2 - (float)1));
    return make_float2(r * p.x, r * p.y);
  }
  if (r2 < (float)0) {
    float r = w / sqrtf(fabsf(ww / r2) - (float)1);
    return make_float2(r * p.x, r * p.y);
  }
  return make_float2(w * p.x, w * p.y);
}

// chaos.wgsl:1908 — one draw, always taken.
P3_INLINE float2 var_epispiral(float2 p, float w, float n, float thickness, float holes, Rng& rng) {
  float th = atan2f(p.y, p.x);
  float d = safe_cos(n * th);
  float u = rng.rand01();
  float inv = (float)1 / d;
  float t = sel(-holes + inv, -holes + (u * thickness) * inv, fabsf(thickness) > (float)1e-30);
  return make_float2(w * t * safe_cos(th), w * t * safe_sin(th));
}

// ---- cellular family (chaos.wgsl:1933-2147) ----------------------------------

// chaos.wgsl:1933 — integer mix of a cell coordinate to two uniforms.
P3_INLINE float2 worley_hash2(int32_t cx, int32_t cy) {
  uint32_t s = (uint32_t)cx * 2654435769u;
  s = s ^ ((uint32_t)cy * 2246822519u);
  s = s ^ (s >> 16);
  s = s * 0x85ebca6bu;
  s = s ^ (s >> 13);
  s = s * 0xc2b2ae35u;
  s = s ^ (s >> 16);
  return make_float2((float)(float)(s & 0xffffu) * (float)(float)(1.0 / 65535.0),
          (float)(float)((s >> 16) & 0xffffu) * (float)(float)(1.0 / 65535.0));
}

// chaos.wgsl:1988 — bubble-wrap lattice.
P3_INLINE float2 var_bwraps(float2 p, float w, float cell, float space, float gain, float itwist, float otwist) {
  if (fabsf(cell) < (float)1e-30) return make_float2(w * p.x, w * p.y);
  float radius = (float)0.5 * (cell / ((float)1 + space * space));
  float g2 = gain * gain / cell + (float)1e-6;
  float mb = g2 * radius;
  if (mb > (float)2) mb = (float)1;
  else mb = mb * ((float)1 / ((mb * mb) / (float)4 + (float)1));
  float rr2 = radius * radius;
  float rfac = radius / smax(mb, (float)1e-30);
  float cx = (floorf(p.x / cell) + (float)0.5) * cell;
  float cy = (floorf(p.y / cell) + (float)0.5) * cell;
  float lx = p.x - cx, ly = p.y - cy;
  if ((lx * lx + ly * ly) > rr2) return make_float2(w * p.x, w * p.y);
  lx = lx * g2;
  ly = ly * g2;
  float rd = rfac / ((lx * lx + ly * ly) / (float)4 + (float)1);
  lx = lx * rd;
  ly = ly * rd;
  float rf = (lx * lx + ly * ly) / smax(rr2, (float)1e-30);
  float th = itwist * ((float)1 - rf) + otwist * rf;
  float st = safe_sin(th), ct = safe_cos(th);
  return make_float2(w * (cx + ct * lx + st * ly), w * (cy - st * lx + ct * ly));
}

// chaos.wgsl:2034 — noise-perturbed cell centre.
P3_INLINE float2 crackle_centre(int32_t cx, int32_t cy, float s, float d) {
  float fx = (float)cx, fy = (float)cy;
  float ex = perlin2d(fx * (float)2.5, fy * (float)2.5);
  float ey = perlin2d(fy * (float)2.5 + (float)30.2, fx * (float)2.5 - (float)12.1);
  return make_float2((fx + d * ex) * s, (fy + d * ey) * s);
}
// chaos.wgsl:2046 — Voronoi boundary ratio of u between cells p and q.
P3_INLINE float crackle_vratio(float2 p, float2 q, float2 u) {
  float dx = p.x - q.x, dy = p.y - q.y;
  float den = dx * dx + dy * dy;
  if (den < (float)1e-30) return (float)1;
  float num = (u.x - q.x) * dx + (u.y - q.y) * dy;
  return (float)2 * num / den;
}
// chaos.wgsl:2073 — four draws (input blur), then a 3x3 Voronoi search.
P3_INLINE float2 var_crackle(float2 p, float w, float cell, float power, float distort, float scale, Rng& rng) {
  if (fabsf(cell) < (float)1e-30) return make_float2(w * p.x, w * p.y);
  float s = cell * (float)0.5;
  float u1 = rng.rand01(), u2 = rng.rand01(), u3 = rng.rand01(), u4 = rng.rand01();
  float blur = (u1 + u2) * (float)0.5 + (u3 - (float)0.5) * (float)0.25;
  float th = (float)2 * kPi * u4;
  float2 u = {blur * sinf(th), blur * cosf(th)};
  const int32_t x0 = f2i(floorf(u.x / s)), y0 = f2i(floorf(u.y / s));
  float2 P[9];
  for (int i = 0; i < 9; ++i)
    P[i] = crackle_centre(iwrap_add(x0, i / 3 - 1), iwrap_add(y0, i % 3 - 1), s, distort);
  int q = 0;
  float best = (float)1e30;
  for (int i = 0; i < 9; ++i) {
    float dx = P[i].x - u.x, dy = P[i].y - u.y;
    float d2 = dx * dx + dy * dy;
    if (d2 < best) { best = d2; q = i; }
  }
  const int32_t xc = iwrap_add(x0, q / 3 - 1), yc = iwrap_add(y0, q % 3 - 1);
  for (int i = 0; i < 9; ++i)
    P[i] = crackle_centre(iwrap_add(xc, i / 3 - 1), iwrap_add(yc, i % 3 - 1), s, distort);
  float L = (float)-1e30;
  for (int i = 0; i < 9; ++i) {
    if (i == 4) continue;
    float ratio = crackle_vratio(P[i], P[4], u);
    if (ratio > L) L = ratio;
  }
  float dxo = u.x - P[4].x, dyo = u.y - P[4].y;
  float ls = L + (float)1e-30;
  float trg = powf(fabsf(ls), power) * scale;
  float Rr = trg / ls;
  return make_float2(w * (P[4].x + dxo * Rr), w * (P[4].y + dyo * Rr));
}

// ---- JWildfire S-tier, first half (chaos.wgsl:2149-2360) ---------------------

// chaos.wgsl:2149 — one draw.
P3_INLINE float2 var_juliaq(float2 p, float w, float power, float divisor, Rng& rng) {
  float ps = sel(power, (float)1, fabsf(power) < (float)1e-30);
  float ap = smax((float)1, fabsf(ps));
  float ip = divisor / ps;
  float hip = (float)0.5 * ip;
  float ip2 = kTau / ps;
  float u = rng.rand01();
  float n = floorf(u * ap);
  float a = atan2f(p.y, p.x) * ip + n * ip2;
  float r = w * powf(p.x * p.x + p.y * p.y, hip);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}

// chaos.wgsl:2169 — one draw picks one of two formulas per half-plane.
P3_INLINE float2 var_glynnia(float2 p, float w, Rng& rng) {
  float v2 = w * (float)0.7071067811865476;
  float r = sqrtf(p.x * p.x + p.y * p.y);
  float coin = rng.rand01();
  float sx = r >= (float)1 ? (float)1 : (float)-1;  // outside the unit disc keeps +x
  if (coin > (float)0.5) {
    float in = r + p.x;
    if (in <= (float)0) return make_float2((float)0, (float)0);
    float d = sqrtf(in);
    if (d == (float)0) return make_float2((float)0, (float)0);
    return make_float2(sx * v2 * d, -(v2 / d) * p.y);
  }
  float d = r + p.x;
  float rad = r * (p.y * p.y + d * d);
  if (rad <= (float)0) return make_float2((float)0, (float)0);
  float dx = sqrtf(rad);
  if (dx == (float)0) return make_float2((float)0, (float)0);
  float rr = w / dx;
  return make_float2(sx * rr * d, rr * p.y);
}

// chaos.wgsl:2213
P3_INLINE float2 var_loonie3(float2 p, float w) {
  float ww = w * w;
  float r2 = (float)2 * ww;
  if (p.x > (float)1e-30) {
    float q = (p.x * p.x + p.y * p.y) / p.x;
    r2 = q * q;
  }
  if (r2 < ww) {
    float r = w * sqrtf(ww / r2 - (float)1);
    return make_float2(r * p.x, r * p.y);
  }
  return make_float2(w * p.x, w * p.y);
}

// chaos.wgsl:2236 — two draws.
P3_INLINE float2 var_falloff(float2 p, float w, float scatter, float mindist, float mx, float my, float x0, float y0, Rng& rng) {
  float rmax = (float)0.04 * scatter;
  float dx = p.x - x0, dy = p.y - y0;
  float d = sqrtf(dx * dx + dy * dy);
  if (d < (float)0) d = (float)0;
  d = (d - mindist) * rmax;
  if (d < (float)0) d = (float)0;
  float u0 = rng.rand01(), u1 = rng.rand01();
  return make_float2(w * (p.x + mx * u0 * d), w * (p.y + my * u1 * d));
}

// chaos.wgsl:2263 — three draws; type 0 (linear), 1 (radial), 2 (gaussian).
P3_INLINE float2 var_falloff2(float2 p, float w, float scatter, float typ, float mx, float my, float x0, float y0, float mindist, Rng& rng) {
  float rmax = (float)0.04 * scatter;
  float dx = p.x - x0, dy = p.y - y0;
  float d = sqrtf(dx * dx + dy * dy);
  if (d < (float)0) d = (float)0;
  d = (d - mindist) * rmax;
  if (d < (float)0) d = (float)0;
  float u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01();
  const int32_t t = f2i(clampv(typ, (float)0, (float)2));
  if (t == 1) {
    float ra = sqrtf(p.x * p.x + p.y * p.y);
    float phi = atan2f(p.y, p.x) + my * d * u1;
    float rr = ra + mx * u0 * d;
    return make_float2(w * rr * safe_cos(phi), w * rr * safe_sin(phi));
  }
  if (t == 2) {
    float sg = d * u1 * kTau;
    float phi = d * u2 * kPi;
    float rad = d * u0;
    float sc = safe_cos(sg);
    return make_float2(w * (p.x + mx * rad * sc * safe_cos(phi)), w * (p.y + my * rad * sc * safe_sin(phi)));
  }
  return make_float2(w * (p.x + mx * u0 * d), w * (p.y + my * u1 * d));
}

// chaos.wgsl:2318 — three draws, gaussian type with optional inversion.
P3_INLINE float2 var_falloff3(float2 p, float w, float scatter, float mx, float my, float x0, float y0, float mindist, float inv, Rng& rng) {
  float rmax = (float)0.04 * scatter;
  float dx = p.x - x0, dy = p.y - y0;
  float radius = sqrtf(dx * dx + dy * dy);
  float base = sel(smax(radius, (float)0), smax((float)1 - radius, (float)0), inv > (float)0.5);
  float dist = smax((base - mindist) * rmax, (float)0);
  float u0 = rng.rand01() - (float)0.5;
  float u1 = rng.rand01() - (float)0.5;
  float u2 = rng.rand01() - (float)0.5;
  float sg = dist * u1 * kTau;
  float phi = dist * u2 * kPi;
  float rad = dist * u0;
  float sc = safe_cos(sg);
  return make_float2(w * (p.x + mx * rad * sc * safe_cos(phi)), w * (p.y + my * rad * sc * safe_sin(phi)));
}

// ---- kaleidoscope / circle family (chaos.wgsl:2362-2527) ---------------------

// chaos.wgsl:2362
P3_INLINE float2 var_collideoscope(float2 p, float w, float ap, float nump) {
  const int32_t ni = imax(1, f2i(nump));
  float nf = (float)ni;
  float knpi = nf / kPi, pikn = kPi / nf;
  float kakn = kPi * ap / nf;
  float a = atan2f(p.y, p.x);
  float r = w * sqrtf(p.x * p.x + p.y * p.y);
  if (a >= (float)0) {
    const int32_t alt = f2i(a * knpi);
    a = (float)alt * pikn + wmod((imod(alt, 2) == 0 ? kakn : -kakn) + a, pikn);
  } else {
    const int32_t alt = f2i(-a * knpi);
    a = -((float)alt * pikn + wmod((imod(alt, 2) != 0 ? -kakn : kakn) - a, pikn));
  }
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}

// Square-to-circle perimeter walk shared by circlize / circlize2
// (chaos.wgsl:2400-2464): returns false when the point is the origin.
P3_INLINE bool square_perimeter(float2 p, float& perim, float& side) {
  float ax = fabsf(p.x), ay = fabsf(p.y);
  if (ax >= ay) {
    perim = p.x >= ay ? ax + p.y : (float)5 * ax - p.y;
    side = ax;
  } else {
    perim = p.y >= ax ? (float)3 * ay - p.x : (float)7 * ay + p.x;
    side = ay;
  }
  return side != (float)0;
}
// chaos.wgsl:2400
P3_INLINE float2 var_circlize(float2 p, float w, float hole) {
  float v4 = w / (kPi * (float)0.25);
  float perim, side;
  if (!square_perimeter(p, perim, side)) return make_float2((float)0, (float)0);
  float r = v4 * side + hole;
  float a = (kPi * (float)0.25) * perim / side - (kPi * (float)0.25);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}
// chaos.wgsl:2434
P3_INLINE float2 var_circlize2(float2 p, float w, float hole) {
  float perim, side;
  if (!square_perimeter(p, perim, side)) return make_float2((float)0, (float)0);
  float r = w * (side + hole);
  float a = (kPi * (float)0.25) * perim / side - (kPi * (float)0.25);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}

// chaos.wgsl:2466 — elliptic-coordinate swirl.
P3_INLINE float2 var_eswirl(float2 p, float w, float inp, float outp) {
  float tmp = p.y * p.y + p.x * p.x + (float)1;
  float tmp2 = (float)2 * p.x;
  float a1 = tmp + tmp2, a2 = tmp - tmp2;
  float s1 = sel((float)0, sqrtf(smax(a1, (float)0)), a1 > (float)0);
  float s2 = sel((float)0, sqrtf(smax(a2, (float)0)), a2 > (float)0);
  float xmax = (s1 + s2) * (float)0.5;
  if (xmax < (float)1) xmax = (float)1;
  float mu = logf(xmax + sqrtf(smax(xmax * xmax - (float)1, (float)0)));
  float t = p.x / xmax;
  if (t > (float)1) t = (float)1;
  if (t < (float)-1) t = (float)-1;
  float nu = acosf(t);
  if (p.y < (float)0) nu = -nu;
  float mus = sel(mu, (float)1e-30, mu == (float)0);
  float nw = nu + mu * outp + inp / mus;
  return make_float2(w * coshf(mu) * safe_cos(nw), w * sinhf(mu) * safe_sin(nw));
}

// chaos.wgsl:2503
P3_INLINE float2 var_petal(float2 p, float w) {
  float cx = safe_cos(p.x), sx = safe_sin(p.x), cy = safe_cos(p.y);
  float a = cx * cy, b = sx * cy;
  float bx = a * a * a, by = b * b * b;
  return make_float2(w * cx * bx, w * cx * by);
}

// ---- Xyrus02 mid-tier, hexes (chaos.wgsl:2529-2760) --------------------------

// chaos.wgsl:2529 — the draw is taken only outside the unit circle.
P3_INLINE float2 var_bcircle(float2 p, float w, float scale, float border, Rng& rng) {
  float bw = fabsf(border);
  if (p.x == (float)0 && p.y == (float)0) return make_float2((float)0, (float)0);
  float x = p.x * scale, y = p.y * scale;
  float r = sqrtf(x * x + y * y);
  if (r <= (float)1) return make_float2(w * x, w * y);
  if (bw == (float)0) return make_float2((float)0, (float)0);
  float ang = atan2f(y, x);
  float u = rng.rand01();
  float om = (float)0.2 * bw * u + (float)1;
  return make_float2(w * om * safe_cos(ang), w * om * safe_sin(ang));
}

// chaos.wgsl:2551 — z / (1 + c1 z + c2 z^2 + c3 z^3), expanded in reals.
P3_INLINE float2 var_curl2(float2 p, float w, float c1, float c2, float c3) {
  float c2d = (float)2 * c2, c3t = (float)3 * c3;
  float x = p.x, y = p.y;
  float x2 = x * x, x3 = x2 * x, y2 = y * y, y3 = y2 * y;
  float re = c3 * x3 - c3t * x * y2 + c2 * x2 - c2 * y2 + c1 * x + (float)1;
  float im = c3t * x2 * y - c3 * y3 + c2d * x * y + c1 * y;
  float r = w / (re * re + im * im);
  return make_float2((x * re + y * im) * r, (y * re - x * im) * r);
}

// chaos.wgsl:2579
P3_INLINE float2 var_murl(float2 p, float w, float cin, float pin) {
  const int32_t pi = f2i(pin);
  float pf = (float)pi;
  float c = sel(cin / (pf - (float)1), cin, pi == 1);
  float p2 = pf / (float)2;
  float vp = w * (c + (float)1);
  float a = atan2f(p.y, p.x) * pf;
  float sa = safe_sin(a), ca = safe_cos(a);
  float r = c * powf(p.x * p.x + p.y * p.y, p2);
  float re = r * ca + (float)1, im = r * sa;
  float r1 = vp / (re * re + im * im + (float)1e-29);
  return make_float2(r1 * (p.x * re + p.y * im), r1 * (p.y * re - p.x * im));
}

// chaos.wgsl:2605
P3_INLINE float2 var_stwins(float2 p, float w, float distort) {
  float x = p.x * w * (float)0.05, y = p.y * w * (float)0.05;
  float x2 = x * x, y2 = y * y;
  float num = (x2 - y2) * safe_sin(kTau * distort * (x + y));
  float sum = x2 + y2;
  float res = num / sel(sum, (float)1, sum == (float)0);
  return make_float2(w * p.x + res, w * p.y + res);
}

// chaos.wgsl:2736 — running max of Voronoi ratios (P relative to centre Q).
P3_INLINE float hex_ratio_step(float2 P, float2 Q, float ux, float uy, float prev) {
  float dx = P.x - Q.x, dy = P.y - Q.y;
  if (dx == (float)0 && dy == (float)0) return sel(prev, (float)1, (float)1 > prev);
  float num = (float)2 * ((ux - Q.x) * dx + (uy - Q.y) * dy);
  float ratio = num / (dx * dx + dy * dy);
  return sel(prev, ratio, ratio > prev);
}
// chaos.wgsl:2725 — ratios against the six neighbours H[1..6] of H[0].
P3_INLINE float hex_ratio_max(const float2* H, float ux, float uy) {
  float m = (float)-1e20;
  for (int i = 1; i < 7; ++i) m = hex_ratio_step(H[i], H[0], ux, uy, m);
  return m;
}
// Hex-lattice centre of cell (i, j) (axial coordinates, cell size s).
P3_INLINE float2 hex_centre(int32_t i, int32_t j, float s) {
  const float sq3 = (float)1.7320508075688772935;
  float fi = (float)i, fj = (float)j;
  return make_float2(((float)1.5 * fi + (float)-1.5 * fj) * s, (sq3 / (float)2 * fi + sq3 / (float)2 * fj) * s);
}
// chaos.wgsl:2637 — hexagonal Voronoi cells with power falloff and rotation.
P3_INLINE float2 var_hexes(float2 p, float w, float cell, float power, float rotate, float scale) {
  if (cell == (float)0) return make_float2((float)0, (float)0);
  const float sq3 = (float)1.7320508075688772935;
  const float ah = (float)(1.0 / 3.0), bh = sq3 / (float)3, ch = (float)(-1.0 / 3.0), dh = sq3 / (float)3;
  float rs = safe_sin(rotate * kTau), rc = safe_cos(rotate * kTau);
  float ux = p.x, uy = p.y, s = cell;
  const int32_t h0 = f2i(floorf((ah * ux + bh * uy) / s));
  const int32_t k0 = f2i(floorf((ch * ux + dh * uy) / s));
  float best = (float)1e30;
  int q = 0;
  for (int di = -1; di < 2; ++di)
    for (int dj = -1; dj < 2; ++dj) {
      float2 c = hex_centre(iwrap_add(h0, di), iwrap_add(k0, dj), s);
      float dx = c.x - ux, dy = c.y - uy;
      float d2 = dx * dx + dy * dy;
      if (d2 < best) { best = d2; q = (di + 1) * 3 + (dj + 1); }
    }
  const int32_t hx = iwrap_add(h0, q / 3 - 1), hy = iwrap_add(k0, q % 3 - 1);
  // centre, then the six neighbours in the reference's order
  const int32_t oi[7] = {0, 0, 1, 1, 0, -1, -1}, oj[7] = {0, 1, 1, 0, -1, -1, 0};
  float2 H[7];
  for (int i = 0; i < 7; ++i) H[i] = hex_centre(iwrap_add(hx, oi[i]), iwrap_add(hy, oj[i]), s);
  float L1 = hex_ratio_max(H, ux, uy);
  float dxo = ux - H[0].x, dyo = uy - H[0].y;
  float trg = powf(L1 + (float)1e-30, power) * scale;
  float vx = dxo * rc + dyo * rs;
  float vy = -dxo * rs + dyo * rc;
  float L2 = hex_ratio_max(H, vx + H[0].x, vy + H[0].y);
  float L = smax(L1, L2);
  float Rr;
  if (L < (float)0.5) Rr = trg / L1;
  else if (L > (float)0.8) Rr = trg / L2;
  else Rr = ((trg / L1) * ((float)0.8 - L) + (trg / L2) * (L - (float)0.5)) / (float)0.3;
  return make_float2(w * (vx * Rr + H[0].x), w * (vy * Rr + H[0].y));
}

// ---- Xyrus02 X-family, blur_circle (chaos.wgsl:2762-3100) --------------------

// chaos.wgsl:2762
P3_INLINE float2 var_xheart(float2 p, float w, float angle, float ratio) {
  float ang = kPi / (float)4 + ((float)0.5 * (kPi / (float)4) * angle);
  float ca = safe_cos(ang), sa = safe_sin(ang);
  float rat = (float)6 + (float)2 * ratio;
  float r24 = p.x * p.x + p.y * p.y + (float)4;
  if (r24 == (float)0) r24 = (float)1;
  float bx = (float)4 / r24, by = rat / r24;
  float xr = ca * (bx * p.x) - sa * (by * p.y);
  float yr = sa * (bx * p.x) + ca * (by * p.y);
  return make_float2(w * xr, w * sel(-yr, yr, xr > (float)0));
}

// chaos.wgsl:2786
P3_INLINE float2 var_xhyperbol(float2 p, float w, float m00, float m01, float m10, float m11, float m20, float m21) {
  float r = (float)1 / (p.x * p.x + p.y * p.y + kEps);
  float x = p.x * r, y = p.y * r;
  float re = m00 * x + m01 * y + m20;
  float im = m10 * x + m11 * y + m21;
  float al = atan2f(im, re) + kTau;
  float sa = safe_sin(al), ca = safe_cos(al);
  float rsq = re * re + im * im;
  float xo = rsq * ca, yo = rsq * sa;
  float ri = w / (xo * xo + yo * yo + kEps);
  return make_float2(xo * ri, yo * ri);
}

// chaos.wgsl:2815
P3_INLINE float2 var_xcurl2(float2 p, float w, float c1, float c2, float c3) {
  float x = p.x, y = p.y;
  float x2 = x * x, y2 = y * y, x3 = x2 * x;
  float re = (float)1 + c1 * x + c2 * (x2 - y2) + c3 * (x3 - (float)3 * x);
  float im = c1 * y + c2 * ((float)2 * x * y) + c3 * ((float)3 * x * y - (float)1);
  float r = w / (re * re + im * im);
  return make_float2((x * re + y * im) * r, (y * re + x * im) * r);
}

// chaos.wgsl:2838 — triangular (trilinear) lattice tiling; two draws, the
// second always taken.
P3_INLINE float2 var_xtrb(float2 p, float w, float pin, float dist, float radius, float width, float ap, float bp, Rng& rng) {
  const int32_t pi = imax(1, f2i(pin));
  float pf = (float)pi;
  float angB = (float)0.047 + ap, angC = (float)0.047 + bp;
  float angA = kPi - angB - angC;
  float sA2 = safe_sin((float)0.5 * angA), cA2 = safe_cos((float)0.5 * angA);
  float sB2 = safe_sin((float)0.5 * angB), cB2 = safe_cos((float)0.5 * angB);
  float sC2 = safe_sin((float)0.5 * angC), cC2 = safe_cos((float)0.5 * angC);
  float sC = safe_sin(angC), cC = safe_cos(angC);
  float as = radius * (sC2 / cC2 + sB2 / cB2);
  float bs = radius * (sC2 / cC2 + sA2 / cA2);
  float cs = radius * (sB2 / cB2 + sA2 / cA2);
  float w1 = (float)1 - width, w2 = (float)2 * width, w3 = (float)1 - width * width;
  float S2 = radius * (as + bs + cs);
  float Ha = S2 / as / (float)6, Hb = S2 / bs / (float)6, Hc = S2 / cs / (float)6;
  float ab = as / bs, ac = as / cs, ba = bs / as, bc = bs / cs, ca = cs / as, cb = cs / bs;
  float S2a = (float)6 * Ha, S2b = (float)6 * Hb, S2c = (float)6 * Hc;
  float S2bc = S2 / (bs + cs) / (float)6, S2ab = S2 / (as + bs) / (float)6, S2ac = S2 / (as + cs) / (float)6;
  const uint32_t absN = (uint32_t)(pi < 0 ? -pi : pi);
  float cN = dist / pf / (float)2;
  float U = p.y + radius;
  float V = p.x * sC - p.y * cC + radius;
  float M = floorf(U / S2a);
  float oal = U - M * S2a;
  float N = floorf(V / S2b);
  float obe = V - N * S2b;
  float oga = S2c - ac * oal - bc * obe;
  float Rn = rng.rand01();
  const bool neg = oga <= (float)0;
  if (neg) { oal = S2a - oal; obe = S2b - obe; oga = -oga; }
  float Al = oal, Be = obe, Ga = oga;
  float Al1 = 0, Be1 = 0, Ga1 = 0, De1 = 0;
  const bool edge = Rn >= w3;
  if (Be < Al) {
    if (Ga < Be) {
      if (edge) { De1 = width * Be; Ga1 = width * Ga; }
      else { Ga1 = w1 * Ga + w2 * Hc * Ga / Be; De1 = w1 * Be + w2 * S2ab * ((float)3 - Ga / Be); }
      Al1 = S2a - ba * De1 - ca * Ga1;
      Be1 = De1;
    } else if (Ga < Al) {
      if (edge) { Ga1 = width * Ga; De1 = width * Be; }
      else { De1 = w1 * Be + w2 * Hb * Be / Ga; Ga1 = w1 * Ga + w2 * S2ac * ((float)3 - Be / Ga); }
      Al1 = S2a - ba * De1 - ca * Ga1;
      Be1 = De1;
    } else {
      if (edge) { Al1 = width * Al; Be1 = width * Be; }
      else { Be1 = w1 * Be + w2 * Hb * Be / Al; Al1 = w1 * Al + w2 * S2ac * ((float)3 - Be / Al); }
    }
  } else {
    if (Ga < Al) {
      if (edge) { De1 = width * Al; Ga1 = width * Ga; }
      else { Ga1 = w1 * Ga + w2 * Hc * Ga / Al; De1 = w1 * Al + w2 * S2ab * ((float)3 - Ga / Al); }
      Be1 = S2b - ab * De1 - cb * Ga1;
      Al1 = De1;
    } else if (Ga < Be) {
      if (edge) { Ga1 = width * Ga; De1 = width * Al; }
      else { De1 = w1 * Al + w2 * Ha * Al / Ga; Ga1 = w1 * Ga + w2 * S2bc * ((float)3 - Al / Ga); }
      Be1 = S2b - ab * De1 - cb * Ga1;
      Al1 = De1;
    } else {
      if (edge) { Be1 = width * Be; Al1 = width * Al; }
      else { Al1 = w1 * Al + w2 * Ha * Al / Be; Be1 = w1 * Be + w2 * S2bc * ((float)3 - Al / Be); }
    }
  }
  (void)Ga1;
  float alpha = Al1, beta = Be1;
  if (neg) { alpha = S2a - alpha; beta = S2b - beta; }
  alpha = alpha + M * S2a;
  beta = beta + N * S2b;
  float inx = (beta - radius + (alpha - radius) * cC) / sC;
  float iny = alpha - radius;
  float u2 = rng.rand01();
  const uint32_t br = absN > 0u ? f2u(u2 * (float)absN) : 0u;
  float ang = (atan2f(iny, inx) + kTau * (float)br) / pf;
  float rp = powf(inx * inx + iny * iny, cN);
  return make_float2(w * w * rp * safe_cos(ang), w * w * rp * safe_sin(ang));
}

// chaos.wgsl:3019 — step to a neighbouring cell along the square spiral.
P3_INLINE float2 var_xyrus_gridout(float2 p, float w) {
  float x = p.x, y = p.y;
  float rx = sel(ceilf(x - (float)0.5), floorf(x + (float)0.5), x >= (float)0);
  float ry = sel(ceilf(y - (float)0.5), floorf(y + (float)0.5), y >= (float)0);
  float dx = 0, dy = 0;
  if (ry <= (float)0) {
    if (rx > (float)0) { if (-ry >= rx) dx = (float)1; else dy = (float)1; }
    else { if (ry <= rx) dx = (float)1; else dy = (float)-1; }
  } else {
    if (rx > (float)0) { if (ry >= rx) dx = (float)-1; else dy = (float)1; }
    else { if (ry > -rx) dx = (float)-1; else dy = (float)-1; }
  }
  return make_float2(w * (x + dx), w * (y + dy));
}

// chaos.wgsl:3053 — two draws; a uniform square mapped onto a disc.
P3_INLINE float2 var_blur_circle(float2 p, float w, float hole, Rng& rng) {
  (void)p;
  float u0 = rng.rand01(), u1 = rng.rand01();
  float x = (float)2 * u0 - (float)1, y = (float)2 * u1 - (float)1;
  float ax = fabsf(x), ay = fabsf(y);
  float s = sel(ay, ax, ax > ay);
  float a = atan2f(y, x);
  float q3 = (float)3 * kPi / (float)4, q1 = kPi / (float)4;
  float ps;
  if (a < -q3) ps = ay;
  else if (a < -q1) ps = (float)2 * s + x;
  else if (a < q1) ps = (float)4 * s + y;
  else if (a < q3) ps = (float)6 * s - x;
  else ps = (float)8 * s - y;
  float r = (w * (float)4 / kPi) * s + hole;
  float phi = q1 * ps / s - kPi;
  return make_float2(r * safe_cos(phi), r * safe_sin(phi));
}

// ---- bipolar2 .. waffle (chaos.wgsl:3102-3451) ----------------------------------

// chaos.wgsl:3102
P3_INLINE float2 var_bipolar2(float2 p, float w, float shift, float a, float b, float c, float d, float e, float f1, float g1, float h) {
  float hp = kPi * (float)0.5, tp = (float)2 / kPi;
  float x2y2 = (p.x * p.x + p.y * p.y) * g1;
  float t = x2y2 + a;
  float x2 = b * p.x;
  float ps = -hp * shift;
  float y = c * atan2f(e * p.y, x2y2 - d) + ps;
  if (y > hp) y = -hp + wmod(y + hp, kPi);
  else if (y < -hp) y = hp - wmod(hp - y, kPi);
  float fn = t + x2, gn = t - x2;
  if (gn == (float)0 || (fn / gn) <= (float)0) return make_float2((float)0, (float)0);
  return make_float2(w * f1 * tp * logf(fn / gn), w * tp * y * h);
}

// chaos.wgsl:3136
P3_INLINE float2 var_bubble2(float2 p, float w, float xs, float ys) {
  float r = w / ((float)0.25 * dot2(p) + (float)1);
  return make_float2(p.x * r * xs, p.y * r * ys);
}

// Inverse hyperbolic family (chaos.wgsl:3155-3237); outputs scaled by 2w/pi.
// chaos.wgsl:3155 — one draw picks the sign.
P3_INLINE float2 var_acosh(float2 p, float w, Rng& rng) {
  float2 s = c_sqr(p);
  float2 d = c_sqrt(make_float2(s.x - (float)1, s.y - (float)0));
  float2 z = c_log(make_float2(p.x + d.x, p.y + d.y));
  float k = w * (float)2 / kPi;
  float sg = sel((float)1, (float)-1, rng.rand01() >= (float)0.5);
  return make_float2(sg * (z.x * k), sg * (z.y * k));
}
// chaos.wgsl:3163
P3_INLINE float2 var_arcsinh(float2 p, float w) {
  float2 s = c_sqr(p);
  float2 d = c_sqrt(make_float2(s.x + (float)1, s.y + (float)0));
  float2 z = c_log(make_float2(p.x + d.x, p.y + d.y));
  float k = w * (float)2 / kPi;
  return make_float2(z.x * k, z.y * k);
}
// chaos.wgsl:3172
P3_INLINE float2 var_arctanh(float2 p, float w) {
  float2 z = c_log(c_div(make_float2(p.x + (float)1, p.y + (float)0), make_float2((float)1 - p.x, (float)0 - p.y)));
  float k = w * (float)2 / kPi;
  return make_float2(z.x * k, z.y * k);
}
// chaos.wgsl:3180
P3_INLINE float2 var_acoth(float2 p, float w) {
  float2 rz = c_recip(p);
  float2 l = c_log(c_div(make_float2(rz.x + (float)1, rz.y + (float)0), make_float2((float)1 - rz.x, (float)0 - rz.y)));
  float k = w * (float)2 / kPi;
  return make_float2(((float)0.5 * l.y) * k, ((float)0.5 * l.x) * k);
}
// chaos.wgsl:3192 — one draw picks the sign.
P3_INLINE float2 var_acosech(float2 p, float w, Rng& rng) {
  float2 rz = c_recip(p);
  float2 s = c_sqr(rz);
  float2 d = c_sqrt(make_float2(s.x - (float)1, s.y - (float)0));
  float2 l = c_log(make_float2(rz.x + d.x, rz.y + d.y));
  float k = w * (float)2 / kPi;
  float fx = l.y * k, fy = l.x * k;
  float sg = sel((float)1, (float)-1, rng.rand01() >= (float)0.5);
  return make_float2(sg * fx, sg * fy);
}
// chaos.wgsl:3206
P3_INLINE float2 var_arcsech2(float2 p, float w) {
  float2 z = c_recip(p);
  float2 zs = c_sqrt(make_float2(z.x - (float)1, z.y - (float)0));
  float2 za = c_sqrt(make_float2(z.x + (float)1, z.y + (float)0));
  float2 m = c_mul(za, zs);
  float2 l = c_log(make_float2(z.x + m.x, z.y + m.y));
  float k = w * (float)2 / kPi;
  float lx = l.x * k, ly = l.y * k;
  const bool neg = ly < (float)0;
  return make_float2(sel(-lx, lx, neg), sel(ly - (float)1, ly + (float)1, neg));
}

// chaos.wgsl:3239
P3_INLINE float2 var_cell2(float2 p, float w, float size, float a, float snx, float sny, float ssx, float ssy) {
  float ss = sel(size, (float)1e-30, fabsf(size) < (float)1e-30);
  float inv = a / ss;
  float cx = floorf(p.x * inv), cy = floorf(p.y * inv);
  float dx = p.x - cx * ss, dy = p.y - cy * ss;
  float sx = cx, sy = cy;
  if (sy >= (float)0) { sy = sy * sny; sx = sx * snx; }
  else { sy = -ssy * sy; sx = sx * ssx; }
  return make_float2(w * (dx + sx * ss), -w * (dy + sy * ss));
}

// chaos.wgsl:3279
P3_INLINE float2 var_curl_sp(float2 p, float w, float powp, float c1, float c2, float sx, float sy) {
  float pw = sel(powp, (float)1e-30, powp == (float)0);
  float pinv = (float)1 / pw;
  float c2x2 = (float)2 * c2;
  float x = powf(fabsf(p.x), pw) * signv(p.x);
  float y = powf(fabsf(p.y), pw) * signv(p.y);
  float d = x * x - y * y;
  float aa = c1 * x + c2 * d;
  float re = sqrtf(aa * aa + sx * sx) * sel((float)-1, (float)1, aa > (float)0) + (float)1;
  float bb = c1 * y + c2x2 * x * y;
  float im = sqrtf(bb * bb + sy * sy) * sel((float)-1, (float)1, bb > (float)0);
  float c = powf(fabsf(re * re + im * im), pinv);
  float r = w / smax(c, (float)1e-30);
  return make_float2((x * re + y * im) * r, (y * re - x * im) * r);
}

// chaos.wgsl:3314
P3_INLINE float2 var_murl2(float2 p, float w, float c, float pf) {
  float p2 = pf * (float)0.5;
  float sp = sel(pf, (float)1e-30, pf == (float)0);
  float invp = (float)1 / sp;
  float cp1 = c + (float)1;
  float vp = w * powf(fabsf(cp1), (float)2 * invp) * sel((float)-1, (float)1, cp1 >= (float)0);
  float a1 = atan2f(p.y, p.x) * sp;
  float r0 = c * powf(fabsf(dot2(p)), p2);
  float re0 = r0 * safe_cos(a1) + (float)1, im0 = r0 * safe_sin(a1);
  float r1 = powf(fabsf(re0 * re0 + im0 * im0), invp);
  float a2 = atan2f(im0, re0) * (float)2 * invp;
  float re1 = r1 * safe_cos(a2), im1 = r1 * safe_sin(a2);
  float rl = vp / smax(r1 * r1, (float)1e-30);
  return make_float2(rl * (p.x * re1 + p.y * im1), rl * (p.y * re1 - p.x * im1));
}

// chaos.wgsl:3348 — two draws; a point on a Lissajous curve.
P3_INLINE float2 var_lissajous(float2 p, float w, float tmin, float tmax, float a, float b, float c, float d, float e, Rng& rng) {
  (void)p;
  float t = (tmax - tmin) * rng.rand01() + tmin;
  float yj = rng.rand01() - (float)0.5;
  float drift = c * t + e * yj;
  return make_float2(w * (safe_sin(a * t + d) + drift), w * (safe_sin(b * t) + drift));
}

// chaos.wgsl:3369 — two draws; a point on a hypotrochoid.
P3_INLINE float2 var_spirograph(float2 p, float w, float a, float b, float d, float tmin, float tmax, float ymin, float ymax, float c1, float c2,
                     Rng& rng) {
  (void)p;
  float t = (tmax - tmin) * rng.rand01() + tmin;
  float yj = (ymax - ymin) * rng.rand01() + ymin;
  float ab = a + b;
  float sb = sel(b, (float)1e-30, fabsf(b) < (float)1e-30);
  float ratio = ab / sb;
  float x1 = ab * safe_cos(t) - c1 * safe_cos(ratio * t);
  float y1 = ab * safe_sin(t) - c2 * safe_sin(ratio * t);
  return make_float2(w * (x1 + d * safe_cos(t) + yj), w * (y1 + d * safe_sin(t) + yj));
}

// chaos.wgsl:3399 — one mode draw, then two or three more per mode
// (operands drawn left to right).
P3_INLINE float2 var_waffle(float2 p, float w, float slices_f, float xth, float yth, float rotation, Rng& rng) {
  (void)p;
  float slices = smax((float)1, floorf(fabsf(slices_f)));
  float inv = (float)1 / slices;
  float vc = w * safe_cos(rotation), vs = w * safe_sin(rotation);
  const uint32_t mode = f2u(smin((float)4, floorf(rng.rand01() * (float)5)));
  float a, r;
  if (mode == 0u) {
    float k0 = floorf(rng.rand01() * slices);
    float u0 = rng.rand01();
    a = (k0 + u0 * xth) * inv;
    float k1 = floorf(rng.rand01() * slices);
    float u1 = rng.rand01();
    r = (k1 + u1 * yth) * inv;
  } else if (mode == 1u) {
    float k0 = floorf(rng.rand01() * slices);
    float u0 = rng.rand01();
    a = (k0 + u0) * inv;
    r = (floorf(rng.rand01() * slices) + yth) * inv;
  } else if (mode == 2u) {
    a = (floorf(rng.rand01() * slices) + xth) * inv;
    float k1 = floorf(rng.rand01() * slices);
    float u1 = rng.rand01();
    r = (k1 + u1) * inv;
  } else if (mode == 3u) {
    a = rng.rand01();
    float k1 = floorf(rng.rand01() * slices);
    float u1 = rng.rand01();
    r = (k1 + yth + u1 * ((float)1 - yth)) * inv;
  } else {
    float k0 = floorf(rng.rand01() * slices);
    float u0 = rng.rand01();
    a = (k0 + xth + u0 * ((float)1 - xth)) * inv;
    r = rng.rand01();
  }
  return make_float2(vc * a + vs * r, -vs * a + vc * r);
}

// ---- glynnSim family .. corners (chaos.wgsl:3453-4111) ------------------------

// chaos.wgsl:3453 — a disc of radius `radius` at angle phi1 filled by two
// draws; outside it, points are inverted into it with probability
// contrast * alpha^|pow| (one draw).
P3_INLINE float2 var_glynnSim1(float2 p, float w, float radius, float radius1, float phi1, float thick, float powp, float contrast,
                    Rng& rng) {
  float a = phi1 * (kPi / (float)180);
  float x1 = radius * safe_cos(a), y1 = radius * safe_sin(a);
  float ap = fabsf(powp);
  float r = sqrtf(dot2(p));
  if (r < radius) {
    float ri = radius1 * (thick + ((float)1 - thick) * rng.rand01());
    float phi = kTau * rng.rand01();
    return make_float2(w * (ri * safe_cos(phi) + x1), w * (ri * safe_sin(phi) + y1));
  }
  float alpha = radius / smax(r, (float)1e-30);
  float x, y;
  if (rng.rand01() > contrast * powf(fabsf(alpha), ap)) { x = p.x; y = p.y; }
  else { float a2 = alpha * alpha; x = a2 * p.x; y = a2 * p.y; }
  float dx = x - x1, dy = y - y1;
  if (dx * dx + dy * dy < radius1 * radius1) {
    float ri = radius1 * (thick + ((float)1 - thick) * rng.rand01());
    float phi = kTau * rng.rand01();
    return make_float2(w * (ri * safe_cos(phi) + x1), w * (ri * safe_sin(phi) + y1));
  }
  return make_float2(w * x, w * y);
}

// chaos.wgsl:3496 — an annular sector between phi1 and phi2.
P3_INLINE float2 var_glynnSim2(float2 p, float w, float radius, float thick, float contrast, float powp, float phi1, float phi2, Rng& rng) {
  float d2r = kPi / (float)180;
  float f1 = phi1 * d2r, f2 = phi2 * d2r;
  float rt = radius + thick;
  float gam = thick * ((float)2 * radius + thick) / smax(rt, (float)1e-30);
  float delta = f2 - f1;
  float ap = fabsf(powp);
  float r = sqrtf(dot2(p));
  if (r < radius) {
    float ri = rt - gam * rng.rand01();
    float ph = f1 + delta * rng.rand01();
    return make_float2(w * (ri * safe_cos(ph)), w * (ri * safe_sin(ph)));
  }
  float alpha = radius / smax(r, (float)1e-30);
  if (rng.rand01() > contrast * powf(fabsf(alpha), ap)) return make_float2(w * p.x, w * p.y);
  float a2 = alpha * alpha;
  return make_float2(w * (a2 * p.x), w * (a2 * p.y));
}

// chaos.wgsl:3527 — two concentric circles picked with probability gamma.
P3_INLINE float2 var_glynnSim3(float2 p, float w, float radius, float thick, float contrast, float powp, Rng& rng) {
  float r1 = radius + thick;
  float r2 = (radius * radius) / smax(r1, (float)1e-30);
  float gam = r1 / smax(r1 + r2, (float)1e-30);
  float ap = fabsf(powp);
  float r = sqrtf(dot2(p));
  if (r < r1) {
    float phi = kTau * rng.rand01();
    float ri = sel(r2, r1, rng.rand01() < gam);
    return make_float2(w * (ri * safe_cos(phi)), w * (ri * safe_sin(phi)));
  }
  float alpha = radius / smax(r, (float)1e-30);
  if (rng.rand01() > contrast * powf(fabsf(alpha), ap)) return make_float2(w * p.x, w * p.y);
  float a2 = alpha * alpha;
  return make_float2(w * (a2 * p.x), w * (a2 * p.y));
}

// chaos.wgsl:3562
P3_INLINE float2 var_flipy(float2 p, float w) {
  return make_float2(w * p.x, w * p.y * sel((float)1, (float)-1, p.x > (float)0));
}

// chaos.wgsl:3571
P3_INLINE float2 var_eclipse(float2 p, float w, float shift) {
  if (!(fabsf(p.y) <= w)) return make_float2(w * p.x, w * p.y);
  float c2 = sqrtf(smax(w * w - p.y * p.y, (float)0));
  float ox;
  if (fabsf(p.x) <= c2) {
    float xs = p.x + shift * w;
    ox = fabsf(xs) >= c2 ? -w * p.x : w * xs;
  } else {
    ox = w * p.x;
  }
  return make_float2(ox, w * p.y);
}

// chaos.wgsl:3596 — barycentric coordinates against basis (a,b), (c,d).
P3_INLINE float2 var_barycentroid(float2 p, float w, float a, float b, float c, float d) {
  float d00 = a * a + b * b, d01 = a * c + b * d, d02 = a * p.x + b * p.y;
  float d11 = c * c + d * d, d12 = c * p.x + d * p.y;
  float den = d00 * d11 - d01 * d01;
  if (fabsf(den) < (float)1e-30) return make_float2(w * p.x, w * p.y);
  float inv = (float)1 / den;
  float u = (d11 * d02 - d01 * d12) * inv;
  float v = (d00 * d12 - d01 * d02) * inv;
  return make_float2(w * (sqrtf(u * u + p.x * p.x) * signv(u)), w * (sqrtf(v * v + p.y * p.y) * signv(v)));
}

// chaos.wgsl:3624 — keep the point on one side of a conic, else (0, 0).
P3_INLINE float2 var_chunk(float2 p, float w, float a, float b, float c, float d, float e, float f, float modep) {
  float aa = w * a, bb = w * b, cc = w * c, dd = w * d, ee = w * e, ff = w * f;
  float r = aa * p.x * p.x + bb * p.x * p.y + cc * p.y * p.y + dd * p.x + ee * p.y + ff;
  const int32_t mode = f2i(modep);
  if (mode == 0 && r <= (float)0) return make_float2(p.x, p.y);
  if (mode == 1 && r > (float)0) return make_float2(p.x, p.y);
  return make_float2((float)0, (float)0);
}

// chaos.wgsl:3668
P3_INLINE float2 var_ennepers(float2 p, float w) {
  float x = p.x, y = p.y;
  return make_float2(w * (x - (x * x * x) / (float)3) + x * y * y, w * (y - (y * y * y) / (float)3) + y * x * x);
}

// chaos.wgsl:3681 — Abramowitz-Stegun 7.1.26 rational approximation.
P3_INLINE float erf_as(float x) {
  const float a1 = (float)0.254829592, a2 = (float)-0.284496736, a3 = (float)1.421413741;
  const float a4 = (float)-1.453152027, a5 = (float)1.061405429, pp = (float)0.3275911;
  float s = sel((float)-1, (float)1, x >= (float)0);
  float ax = fabsf(x);
  float t = (float)1 / ((float)1 + pp * ax);
  float poly = (((((a5 * t) + a4) * t + a3) * t + a2) * t + a1) * t;
  return s * ((float)1 - poly * expf(-ax * ax));
}
// chaos.wgsl:3695
P3_INLINE float2 var_erf(float2 p, float w) { return make_float2(w * erf_as(p.x), w * erf_as(p.y)); }

// chaos.wgsl:3703
P3_INLINE float2 var_circus(float2 p, float w, float scale) {
  float s1 = sel((float)1 / scale, (float)1e30, scale == (float)0);
  float r = sqrtf(p.x * p.x + p.y * p.y);
  float a = atan2f(p.y, p.x);
  float sa = sinf(a), ca = cosf(a);
  float rs = sel(r * s1, r * scale, r <= (float)1);
  return make_float2(w * rs * ca, w * rs * sa);
}

// chaos.wgsl:3722 — one draw where the disc and the corner region overlap.
P3_INLINE float2 var_asteria(float2 p, float w, float alpha, Rng& rng) {
  float sa = safe_sin(kPi * alpha), ca = safe_cos(kPi * alpha);
  float x0 = w * p.x, y0 = w * p.y;
  float r = x0 * x0 + y0 * y0;
  float xx = (fabsf(x0) - (float)1) * (fabsf(x0) - (float)1);
  float yy = (fabsf(y0) - (float)1) * (fabsf(y0) - (float)1);
  float r2 = sqrtf(yy + xx);
  const bool inside = r < (float)1;
  bool keep;
  if (inside && r2 < (float)1) keep = rng.rand01() > (float)0.35;
  else keep = !inside;
  if (keep) return make_float2(x0, y0);
  float rx = x0 * ca - y0 * sa, ry = x0 * sa + y0 * ca;
  float den = sqrtf(smax((float)1 - ry * ry, (float)1e-30));
  float in = smax((float)1 - (-fabsf(ry) + (float)1) * (-fabsf(ry) + (float)1), (float)0);
  float nx = rx / den * ((float)1 - sqrtf(in));
  return make_float2(nx * ca + ry * sa, -nx * sa + ry * ca);
}

// chaos.wgsl:3769
P3_INLINE float2 var_clifford_js(float2 p, float w, float a, float b, float c, float d) {
  return make_float2(w * (safe_sin(a * p.y) + c * safe_cos(a * p.x)),
          w * (safe_sin(b * p.x) + d * safe_cos(b * p.y)));
}

// chaos.wgsl:3785
P3_INLINE float2 var_devil_warp(float2 p, float w, float a, float b, float effect, float warp, float rmin, float rmax) {
  (void)w;
  float x = p.x, y = p.y;
  float r2 = (float)1 / smax(x * x + y * y, (float)1e-30);
  float ba = x * x + r2 * b * y * y;
  float bb = y * y + r2 * a * x * x;
  float pa = sel((float)0, powf(ba, warp), ba > (float)0);
  float pb = sel((float)0, powf(bb, warp), bb > (float)0);
  float r = clampv(pa - pb, rmin, rmax);
  r = effect * r;
  return make_float2(x * ((float)1 + r), y * ((float)1 + r));
}

// chaos.wgsl:3808 — integer lattice noise in [0, 1] (i32 arithmetic wraps).
P3_INLINE float voron_noise(int32_t x) {
  uint32_t s = ((uint32_t)x << 13) ^ (uint32_t)x;
  uint32_t r = (s * (s * s * 15731u + 789221u) + 1376312589u) & 0x7fffffffu;
  return (float)(float)(int32_t)r * (float)(float)(1.0 / 2147483647.0);
}
// chaos.wgsl:3815 — nearest of up to 26 hashed feature points per cell.
P3_INLINE float2 var_voron(float2 p, float w, float k, float step, float num_f, float xs_f, float ys_f) {
  const int32_t xs = f2i(xs_f), ys = f2i(ys_f);
  float ss = sel(step, (float)1e-30, step == (float)0);
  const int32_t M = f2i(floorf(p.x / ss)), N = f2i(floorf(p.y / ss));
  float rmin = (float)20, X0 = 0, Y0 = 0;
  for (int32_t i = -1; i < 2; ++i) {
    const int32_t M1 = iwrap_add(M, i);
    for (int32_t j = -1; j < 2; ++j) {
      const int32_t N1 = iwrap_add(N, j);
      const int32_t h = iwrap_add(iwrap_add(iwrap_mul(19, M1), iwrap_mul(257, N1)), xs);
      const int32_t kx = iwrap_add(1, f2i(floorf(num_f * voron_noise(h))));
      const int32_t cnt = imin(imax(kx, 1), 26);
      for (int32_t l = 0; l < cnt; ++l) {
        const int32_t hx = iwrap_add(iwrap_add(iwrap_add(l, iwrap_mul(64, M1)), iwrap_mul(15, N1)), xs);
        const int32_t hy = iwrap_add(iwrap_add(iwrap_add(l, iwrap_mul(21, M1)), iwrap_mul(33, N1)), ys);
        float X = (voron_noise(hx) + (float)M1) * step;
        float Y = (voron_noise(hy) + (float)N1) * step;
        float ox = p.x - X, oy = p.y - Y;
        float r = sqrtf(ox * ox + oy * oy);
        if (r < rmin) { rmin = r; X0 = X; Y0 = Y; }
      }
    }
  }
  return make_float2(w * (k * (p.x - X0) + X0), w * (k * (p.y - Y0) + Y0));
}

// chaos.wgsl:3865
P3_INLINE float2 var_henon(float2 p, float w, float a, float b, float c) {
  return make_float2(w * (c - a * p.x * p.x + p.y), w * b * p.x);
}

// chaos.wgsl:3874
P3_INLINE float2 var_atan(float2 p, float w, float modep, float stretch) {
  float nrm = w * (float)2 / kPi;
  const int32_t mode = f2i(modep);
  if (mode == 0) return make_float2(w * p.x, nrm * atanf(stretch * p.y));
  if (mode == 1) return make_float2(nrm * atanf(stretch * p.x), w * p.y);
  if (mode == 2) return make_float2(nrm * atanf(stretch * p.x), nrm * atanf(stretch * p.y));
  return make_float2((float)0, (float)0);
}

// chaos.wgsl:3896
P3_INLINE float2 var_cardioid(float2 p, float w, float a) {
  float th = atan2f(p.y, p.x);
  float rs = p.x * p.x + p.y * p.y + safe_sin(a * th) + (float)1;
  float r = w * sqrtf(smax(rs, (float)0));
  return make_float2(r * cosf(th), r * sinf(th));
}

// chaos.wgsl:3910 — one draw: a point on Fay's chrysanthemum curve.
P3_INLINE float2 var_chrysanthemum(float2 p, float w, Rng& rng) {
  (void)p;
  float u = (float)21 * kPi * rng.rand01();
  float p4 = sinf((float)17 * u / (float)3);
  float p8 = sinf((float)2 * cosf((float)3 * u) - (float)28 * u);
  float p44 = p4 * p4 * p4 * p4;
  float p82 = p8 * p8;
  float p88 = p82 * p82 * p82 * p82;
  float rr = (float)5 * ((float)1 + sinf((float)11 * u / (float)5)) - (float)4 * p44 * p88;
  float r = w * (float)0.1 * rr;
  return make_float2(r * cosf(u), r * sinf(u));
}

// chaos.wgsl:3928 — bipolar coordinates with angular folding.
P3_INLINE float2 var_bcollide(float2 p, float w, float nump, float a) {
  float num = smax((float)1, nump);
  float bnpi = num / kPi, pibn = kPi / num, bab = kPi * a / num;
  float xp = p.x + (float)1, xm = p.x - (float)1, y2 = p.y * p.y;
  float tau = (float)0.5 * (logf(smax(xp * xp + y2, (float)1e-30)) - logf(smax(xm * xm + y2, (float)1e-30)));
  float sr = kPi - atan2f(p.y, xp) - atan2f(p.y, (float)1 - p.x);
  const int32_t alt = f2i(sr * bnpi);
  float off = sel(-bab, bab, (alt & 1) == 0);
  float folded = sr + off - floorf((sr + off) / pibn) * pibn;
  float sg = (float)alt * pibn + folded;
  float tmp = coshf(tau) - safe_cos(sg);
  float ts = sel(tmp, (float)1e-30, fabsf(tmp) < (float)1e-30);
  return make_float2(w * sinhf(tau) / ts, w * safe_sin(sg) / ts);
}

// chaos.wgsl:3969
P3_INLINE float2 var_bsplit(float2 p, float w, float sx, float sy) {
  float ax = p.x + sx;
  float sn = sinf(ax);
  if (fabsf(sn) < (float)1e-6) return make_float2((float)0, (float)0);
  float cs = cosf(ax);
  float tn = sn / cs;
  float tsafe = sel(tn, signv(tn) * (float)1e6, fabsf(cs) < (float)1e-6);
  return make_float2(w / tsafe * safe_cos(p.y + sy), w / sn * ((float)-1 * p.y + sy));
}

// chaos.wgsl:3990
P3_INLINE float2 var_bulge(float2 p, float w, float n) {
  float rs = smax(sqrtf(p.x * p.x + p.y * p.y), (float)1e-30);
  float sc = powf(rs, n) / rs;
  return make_float2(w * p.x * sc, w * p.y * sc);
}

// chaos.wgsl:4012 — two draws, always taken.
P3_INLINE float2 var_checks(float2 p, float w, float cx, float cy, float size, float rnd, Rng& rng) {
  float cs = (float)1 / (size + (float)1e-6);
  const int32_t par = iwrap_add(f2i(rint_even(p.x * cs)), f2i(rint_even(p.y * cs)));
  float rnx = rnd * rng.rand01();
  float rny = rnd * rng.rand01();
  float dx, dy;
  if ((par & 1) == 0) { dx = -cx + rnx; dy = -cy; }
  else { dx = cx; dy = cy + rny; }
  return make_float2(w * (p.x + dx), w * (p.y + dy));
}

// chaos.wgsl:4046 — one draw plus a sin-hash of the position.
P3_INLINE float2 var_circular(float2 p, float w, float angle, float seed, Rng& rng) {
  float ca = angle * kPi / (float)180;
  float ar = sinf(p.x * (float)12.9898 + p.y * (float)78.233 + seed) * (float)43758.5453;
  float aux = ar - truncf(ar);
  float rnd = ((float)2 * (rng.rand01() + aux) - (float)2) * ca;
  float rad = sqrtf(p.x * p.x + p.y * p.y);
  float ang = atan2f(p.y, p.x);
  return make_float2(w * cosf(ang + rnd) * rad, w * sinf(ang + rnd) * rad);
}
// chaos.wgsl:4066
P3_INLINE float2 var_circular2(float2 p, float w, float angle, float seed, float xx, float yy, Rng& rng) {
  float ca = angle * kPi / (float)180;
  float ar = sinf(p.x * xx + p.y * yy + seed) * (float)43758.5453;
  float aux = ar - truncf(ar);
  float rnd = ((float)2 * (rng.rand01() + aux) - (float)2) * ca;
  float rad = sqrtf(p.x * p.x + p.y * p.y);
  float ang = atan2f(p.y, p.x);
  return make_float2(w * cosf(ang + rnd) * rad, w * sinf(ang + rnd) * rad);
}

// chaos.wgsl:4081
P3_INLINE float2 var_corners(float2 p, float w, float cx, float cy, float mx, float my, float xp, float yp, float xyp, float logm, float logb) {
  float xs = p.x * p.x, ys = p.y * p.y;
  float ex, ey;
  if (logm == (float)0) {
    ex = powf(smax(xs, (float)0), xp + xyp) * mx;
    ey = powf(smax(ys, (float)0), yp + xyp) * my;
  } else {
    float lb = logf(smax(fabsf(logb), (float)1.000001));
    ex = powf(logf((xs * mx) + (float)3) / lb, xp + (float)2.25 + xyp) - (float)1.33;
    ey = powf(logf((ys * my) + (float)3) / lb, yp + (float)2.25 + xyp) - (float)1.33;
  }
  return make_float2(sel(-w * ex - cx, w * ex + cx, p.x > (float)0), sel(-w * ey - cy, w * ey + cy, p.y > (float)0));
}

// ---- circleblur .. twoface (chaos.wgsl:4113-4812) -----------------------------

// chaos.wgsl:4113 — two draws: a uniform point in the disc.
P3_INLINE float2 var_circleblur(float2 p, float w, Rng& rng) {
  (void)p;
  float rad = sqrtf(rng.rand01());
  float a = rng.rand01() * (float)2 * kPi;
  return make_float2(w * cosf(a) * rad, w * sinf(a) * rad);
}

// chaos.wgsl:4130 — Binet-form Fibonacci spiral.
P3_INLINE float2 var_fibonacci2(float2 p, float w, float sc, float sc2) {
  const float ffive = (float)0.447213595, lnphi = (float)0.481211825;
  float a = p.y * lnphi;
  float s1 = safe_sin(a), c1 = safe_cos(a);
  float b = (p.x * kPi + p.y * lnphi) * (float)-1;
  float s2 = safe_sin(b), c2 = safe_cos(b);
  float e1 = sc * expf(sc2 * (p.x * lnphi));
  float e2 = sc * expf(sc2 * ((p.x * lnphi - p.y * kPi) * (float)-1));
  return make_float2(w * (e1 * c1 - e2 * c2) * ffive, w * (e1 * s1 - e2 * s2) * ffive);
}

// Hyperbolic {p, q} tiling radius (chaos.wgsl:4152-4220); `form` picks the
// reference's two algebraically equal spellings of r^2.
P3_INLINE float hypertile_r(float pp, float q, bool form1, float& pa) {
  float pi_ = smax((float)3, pp), qs = smax((float)3, q);
  pa = (float)2 * kPi / pi_;
  float r2;
  if (!form1) {
    float qa = (float)2 * kPi / qs;
    float den = cosf(pa) + cosf(qa);
    r2 = sel((float)1, ((float)1 - cosf(pa)) / den + (float)1, fabsf(den) > (float)1e-30);
  } else {
    float cp = cosf(pa);
    float den = cp + cosf((float)2 * kPi / qs);
    r2 = sel((float)1, (float)1 - (cp - (float)1) / den, fabsf(den) > (float)1e-30);
  }
  return sel((float)1, (float)1 / sqrtf(smax(r2, (float)1e-30)), r2 > (float)0);
}
// chaos.wgsl:4152 — Moebius step of the tiling group by the nth rotation.
P3_INLINE float2 var_hypertile(float2 p, float w, float pp, float q, float n) {
  float pa;
  float r = hypertile_r(pp, q, false, pa);
  float an = n * pa;
  float re = r * safe_cos(an), im = r * safe_sin(an);
  float a = p.x + re, b = p.y - im;
  float c = re * p.x - im * p.y + (float)1, d = re * p.y + im * p.x;
  float vr = w / smax(c * c + d * d, (float)1e-30);
  return make_float2(vr * (a * c + b * d), vr * (b * c - a * d));
}
// chaos.wgsl:4176 — one draw picks one of ten rotations.
P3_INLINE float2 var_hypertile1(float2 p, float w, float pp, float q, Rng& rng) {
  float pa;
  float r = hypertile_r(pp, q, true, pa);
  float rpa = floorf(rng.rand01() * (float)10) * pa;
  float re = r * cosf(rpa), im = r * sinf(rpa);
  float a = p.x + re, b = p.y - im;
  float c = re * p.x - im * p.y + (float)1, d = re * p.y + im * p.x;
  float vr = w / smax(c * c + d * d, (float)1e-30);
  return make_float2(vr * (a * c + b * d), vr * (b * c - a * d));
}
// chaos.wgsl:4200 — one draw picks the rotation applied after the step.
P3_INLINE float2 var_hypertile2(float2 p, float w, float pp, float q, Rng& rng) {
  float pa;
  float r = hypertile_r(pp, q, true, pa);
  float a = p.x + r, b = p.y;
  float c = r * p.x + (float)1, d = r * p.y;
  float xx = a * c + b * d, yy = b * c - a * d;
  float vr = w / smax(c * c + d * d, (float)1e-30);
  float rpa = floorf(rng.rand01() * (float)32767) * pa;
  float ca = safe_cos(rpa), sa = safe_sin(rpa);
  return make_float2(vr * (xx * ca + yy * sa), vr * (yy * ca - xx * sa));
}

// chaos.wgsl:4226
P3_INLINE float2 var_idisc(float2 p, float w) {
  float r = sqrtf(p.x * p.x + p.y * p.y);
  float a = kPi / (r + (float)1);
  float v = atan2f(p.y, p.x) * w / kPi;
  return make_float2(v * cosf(a), v * sinf(a));
}

// chaos.wgsl:4243
P3_INLINE float2 var_hole(float2 p, float w, float a, float insidep) {
  float al = atan2f(p.y, p.x);
  float delta = powf(smax(al / kPi + (float)1, (float)1e-30), a);
  float ss = p.x * p.x + p.y * p.y;
  float r = sel(w * sqrtf(smax(ss + delta, (float)0)), w * delta / smax(ss + delta, (float)1e-30),
            f2i(insidep) != 0);
  return make_float2(r * cosf(al), r * sinf(al));
}

// chaos.wgsl:4260 — cos/sin of 45 radians, folded to f32 constants.
P3_INLINE float2 var_kaleidoscope(float2 p, float w, float pull, float rotate, float lineup, float offx, float offy) {
  const float c45 = (float)0.5253219888177297, s45 = (float)0.8509035245341184;
  float ox = w * (((rotate * p.x) * c45 - p.y * s45 + lineup) + offx);
  float oy = sel(w * ((rotate * p.y) * c45 + p.x * s45 - pull - lineup),
             w * (((rotate * p.y) * c45 + p.x * s45 + pull + lineup) + offy), p.y > (float)0);
  return make_float2(ox, oy);
}

// chaos.wgsl:4275
P3_INLINE float2 var_layered_spiral(float2 p, float w, float radius) {
  float a = p.x * radius;
  float t = p.x * p.x + p.y * p.y + (float)1e-30;
  return make_float2(w * a * safe_cos(t), w * a * safe_sin(t));
}

// chaos.wgsl:4283
P3_INLINE float2 var_linear_t(float2 p, float w, float px, float py) {
  float sx = sel((float)-1, (float)1, p.x >= (float)0), sy = sel((float)-1, (float)1, p.y >= (float)0);
  return make_float2(w * sx * powf(smax(fabsf(p.x), (float)1e-30), px),
          w * sy * powf(smax(fabsf(p.y), (float)1e-30), py));
}

// chaos.wgsl:4296 — one draw along a fixed direction.
P3_INLINE float2 var_line(float2 p, float w, float delta, float phi, Rng& rng) {
  (void)p;
  float cd = safe_cos(delta * kPi), sd = safe_sin(delta * kPi);
  float cp = safe_cos(phi * kPi), sp = safe_sin(phi * kPi);
  float ux = cd * cp, uy = sd * cp, uz = sp;
  float r = sqrtf(smax(ux * ux + uy * uy + uz * uz, (float)1e-30));
  ux = ux / r;
  uy = uy / r;
  float t = rng.rand01() * w;
  return make_float2(ux * t, uy * t);
}

// chaos.wgsl:4323
P3_INLINE float2 var_ovoid(float2 p, float w, float px, float py) {
  float r = w / (p.x * p.x + p.y * p.y + (float)1e-6);
  return make_float2(p.x * r * px, p.y * r * py);
}

// chaos.wgsl:4331 — one draw picks the root.
P3_INLINE float2 var_phoenix_julia(float2 p, float w, float power, float dist, float xd, float yd, Rng& rng) {
  float ps = sel(power, (float)1, power == (float)0);
  float inv_n = dist / ps, inv2 = (float)2 * kPi / ps;
  float cn = dist / ((float)2 * ps);
  float px = p.x * (xd + (float)1), py = p.y * (yd + (float)1);
  float k = floorf(rng.rand01() * (float)32768);
  float a = atan2f(py, px) * inv_n + k * inv2;
  float r = w * powf(smax(p.x * p.x + p.y * p.y, (float)1e-30), cn);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}

// chaos.wgsl:4359
P3_INLINE float2 var_unpolar(float2 p, float w) {
  float v = (w / kPi) * (float)0.5;
  float r = expf(p.y);
  return make_float2(v * r * safe_sin(p.x), v * r * safe_cos(p.x));
}

// chaos.wgsl:4368
P3_INLINE float2 var_shredrad(float2 p, float w, float n, float widthp) {
  float width = clampv(widthp, (float)-1, (float)1);
  float alpha = (float)2 * kPi / smax((float)0.001, n);
  float ang = atan2f(p.y, p.x);
  float rad = sqrtf(p.x * p.x + p.y * p.y);
  float xa = (ang + (float)3 * kPi + alpha * (float)0.5) / alpha;
  float za = ((xa - floorf(xa)) * width + floorf(xa)) * alpha - kPi - alpha * (float)0.5 * width;
  return make_float2(w * rad * safe_cos(za), w * rad * safe_sin(za));
}

// chaos.wgsl:4395 — one draw: a seed of Vogel's sunflower model.
P3_INLINE float2 var_vogel(float2 p, float w, float n, float scale, Rng& rng) {
  const float phi = (float)1.61803398874989;
  float k2 = (float)2 * kPi / (phi * phi);
  float i = floorf(rng.rand01() * smax((float)1, n)) + (float)1;
  float a = i * k2;
  float r = w * (sqrtf(p.x * p.x + p.y * p.y) + sqrtf(i));
  return make_float2(r * (safe_cos(a) + scale * p.x), r * (safe_sin(a) + scale * p.y));
}

// chaos.wgsl:4414 — one draw only in dual mode inside the unit disc.
P3_INLINE float2 var_yin_yang(float2 p, float w, float radius, float ang1, float ang2, float dualp, float outp, Rng& rng) {
  float sa = safe_sin(kPi * ang1), ca = safe_cos(kPi * ang1);
  float sb = safe_sin(kPi * ang2), cb = safe_cos(kPi * ang2);
  const int32_t dual = f2i(dualp), outside = f2i(outp);
  float x = p.x, y = p.y;
  float inv = (float)1, RR = radius;
  if (x * x + y * y < (float)1) {
    float nx = x * ca - y * sa, ny = x * sa + y * ca;
    if (dual == 1 && rng.rand01() > (float)0.5) {
      inv = (float)-1;
      RR = (float)1 - radius;
      nx = x * cb - y * sb;
      ny = x * sb + y * cb;
    }
    x = nx;
    y = ny;
    if (y > (float)0) {
      float t = sqrtf(smax((float)1 - y * y, (float)0));
      float k = x / smax(fabsf(t), (float)1e-30) * signv(t);
      float t1 = (t - (float)0.5) * (float)2;
      float al = ((float)1 - k) * (float)0.5;
      float be = (float)1 - al;
      float dx = al * (RR - (float)1);
      float k1 = al * RR + be * (float)1;
      return make_float2(w * (t1 * k1 + dx) * inv, w * sqrtf(smax((float)1 - t1 * t1, (float)0)) * k1 * inv);
    }
    return make_float2(w * (x * ((float)1 - RR) + RR) * inv, w * (y * ((float)1 - RR)) * inv);
  }
  if (outside == 1) return make_float2(w * p.x, w * p.y);
  return make_float2((float)0, (float)0);
}

// chaos.wgsl:4466 — one draw picks the winding.
P3_INLINE float2 var_squish(float2 p, float w, float powp, Rng& rng) {
  float power = smax((float)2, powp);
  float ip = (float)1 / power;
  float ax = fabsf(p.x), ay = fabsf(p.y);
  float s, pp;
  if (ax > ay) { s = ax; pp = sel((float)4 * s - p.y, p.y, p.x > (float)0); }
  else { s = ay; pp = sel((float)6 * s + p.x, (float)2 * s - p.x, p.y > (float)0); }
  float rot = floorf(power * rng.rand01());
  pp = ip * (pp + (float)8 * s * rot);
  if (pp <= (float)1 * s) return make_float2(w * s, w * pp);
  if (pp <= (float)3 * s) return make_float2(w * ((float)2 * s - pp), w * s);
  if (pp <= (float)5 * s) return make_float2(-w * s, w * ((float)4 * s - pp));
  if (pp <= (float)7 * s) return make_float2(-w * ((float)6 * s - pp), -w * s);
  return make_float2(w * s, -w * ((float)8 * s - pp));
}

// chaos.wgsl:4501
P3_INLINE float2 var_target(float2 p, float w, float even, float odd, float size) {
  float h = (float)0.5 * size;
  float a = atan2f(p.y, p.x);
  float r = sqrtf(p.x * p.x + p.y * p.y);
  float t = logf(smax(r, (float)1e-30));
  if (t < (float)0) t = t - h;
  float at = fabsf(t);
  float ss = smax(fabsf(size), (float)1e-30);
  t = at - floorf(at / ss) * ss;
  a = a + sel(odd, even, t < h);
  return make_float2(r * safe_cos(a) * w, r * safe_sin(a) * w);
}

// chaos.wgsl:4530
P3_INLINE float2 var_funnel(float2 p, float w, float effect) {
  float cx = safe_cos(p.x), cy = safe_cos(p.y);
  float secx = (float)1 / sel(cx, (float)1e-30, fabsf(cx) < (float)1e-30);
  float secy = (float)1 / sel(cy, (float)1e-30, fabsf(cy) < (float)1e-30);
  float off = effect * kPi;
  return make_float2(w * safe_tanh(p.x) * (secx + off), w * safe_tanh(p.y) * (secy + off));
}

// chaos.wgsl:4550
P3_INLINE float2 var_holesq(float2 p, float w) {
  float x = w * p.x, y = w * p.y;
  float fx = fabsf(x), fy = fabsf(y);
  if (fx + fy > (float)1) return make_float2(x, y);
  if (fx > fy) return make_float2(sel((x + fy - (float)1) * (float)0.5, (x - fy + (float)1) * (float)0.5, x >= (float)0), y);
  return make_float2(x, sel((y + fx - (float)1) * (float)0.5, (y - fx + (float)1) * (float)0.5, y >= (float)0));
}

// chaos.wgsl:4568 — ten hole shapes.
P3_INLINE float2 var_hole2(float2 p, float w, float a, float b, float c, float d, float insidep, float shapep) {
  float rho = p.x * p.x + p.y * p.y;
  float th = atan2f(p.y, p.x) * d;
  float delta = powf(smax(th / kPi + (float)1, (float)1e-30), a) * c;
  float r1;
  switch (f2i(shapep)) {
    case 0: r1 = sqrtf(smax(rho, (float)0)) + delta; break;
    case 1: r1 = sqrtf(smax(rho + delta, (float)0)); break;
    case 2: r1 = sqrtf(smax(rho + safe_sin(b * th) + delta, (float)0)); break;
    case 3: r1 = sqrtf(smax(rho + safe_sin(th) + delta, (float)0)); break;
    case 4: r1 = sqrtf(smax(rho + th * th - delta + (float)1, (float)0)); break;
    case 5: r1 = sqrtf(smax(rho + fabsf(safe_tan(th)) + delta, (float)0)); break;
    case 6: r1 = sqrtf(smax(rho * ((float)1 + safe_sin(b * th)) + delta, (float)0)); break;
    case 7: r1 = sqrtf(smax(rho + fabsf(safe_sin((float)0.5 * b * th)) + delta, (float)0)); break;
    case 8: r1 = sqrtf(smax(rho + safe_sin(kPi * safe_sin(b * th)) + delta, (float)0)); break;
    case 9:
      r1 = sqrtf(smax(
          rho + (safe_sin(b * th) + safe_sin((float)2 * b * th + kPi * (float)0.5)) * (float)0.5 + delta, (float)0));
      break;
    default: r1 = (float)1; break;
  }
  float rf = sel(w * r1, w / smax(r1, (float)1e-30), f2i(insidep) != 0);
  return make_float2(rf * safe_cos(th), rf * safe_sin(th));
}

// chaos.wgsl:4604 — one draw picks one of four inversion centres.
P3_INLINE float2 var_lace_js(float2 p, float w, Rng& rng) {
  const float r = (float)2, h3 = (float)0.8660254037844386;
  float r0 = sqrtf(p.x * p.x + p.y * p.y);
  float u = rng.rand01();
  float x, y;
  if (u > (float)0.75) {
    float th = atan2f(p.y, p.x - (float)1);
    y = -r0 * cosf(th) / r + (float)1;
    x = -r0 * sinf(th) / r;
  } else if (u > (float)0.5) {
    float th = atan2f(p.y - h3, p.x + (float)0.5);
    y = -r0 * cosf(th) / r - (float)0.5;
    x = -r0 * sinf(th) / r + h3;
  } else if (u > (float)0.25) {
    float th = atan2f(p.y + h3, p.x + (float)0.5);
    y = -r0 * cosf(th) / r - (float)0.5;
    x = -r0 * sinf(th) / r - h3;
  } else {
    float th = atan2f(p.y, p.x);
    y = -r0 * cosf(th) / r;
    x = -r0 * sinf(th) / r;
  }
  return make_float2(w * x, w * y);
}

// chaos.wgsl:4633 — modes 0/1 draw once for the sign of the root.
P3_INLINE float2 var_julia_outside(float2 p, float w, float rediv, float imdiv, float modep, Rng& rng) {
  const int32_t mode = f2i(modep);
  const bool sq = mode == 0 || mode == 2, rt = mode == 0 || mode == 1;
  float2 z = p, z2 = p;
  if (sq) z = c_sqrt(z);
  z = make_float2(z.x + (float)1, z.y);
  if (sq) z = c_sqr(z);
  if (sq) z2 = c_sqrt(z2);
  z2 = make_float2(z2.x - (float)1, z2.y);
  if (sq) z2 = c_sqr(z2);
  z = c_div(z, z2);
  if (rt) z = c_sqrt(z);
  z = c_div(z, make_float2(rediv, imdiv));
  if (rt) {
    float sg = sel((float)-1, (float)1, rng.rand01() < (float)0.5);
    return make_float2(w * sg * z.x, w * sg * z.y);
  }
  return make_float2(w * z.x, w * z.y);
}

// chaos.wgsl:4680 — a different map per quadrant.
P3_INLINE float2 var_fourth(float2 p, float w, float spin, float space, float twist, float offx, float offy) {
  float ww = w * w;
  if (p.x > (float)0 && p.y > (float)0) {
    float th = atan2f(p.y, p.x);
    float r = (float)1 / smax(sqrtf(p.x * p.x + p.y * p.y), (float)1e-30);
    return make_float2(w * r * cosf(th), w * r * sinf(th));
  }
  if (p.x > (float)0 && p.y < (float)0) {
    float r2 = p.x * p.x + p.y * p.y;
    if (r2 < ww) {
      float r = w * sqrtf(smax(ww / smax(r2, (float)1e-30) - (float)1, (float)0));
      return make_float2(r * p.x, r * p.y);
    }
    return make_float2(w * p.x, w * p.y);
  }
  if (p.x < (float)0 && p.y > (float)0) {
    float xx = p.x - offx, yy = p.y + offy;
    float r0 = sqrtf(xx * xx + yy * yy);
    if (r0 < w) {
      float th = atan2f(yy, xx) + spin + twist * (w - r0);
      float r = w * r0;
      return make_float2(r * safe_cos(th) + offx, r * safe_sin(th) - offy);
    }
    float r = w * ((float)1 + space / smax(r0, (float)1e-30));
    return make_float2(r * xx + offx, r * yy - offy);
  }
  return make_float2(w * p.x, w * p.y);
}

// chaos.wgsl:4728
P3_INLINE float2 var_pulse(float2 p, float w, float fx, float fy, float sx, float sy) {
  return make_float2(w * (p.x + sx * safe_sin(p.x * fx)), w * (p.y + sy * safe_sin(p.y * fy)));
}

// chaos.wgsl:4738
P3_INLINE float2 var_rays1(float2 p, float w) {
  float t = p.x * p.x + p.y * p.y;
  float tv = safe_tan(sqrtf(smax(t, (float)1e-30)));
  float it = (float)1 / sel(tv, (float)1e-30, fabsf(tv) < (float)1e-30);
  float u = it + w * ((float)2 / kPi) * ((float)2 / kPi);
  return make_float2(w * u * t / sel(p.x, (float)1e-30, p.x == (float)0), w * u * t / sel(p.y, (float)1e-30, p.y == (float)0));
}
// chaos.wgsl:4751
P3_INLINE float2 var_rays2(float2 p, float w) {
  float t = p.x * p.x + p.y * p.y;
  float ts = smax(t, (float)1e-30);
  float in = (ts + (float)1e-6) * safe_tan((float)1 / ts + (float)1e-6);
  float ci = safe_cos(in);
  float u = (float)1 / sel(ci, (float)1e-30, fabsf(ci) < (float)1e-30);
  float co = w / (float)10;
  return make_float2(co * u * t / sel(p.x, (float)1e-30, p.x == (float)0), co * u * t / sel(p.y, (float)1e-30, p.y == (float)0));
}
// chaos.wgsl:4766
P3_INLINE float2 var_rays3(float2 p, float w) {
  float t = p.x * p.x + p.y * p.y;
  float ts = smax(t, (float)1e-30);
  float in = safe_sin(t * t + (float)1e-6) * safe_sin((float)1 / (ts * ts) + (float)1e-6);
  float den = sqrtf(smax(safe_cos(in), (float)1e-30));
  float u = (float)1 / smax(den, (float)1e-30);
  float co = w / (float)10;
  return make_float2(co * u * safe_cos(t) * t / sel(p.x, (float)1e-30, p.x == (float)0),
          co * u * safe_tan(t) * t / sel(p.y, (float)1e-30, p.y == (float)0));
}

// chaos.wgsl:4786
P3_INLINE float2 var_tancos(float2 p, float w) {
  float d1 = (float)1e-6 + p.x * p.x + p.y * p.y;
  float d2 = w / d1;
  return make_float2(d2 * safe_tanh(d1) * (float)2 * p.x, d2 * safe_cos(d1) * (float)2 * p.y);
}

// chaos.wgsl:4802
P3_INLINE float2 var_twoface(float2 p, float w) {
  float v = w;
  if (p.x > (float)0) v = v / smax(p.x * p.x + p.y * p.y, (float)1e-30);
  return make_float2(v * p.x, v * p.y);
}

// ---- elliptic / bipolar family .. circlecrop (chaos.wgsl:4814-5363) -----------

// Elliptic coordinates (mu, nu) of p shared by e_julia / e_collide / e_mod
// (chaos.wgsl:4814-4936): returns xmax = cosh(mu) (>= 1) and nu in [0, pi].
P3_INLINE float ell_xmax(float x, float r2) {
  float tmp = r2 + (float)1, tmp2 = (float)2 * x;
  float sa = sqrtf(smax(tmp + tmp2, (float)0)), sb = sqrtf(smax(tmp - tmp2, (float)0));
  float xmax = (sa + sb) * (float)0.5;
  return xmax < (float)1 ? (float)1 : xmax;
}
P3_INLINE float ell_nu(float x, float xmax) {
  float t = x / xmax;
  if (t > (float)1) t = (float)1;
  if (t < (float)-1) t = (float)-1;
  return acosf(t);
}

// chaos.wgsl:4814 — one draw picks the root.
P3_INLINE float2 var_e_julia(float2 p, float w, float powp, Rng& rng) {
  float ps = sel(powp, (float)1, powp == (float)0);
  float pa = fabsf(ps);
  float r2 = p.y * p.y + p.x * p.x;
  float x = p.x;
  if (!(ps > (float)0)) {
    r2 = (float)1 / smax(r2, (float)1e-30);
    x = p.x * r2;
  }
  float xmax = ell_xmax(x, r2);
  float mu0 = acoshf(xmax);
  float nu = ell_nu(x, xmax);
  if (p.y < (float)0) nu = -nu;
  float k = floorf(rng.rand01() * pa);
  nu = nu / ps + ((float)2 * kPi / ps) * k;
  float mu = mu0 / ps;
  return make_float2(w * coshf(mu) * safe_cos(nu), w * sinhf(mu) * safe_sin(nu));
}

// chaos.wgsl:4857 — one draw when filled > 0, a second when it fills.
P3_INLINE float2 var_cannabis_curve_wf(float2 p, float w, float filled, Rng& rng) {
  float a = atan2f(p.y, p.x);
  float r = ((float)1 + (float)0.9 * cosf((float)8 * a)) * ((float)1 + (float)0.1 * cosf((float)24 * a)) *
        ((float)0.9 + (float)0.1 * cosf((float)200 * a)) * ((float)1 + sinf(a));
  a = a + kPi * (float)0.5;
  if (filled > (float)0 && filled > rng.rand01()) r = r * rng.rand01();
  return make_float2(w * sinf(a) * r, w * cosf(a) * r);
}

// chaos.wgsl:4868
P3_INLINE float2 var_e_collide(float2 p, float w, float nump, float a) {
  float num = smax((float)1, nump);
  float knpi = num / kPi, pikn = kPi / num, ka = kPi * a / num;
  float xmax = ell_xmax(p.x, p.y * p.y + p.x * p.x);
  float nu = ell_nu(p.x, xmax);
  const int32_t alt = f2i(nu * knpi);
  float off = sel(-ka, ka, (alt & 1) == 0);
  float arg = nu + off;
  nu = (float)alt * pikn + (arg - floorf(arg / pikn) * pikn);
  if (p.y <= (float)0) nu = -nu;
  return make_float2(w * xmax * cosf(nu),
          w * sqrtf(smax(xmax - (float)1, (float)0)) * sqrtf(xmax + (float)1) * sinf(nu));
}

// chaos.wgsl:4897
P3_INLINE float2 var_e_mod(float2 p, float w, float radius, float distance) {
  float xmax = ell_xmax(p.x, p.y * p.y + p.x * p.x);
  float mu = acoshf(xmax);
  float nu = ell_nu(p.x, xmax);
  if (p.y < (float)0) nu = -nu;
  if (mu < radius && -mu < radius) {
    float rs = smax(radius, (float)1e-30);
    float two = (float)2 * rs;
    if (nu > (float)0) mu = wmod(mu + rs + distance * rs, two) - rs;
    else mu = wmod(mu - rs - distance * rs, two) + rs;
  }
  return make_float2(w * coshf(mu) * cosf(nu), w * sinhf(mu) * sinf(nu));
}

// Tiled band shared by intersection / parallel: the coordinate folded into
// [-m1, m1) with period m2*m1 (chaos.wgsl:4938-5204, `%` spelling).
P3_INLINE float band_fold_mod(float v, float m1, float period) {
  if (v > m1) return -m1 + wmod(v + m1, smax(period, (float)1e-30));
  if (v < -m1) return m1 - wmod(m1 - v, smax(period, (float)1e-30));
  return v;
}
// The same fold written with floor (parallel).
P3_INLINE float band_fold_floor(float v, float m1, float period) {
  float rs = smax(period, (float)1e-30);
  if (v > m1) { float a = v + m1; return -m1 + (a - floorf(a / rs) * rs); }
  if (v < -m1) { float a = m1 - v; return m1 - (a - floorf(a / rs) * rs); }
  return v;
}
// Random integer step along one axis: three draws (branch already taken).
P3_INLINE float band_jump(float width, Rng& rng) {
  float x = sel(-width, width, rng.rand01() < (float)0.5);
  return rint_even(x * logf(smax(rng.rand01(), (float)1e-30)));
}

// chaos.wgsl:4938 — one branch draw, then two more.
P3_INLINE float2 var_intersection(float2 p, float w, float xwidth, float xtile, float xm1, float xm2, float xheight, float yheight,
                       float ytile, float ym1, float ym2, float ywidth, Rng& rng) {
  if (w == (float)0) return make_float2((float)0, (float)0);
  float xr1 = xm2 * xm1, yr1 = ym2 * ym1;
  if (rng.rand01() < (float)0.5) {
    float ox = xtile * (p.x + band_jump(xwidth, rng));
    return make_float2(ox, xheight * band_fold_mod(p.y, xm1, xr1));
  }
  float oy = ytile * (p.y + band_jump(yheight, rng));
  return make_float2(ywidth * band_fold_mod(p.x, ym1, yr1), oy);
}

// chaos.wgsl:4985
P3_INLINE float2 var_inv_squircular(float2 p, float w) {
  if (w == (float)0) return make_float2((float)0, (float)0);
  const float sq2 = (float)1.41421356237;
  float u = p.x, v = p.y;
  float ri = u * u + v * v;
  float r2 = sqrtf(smax(ri * (w * w * ri - (float)4 * u * u * v * v) / w, (float)0));
  float r = sqrtf(smax(ri - r2, (float)0)) / sq2;
  return make_float2(r / sel(u, (float)1e-30, u == (float)0), r / sel(v, (float)1e-30, v == (float)0));
}

// chaos.wgsl:5008
P3_INLINE float2 var_lozi(float2 p, float w, float a, float b, float c) {
  return make_float2(w * (c - a * fabsf(p.x) + p.y), w * b * p.x);
}

// chaos.wgsl:5014
P3_INLINE float2 var_hypershift(float2 p, float w, float shift, float stretch) {
  float scale = (float)1 - shift * shift;
  float r1 = (float)1 / smax(p.x * p.x + p.y * p.y, (float)1e-30);
  float x = r1 * p.x + shift, y = r1 * p.y;
  float rad = w * scale / smax(x * x + y * y, (float)1e-30);
  return make_float2(rad * x + shift, rad * y * stretch);
}

// chaos.wgsl:5026 — offset from the nearest hexagon centre (cube rounding).
P3_INLINE float2 var_hex_modulus(float2 p, float w, float size) {
  const float h32 = (float)0.8660254037844386, sq3 = (float)1.7320508075688772;
  float hs = h32 / smax(size, (float)1e-30);
  float wt = w / h32;
  float X = p.x * hs, Y = p.y * hs;
  float x = (float)0.5773502691896258 * X - Y / (float)3;
  float z = (float)2 * Y / (float)3;
  float y = -x - z;
  float rx = rint_even(x), ry = rint_even(y), rz = rint_even(z);
  float dx = fabsf(rx - x), dy = fabsf(ry - y), dz = fabsf(rz - z);
  if (dx > dy && dx > dz) rx = -ry - rz;
  else if (dy > dz) ry = -rx - rz;
  else rz = -rx - ry;
  (void)ry;
  float fx = sq3 * rx + h32 * rz, fy = (float)1.5 * rz;
  return make_float2((X - fx) * wt, (Y - fy) * wt);
}

// chaos.wgsl:5067 — one draw decides whether the point is pushed to a border.
P3_INLINE float2 var_boarders2(float2 p, float w, float c, float left, float right, Rng& rng) {
  float ac = smax(fabsf(c), (float)1e-6);
  float al = smax(fabsf(left), (float)1e-6), ar = smax(fabsf(right), (float)1e-6);
  float cl = ac * al, cr = ac + ac * ar;
  float rx = rint_even(p.x), ry = rint_even(p.y);
  float ox = p.x - rx, oy = p.y - ry;
  if (rng.rand01() >= cr) return make_float2(w * (ox * ac + rx), w * (oy * ac + ry));
  float sx = sel(ox, (float)1e-30, ox == (float)0), sy = sel(oy, (float)1e-30, oy == (float)0);
  if (fabsf(ox) >= fabsf(oy)) {
    if (ox >= (float)0) return make_float2(w * (ox * ac + rx + cl), w * (oy * ac + ry + cl * oy / sx));
    return make_float2(w * (ox * ac + rx - cl), w * (oy * ac + ry - cl * oy / sx));
  }
  if (oy >= (float)0) return make_float2(w * (ox * ac + rx + ox / sy * cl), w * (oy * ac + ry + cl));
  return make_float2(w * (ox * ac + rx - ox / sy * cl), w * (oy * ac + ry - cl));
}

// Bipolar coordinates (tau, sigma) of p (chaos.wgsl:5107-5145).
P3_INLINE float bip_tau(float2 p) {
  float xp = p.x + (float)1, xm = p.x - (float)1, y2 = p.y * p.y;
  return (float)0.5 * (logf(smax(xp * xp + y2, (float)1e-30)) - logf(smax(xm * xm + y2, (float)1e-30)));
}
P3_INLINE float bip_sigma(float2 p) {
  return kPi - atan2f(p.y, p.x + (float)1) - atan2f(p.y, (float)1 - p.x);
}
// chaos.wgsl:5107
P3_INLINE float2 var_b_mod(float2 p, float w, float radius, float distance) {
  float tau = bip_tau(p);
  float sg = bip_sigma(p);
  if (tau < radius && -tau < radius) {
    float rs = smax(radius, (float)1e-30);
    float two = (float)2 * rs;
    float arg = tau + rs + distance * rs;
    tau = arg - floorf(arg / two) * two - rs;
  }
  float tmp = coshf(tau) - cosf(sg);
  float ts = sel(tmp, (float)1e-30, fabsf(tmp) < (float)1e-30);
  return make_float2(w * sinhf(tau) / ts, w * sinf(sg) / ts);
}
// chaos.wgsl:5127 — one draw picks the root.
P3_INLINE float2 var_b_transform(float2 p, float w, float rotate, float powp, float move, float split, Rng& rng) {
  float power = smax((float)1, powp);
  float tau = bip_tau(p) / power + move;
  float sg = bip_sigma(p) + rotate;
  float k = floorf(rng.rand01() * power);
  sg = sg / power + ((float)2 * kPi / power) * k;
  tau = tau + sel(-split, split, p.x >= (float)0);
  float tmp = coshf(tau) - safe_cos(sg);
  float ts = sel(tmp, (float)1e-30, fabsf(tmp) < (float)1e-30);
  return make_float2(w * sinhf(tau) / ts, w * safe_sin(sg) / ts);
}

// chaos.wgsl:5147 — like intersection, with two horizontal band sets.
P3_INLINE float2 var_parallel(float2 p, float w, float x1w, float x1t, float x1m1, float x1m2, float x1h, float x1mv, float x2w, float x2t, float x2m1,
                   float x2m2, Rng& rng) {
  const float x2h = (float)0.5, x2mv = (float)1;
  if (w == (float)0) return make_float2((float)0, (float)0);
  float xr1 = x1m2 * x1m1, xr2 = x2m2 * x2m1;
  if (rng.rand01() < (float)0.5) {
    float ox = x1t * (p.x + band_jump(x1w, rng));
    return make_float2(ox, x1h * band_fold_floor(p.y, x1m1, xr1) + w * x1mv);
  }
  float ox = x2t * (p.x + band_jump(x2w, rng));
  return make_float2(ox, x2h * band_fold_floor(p.y, x2m1, xr2) - w * x2mv);
}

// chaos.wgsl:5206
P3_INLINE float2 var_waves3(float2 p, float w, float sx, float sy, float fx, float fy, float sxf, float syf) {
  float sxx = (float)0.5 * sx * ((float)1 + safe_sin(p.y * sxf));
  float syy = (float)0.5 * sy * ((float)1 + safe_sin(p.x * syf));
  return make_float2(w * (p.x + safe_sin(p.y * fx) * sxx), w * (p.y + safe_sin(p.x * fy) * syy));
}

// chaos.wgsl:5219 — per-row sin-hash amplitude.
P3_INLINE float2 var_waves4(float2 p, float w, float sx, float sy, float fx, float fy, float contp, float yfact) {
  float cell = floorf(p.y * fx / kTau);
  float ax = sinf(cell * (float)12.9898 + cell * (float)78.233 + (float)1 + p.y * (float)0.001 * yfact) * (float)43758.5453;
  ax = ax - truncf(ax);
  if (contp > (float)0.5) ax = sel((float)0, (float)1, ax > (float)0.5);
  return make_float2(w * (p.x + safe_sin(p.y * fx) * ax * ax * sx), w * (p.y + safe_sin(p.x * fy) * sy));
}

// chaos.wgsl:5237
P3_INLINE float2 var_scry2(float2 p, float w, float sides_f, float star, float circle) {
  const int32_t sides = imin(imax(f2i(sides_f), 1), 50);
  float a = kTau / (float)sides;
  float sa = safe_sin(a), ca = safe_cos(a);
  float sins = safe_sin(-star * kPi * (float)0.5), coss = safe_cos(star * kPi * (float)0.5);
  float sinc = safe_sin(circle * kPi * (float)0.5), cosc = safe_cos(circle * kPi * (float)0.5);
  float xr = p.x, yr = p.y;
  float r2 = xr * coss + fabsf(yr) * sins;
  float cr = sqrtf(xr * xr + yr * yr);
  int32_t iters = 0;
  for (int32_t i = 0; i < sides - 1; ++i) {
    float nx = xr * ca - yr * sa;
    yr = xr * sa + yr * ca;
    xr = nx;
    r2 = smax(r2, xr * coss + fabsf(yr) * sins);
    iters = i + 1;
  }
  r2 = r2 * cosc + cr * sinc;
  float r1 = r2;
  if (iters > 1) r2 = r2 * r2;
  else r2 = fabsf(r2) * r2;
  if (w == (float)0) return make_float2((float)0, (float)0);
  float d = r1 * (r2 + (float)1 / w);
  if (d == (float)0) return make_float2((float)0, (float)0);
  float r = (float)1 / d;
  return make_float2(p.x * r, p.y * r);
}

// chaos.wgsl:5281
P3_INLINE float2 var_ennepers2(float2 p, float w, float ap, float bp, float cp) {
  float x = p.x, y = p.y;
  float r2 = (float)1 / smax(x * x + y * y, (float)1e-30);
  float dxy = (ap * x) * (ap * x) - (bp * y) * (bp * y);
  return make_float2(w * x * (ap * ap - dxy * r2 - cp * sqrtf(fabsf(x))),
          w * y * (bp * bp - dxy * r2 - cp * sqrtf(fabsf(y))));
}

// chaos.wgsl:5294 — one draw picks one of three Moebius maps.
P3_INLINE float2 var_apollony(float2 p, float w, Rng& rng) {
  const float sq3 = (float)1.7320508075688772;
  float dx = (float)1 + sq3 - p.x;
  float den = smax(dx * dx + p.y * p.y, (float)1e-30);
  float a0 = (float)3 * dx / den - ((float)1 + sq3) / ((float)2 + sq3);
  float b0 = (float)3 * p.y / den;
  float ab = smax(a0 * a0 + b0 * b0, (float)1e-30);
  float fx = a0 / ab, fy = -b0 / ab;
  const int32_t br = imod(f2i(floorf((float)4 * rng.rand01())), 3);
  float x, y;
  if (br == 0) { x = a0; y = b0; }
  else if (br == 1) { x = -fx * (float)0.5 - fy * sq3 * (float)0.5; y = fx * sq3 * (float)0.5 - fy * (float)0.5; }
  else { x = -fx * (float)0.5 + fy * sq3 * (float)0.5; y = -fx * sq3 * (float)0.5 - fy * (float)0.5; }
  return make_float2(w * x, w * y);
}

// chaos.wgsl:5326 — one draw (scatter band), always taken.
P3_INLINE float2 var_circlecrop(float2 p, float w, float radius, float cx, float cy, float scatter, float zerop, Rng& rng) {
  float ca = clampv(scatter, (float)-1, (float)1);
  float dx = p.x - cx, dy = p.y - cy;
  float rad = sqrtf(dx * dx + dy * dy);
  float ang = atan2f(dy, dx);
  float rdc = radius + (rng.rand01() * (float)0.5 * ca);
  const bool esc = rad > radius, zero = zerop > (float)0.5;
  if (zero && esc) return make_float2((float)0, (float)0);
  if (!zero && esc) return make_float2(w * rdc * cosf(ang) + cx, w * rdc * sinf(ang) + cy);
  return make_float2(w * dx + cx, w * dy + cy);
}

// ---- conformal / escape-time maps (chaos.wgsl:5365-5520) ----------------------

// chaos.wgsl:5365 — z^k by binary exponentiation.
P3_INLINE float2 c_pow_int(float2 z, int32_t k) {
  float2 res{(float)1, (float)0}, base = z;
  for (int32_t e = k; e > 0; e >>= 1) {
    if ((e & 1) == 1) res = c_mul(res, base);
    base = c_sqr(base);
  }
  return res;
}
// chaos.wgsl:5383 — one Newton step for z^n = 1.
P3_INLINE float2 var_newton(float2 p, float w, float nin) {
  const int32_t n = imin(imax(f2i(nin + (float)0.5), 2), 8);
  float2 zn = c_pow_int(p, n), zm = c_pow_int(p, n - 1);
  float nf = (float)n, nm1 = (float)(n - 1);
  float2 num{nm1 * zn.x + (float)1, nm1 * zn.y};
  float2 den{nf * zm.x, nf * zm.y};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_div(num, den);
  return make_float2(w * q.x, w * q.y);
}
// chaos.wgsl:5438
P3_INLINE float2 var_burning_ship(float2 p, float w, float cx, float cy) {
  float2 s = c_sqr(make_float2(fabsf(p.x), fabsf(p.y)));
  return make_float2(w * (s.x + cx), w * (s.y + cy));
}
// chaos.wgsl:5444
P3_INLINE float2 var_magnet1(float2 p, float w, float cx, float cy) {
  float2 s = c_sqr(p);
  float2 num{s.x + cx - (float)1, s.y + cy - (float)0};
  float2 den{(float)2 * p.x + cx - (float)2, (float)2 * p.y + cy - (float)0};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_sqr(c_div(num, den));
  return make_float2(w * q.x, w * q.y);
}
// chaos.wgsl:5455 — relaxed Newton step for z^3 = 1 plus c.
P3_INLINE float2 var_nova(float2 p, float w, float cx, float cy, float relax) {
  float2 z3 = c_pow_int(p, 3), z2 = c_sqr(p);
  float2 num{z3.x - (float)1, z3.y - (float)0};
  float2 den{(float)3 * z2.x, (float)3 * z2.y};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_div(num, den);
  float sx = relax * q.x, sy = relax * q.y;
  return make_float2(w * (p.x - sx + cx), w * (p.y - sy + cy));
}
// chaos.wgsl:5467 — Halley step for z^3 = 1 plus c.
P3_INLINE float2 var_halley(float2 p, float w, float cx, float cy) {
  float2 z3 = c_pow_int(p, 3), z2 = c_sqr(p);
  float2 f{z3.x - (float)1, z3.y - (float)0};
  float2 fp{(float)3 * z2.x, (float)3 * z2.y};
  float2 fpp{(float)6 * p.x, (float)6 * p.y};
  float2 m = c_mul(f, fp);
  float2 num{(float)2 * m.x, (float)2 * m.y};
  float2 s = c_sqr(fp), m2 = c_mul(f, fpp);
  float2 den{(float)2 * s.x - m2.x, (float)2 * s.y - m2.y};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_div(num, den);
  return make_float2(w * (p.x - q.x + cx), w * (p.y - q.y + cy));
}
// chaos.wgsl:5506 — Blaschke factor z (z - a) / (1 - conj(a) z).
P3_INLINE float2 var_blaschke(float2 p, float w, float ax, float ay) {
  float2 num = c_mul(p, make_float2(p.x - ax, p.y - ay));
  float2 m = c_mul(make_float2(ax, -ay), p);
  float2 den{(float)1 - m.x, (float)0 - m.y};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_div(num, den);
  return make_float2(w * q.x, w * q.y);
}

// chaos.wgsl:5518 — Cayley map (z - is) / (z + is).
P3_INLINE float2 var_cayley(float2 p, float w, float s) {
  float2 den{p.x + (float)0, p.y + s};
  if (den.x * den.x + den.y * den.y < (float)1e-20) return make_float2(w * p.x, w * p.y);
  float2 q = c_div(make_float2(p.x - (float)0, p.y - s), den);
  return make_float2(w * q.x, w * q.y);
}

// chaos.wgsl:5532 — Lanczos (g = 7, 9 terms) complex Gamma with reflection.
P3_INLINE float2 var_complex_gamma(float2 p, float w, float scale) {
  const float sq2pi = (float)2.5066282746310002;
  const float kC[8] = {(float)(float)676.5203681218851,     (float)(float)-1259.1392167224028,
                   (float)(float)771.32342877765313,    (float)(float)-176.61502916214059,
                   (float)(float)12.507343278686905,    (float)(float)-0.13857109526572012,
                   (float)(float)9.9843695780195716e-6, (float)(float)1.5056327351493116e-7};
  float2 z = p;
  const bool refl = z.x < (float)0.5;
  if (refl) z = make_float2((float)1 - z.x, -z.y);
  float2 x{z.x - (float)1, z.y - (float)0};
  float2 A{(float)0.99999999999980993, (float)0};
  for (int i = 0; i < 8; ++i) {
    float2 d = c_div(make_float2(kC[i], (float)0), make_float2(x.x + (float)(i + 1), x.y + (float)0));
    A = make_float2(A.x + d.x, A.y + d.y);
  }
  float2 t{x.x + ((float)7 + (float)0.5), x.y + (float)0};
  float2 tp = c_pow(t, make_float2(x.x + (float)0.5, x.y + (float)0));
  float2 en = c_exp(make_float2(-t.x, -t.y));
  float2 m = c_mul(tp, c_mul(en, A));
  float2 res{sq2pi * m.x, sq2pi * m.y};
  if (refl) {
    float2 sp = c_sin(make_float2((float)3.14159265 * p.x, (float)3.14159265 * p.y));
    res = c_div(c_div(make_float2((float)3.14159265, (float)0), sp), res);
  }
  float k = w * scale;
  return make_float2(k * res.x, k * res.y);
}

// chaos.wgsl:5581 — Halley iterations for W e^W = z (1..4 steps).
P3_INLINE float2 var_lambert_w(float2 p, float w, float itin) {
  const int32_t iters = imin(imax(f2i(itin + (float)0.5), 1), 4);
  float2 wn;
  if (len2(p) < (float)2) {
    wn = c_log(make_float2((float)1 + p.x, (float)0 + p.y));
  } else {
    float2 lz = c_log(p), llz = c_log(lz);
    wn = make_float2(lz.x - llz.x, lz.y - llz.y);
  }
  for (int32_t i = 0; i < iters; ++i) {
    float2 ew = c_exp(wn);
    float2 we = c_mul(wn, ew);
    float2 f{we.x - p.x, we.y - p.y};
    float2 w1{wn.x + (float)1, wn.y + (float)0}, w2{wn.x + (float)2, wn.y + (float)0};
    float2 t1{(float)2 * w1.x, (float)2 * w1.y};
    if (t1.x * t1.x + t1.y * t1.y < (float)1e-20) break;
    float2 in = c_div(c_mul(w2, f), t1);
    float2 fp = c_mul(ew, w1);
    float2 den{fp.x - in.x, fp.y - in.y};
    if (den.x * den.x + den.y * den.y < (float)1e-20) break;
    float2 q = c_div(f, den);
    wn = make_float2(wn.x - q.x, wn.y - q.y);
  }
  return make_float2(w * wn.x, w * wn.y);
}

// ---- map projections (chaos.wgsl:5621-5669) -----------------------------------

// chaos.wgsl:5621
P3_INLINE float2 var_mercator(float2 p, float w) {
  float lat = clampv(p.y, (float)-1.5, (float)1.5);
  return make_float2(w * p.x, w * logf(fabsf(safe_tan((float)0.78539816 + lat * (float)0.5)) + (float)1e-6));
}
// chaos.wgsl:5628
P3_INLINE float2 var_lambert(float2 p, float w) {
  float k = sqrtf((float)2 / ((float)1 + safe_cos(p.y) * safe_cos(p.x) + (float)1e-6));
  return make_float2(w * (k * safe_cos(p.y) * safe_sin(p.x)), w * (k * safe_sin(p.y)));
}
// chaos.wgsl:5635 — four Newton steps for the auxiliary angle.
P3_INLINE float2 var_mollweide(float2 p, float w) {
  float t = p.y;
  float tgt = (float)3.14159265 * safe_sin(p.y);
  for (int i = 0; i < 4; ++i) {
    float s2 = safe_sin((float)2 * t), c2 = safe_cos((float)2 * t);
    float f = (float)2 * t + s2 - tgt;
    float df = (float)2 + (float)2 * c2;
    if (fabsf(df) < (float)1e-6) break;
    t = t - f / df;
  }
  return make_float2(w * ((float)0.9003163 * p.x * safe_cos(t)), w * ((float)1.4142135 * safe_sin(t)));
}
// chaos.wgsl:5652
P3_INLINE float2 var_hammer(float2 p, float w) {
  float z = sqrtf((float)1 + safe_cos(p.y) * safe_cos(p.x * (float)0.5) + (float)1e-6);
  return make_float2(w * (((float)2.828427 * safe_cos(p.y) * safe_sin(p.x * (float)0.5)) / z),
          w * (((float)1.4142135 * safe_sin(p.y)) / z));
}
// chaos.wgsl:5660
P3_INLINE float2 var_stereographic(float2 p, float w) {
  float k = (float)2 / ((float)1 + safe_cos(p.y) * safe_cos(p.x) + (float)1e-6);
  return make_float2(w * (k * safe_cos(p.y) * safe_sin(p.x)), w * (k * safe_sin(p.y)));
}

// ---- one-step attractor maps, folds, toral maps (chaos.wgsl:5671-5808) ----------

// chaos.wgsl:5671 — Chirikov standard map.
P3_INLINE float2 var_standard_map(float2 p, float w, float k) {
  float xp = p.x + k * safe_sin(p.y);
  return make_float2(w * xp, w * (p.y + xp));
}
// chaos.wgsl:5678
P3_INLINE float2 var_de_jong(float2 p, float w, float a, float b, float c, float d) {
  return make_float2(w * (safe_sin(a * p.y) - safe_cos(b * p.x)), w * (safe_sin(c * p.x) - safe_cos(d * p.y)));
}
// chaos.wgsl:5685
P3_INLINE float2 var_ikeda(float2 p, float w, float u) {
  float t = (float)0.4 - (float)6 / ((float)1 + p.x * p.x + p.y * p.y);
  float st = safe_sin(t), ct = safe_cos(t);
  return make_float2(w * ((float)1 + u * (p.x * ct - p.y * st)), w * (u * (p.x * st + p.y * ct)));
}
// chaos.wgsl:5699 — reflect each coordinate into [-limit, limit].
P3_INLINE float box_fold1(float v, float lim) {
  if (v > lim) return (float)2 * lim - v;
  if (v < -lim) return (float)-2 * lim - v;
  return v;
}
P3_INLINE float2 var_box_fold(float2 p, float w, float limit) {
  return make_float2(w * box_fold1(p.x, limit), w * box_fold1(p.y, limit));
}
// Mandelbox sphere-fold scale (chaos.wgsl:5710-5745).
P3_INLINE float sphere_scale(float r2, float rmin, float rmax) {
  float a = rmin * rmin, b = rmax * rmax;
  if (r2 < a) return b / a;
  if (r2 < b) return b / smax(r2, (float)1e-6);
  return (float)1;
}
// chaos.wgsl:5710
P3_INLINE float2 var_sphere_fold(float2 p, float w, float rmin, float rmax) {
  float sc = sphere_scale(p.x * p.x + p.y * p.y, rmin, rmax);
  return make_float2(w * (p.x * sc), w * (p.y * sc));
}
// chaos.wgsl:5724
P3_INLINE float2 var_mandelbox_step(float2 p, float w, float scale, float rmin, float rmax, float cx, float cy) {
  float x = box_fold1(p.x, (float)1), y = box_fold1(p.y, (float)1);
  float sf = sphere_scale(x * x + y * y, rmin, rmax);
  return make_float2(w * (x * sf * scale + cx), w * (y * sf * scale + cy));
}
// chaos.wgsl:5747 — mirror the angle into one wedge of n.
P3_INLINE float2 var_kifs_fold(float2 p, float w, float n, float offset) {
  float r = len2(p);
  if (r < (float)1e-6) return make_float2((float)0, (float)0);
  float a = atan2f(p.y, p.x) - offset;
  float th = (float)6.283185307179586 / smax((float)1, fabsf(n));
  a = a - th * floorf(a / th);
  if (a > th * (float)0.5) a = th - a;
  a = a + offset;
  return make_float2(w * (r * safe_cos(a)), w * (r * safe_sin(a)));
}
// chaos.wgsl:5765
P3_INLINE float2 var_arnold_cat(float2 p, float w) {
  float xp = p.x * (float)2 + p.y, yp = p.x + p.y;
  return make_float2(w * (fract(xp + (float)0.5) - (float)0.5), w * (fract(yp + (float)0.5) - (float)0.5));
}
// chaos.wgsl:5772
P3_INLINE float2 var_bakers_map(float2 p, float w) {
  float x = fract(p.x + (float)0.5), y = fract(p.y + (float)0.5);
  float xp, yp;
  if (x < (float)0.5) { xp = (float)2 * x; yp = y * (float)0.5; }
  else { xp = (float)2 * x - (float)1; yp = y * (float)0.5 + (float)0.5; }
  return make_float2(w * (xp - (float)0.5), w * (yp - (float)0.5));
}
// chaos.wgsl:5788
P3_INLINE float2 var_tent_map(float2 p, float w) {
  float x = fract(p.x + (float)0.5), y = fract(p.y + (float)0.5);
  float xp = (float)1 - fabsf((float)1 - (float)2 * x), yp = (float)1 - fabsf((float)1 - (float)2 * y);
  return make_float2(w * (xp - (float)0.5), w * (yp - (float)0.5));
}
// chaos.wgsl:5797
P3_INLINE float2 var_logistic_map(float2 p, float w, float r) {
  float x = fract(p.x + (float)0.5), y = fract(p.y + (float)0.5);
  return make_float2(w * (r * x * ((float)1 - x) - (float)0.5), w * (r * y * ((float)1 - y) - (float)0.5));
}

// ---- plane curves, flows, spirals, series (chaos.wgsl:5810-5969) ----------------

// Polar point (r, theta) scaled by w.
P3_INLINE float2 polar_w(float w, float r, float th) { return make_float2(w * (r * safe_cos(th)), w * (r * safe_sin(th))); }

// chaos.wgsl:5810
P3_INLINE float2 var_superellipse(float2 p, float w, float a, float b, float n) {
  float th = atan2f(p.y, p.x);
  float sa = smax(fabsf(a), (float)1e-4), sb = smax(fabsf(b), (float)1e-4);
  float sn = smax(n, (float)0.01);
  float c = fabsf(safe_cos(th) / sa), s = fabsf(safe_sin(th) / sb);
  float r = powf(smax(powf(c, sn) + powf(s, sn), (float)1e-6), (float)-1 / sn);
  return polar_w(w, r, th);
}
// chaos.wgsl:5822
P3_INLINE float2 var_limacon(float2 p, float w, float a, float b) {
  float th = atan2f(p.y, p.x);
  return polar_w(w, b + a * safe_cos(th), th);
}
// chaos.wgsl:5829
P3_INLINE float2 var_epicycloid(float2 p, float w, float k) {
  float t = atan2f(p.y, p.x);
  float k1 = k + (float)1;
  return make_float2(w * (k1 * safe_cos(t) - safe_cos(k1 * t)), w * (k1 * safe_sin(t) - safe_sin(k1 * t)));
}
// chaos.wgsl:5838
P3_INLINE float2 var_catenary(float2 p, float w, float a) {
  float sa = smax(fabsf(a), (float)1e-4) * signv(a + (float)1e-8);
  return make_float2(w * p.x, w * (sa * coshf(p.x / sa)));
}
// chaos.wgsl:5845
P3_INLINE float2 var_tractrix(float2 p, float w) {
  return make_float2(w * (p.x - safe_tanh(p.x)), w * ((float)1 / coshf(p.x)));
}
// chaos.wgsl:5852
P3_INLINE float2 var_tinkerbell(float2 p, float w, float a, float b, float c, float d) {
  return make_float2(w * (p.x * p.x - p.y * p.y + a * p.x + b * p.y), w * ((float)2 * p.x * p.y + c * p.x + d * p.y));
}
// chaos.wgsl:5859 — one Euler step of the forced Duffing oscillator.
P3_INLINE float2 var_duffing(float2 p, float w, float h, float delta, float gamma, float omega) {
  float xp = p.x + h * p.y;
  float yp = p.y + h * (p.x - p.x * p.x * p.x - delta * p.y + gamma * safe_cos(omega * p.x));
  return make_float2(w * xp, w * yp);
}
// chaos.wgsl:5868
P3_INLINE float2 var_vanderpol(float2 p, float w, float h, float mu) {
  return make_float2(w * (p.x + h * p.y), w * (p.y + h * (mu * ((float)1 - p.x * p.x) * p.y - p.x)));
}
// chaos.wgsl:5875 — Euler step with |p| as the third coordinate.
P3_INLINE float2 var_rossler(float2 p, float w, float h, float a) {
  float z = len2(p);
  return make_float2(w * (p.x + h * (-p.y - z)), w * (p.y + h * (p.x + a * p.y)));
}
// chaos.wgsl:5883
P3_INLINE float2 var_droste(float2 p, float w, float s) {
  float r = smax(len2(p), (float)1e-6);
  float th = atan2f(p.y, p.x);
  float k = logf(smax(s, (float)1e-4)) / (float)6.283185307179586;
  float re = logf(r) - k * th;
  float im = th + k * logf(r);
  float er = expf(re);
  return make_float2(w * (er * safe_cos(im)), w * (er * safe_sin(im)));
}
// chaos.wgsl:5896
P3_INLINE float2 var_logspiral(float2 p, float w, float a, float k) {
  float th = atan2f(p.y, p.x);
  return polar_w(w, a * expf(k * th), th);
}
// Angle in (0, 2pi], floored at 1e-6 (chaos.wgsl:5903-5925).
P3_INLINE float pos_angle(float2 p) {
  float t = atan2f(p.y, p.x);
  return smax(sel(t, t + (float)6.283185307179586, t < (float)0), (float)1e-6);
}
// chaos.wgsl:5903
P3_INLINE float2 var_fermat_spiral(float2 p, float w, float a) {
  float th = pos_angle(p);
  return polar_w(w, a * sqrtf(th), th);
}
// chaos.wgsl:5911
P3_INLINE float2 var_lituus(float2 p, float w, float a) {
  float th = pos_angle(p);
  return polar_w(w, a / sqrtf(th), th);
}
// chaos.wgsl:5919
P3_INLINE float2 var_hyperbolic_spiral(float2 p, float w, float a) {
  float th = pos_angle(p);
  return polar_w(w, a / th, th);
}
// chaos.wgsl:5927 — truncated Weierstrass cosine series.
P3_INLINE float2 var_weierstrass(float2 p, float w, float a, float b, float terms, float amp) {
  float wx = 0, wy = 0, ap = 1, bp = 1;
  const uint32_t n = f2u(clampv(terms, (float)1, (float)16));
  for (uint32_t i = 0; i < n; ++i) {
    wx += ap * safe_cos(bp * (float)3.1415926535 * p.x);
    wy += ap * safe_cos(bp * (float)3.1415926535 * p.y);
    ap *= a;
    bp *= b;
  }
  return make_float2(w * (p.x + amp * wx), w * (p.y + amp * wy));
}
// chaos.wgsl:5943 — truncated Takagi (blancmange) series.
P3_INLINE float2 var_takagi(float2 p, float w, float terms, float amp) {
  float tx = 0, ty = 0, p2 = 1;
  const uint32_t n = f2u(clampv(terms, (float)1, (float)16));
  for (uint32_t i = 0; i < n; ++i) {
    float xs = p2 * p.x, ys = p2 * p.y;
    tx += fabsf(xs - floorf(xs + (float)0.5)) / p2;
    ty += fabsf(ys - floorf(ys + (float)0.5)) / p2;
    p2 *= (float)2;
  }
  return make_float2(w * (p.x + amp * tx), w * (p.y + amp * ty));
}
// chaos.wgsl:5959
P3_INLINE float2 var_cantor_stairs(float2 p, float w, float terms, float amp) {
  float cx = p.x, cy = p.y;
  const uint32_t n = f2u(clampv(terms, (float)1, (float)8));
  for (uint32_t i = 0; i < n; ++i) {
    cx = (cx + safe_sin(cx * (float)6.2831853)) * (float)0.5;
    cy = (cy + safe_sin(cy * (float)6.2831853)) * (float)0.5;
  }
  return make_float2(w * (p.x + amp * cx), w * (p.y + amp * cy));
}

// ---- billiards (chaos.wgsl:5971-6327) -------------------------------------------

// Specular reflection of v about unit normal n.
P3_INLINE float2 reflect2(float vx, float vy, float nx, float ny) {
  float d = vx * nx + vy * ny;
  return make_float2(vx - (float)2 * d * nx, vy - (float)2 * d * ny);
}

// chaos.wgsl:5971 — one step of length `step` inside a circular table.
P3_INLINE float2 var_billiard_circle(float2 p, float w, float radius, float step, float angle) {
  float rc = smax(radius, (float)1.0e-5);
  float px = p.x, py = p.y;
  float r0 = len2(p);
  if (r0 > rc) { px = (px / r0) * rc; py = (py / r0) * rc; r0 = rc; }
  float bx = (float)1, by = (float)0;
  if (r0 > (float)1.0e-6) { bx = px / r0; by = py / r0; }
  float ca = safe_cos(angle), sa = safe_sin(angle);
  float vx = bx * ca - by * sa, vy = bx * sa + by * ca;
  float dpv = px * vx + py * vy;
  float disc = dpv * dpv + rc * rc - (px * px + py * py);
  float th = -dpv + sqrtf(smax(disc, (float)0));
  float fx = px + step * vx, fy = py + step * vy;
  if (th > (float)0 && th < step) {
    float hx = px + th * vx, hy = py + th * vy;
    float2 r = reflect2(vx, vy, -hx / rc, -hy / rc);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return make_float2(w * fx, w * fy);
}

// Exit time of the ray p + t v through the circle of radius R centred at
// (cx, 0); negative when the discriminant is (chaos.wgsl:6069-6100).
P3_INLINE bool circle_exit(float px, float py, float vx, float vy, float cx, float R, float& t) {
  float dx = px - cx;
  float dpv = dx * vx + py * vy;
  float disc = dpv * dpv + R * R - (dx * dx + py * py);
  if (!(disc >= (float)0)) return false;
  t = -dpv + sqrtf(disc);
  return true;
}

// chaos.wgsl:6014 — stadium: two flat walls and two semicircular caps.
P3_INLINE float2 var_billiard_stadium(float2 p, float w, float width, float height, float step, float angle) {
  float R = smax(height / (float)2, (float)1.0e-5), w2 = smax(width / (float)2, (float)1.0e-5);
  float px = p.x, py = p.y;
  if (fabsf(px) > w2) {
    float cx = sel(-w2, w2, px > (float)0);
    float dx = px - cx;
    float dist = sqrtf(dx * dx + py * py);
    if (dist > R) { px = cx + (dx / dist) * R; py = (py / dist) * R; }
  } else if (fabsf(py) > R) {
    py = sel(-R, R, py > (float)0);
  }
  float vx = safe_cos(angle), vy = safe_sin(angle);
  float th = (float)1.0e10;
  int wall = 0;  // 1 top, 2 bottom, 3 right cap, 4 left cap
  if (vy > (float)0) {
    float t = (R - py) / vy, x = px + t * vx;
    if (t >= (float)0 && x >= -w2 && x <= w2 && t < th) { th = t; wall = 1; }
  }
  if (vy < (float)0) {
    float t = (-R - py) / vy, x = px + t * vx;
    if (t >= (float)0 && x >= -w2 && x <= w2 && t < th) { th = t; wall = 2; }
  }
  float t;
  if (circle_exit(px, py, vx, vy, w2, R, t)) {
    float x = px + t * vx;
    if (t >= (float)0 && x >= w2 - (float)1.0e-5 && t < th) { th = t; wall = 3; }
  }
  if (circle_exit(px, py, vx, vy, -w2, R, t)) {
    float x = px + t * vx;
    if (t >= (float)0 && x <= -w2 + (float)1.0e-5 && t < th) { th = t; wall = 4; }
  }
  float fx = px + step * vx, fy = py + step * vy;
  if (th > (float)0 && th < step) {
    float hx = px + th * vx, hy = py + th * vy;
    float nx = 0, ny = 0;
    if (wall == 1) ny = (float)-1;
    else if (wall == 2) ny = (float)1;
    else if (wall == 3) {
      float rl = smax(len2(make_float2(hx - w2, hy)), (float)1.0e-5);
      nx = -(hx - w2) / rl;
      ny = -hy / rl;
    } else if (wall == 4) {
      float rl = smax(len2(make_float2(hx + w2, hy)), (float)1.0e-5);
      nx = -(hx + w2) / rl;
      ny = -hy / rl;
    }
    float2 r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return make_float2(w * fx, w * fy);
}

// chaos.wgsl:6121 — square table with a central circular scatterer.
P3_INLINE float2 var_billiard_sinai(float2 p, float w, float length, float radius, float step, float angle) {
  float L2 = smax(length / (float)2, (float)1.0e-5), R = smax(radius, (float)1.0e-5);
  float px = p.x, py = p.y;
  if (fabsf(px) > L2) px = sel(-L2, L2, px > (float)0);
  if (fabsf(py) > L2) py = sel(-L2, L2, py > (float)0);
  float r0 = sqrtf(px * px + py * py);
  if (r0 < R) {
    if (r0 > (float)1.0e-6) { px = (px / r0) * R; py = (py / r0) * R; }
    else {
      float th = (float)hash01(__float_as_uint((float)p.x) ^ __float_as_uint((float)p.y)) * (float)6.283185307179586;
      px = R * cosf(th);
      py = R * sinf(th);
    }
  }
  float vx = safe_cos(angle), vy = safe_sin(angle);
  float th = (float)1.0e10;
  int wall = 0;  // 1 left, 2 right, 3 bottom, 4 top, 5 scatterer
  if (vx < (float)0) {
    float t = (-L2 - px) / vx, y = py + t * vy;
    if (t >= (float)0 && y >= -L2 && y <= L2 && t < th) { th = t; wall = 1; }
  }
  if (vx > (float)0) {
    float t = (L2 - px) / vx, y = py + t * vy;
    if (t >= (float)0 && y >= -L2 && y <= L2 && t < th) { th = t; wall = 2; }
  }
  if (vy < (float)0) {
    float t = (-L2 - py) / vy, x = px + t * vx;
    if (t >= (float)0 && x >= -L2 && x <= L2 && t < th) { th = t; wall = 3; }
  }
  if (vy > (float)0) {
    float t = (L2 - py) / vy, x = px + t * vx;
    if (t >= (float)0 && x >= -L2 && x <= L2 && t < th) { th = t; wall = 4; }
  }
  {
    float dpv = px * vx + py * vy;
    float disc = dpv * dpv - ((px * px + py * py) - R * R);
    if (disc >= (float)0) {
      float t = -dpv - sqrtf(disc);
      if (t >= (float)0 && t < th) { th = t; wall = 5; }
    }
  }
  float fx = px + step * vx, fy = py + step * vy;
  if (th > (float)0 && th < step) {
    float hx = px + th * vx, hy = py + th * vy;
    float nx = 0, ny = 0;
    if (wall == 1) nx = (float)1;
    else if (wall == 2) nx = (float)-1;
    else if (wall == 3) ny = (float)1;
    else if (wall == 4) ny = (float)-1;
    else if (wall == 5) { nx = hx / R; ny = hy / R; }
    float2 r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return make_float2(w * fx, w * fy);
}

// chaos.wgsl:6227 — regular polygon table (3..12 sides).
P3_INLINE float2 var_billiard_polygon(float2 p, float w, float sides_f, float radius, float step, float angle) {
  const uint32_t n = f2u(clampv(sides_f + (float)0.5, (float)3, (float)12));
  const float two_pi = (float)(2.0 * 3.141592653589793);
  float R = smax(radius, (float)1.0e-5);
  float inr = R * cosf((float)3.141592653589793 / (float)n);
  float px = p.x, py = p.y;
  float r0 = sqrtf(px * px + py * py);
  if (r0 > inr) {
    float sc = inr / smax(r0, (float)1.0e-6);
    px = px * sc;
    py = py * sc;
  }
  float vx = safe_cos(angle), vy = safe_sin(angle);
  float th = (float)1.0e10;
  int32_t seg = -1;
  float ax = R * safe_cos((float)0), ay = R * safe_sin((float)0);
  for (uint32_t k = 0; k < n; ++k) {
    float ph = (two_pi * (float)(k + 1u)) / (float)n;
    float bx = R * safe_cos(ph), by = R * safe_sin(ph);
    float dx = bx - ax, dy = by - ay;
    float det = -vx * dy + vy * dx;
    if (fabsf(det) > (float)1.0e-6) {
      float t = (-(ax - px) * dy + (ay - py) * dx) / det;
      float u = (vx * (ay - py) - vy * (ax - px)) / det;
      if (t >= (float)0 && u >= (float)0 && u <= (float)1 && t < th) { th = t; seg = (int32_t)k; }
    }
    ax = bx;
    ay = by;
  }
  float fx = px + step * vx, fy = py + step * vy;
  if (th > (float)0 && th < step && seg != -1) {
    float hx = px + th * vx, hy = py + th * vy;
    float p0 = (two_pi * (float)seg) / (float)n, p1 = (two_pi * (float)(seg + 1)) / (float)n;
    float sx = R * safe_cos(p0), sy = R * safe_sin(p0);
    float ex_ = R * safe_cos(p1), ey_ = R * safe_sin(p1);
    float ex = ex_ - sx, ey = ey_ - sy;
    float el = smax(sqrtf(ex * ex + ey * ey), (float)1.0e-5);
    float nx = -ey / el, ny = ex / el;
    float mx = (float)0.5 * (sx + ex_), my = (float)0.5 * (sy + ey_);
    if (nx * mx + ny * my > (float)0) { nx = -nx; ny = -ny; }
    float2 r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return make_float2(w * fx, w * fy);
}

// ---- physics warps (chaos.wgsl:6329-6433) ---------------------------------------

// chaos.wgsl:6329 — hyperbolic rotation along a rotated axis.
P3_INLINE float2 var_lorentz_boost(float2 p, float w, float rapidity, float angle) {
  float c = safe_cos(angle), s = safe_sin(angle);
  float rx = p.x * c + p.y * s, ry = -p.x * s + p.y * c;
  float ch = coshf(rapidity), sh = sinhf(rapidity);
  float xp = rx * ch + ry * sh, yp = rx * sh + ry * ch;
  return make_float2(w * (xp * c - yp * s), w * (xp * s + yp * c));
}
// chaos.wgsl:6355
P3_INLINE float2 var_schwarzschild_lensing(float2 p, float w, float mass, float eps) {
  float b = len2(p) + smax(eps, (float)1e-6);
  float al = mass / b;
  float ca = safe_cos(al), sa = safe_sin(al);
  return make_float2(w * (p.x * ca - p.y * sa), w * (p.x * sa + p.y * ca));
}
// chaos.wgsl:6372 — Euler step along the field of a +/- charge pair.
P3_INLINE float2 var_field_dipole(float2 p, float w, float charge, float sep, float step, float angle) {
  float d = sep * (float)0.5;
  float dx = safe_cos(angle), dy = safe_sin(angle);
  float ppx = p.x - d * dx, ppy = p.y - d * dy;
  float npx = p.x - -d * dx, npy = p.y - -d * dy;
  float d3p = powf(ppx * ppx + ppy * ppy + (float)1e-4, (float)1.5);
  float d3n = powf(npx * npx + npy * npy + (float)1e-4, (float)1.5);
  float ex = charge * (ppx / d3p - npx / d3n), ey = charge * (ppy / d3p - npy / d3n);
  return make_float2(w * (p.x + step * ex), w * (p.y + step * ey));
}
// chaos.wgsl:6391 — pull toward N magnets on a circle, with damping.
P3_INLINE float2 var_magnetic_pendulum_pos(float2 p, float w, float magnets, float radius, float strength, float damping) {
  const uint32_t n = f2u(clampv(magnets, (float)3, (float)6));
  float inv = (float)1 / (float)n;
  float fx = 0, fy = 0;
  for (uint32_t i = 0; i < n; ++i) {
    float th = (float)6.283185307179586 * (float)i * inv;
    float rx = radius * safe_cos(th) - p.x, ry = radius * safe_sin(th) - p.y;
    float pw = powf(rx * rx + ry * ry + (float)1e-4, (float)1.5);
    fx = fx + rx / pw;
    fy = fy + ry / pw;
  }
  return make_float2(w * (p.x + strength * fx - damping * p.x), w * (p.y + strength * fy - damping * p.y));
}

// ---- modular forms (chaos.wgsl:462-522, 6435-6498) --------------------------------

P3_INLINE float2 upper_half(float2 p, float im_floor) {
  float im = fabsf(p.y) + smax(im_floor, (float)0.02);
  return make_float2(p.x, smax(im, (float)0.006435));
}
P3_INLINE float2 nome(float2 tau) { return c_exp(make_float2(-kPi * tau.y, kPi * tau.x)); }
// Jacobi theta series in the nome (8 terms; powers by recurrence).
P3_INLINE float2 theta3(float2 q) {
  float2 q2 = c_sqr(q), qn = q, odd = c_mul(q2, q), sum = q;
  for (int n = 2; n <= 8; ++n) {
    qn = c_mul(qn, odd);
    sum = make_float2(sum.x + qn.x, sum.y + qn.y);
    odd = c_mul(odd, q2);
  }
  return make_float2((float)1 + (float)2 * sum.x, (float)0 + (float)2 * sum.y);
}
P3_INLINE float2 theta2(float2 q) {
  float2 q2 = c_sqr(q), term{(float)1, (float)0}, ev = q2, sum = term;
  for (int n = 1; n <= 7; ++n) {
    term = c_mul(term, ev);
    sum = make_float2(sum.x + term.x, sum.y + term.y);
    ev = c_mul(ev, q2);
  }
  float2 m = c_mul(c_pow(q, make_float2((float)0.25, (float)0)), sum);
  return make_float2((float)2 * m.x, (float)2 * m.y);
}
// log(1 + |z|) z / |z| of z clamped to +/-1e18 per component.
P3_INLINE float2 log_compress(float2 z, float w) {
  float2 c{clampv(z.x, (float)-1e18, (float)1e18), clampv(z.y, (float)-1e18, (float)1e18)};
  float mag = len2(c);
  float sc = logf((float)1 + mag) / smax(mag, (float)1e-12);
  return make_float2(w * sc * c.x, w * sc * c.y);
}
// chaos.wgsl:6435
P3_INLINE float2 var_jacobi_theta(float2 p, float w, float imf) {
  float2 t = theta3(nome(upper_half(p, imf)));
  return make_float2(w * t.x, w * t.y);
}
// chaos.wgsl:6442 — modular lambda (theta2 / theta3)^4.
P3_INLINE float2 lambda_of(float2 p, float imf) {
  float2 q = nome(upper_half(p, imf));
  float2 r = c_div(theta2(q), theta3(q));
  return c_sqr(c_sqr(r));
}
P3_INLINE float2 var_modular_lambda(float2 p, float w, float imf) {
  float2 l = lambda_of(p, imf);
  return make_float2(w * l.x, w * l.y);
}
// chaos.wgsl:6453 — Klein j from lambda, log-compressed.
P3_INLINE float2 var_klein_j(float2 p, float w, float imf) {
  float2 lam = lambda_of(p, imf);
  float2 l2 = c_sqr(lam);
  float2 a{(float)1 - lam.x + l2.x, (float)0 - lam.y + l2.y};
  float2 a3 = c_pow_int(a, 3);
  float2 num{(float)256 * a3.x, (float)256 * a3.y};
  float2 oml{(float)1 - lam.x, (float)0 - lam.y};
  float2 den = c_mul(l2, c_sqr(oml));
  return log_compress(c_div(num, den), w);
}
// chaos.wgsl:6479 — Weierstrass P over a 5x5 lattice patch.
P3_INLINE float2 var_weierstrass_p(float2 p, float w, float o1r, float o1i, float o2r, float o2i) {
  float2 acc = c_recip(c_sqr(p));
  for (int m = -2; m <= 2; ++m)
    for (int n = -2; n <= 2; ++n) {
      if (m == 0 && n == 0) continue;
      float ox = (float)m * o1r + (float)n * o2r, oy = (float)m * o1i + (float)n * o2i;
      float2 a = c_recip(c_sqr(make_float2(p.x - ox, p.y - oy)));
      float2 b = c_recip(c_sqr(make_float2(ox, oy)));
      acc = make_float2(acc.x + a.x - b.x, acc.y + a.y - b.y);
    }
  return log_compress(acc, w);
}
// chaos.wgsl:6500 — fractional part of 1/t, |t| floored at 1e-4.
P3_INLINE float gauss_frac(float t) {
  float at = smax(fabsf(t), (float)1.0e-4);
  float r = sel((float)1, (float)-1, t < (float)0) / at;
  return r - floorf(r);
}
P3_INLINE float2 var_gauss_map(float2 p, float w) { return make_float2(w * gauss_frac(p.x), w * gauss_frac(p.y)); }

// ---- lattice value noise, special functions, digit codes (chaos.wgsl:6518-6870) ---

P3_INLINE float vnoise_corner(int32_t ix, int32_t iy) {
  uint32_t ux = (uint32_t)ix * 0x9e3779b1u, uy = (uint32_t)iy * 0x85ebca77u;
  return (float)hash01(ux ^ uy) * (float)2 - (float)1;
}
P3_INLINE float vnoise_smooth(float t) { return t * t * ((float)3 - (float)2 * t); }
// Value noise with x cells wrapped modulo `period` (period 0: no wrap).
P3_INLINE float value_noise2(float px, float py, int32_t period) {
  float fx = floorf(px), fy = floorf(py);
  const int32_t ix = f2i(fx), iy = f2i(fy);
  float ux = vnoise_smooth(px - fx), uy = vnoise_smooth(py - fy);
  int32_t i0 = ix, i1 = iwrap_add(ix, 1);
  if (period > 0) {
    i0 = imod(i0, period); if (i0 < 0) i0 += period;
    i1 = imod(i1, period); if (i1 < 0) i1 += period;
  }
  const int32_t j1 = iwrap_add(iy, 1);
  float bottom = mixv(vnoise_corner(i0, iy), vnoise_corner(i1, iy), ux);
  float top = mixv(vnoise_corner(i0, j1), vnoise_corner(i1, j1), ux);
  return mixv(bottom, top, uy);
}
// chaos.wgsl:6575 — fBm of angle-periodic value noise, polar/iso blend.
P3_INLINE float lich_field(float2 p, float freq, float branches, float radial, int32_t octaves) {
  float r = len2(p);
  float ang = atan2f(p.y, p.x) / kTau;
  float u = mixv(p.x * freq, (ang + (float)0.5) * branches, radial);
  float v = mixv(p.y * freq, r * freq, radial);
  int32_t period = imax(f2i(rint_even(branches)), 1);
  float sum = 0, amp = 1, tot = 0;
  for (int32_t k = 0; k < octaves; ++k) {
    sum = sum + amp * value_noise2(u, v, period);
    tot = tot + amp;
    amp = amp * (float)0.5;
    u = u * (float)2;
    v = v * (float)2;
    period = iwrap_mul(period, 2);
  }
  return sum / tot;
}
// chaos.wgsl:6600 — Abramowitz-Stegun 9.4.1 / 9.4.3 polynomial J0.
P3_INLINE float bessel_j0(float x) {
  float ax = fabsf(x);
  if (ax < (float)3) {
    float y = (x * x) / (float)9;
    return (float)1 + y * ((float)-2.2499997 + y * ((float)1.2656208 + y * ((float)-0.3163866 +
                  y * ((float)0.0444479 + y * ((float)-0.0039444 + y * (float)0.0002100)))));
  }
  float z = (float)3 / ax, y = z * z;
  float amp = (float)0.79788456 + y * ((float)-0.00000077 + y * ((float)-0.00552740 + y * ((float)-0.00009512 +
          y * ((float)0.00137237 + y * ((float)-0.00072805 + y * (float)0.00014476)))));
  float ph = ax - (float)0.78539816 + y * ((float)-0.04166397 + y * ((float)-0.00003954 +
         y * ((float)0.00262573 + y * ((float)-0.00054125 + y * ((float)-0.00029333 + y * (float)0.00013558)))));
  return (amp / sqrtf(ax)) * safe_cos(ph);
}
// chaos.wgsl:6617 — Airy Ai: asymptotic tails, Maclaurin series between.
P3_INLINE float airy_ai(float x) {
  if (x > (float)4) {
    float xi = (float)(2.0 / 3.0) * powf(x, (float)1.5);
    return expf(-xi) / ((float)(2.0 * 1.7724539) * powf(x, (float)0.25));
  }
  if (x < (float)-5) {
    float ax = -x;
    float xi = (float)(2.0 / 3.0) * powf(ax, (float)1.5);
    return safe_sin(xi + (float)0.78539816) / ((float)1.7724539 * powf(ax, (float)0.25));
  }
  const float c1 = (float)0.355028053887817, c2 = (float)0.258819403792807;
  float x3 = x * x * x;
  float f = 1, tf = 1, g = x, tg = x;
  for (int k = 1; k < 12; ++k) {
    float kf = (float)k;
    tf = tf * (x3 / (((float)3 * kf - (float)1) * ((float)3 * kf)));
    f = f + tf;
    tg = tg * (x3 / (((float)3 * kf) * ((float)3 * kf + (float)1)));
    g = g + tg;
  }
  return c1 * f - c2 * g;
}
// chaos.wgsl:6643 — Struve H1 power series (16 terms), x clamped to +/-10.
P3_INLINE float struve_h1(float xin) {
  float x = clampv(xin, (float)-10, (float)10);
  float h = (float)0.5 * x, h2 = h * h;
  float term = h2 / (float)(0.8862269 * 1.3293404);
  float sum = term;
  for (int m = 1; m < 16; ++m) {
    float mf = (float)m;
    term = term * (-(h2)) / ((mf + (float)0.5) * (mf + (float)1.5));
    sum = sum + term;
  }
  return sum;
}
// chaos.wgsl:6658-6675 — fixed-point cell index of a coordinate in
// [-extent, extent) (wrapped) and the cell centre of an index.
P3_INLINE uint32_t bits_mask(uint32_t bits) { return (1u << (bits & 31u)) - 1u; }
P3_INLINE uint32_t fp_encode(float c, float extent, uint32_t bits) {
  float s = smax(extent, (float)1.0e-4);
  float levels = (float)(float)(1u << (bits & 31u));
  float norm = (c + s) / ((float)2 * s);
  float folded = norm - floorf(norm);
  return umin(f2u(folded * levels), bits_mask(bits));
}
P3_INLINE float fp_decode(uint32_t i, float extent, uint32_t bits) {
  float s = smax(extent, (float)1.0e-4);
  float levels = (float)(float)(1u << (bits & 31u));
  float unit = ((float)(float)i + (float)0.5) / levels;
  return unit * (float)2 * s - s;
}
// chaos.wgsl:6677 — Chebyshev T_n by the three-term recurrence.
P3_INLINE float cheb_T(float x, int32_t n) {
  if (n <= 0) return (float)1;
  if (n == 1) return x;
  float a = 1, b = x, t = x;
  for (int32_t k = 2; k <= n; ++k) { t = (float)2 * x * b - a; a = b; b = t; }
  return t;
}
// chaos.wgsl:6690 — Legendre P_n (Bonnet recurrence).
P3_INLINE float legendre_P(float x, int32_t n) {
  if (n <= 0) return (float)1;
  if (n == 1) return x;
  float a = 1, b = x, t = x;
  for (int32_t k = 1; k < n; ++k) {
    float fk = (float)k;
    t = (((float)2 * fk + (float)1) * x * b - fk * a) / (fk + (float)1);
    a = b;
    b = t;
  }
  return t;
}
// chaos.wgsl:6706 — associated Legendre P_l^m (Condon-Shortley phase).
P3_INLINE float assoc_legendre(int32_t l, int32_t m, float u) {
  float pmm = 1;
  if (m > 0) {
    float s = sqrtf(smax((float)1 - u * u, (float)0));
    float fact = 1;
    for (int32_t i = 1; i <= m; ++i) { pmm = pmm * (-fact) * s; fact = fact + (float)2; }
  }
  if (l == m) return pmm;
  float pmm1 = u * ((float)2 * (float)m + (float)1) * pmm;
  if (l == m + 1) return pmm1;
  float pll = 0;
  for (int32_t ll = m + 2; ll <= l; ++ll) {
    float fl = (float)ll;
    pll = (((float)2 * fl - (float)1) * u * pmm1 - (fl + (float)m - (float)1) * pmm) / (fl - (float)m);
    pmm = pmm1;
    pmm1 = pll;
  }
  return pll;
}
// chaos.wgsl:6728 — |Y_l^m(theta)| without normalisation, capped at 4.
P3_INLINE float sph_harmonic(float th, int32_t l, int32_t m) {
  const int32_t mm = m < 0 ? -m : m;
  float y = assoc_legendre(l, mm, safe_cos(th)) * safe_cos((float)mm * th);
  return clampv(fabsf(y), (float)0, (float)4);
}
// chaos.wgsl:6742 — Chladni plate field F and its gradient.
P3_INLINE void chladni(float x, float y, float n, float m, float& F, float& Fx, float& Fy) {
  float an = kPi * n, am = kPi * m;
  float cnx = safe_cos(an * x), snx = safe_sin(an * x);
  float cmy = safe_cos(am * y), smy = safe_sin(am * y);
  float cmx = safe_cos(am * x), smx = safe_sin(am * x);
  float cny = safe_cos(an * y), sny = safe_sin(an * y);
  F = cnx * cmy - cmx * cny;
  Fx = -an * snx * cmy + am * smx * cny;
  Fy = -am * cnx * smy + an * cmx * sny;
}
// chaos.wgsl:6760 — generalised Laguerre L_k^alpha.
P3_INLINE float assoc_laguerre(int32_t k, float alpha, float x) {
  if (k <= 0) return (float)1;
  float a = 1, b = (float)1 + alpha - x;
  if (k == 1) return b;
  for (int32_t j = 1; j < k; ++j) {
    float fj = (float)j;
    float c = (((float)2 * fj + (float)1 + alpha - x) * b - (fj + alpha) * a) / (fj + (float)1);
    a = b;
    b = c;
  }
  return b;
}
// chaos.wgsl:6773 — hydrogen radial density rho^(2l) e^-rho (L^(2l+1)_(n-l-1))^2.
P3_INLINE float radial_psi2(int32_t n, int32_t l, float rho) {
  float rf = smax(rho, (float)0);
  float lag = assoc_laguerre(n - l - 1, (float)2 * (float)l + (float)1, rf);
  float rp = 1;
  if (l > 0)
    for (int32_t j = 0; j < 2 * l; ++j) rp = rp * rf;
  return rp * expf(-rf) * (lag * lag);
}
// chaos.wgsl:6796 — Hastings polynomial approximations of E(m), K(m).
P3_INLINE float elliptic_E(float m) {
  float m1 = clampv((float)1 - m, (float)0, (float)1);
  float l = logf((float)1 / smax(m1, (float)1.0e-6));
  float a = (float)1 + m1 * ((float)0.4630151 + m1 * (float)0.1077812);
  float b = m1 * ((float)0.2452727 + m1 * (float)0.0412496);
  return a + b * l;
}
P3_INLINE float elliptic_K(float m) {
  float m1 = clampv((float)1 - m, (float)1.0e-3, (float)1);
  float l = logf((float)1 / m1);
  float a = (float)1.3862944 + m1 * ((float)0.1119723 + m1 * (float)0.0725296);
  float b = (float)0.5 + m1 * ((float)0.1213478 + m1 * (float)0.0288729);
  return a + b * l;
}
// chaos.wgsl:6813 — Winitzki's closed-form inverse error function.
P3_INLINE float erfinv_w(float xin) {
  float x = clampv(xin, (float)-0.999999, (float)0.999999);
  const float a = (float)0.147;
  float l1 = logf((float)1 - x * x);
  float t1 = (float)2 / (kPi * a) + (float)0.5 * l1;
  float in = sqrtf(smax(t1 * t1 - l1 / a, (float)0)) - t1;
  return signv(x) * sqrtf(smax(in, (float)0));
}
// chaos.wgsl:6825-6870 — base-3 digit codes (Peano curve).
P3_INLINE uint32_t pow3u(uint32_t n) {
  uint32_t p = 1u;
  for (uint32_t i = 0; i < n; ++i) p = p * 3u;
  return p;
}
P3_INLINE uint32_t tri_encode(float c, float extent, uint32_t trits) {
  float s = smax(extent, (float)1.0e-4);
  const uint32_t cells = pow3u(trits);
  float norm = (c + s) / ((float)2 * s);
  float folded = norm - floorf(norm);
  return umin(f2u(folded * (float)(float)cells), cells - 1u);
}
P3_INLINE float tri_decode(uint32_t i, float extent, uint32_t trits) {
  float s = smax(extent, (float)1.0e-4);
  float levels = (float)(float)pow3u(trits);
  return (((float)(float)i + (float)0.5) / levels) * (float)2 * s - s;
}
P3_INLINE uint32_t peano_scramble(uint32_t idx, uint32_t trits) {
  bool flip = false;
  uint32_t out = 0u;
  for (uint32_t k = trits; k > 0u;) {
    --k;
    const uint32_t place = pow3u(k);
    const uint32_t d = (idx / place) % 3u;
    const uint32_t e = flip ? 2u - d : d;
    out = out + e * place;
    if ((e & 1u) == 1u) flip = !flip;
  }
  return out;
}

// ---- special-function radial profiles .. aperiodic (chaos.wgsl:6872-7194) -------

// chaos.wgsl:6872
P3_INLINE float2 var_bessel_j0(float2 p, float w, float freq) {
  float j = bessel_j0(freq * len2(p));
  return make_float2(w * (j * p.x), w * (j * p.y));
}
// chaos.wgsl:6879
P3_INLINE float2 var_airy_radial(float2 p, float w, float scale, float shift) {
  float a = airy_ai(scale * (len2(p) - shift));
  float k = (float)3 * a;
  return make_float2(w * (k * p.x), w * (k * p.y));
}
// chaos.wgsl:6886 — rational approximation of the Fresnel integrals.
P3_INLINE float2 var_cornu_spiral(float2 p, float w, float freq) {
  float t = freq * p.x;
  float s = fabsf(t);
  float sg = sel((float)1, (float)-1, t < (float)0);
  float f = ((float)1 + (float)0.926 * s) / ((float)2 + (float)1.792 * s + (float)3.104 * (s * s));
  float g = (float)1 / ((float)2 + (float)4.142 * s + (float)3.492 * (s * s) + (float)6.670 * (s * s * s));
  float a = (kPi * (s * s)) / (float)2;
  float sa = safe_sin(a), ca = safe_cos(a);
  float ci = (float)0.5 + f * sa - g * ca;
  float si = (float)0.5 - f * ca - g * sa;
  return make_float2(w * (sg * ci), w * (sg * si + (float)0.25 * p.y));
}
// chaos.wgsl:6902
P3_INLINE float2 var_struve_h1(float2 p, float w, float freq) {
  float k = (float)0.6 * struve_h1(freq * len2(p));
  return make_float2(w * (k * p.x), w * (k * p.y));
}
// chaos.wgsl:6911 — softened inverse-square pull toward mass m at c.
P3_INLINE float2 nbody_pull(float2 p, float cx, float cy, float m, float g, float eps) {
  float dx = cx - p.x, dy = cy - p.y;
  float r2 = (dx * dx + dy * dy) + eps;
  float inv = g * m / (r2 * sqrtf(r2));
  return make_float2(inv * dx, inv * dy);
}
// chaos.wgsl:6917
P3_INLINE float2 var_nbody_lensing(float2 p, float w, float c1x, float c1y, float c2x, float c2y, float m1, float m2, float g, float eps) {
  float e = smax(eps, (float)5.0e-3);
  float2 a = nbody_pull(p, c1x, c1y, m1, g, e), b = nbody_pull(p, c2x, c2y, m2, g, e);
  float dx = a.x + b.x, dy = a.y + b.y;
  return make_float2(w * (p.x + dx), w * (p.y + dy));
}
// chaos.wgsl:6926 — displacement along the curl of a value-noise stream function.
P3_INLINE float2 var_curl_noise(float2 p, float w, float freq, float amp) {
  float f = smax(freq, (float)1.0e-3);
  float sx = p.x * f, sy = p.y * f;
  const float h = (float)1.0e-2;
  float dx = (value_noise2(sx + h, sy, 0) - value_noise2(sx - h, sy, 0)) / ((float)2 * h);
  float dy = (value_noise2(sx, sy + h, 0) - value_noise2(sx, sy - h, 0)) / ((float)2 * h);
  float ux = clampv(dy, (float)-8, (float)8), uy = clampv(-dx, (float)-8, (float)8);
  return make_float2(w * (p.x + amp * ux), w * (p.y + amp * uy));
}
// chaos.wgsl:6944 — one Newton step toward the zero set of a noise field.
P3_INLINE float2 var_lichtenberg(float2 p, float w, float freqin, float brin, float radin, float detin, float strin) {
  const float H = (float)1.0e-2, R_EPS = (float)1.0e-4, G_EPS = (float)1.0e-3, MAX_STEP = (float)2;
  float freq = smax(freqin, (float)1.0e-3);
  float br = clampv(rint_even(brin), (float)1, (float)16);
  float rad = clampv(radin, (float)0, (float)1);
  const int32_t oct = f2i(clampv(rint_even(detin), (float)1, (float)4));
  float str = smax(strin, (float)0);
  if (len2(p) < R_EPS) return make_float2(w * p.x, w * p.y);
  float f0 = lich_field(p, freq, br, rad, oct);
  float fxp = lich_field(make_float2(p.x + H, p.y + (float)0), freq, br, rad, oct);
  float fxm = lich_field(make_float2(p.x - H, p.y - (float)0), freq, br, rad, oct);
  float fyp = lich_field(make_float2(p.x + (float)0, p.y + H), freq, br, rad, oct);
  float fym = lich_field(make_float2(p.x - (float)0, p.y - H), freq, br, rad, oct);
  float gx = (fxp - fxm) / ((float)2 * H), gy = (fyp - fym) / ((float)2 * H);
  float g2 = smax(gx * gx + gy * gy, G_EPS);
  float k = -(f0 / g2);
  float stx = k * gx, sty = k * gy;
  float sl = sqrtf(stx * stx + sty * sty);
  if (sl > MAX_STEP) { float c = MAX_STEP / sl; stx = stx * c; sty = sty * c; }
  return make_float2(w * (p.x + str * stx), w * (p.y + str * sty));
}
P3_INLINE float cop_sigmoid(float t) { return (float)1 / ((float)1 + expf(-t)); }
// chaos.wgsl:6981 — Gaussian copula: y replaced by its rho-correlated normal.
P3_INLINE float2 var_copula_gaussian(float2 p, float w, float strength, float rhoin) {
  float s = smax(strength, (float)1.0e-3);
  const float hi = (float)(1.0 - 1.0e-4);
  float u = clampv(cop_sigmoid(s * p.x), (float)1.0e-4, hi);
  float v = clampv(cop_sigmoid(s * p.y), (float)1.0e-4, hi);
  float zx = (float)1.4142135 * erfinv_w((float)2 * u - (float)1);
  float zy = (float)1.4142135 * erfinv_w((float)2 * v - (float)1);
  float rho = clampv(rhoin, (float)-0.95, (float)0.95);
  float zyp = rho * zx + sqrtf(smax((float)1 - rho * rho, (float)0)) * zy;
  return make_float2(w * p.x, w * (zyp / s));
}
// chaos.wgsl:6998 — Clayton copula conditional quantile.
P3_INLINE float2 var_copula_clayton(float2 p, float w, float strength, float thin) {
  float s = smax(strength, (float)1.0e-3);
  const float hi = (float)(1.0 - 1.0e-4);
  float u = clampv(cop_sigmoid(s * p.x), (float)1.0e-4, hi);
  float v = clampv(cop_sigmoid(s * p.y), (float)1.0e-4, hi);
  float th = clampv(thin, (float)0.05, (float)8);
  float ua = powf(u, -th), va = powf(v, -th);
  float base = smax(ua + va - (float)1, (float)1.0e-6);
  float h = powf(u, -th - (float)1) * powf(base, (float)-1 - (float)1 / th);
  float up = clampv(h, (float)1.0e-4, hi);
  float yp = clampv(logf(up / ((float)1 - up)), (float)-12, (float)12) / s;
  return make_float2(w * p.x, w * yp);
}
// chaos.wgsl:7021 — truncated series of the disc-to-polygon map.
P3_INLINE float2 var_schwarz_christoffel(float2 p, float w, float sides) {
  float n = smax(rint_even(sides), (float)3);
  const int32_t ni = f2i(n);
  float r = len2(p);
  float2 z = p;
  if (r > (float)0.999) { float k = (float)0.999 / smax(r, (float)1.0e-12); z = make_float2(p.x * k, p.y * k); }
  float a = (float)2 / n;
  float ax = 0, ay = 0, coef = 1;
  for (int32_t k = 0; k < 10; ++k) {
    const int32_t pi = iwrap_add(iwrap_mul(ni, k), 1);
    float2 zp = c_pow_int(z, pi);
    float c = coef / (float)pi;
    ax = ax + c * zp.x;
    ay = ay + c * zp.y;
    coef = coef * (a + (float)k) / (float)(k + 1);
  }
  return make_float2(w * ax, w * ay);
}
// chaos.wgsl:7047 — Doyle-style log-spiral shear with a ripple.
P3_INLINE float2 var_doyle(float2 p, float w, float dp, float dq) {
  float pp = smax(fabsf(dp), (float)1);
  float2 L = c_log(p);
  float u = L.x, v = L.y;
  float pitch = dq / pp;
  float us = u + pitch * v;
  float rip = (float)0.12 * (safe_sin(pp * v) + safe_cos(pp * v - kTau * us));
  float vs = v + pitch * rip;
  float2 e = c_exp(make_float2(us + rip, vs));
  return make_float2(w * e.x, w * e.y);
}
// chaos.wgsl:7070 — gradient of an n-fold sum of plane waves.
P3_INLINE float2 var_quasicrystal(float2 p, float w, float symin, float freqin) {
  const int32_t n = f2i(clampv(rint_even(symin), (float)2, (float)12));
  float freq = smax(freqin, (float)1.0e-3);
  float gx = 0, gy = 0;
  for (int32_t k = 0; k < n; ++k) {
    float th = kPi * (float)k / (float)n;
    float dx = cosf(th), dy = sinf(th);
    float d = freq * (p.x * dx + p.y * dy);
    float s = safe_sin(d);
    gx = gx - s * dx;
    gy = gy - s * dy;
  }
  gx = gx / (float)n;
  gy = gy / (float)n;
  return make_float2(w * (p.x + (float)0.4 * gx), w * (p.y + (float)0.4 * gy));
}
// chaos.wgsl:7093 — de Bruijn pentagrid cell vertex.
P3_INLINE float2 var_penrose(float2 p, float w, float scalein, float offset) {
  float scale = smax(scalein, (float)1.0e-3);
  float vx = 0, vy = 0;
  for (int j = 0; j < 5; ++j) {
    float th = kTau * (float)j / (float)5;
    float dx = cosf(th), dy = sinf(th);
    float kk = floorf(scale * (p.x * dx + p.y * dy) + offset);
    vx = vx + kk * dx;
    vy = vy + kk * dy;
  }
  return make_float2(w * vx / scale, w * vy / scale);
}
// chaos.wgsl:7115 — analytic interpolation of the Collatz map.
P3_INLINE float2 var_collatz(float2 p, float w, float scalein, float shift) {
  float scale = smax(fabsf(scalein), (float)1.0e-3);
  float2 z{p.x * scale + shift, p.y * scale};
  const float hp = kPi * (float)0.5;
  float2 arg{z.x * hp, z.y * hp};
  float2 sn = c_sin(arg), cs = c_sin(make_float2(arg.x + hp, arg.y + (float)0));
  float2 s2 = c_mul(sn, sn), c2 = c_mul(cs, cs);
  float2 ev = c_mul(make_float2((float)0.5 * z.x, (float)0.5 * z.y), c2);
  float2 od = c_mul(make_float2((float)0.5 * ((float)3 * z.x + (float)1), (float)0.5 * ((float)3 * z.y + (float)0)), s2);
  return make_float2(w * (ev.x + od.x), w * (ev.y + od.y));
}
// chaos.wgsl:7137 — digamma by recurrence to z + 6 and the asymptotic series.
P3_INLINE float2 var_digamma(float2 p, float w, float scalein, float shift) {
  float scale = smax(fabsf(scalein), (float)1.0e-3);
  float2 z{p.x * scale + shift, p.y * scale};
  float2 sum{(float)0, (float)0};
  for (int k = 0; k < 6; ++k) {
    float2 r = c_recip(z);
    sum = make_float2(sum.x + r.x, sum.y + r.y);
    z = make_float2(z.x + (float)1, z.y + (float)0);
  }
  float2 lz = c_log(z), zi = c_recip(z);
  float2 zi2 = c_mul(zi, zi), zi4 = c_mul(zi2, zi2);
  const float k12 = (float)(1.0 / 12.0), k120 = (float)(1.0 / 120.0);
  float px = lz.x - (float)0.5 * zi.x - k12 * zi2.x + k120 * zi4.x;
  float py = lz.y - (float)0.5 * zi.y - k12 * zi2.y + k120 * zi4.y;
  return make_float2(w * (px - sum.x), w * (py - sum.y));
}
// chaos.wgsl:7160 — Sprott quadratic map (ten coefficients).
P3_INLINE float2 var_sprott_poly(float2 p, float w, const float* c) {
  float x = p.x, y = p.y;
  float x2 = x * x, xy = x * y, y2 = y * y;
  return make_float2(w * (c[0] * x + c[1] * x2 + c[2] * xy + c[3] * y + c[4] * y2),
          w * (c[5] * x + c[6] * x2 + c[7] * xy + c[8] * y + c[9] * y2));
}
// chaos.wgsl:7175 — Martin's hopalong map.
P3_INLINE float2 var_hopalong(float2 p, float w, float a, float b, float c) {
  return make_float2(w * (p.y - signv(p.x) * sqrtf(fabsf(b * p.x - c))), w * (a - p.x));
}
// chaos.wgsl:7184
P3_INLINE float gm_G(float x, float a) { return a * x + (float)2 * ((float)1 - a) * x * x / ((float)1 + x * x); }
P3_INLINE float2 var_gumowski_mira(float2 p, float w, float a, float b) {
  float xp = b * p.y + gm_G(p.x, a);
  float yp = -p.x + gm_G(xp, a);
  return make_float2(w * xp, w * yp);
}

// ---- polynomials, digit codes, optics, waves, orbitals (chaos.wgsl:7196-7466) ----

// Integer order round(v) clamped to [lo, hi].
P3_INLINE int32_t order_of(float v, float lo, float hi) { return f2i(clampv(rint_even(v), lo, hi)); }
// chaos.wgsl:7196
P3_INLINE float2 var_chebyshev(float2 p, float w, float ox, float oy) {
  const int32_t nx = order_of(ox, (float)0, (float)12), ny = order_of(oy, (float)0, (float)12);
  return make_float2(w * cheb_T(clampv(p.x, (float)-1, (float)1), nx), w * cheb_T(clampv(p.y, (float)-1, (float)1), ny));
}
// chaos.wgsl:7204
P3_INLINE float2 var_legendre(float2 p, float w, float ox, float oy) {
  const int32_t nx = order_of(ox, (float)0, (float)12), ny = order_of(oy, (float)0, (float)12);
  return make_float2(w * legendre_P(clampv(p.x, (float)-1, (float)1), nx), w * legendre_P(clampv(p.y, (float)-1, (float)1), ny));
}
// Radial rescale of p to length rp (r > 0).
P3_INLINE float2 radial_to(float2 p, float w, float r, float rp) {
  float ir = (float)1 / r;
  return make_float2(w * (p.x * ir * rp), w * (p.y * ir * rp));
}
// chaos.wgsl:7214
P3_INLINE float2 var_spherical_harmonic(float2 p, float w, float dl, float om, float amount) {
  float r = len2(p);
  if (r < (float)1.0e-6) return make_float2((float)0, (float)0);
  float th = atan2f(p.y, p.x);
  const int32_t l = order_of(dl, (float)0, (float)6);
  const int32_t m = f2i(clampv(rint_even(om), (float)0, (float)l));
  float y = sph_harmonic(th, l, m);
  float amt = clampv(amount, (float)0, (float)2);
  float sc = ((float)1 - amt) + amt * y;
  return radial_to(p, w, r, r * clampv(sc, (float)-3, (float)3));
}
// chaos.wgsl:7230 — random-phase angular Fourier envelope.
P3_INLINE float2 var_fourier_warp(float2 p, float w, float harmonics, float amp, float seed) {
  float r = len2(p);
  if (r < (float)1.0e-6) return make_float2((float)0, (float)0);
  float th = atan2f(p.y, p.x);
  const int32_t n = order_of(harmonics, (float)1, (float)8);
  float amt = clampv(amp, (float)0, (float)0.9);
  const uint32_t sb = f2u(clampv(seed, (float)0, (float)1024));
  float acc = 0;
  for (int32_t k = 1; k <= n; ++k) {
    float fk = (float)k;
    float phi = (float)hash01(sb * 2654435761u + (uint32_t)k * 40503u) * kTau;
    acc = acc + safe_cos(fk * th + phi) / fk;
  }
  float env = (float)1 + amt * acc;
  return radial_to(p, w, r, r * clampv(env, (float)0.05, (float)3));
}
P3_INLINE uint32_t reverse_bits_n(uint32_t x, uint32_t bits) {
  uint32_t r = 0u;
  for (uint32_t k = 0; k < bits; ++k) { r = (r << 1) | (x & 1u); x >>= 1; }
  return r;
}
// chaos.wgsl:7264 — van der Corput digit reversal of the fixed-point cell.
P3_INLINE float2 var_radical_inverse(float2 p, float w, float extent, float bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (float)2, (float)24));
  uint32_t rx = reverse_bits_n(fp_encode(p.x, extent, bits), bits) & bits_mask(bits);
  uint32_t ry = reverse_bits_n(fp_encode(p.y, extent, bits), bits) & bits_mask(bits);
  return make_float2(w * fp_decode(rx, extent, bits), w * fp_decode(ry, extent, bits));
}
// chaos.wgsl:7276
P3_INLINE float2 var_gray_code(float2 p, float w, float extent, float bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (float)2, (float)24));
  const uint32_t m = bits_mask(bits);
  uint32_t ix = fp_encode(p.x, extent, bits), iy = fp_encode(p.y, extent, bits);
  uint32_t gx = (ix ^ (ix >> 1)) & m, gy = (iy ^ (iy >> 1)) & m;
  return make_float2(w * fp_decode(gx, extent, bits), w * fp_decode(gy, extent, bits));
}
P3_INLINE uint32_t part1by1(uint32_t x, uint32_t bits) {
  uint32_t r = 0u;
  for (uint32_t i = 0; i < bits; ++i) r = r | (((x >> (i & 31u)) & 1u) << ((i << 1) & 31u));
  return r;
}
// chaos.wgsl:7301 — interleave, then split back into two coordinates.
P3_INLINE float2 var_morton_zorder(float2 p, float w, float extent, float bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (float)2, (float)16));
  const uint32_t m = bits_mask(bits);
  uint32_t ix = fp_encode(p.x, extent, bits), iy = fp_encode(p.y, extent, bits);
  uint32_t code = part1by1(ix, bits) | (part1by1(iy, bits) << 1);
  uint32_t nx = code & m, ny = (code >> (bits & 31u)) & m;
  return make_float2(w * fp_decode(nx, extent, bits), w * fp_decode(ny, extent, bits));
}
// chaos.wgsl:7316 — refract the radial direction at a horizontal interface.
P3_INLINE float2 var_snell_refraction(float2 p, float w, float nratio, float strength) {
  float r = len2(p);
  if (r < (float)1.0e-6) return make_float2(w * p.x, w * p.y);
  float dx = p.x / r, dy = p.y / r;
  float nr = clampv(nratio, (float)0.05, (float)5);
  float s2 = nr * dx;
  float ox, oy;
  if (fabsf(s2) >= (float)1) { ox = dx; oy = -dy; }
  else {
    float c2 = sqrtf(smax((float)1 - s2 * s2, (float)0));
    ox = s2;
    oy = sel(-c2, c2, dy >= (float)0);
  }
  float s = clampv(strength, (float)0, (float)1);
  float bx = mixv(dx, ox, s), by = mixv(dy, oy, s);
  float bl = smax(sqrtf(bx * bx + by * by), (float)1.0e-6);
  bx = bx / bl;
  by = by / bl;
  return make_float2(w * r * bx, w * r * by);
}
// chaos.wgsl:7345
P3_INLINE float2 var_grin_lens(float2 p, float w, float focal, float eps) {
  float f = smax(fabsf(focal), (float)0.05), e = smax(fabsf(eps), (float)1.0e-3);
  float r = len2(p);
  if (r < (float)1.0e-6) return make_float2(w * p.x, w * p.y);
  float soft = (e * e) / (r * r + e * e);
  float pull = (r / f) * soft;
  float dx = p.x / r, dy = p.y / r;
  float nr = smax(r - pull, -r);
  return make_float2(w * nr * dx, w * nr * dy);
}
// chaos.wgsl:7359
P3_INLINE float2 var_caustic_fold(float2 p, float w, float freq, float amp, float phase) {
  float k = clampv(freq, (float)0, (float)16), a = clampv(amp, (float)0, (float)2);
  float gx = -safe_sin(k * p.x + phase), gy = -safe_sin(k * p.y + phase);
  return make_float2(w * clampv(p.x + a * gx, (float)-8, (float)8), w * clampv(p.y + a * gy, (float)-8, (float)8));
}
// chaos.wgsl:7377 — unit step down the Chladni field toward its nodal lines.
P3_INLINE float2 var_chladni(float2 p, float w, float nin, float min_, float step) {
  float n = clampv(floorf(nin + (float)0.5), (float)1, (float)12), m = clampv(floorf(min_ + (float)0.5), (float)1, (float)12);
  float F, gx, gy;
  chladni(p.x, p.y, n, m, F, gx, gy);
  float s = sel((float)1, (float)-1, F < (float)0);
  gx = s * gx;
  gy = s * gy;
  float inv = (float)1 / (sqrtf(gx * gx + gy * gy) + (float)1.0e-4);
  return make_float2(w * (p.x - step * gx * inv), w * (p.y - step * gy * inv));
}
// chaos.wgsl:7395 — up to four standing-wave modes (inactive ones gated by 0).
P3_INLINE float2 var_standing_wave(float2 p, float w, float modes, float freq, float amp, float decay) {
  float kmax = clampv(floorf(modes + (float)0.5), (float)1, (float)4);
  float d = clampv(decay, (float)0, (float)0.999);
  float dx = 0, dy = 0, ak = amp;
  for (int k = 1; k <= 4; ++k) {
    float kf = (float)k * kPi * freq;
    float gate = sel((float)0, (float)1, (float)k <= kmax);
    dx = dx + gate * ak * safe_sin(kf * p.x) * safe_sin(kf * p.y);
    dy = dy + gate * ak * safe_sin(kf * p.y) * safe_cos(kf * p.x);
    ak = ak * d;
  }
  return make_float2(w * (p.x + dx), w * (p.y + dy));
}
// chaos.wgsl:7413
P3_INLINE float2 var_moire(float2 p, float w, float freq, float beat, float angle, float amp) {
  float ca = safe_cos(angle), sa = safe_sin(angle);
  float u = p.x * ca + p.y * sa, v = -p.x * sa + p.y * ca;
  float f2 = freq + beat;
  float gx = safe_sin(freq * kPi * p.x) * safe_sin(f2 * kPi * u);
  float gy = safe_sin(freq * kPi * p.y) * safe_sin(f2 * kPi * v);
  return make_float2(w * (p.x + amp * gx), w * (p.y + amp * gy));
}
// Radius 1.5 a / (1 + a) along p's direction ((1, 0) at the origin).
P3_INLINE float2 shell_out(float2 p, float w, float r, float amp) {
  float ro = (float)1.5 * (amp / ((float)1 + amp));
  float dx = (float)1, dy = (float)0;
  if (r > (float)1.0e-6) { dx = p.x / r; dy = p.y / r; }
  return make_float2(w * ro * dx, w * ro * dy);
}
// chaos.wgsl:7428
P3_INLINE float2 var_radial_shell(float2 p, float w, float nin, float lin, float shell) {
  const int32_t n = imin(imax(f2i(rint_even(nin)), 1), 4);
  const int32_t l = imin(imax(f2i(rint_even(lin)), 0), n - 1);
  float r = len2(p);
  float rho = ((float)2 * r) / ((float)n * smax(shell, (float)1.0e-3));
  return shell_out(p, w, r, radial_psi2(n, l, rho));
}
// chaos.wgsl:7444
P3_INLINE float2 var_hydrogen_orbital(float2 p, float w, float nin, float lin, float min_, float shell, float lobe) {
  const int32_t n = imin(imax(f2i(rint_even(nin)), 1), 4);
  const int32_t l = imin(imax(f2i(rint_even(lin)), 0), n - 1);
  const int32_t m = imin(imax(f2i(rint_even(min_)), -l), l);
  float r = len2(p);
  float rho = ((float)2 * r) / ((float)n * smax(shell, (float)1.0e-3));
  float rad = radial_psi2(n, l, rho);
  float yl = sph_harmonic(atan2f(p.y, p.x), l, m);
  float amp = mixv(rad, rad * (yl * yl), clampv(lobe, (float)0, (float)1));
  return shell_out(p, w, r, amp);
}

// ---- distributions, wavelets, celestial, knots, follow-ons (chaos.wgsl:7467-7729) --

// u = r^2 / (1 + r^2) clamped to [1e-4, 1 - 1e-4].
P3_INLINE float cdf_u(float r2) {
  return clampv(r2 / ((float)1 + r2), (float)1.0e-4, (float)(1.0 - 1.0e-4));
}
// w * rp along p / |p| ((0, 0) when |p| < 1e-9).
P3_INLINE float2 along(float2 p, float w, float r, float rp) {
  if (r < (float)1.0e-9) return make_float2(w * rp * (float)0, w * rp * (float)0);
  float m = smax(r, (float)1.0e-12);
  return make_float2(w * rp * (p.x / m), w * rp * (p.y / m));
}
// chaos.wgsl:7467
P3_INLINE float2 var_weibull_cdf(float2 p, float w, float lambda, float k) {
  float r2 = dot2(p), r = sqrtf(r2);
  float u = cdf_u(r2);
  float kk = smax(fabsf(k), (float)0.05);
  return along(p, w, r, lambda * powf(-logf((float)1 - u), (float)1 / kk));
}
// chaos.wgsl:7477
P3_INLINE float2 var_logistic_cdf(float2 p, float w, float mu, float s) {
  float r2 = dot2(p), r = sqrtf(r2);
  float u = cdf_u(r2);
  float lg = clampv(logf(u / ((float)1 - u)), (float)-12, (float)12);
  return along(p, w, r, smax(mu + s * lg, (float)0));
}
// chaos.wgsl:7490
P3_INLINE float2 var_cauchy_cdf(float2 p, float w, float gamma, float tail) {
  float r2 = dot2(p), r = sqrtf(r2);
  float u = cdf_u(r2);
  float raw = gamma * safe_tan(kPi * (u - (float)0.5));
  float cap = smax(tail, (float)0.01);
  return along(p, w, r, clampv(raw, -cap, cap));
}
// chaos.wgsl:7502
P3_INLINE float2 var_pareto_cdf(float2 p, float w, float xm, float alpha, float tail) {
  float r2 = dot2(p), r = sqrtf(r2);
  float u = cdf_u(r2);
  float a = smax(fabsf(alpha), (float)0.05);
  float raw = xm * powf((float)1 - u, (float)-1 / a);
  return along(p, w, r, smin(raw, smax(tail, (float)0.01)));
}
// Radial scale (1 + amp psi) applied to p.
P3_INLINE float2 scaled(float2 p, float w, float sc) { return make_float2(w * sc * p.x, w * sc * p.y); }
// chaos.wgsl:7517
P3_INLINE float2 var_morlet(float2 p, float w, float freq, float sigma, float amp) {
  float r = len2(p);
  float t = r / smax(sigma, (float)0.05);
  float psi = expf((float)-0.5 * smin(t * t, (float)80)) * safe_cos(freq * r);
  return scaled(p, w, (float)1 + amp * psi);
}
// chaos.wgsl:7527
P3_INLINE float2 var_mexican_hat(float2 p, float w, float sigma, float amp) {
  float r = len2(p);
  float t = r / smax(sigma, (float)0.05);
  float t2 = smin(t * t, (float)80);
  float psi = ((float)1 - t2) * expf((float)-0.5 * t2);
  return scaled(p, w, (float)1 + amp * psi);
}
// chaos.wgsl:7537
P3_INLINE float2 var_chirp(float2 p, float w, float alpha, float amp, float decay) {
  float r2 = dot2(p);
  float psi = expf(-decay * smin(r2, (float)160)) * safe_sin(alpha * r2);
  return scaled(p, w, (float)1 + amp * psi);
}
// chaos.wgsl:7548 — two Newton steps of Kepler's equation.
P3_INLINE float2 var_kepler_orbit(float2 p, float w, float ein, float scale) {
  float e = clampv(ein, (float)0, (float)0.95);
  float M = atan2f(p.y, p.x);
  float a = scale * (len2(p) + (float)1.0e-6);
  float E = M;
  for (int i = 0; i < 2; ++i) {
    float s = safe_sin(E), c = safe_cos(E);
    E = E - (E - e * s - M) / smax((float)1 - e * c, (float)0.05);
  }
  float b = sqrtf(smax((float)1 - e * e, (float)0));
  return make_float2(w * (a * (safe_cos(E) - e)), w * (a * b * safe_sin(E)));
}
// chaos.wgsl:7567 — Euler step in the rotating-frame restricted 3-body field.
P3_INLINE float2 var_restricted_3body(float2 p, float w, float muin, float step, float coriolis) {
  float mu = clampv(muin, (float)0.01, (float)0.5);
  float r1x = p.x - -mu, r1y = p.y - (float)0;
  float r2x = p.x - ((float)1 - mu), r2y = p.y - (float)0;
  float d1 = powf(r1x * r1x + r1y * r1y + (float)1.0e-3, (float)1.5);
  float d2 = powf(r2x * r2x + r2y * r2y + (float)1.0e-3, (float)1.5);
  float gx = p.x - ((float)1 - mu) * (r1x / d1) - mu * (r2x / d2);
  float gy = p.y - ((float)1 - mu) * (r1y / d1) - mu * (r2y / d2);
  float dx = step * (gx + coriolis * -gy), dy = step * (gy + coriolis * gx);
  float dl = sqrtf(dx * dx + dy * dy);
  if (dl > (float)2) { float c = (float)2 / dl; dx = dx * c; dy = dy * c; }
  return make_float2(w * (p.x + dx), w * (p.y + dy));
}
// chaos.wgsl:7585 — Hill's epicyclic solution advanced by `phase`.
P3_INLINE float2 var_hill_epicyclic(float2 p, float w, float kappa, float phase, float shear) {
  float k = clampv(kappa, (float)0.2, (float)2);
  float phi = phase * k;
  float cf = safe_cos(phi), sf = safe_sin(phi);
  float xo = p.x * cf - (sf / k) * p.y;
  float yo = ((float)2 * k) * sf * p.x + p.y * cf;
  return make_float2(w * xo, w * (yo + shear * p.x));
}
// chaos.wgsl:7598
P3_INLINE float2 var_torus_knot(float2 p, float w, float pp, float qq, float radius, float tube) {
  float phi = atan2f(p.y, p.x);
  float pf = smax(rint_even(pp), (float)1), qf = smax(rint_even(qq), (float)1);
  float amp = tube * clampv(len2(p), (float)0, (float)1);
  float rr = radius + amp * safe_cos(qf * phi);
  return make_float2(w * (rr * safe_cos(pf * phi)), w * (rr * safe_sin(pf * phi)));
}
// chaos.wgsl:7612 — alternate angular lanes twisted in opposite senses.
P3_INLINE float2 var_braid_warp(float2 p, float w, float strands, float twist, float crossings) {
  float n = smax(rint_even(strands), (float)2), cr = smax(rint_even(crossings), (float)1);
  float r = len2(p);
  float th = atan2f(p.y, p.x);
  float a = (th + kPi) / kTau;
  float lane = floorf(a * n);
  float frac = a * n - lane;
  float sg = sel((float)1, (float)-1, (f2i(lane) & 1) == 1);
  float weave = safe_sin(kPi * frac);
  float dth = twist * sg * weave * safe_sin(cr * kPi * frac) * clampv(r, (float)0, (float)1);
  float nt = th + dth;
  return make_float2(w * r * safe_cos(nt), w * r * safe_sin(nt));
}
// chaos.wgsl:7638
P3_INLINE float2 var_airy_caustic(float2 p, float w, float scale, float r0, float amp) {
  float a = airy_ai(scale * (len2(p) - r0));
  float env = clampv((float)1 + amp * a, (float)0.05, (float)3);
  return make_float2(w * (env * p.x), w * (env * p.y));
}
// chaos.wgsl:7650
P3_INLINE float2 var_elliptic_E(float2 p, float w, float scale) {
  float r2 = dot2(p), r = sqrtf(r2);
  float e = elliptic_E(r2 / ((float)1 + r2));
  float k = scale * e;
  if (r < (float)1.0e-9) return make_float2(w * k * (float)0, w * k * (float)0);
  float m = smax(r, (float)1.0e-12);
  return make_float2(w * k * (p.x / m), w * k * (p.y / m));
}
// chaos.wgsl:7661
P3_INLINE float2 var_elliptic_K(float2 p, float w, float scale, float tail) {
  float r2 = dot2(p), r = sqrtf(r2);
  float k = elliptic_K(r2 / ((float)1 + r2));
  return along(p, w, r, smin(scale * k, smax(tail, (float)0.01)));
}
// chaos.wgsl:7676
P3_INLINE float2 var_gaussian_cdf(float2 p, float w, float mu, float sigma) {
  float r2 = dot2(p), r = sqrtf(r2);
  float q = erfinv_w((float)2 * cdf_u(r2) - (float)1);
  return along(p, w, r, smax(mu + sigma * (float)1.4142135 * q, (float)0));
}
// chaos.wgsl:7689
P3_INLINE float2 var_levy_cdf(float2 p, float w, float c, float tail) {
  float r2 = dot2(p), r = sqrtf(r2);
  float e = erfinv_w((float)1 - cdf_u(r2));
  float raw = smax(c, (float)1.0e-4) / smax((float)2 * e * e, (float)1.0e-6);
  return along(p, w, r, smin(raw, smax(tail, (float)0.01)));
}
// chaos.wgsl:7707
P3_INLINE float2 var_peano(float2 p, float w, float extent, float trits_f) {
  const uint32_t t = f2u(clampv(trits_f, (float)2, (float)15));
  uint32_t sx = peano_scramble(tri_encode(p.x, extent, t), t);
  uint32_t sy = peano_scramble(tri_encode(p.y, extent, t), t);
  return make_float2(w * tri_decode(sx, extent, t), w * tri_decode(sy, extent, t));
}


// ---- direct-colour side (chaos.wgsl:1682-1780, 5408, 5484, 6410) --------------

// chaos.wgsl:1696 — HSL to RGB (h wrapped, s and l clamped).
P3_INLINE float3 hsl_rgb(float h_in, float s_in, float l_in) {
  float h = fract(h_in);
  float s = clampv(s_in, 0.0f, 1.0f), l = clampv(l_in, 0.0f, 1.0f);
  float c = (1.0f - fabsf(2.0f * l - 1.0f)) * s;
  float h6 = h * 6.0f;
  float x = c * (1.0f - fabsf((h6 - 2.0f * floorf(h6 * 0.5f)) - 1.0f));
  float m = l - c * 0.5f;
  float r, g, b;
  if (h6 < 1.0f) { r = c; g = x; b = 0.0f; }
  else if (h6 < 2.0f) { r = x; g = c; b = 0.0f; }
  else if (h6 < 3.0f) { r = 0.0f; g = c; b = x; }
  else if (h6 < 4.0f) { r = 0.0f; g = x; b = c; }
  else if (h6 < 5.0f) { r = x; g = 0.0f; b = c; }
  else { r = c; g = 0.0f; b = x; }
  return make_float3(r + m, g + m, b + m);
}
// chaos.wgsl:1682
P3_INLINE float3 dc_linear_color(float2 p) {
  return make_float3(clampv(0.5f + 0.5f * p.x, 0.0f, 1.0f), clampv(0.5f + 0.5f * p.y, 0.0f, 1.0f),
                     clampv(0.5f - 0.25f * (p.x + p.y), 0.0f, 1.0f));
}
// chaos.wgsl:1721
P3_INLINE float3 dc_perlin_color(float2 p, float scale, float octaves, float seed) {
  float n = perlin_fbm(p.x, p.y, octaves, smax(scale, 1e-6f));
  return hsl_rgb(fract(0.5f + 0.5f * n + seed), 1.0f, 0.55f);
}
// chaos.wgsl:1732 — hashed colour per grid cell.
P3_INLINE float3 dc_gridout_color(float2 p, float cells) {
  float n = smax(cells, 1.0f);
  const int32_t cx = f2i(floorf(p.x * n)), cy = f2i(floorf(p.y * n));
  const uint32_t mixed = (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u)) + 0x9e3779b9u;
  return make_float3(hash01(mixed), hash01(mixed ^ 0xdeadbeefu), hash01(mixed ^ 0x13579bdfu));
}
// chaos.wgsl:1756
P3_INLINE float3 dc_cylinder_color(float2 p) {
  float hue = fract(0.5f + 0.5f * safe_sin(p.x));
  float lit = clampv(0.5f + 0.25f * safe_tanh(p.y * 0.5f), 0.2f, 0.85f);
  return hsl_rgb(hue, 0.9f, lit);
}
// chaos.wgsl:5408 — hue of the root of unity nearest the Newton image.
P3_INLINE float3 newton_color(float2 p, float nin) {
  const int32_t n = imin(imax(f2i(nin + 0.5f), 2), 8);
  float2 z = var_newton(p, 1.0f, (float)n);
  int32_t best = 0;
  float bd = 1.0e30f;
  const float step = 6.2831853f / (float)n;
  for (int32_t k = 0; k < n; ++k) {
    float ang = step * (float)k;
    float dx = z.x - cosf(ang), dy = z.y - sinf(ang);
    float d2 = dx * dx + dy * dy;
    if (d2 < bd) { bd = d2; best = k; }
  }
  return hsl_rgb((float)best / (float)n, 1.0f, 0.55f);
}
// chaos.wgsl:6410 — hue of the nearest magnet.
P3_INLINE float3 magnet_color(float2 p, float magnets, float radius) {
  const uint32_t n = f2u(clampv(magnets, 3.0f, 6.0f));
  float inv = 1.0f / (float)n;
  float bd = 1e30f;
  uint32_t best = 0u;
  for (uint32_t i = 0; i < n; ++i) {
    float th = 6.283185307179586f * (float)i * inv;
    float rx = radius * safe_cos(th) - p.x, ry = radius * safe_sin(th) - p.y;
    float d2 = rx * rx + ry * ry;
    if (d2 < bd) { bd = d2; best = i; }
  }
  return hsl_rgb((float)best * inv, 1.0f, 0.55f);
}
// chaos.wgsl:5484 — smooth escape-time band of up to 12 iterations.
P3_INLINE float3 escape_color(float2 p, int idx, float cx, float cy, float relax) {
  float2 z = p;
  int32_t it = 0;
  for (int32_t i = 0; i < 12; ++i) {
    if (idx == 310) z = var_burning_ship(z, 1.0f, cx, cy);
    else if (idx == 311) z = var_magnet1(z, 1.0f, cx, cy);
    else if (idx == 312) z = var_nova(z, 1.0f, cx, cy, relax);
    else z = var_halley(z, 1.0f, cx, cy);
    it = i + 1;
    if (z.x * z.x + z.y * z.y > 256.0f) break;
  }
  float r = smax(len2(z), 1.0001f);
  float mu = (float)it + 1.0f - log2f(smax(logf(r), 1.0e-6f));
  return hsl_rgb(fract(mu * 0.12f), 1.0f, 0.55f);
}

// ---- dispatch, catalog indices 99..325 (chaos.wgsl:7848-8106) --------------------
// Entries 99..101 are colour-only (no position contribution).
P3_INLINE float2 apply_variation_ext(int idx, float2 p, xrec_t xf, int k, float w, Rng& rng) {
#define P(i) var_param(xf, k, i)
  switch (idx) {
    case 99: case 100: case 101: return make_float2(0.0f, 0.0f);
    case 102: return var_dc_cylinder_pos(p, w);
    case 103: return var_cpow2(p, w, P(0), P(1), P(2), P(3), rng);
    case 104: return var_cpow3(p, w, P(0), P(1), P(2), P(3), rng);
    case 105: return var_loonie2(p, w, P(0), P(1), P(2));
    case 106: return var_epispiral(p, w, P(0), P(1), P(2), rng);
    case 107: return var_bwraps(p, w, P(0), P(1), P(2), P(3), P(4));
    case 108: return var_crackle(p, w, P(0), P(1), P(2), P(3), rng);
    case 109: return var_juliaq(p, w, P(0), P(1), rng);
    case 110: return var_glynnia(p, w, rng);
    case 111: return var_loonie3(p, w);
    case 112: return var_falloff(p, w, P(0), P(1), P(2), P(3), P(4), P(5), rng);
    case 113: return var_falloff2(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), rng);
    case 114: return var_falloff3(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), rng);
    case 115: return var_collideoscope(p, w, P(0), P(1));
    case 116: return var_circlize(p, w, P(0));
    case 117: return var_circlize2(p, w, P(0));
    case 118: return var_eswirl(p, w, P(0), P(1));
    case 119: return var_petal(p, w);
    case 120: return var_bcircle(p, w, P(0), P(1), rng);
    case 121: return var_curl2(p, w, P(0), P(1), P(2));
    case 122: return var_murl(p, w, P(0), P(1));
    case 123: return var_stwins(p, w, P(0));
    case 124: return var_hexes(p, w, P(0), P(1), P(2), P(3));
    case 125: return var_xheart(p, w, P(0), P(1));
    case 126: return var_xhyperbol(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 127: return var_xcurl2(p, w, P(0), P(1), P(2));
    case 128: return var_xtrb(p, w, P(0), P(1), P(2), P(3), P(4), P(5), rng);
    case 129: return var_xyrus_gridout(p, w);
    case 130: return var_blur_circle(p, w, P(0), rng);
    case 131: return var_bipolar2(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8));
    case 132: return var_bubble2(p, w, P(0), P(1));
    case 133: return var_acosh(p, w, rng);
    case 134: return var_arcsinh(p, w);
    case 135: return var_arctanh(p, w);
    case 136: return var_acoth(p, w);
    case 137: return var_acosech(p, w, rng);
    case 138: return var_arcsech2(p, w);
    case 139: return var_cell2(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 140: return var_curl_sp(p, w, P(0), P(1), P(2), P(3), P(4));
    case 141: return var_murl2(p, w, P(0), P(1));
    case 142: return var_lissajous(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), rng);
    case 143: return var_spirograph(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8), rng);
    case 144: return var_waffle(p, w, P(0), P(1), P(2), P(3), rng);
    case 145: return var_glynnSim1(p, w, P(0), P(1), P(2), P(3), P(4), P(5), rng);
    case 146: return var_glynnSim2(p, w, P(0), P(1), P(2), P(3), P(4), P(5), rng);
    case 147: return var_glynnSim3(p, w, P(0), P(1), P(2), P(3), rng);
    case 148: return var_flipy(p, w);
    case 149: return var_eclipse(p, w, P(0));
    case 150: return var_barycentroid(p, w, P(0), P(1), P(2), P(3));
    case 151: return var_chunk(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6));
    case 152: return var_ennepers(p, w);
    case 153: return var_erf(p, w);
    case 154: return var_circus(p, w, P(0));
    case 155: return var_asteria(p, w, P(0), rng);
    case 156: return var_clifford_js(p, w, P(0), P(1), P(2), P(3));
    case 157: return var_devil_warp(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 158: return var_voron(p, w, P(0), P(1), P(2), P(3), P(4));
    case 159: return var_henon(p, w, P(0), P(1), P(2));
    case 160: return var_atan(p, w, P(0), P(1));
    case 161: return var_cardioid(p, w, P(0));
    case 162: return var_chrysanthemum(p, w, rng);
    case 163: return var_bcollide(p, w, P(0), P(1));
    case 164: return var_bsplit(p, w, P(0), P(1));
    case 165: return var_bulge(p, w, P(0));
    case 166: return var_checks(p, w, P(0), P(1), P(2), P(3), rng);
    case 167: return var_circular(p, w, P(0), P(1), rng);
    case 168: return var_circular2(p, w, P(0), P(1), P(2), P(3), rng);
    case 169: return var_corners(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8));
    case 170: return var_circleblur(p, w, rng);
    case 171: return var_fibonacci2(p, w, P(0), P(1));
    case 172: return var_hypertile(p, w, P(0), P(1), P(2));
    case 173: return var_hypertile1(p, w, P(0), P(1), rng);
    case 174: return var_hypertile2(p, w, P(0), P(1), rng);
    case 175: return var_idisc(p, w);
    case 176: return var_hole(p, w, P(0), P(1));
    case 177: return var_kaleidoscope(p, w, P(0), P(1), P(2), P(3), P(4));
    case 178: return var_layered_spiral(p, w, P(0));
    case 179: return var_linear_t(p, w, P(0), P(1));
    case 180: return var_line(p, w, P(0), P(1), rng);
    case 181: return var_ovoid(p, w, P(0), P(1));
    case 182: return var_phoenix_julia(p, w, P(0), P(1), P(2), P(3), rng);
    case 183: return var_unpolar(p, w);
    case 184: return var_shredrad(p, w, P(0), P(1));
    case 185: return var_vogel(p, w, P(0), P(1), rng);
    case 186: return var_yin_yang(p, w, P(0), P(1), P(2), P(3), P(4), rng);
    case 187: return var_squish(p, w, P(0), rng);
    case 188: return var_target(p, w, P(0), P(1), P(2));
    case 189: return var_funnel(p, w, P(0));
    case 190: return var_holesq(p, w);
    case 191: return var_hole2(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 192: return var_lace_js(p, w, rng);
    case 193: return var_julia_outside(p, w, P(0), P(1), P(2), rng);
    case 194: return var_fourth(p, w, P(0), P(1), P(2), P(3), P(4));
    case 195: return var_pulse(p, w, P(0), P(1), P(2), P(3));
    case 196: return var_rays1(p, w);
    case 197: return var_rays2(p, w);
    case 198: return var_rays3(p, w);
    case 199: return var_tancos(p, w);
    case 200: return var_twoface(p, w);
    case 201: return var_e_julia(p, w, P(0), rng);
    case 202: return var_cannabis_curve_wf(p, w, P(0), rng);
    case 203: return var_e_collide(p, w, P(0), P(1));
    case 204: return var_e_mod(p, w, P(0), P(1));
    case 205: return var_intersection(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8), P(9), rng);
    case 206: return var_inv_squircular(p, w);
    case 207: return var_lozi(p, w, P(0), P(1), P(2));
    case 208: return var_hypershift(p, w, P(0), P(1));
    case 209: return var_hex_modulus(p, w, P(0));
    case 210: return var_boarders2(p, w, P(0), P(1), P(2), rng);
    case 211: return var_b_mod(p, w, P(0), P(1));
    case 212: return var_b_transform(p, w, P(0), P(1), P(2), P(3), rng);
    case 213: return var_parallel(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8), P(9), rng);
    case 214: return var_waves3(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 215: return var_waves4(p, w, P(0), P(1), P(2), P(3), P(4), P(5));
    case 216: return var_scry2(p, w, P(0), P(1), P(2));
    case 217: return var_ennepers2(p, w, P(0), P(1), P(2));
    case 218: return var_apollony(p, w, rng);
    case 219: return var_circlecrop(p, w, P(0), P(1), P(2), P(3), P(4), rng);
    case 220: return var_newton(p, w, P(0));
    case 221: return var_blaschke(p, w, P(0), P(1));
    case 222: return var_cayley(p, w, P(0));
    case 223: return var_complex_gamma(p, w, P(0));
    case 224: return var_lambert_w(p, w, P(0));
    case 225: return var_mercator(p, w);
    case 226: return var_lambert(p, w);
    case 227: return var_mollweide(p, w);
    case 228: return var_hammer(p, w);
    case 229: return var_stereographic(p, w);
    case 230: return var_standard_map(p, w, P(0));
    case 231: return var_de_jong(p, w, P(0), P(1), P(2), P(3));
    case 232: return var_ikeda(p, w, P(0));
    case 233: return var_box_fold(p, w, P(0));
    case 234: return var_sphere_fold(p, w, P(0), P(1));
    case 235: return var_mandelbox_step(p, w, P(0), P(1), P(2), P(3), P(4));
    case 236: return var_kifs_fold(p, w, P(0), P(1));
    case 237: return var_arnold_cat(p, w);
    case 238: return var_bakers_map(p, w);
    case 239: return var_tent_map(p, w);
    case 240: return var_logistic_map(p, w, P(0));
    case 241: return var_superellipse(p, w, P(0), P(1), P(2));
    case 242: return var_limacon(p, w, P(0), P(1));
    case 243: return var_epicycloid(p, w, P(0));
    case 244: return var_catenary(p, w, P(0));
    case 245: return var_tractrix(p, w);
    case 246: return var_tinkerbell(p, w, P(0), P(1), P(2), P(3));
    case 247: return var_duffing(p, w, P(0), P(1), P(2), P(3));
    case 248: return var_vanderpol(p, w, P(0), P(1));
    case 249: return var_rossler(p, w, P(0), P(1));
    case 250: return var_droste(p, w, P(0));
    case 251: return var_logspiral(p, w, P(0), P(1));
    case 252: return var_fermat_spiral(p, w, P(0));
    case 253: return var_lituus(p, w, P(0));
    case 254: return var_hyperbolic_spiral(p, w, P(0));
    case 255: return var_weierstrass(p, w, P(0), P(1), P(2), P(3));
    case 256: return var_takagi(p, w, P(0), P(1));
    case 257: return var_cantor_stairs(p, w, P(0), P(1));
    case 258: return var_billiard_circle(p, w, P(0), P(1), P(2));
    case 259: return var_billiard_stadium(p, w, P(0), P(1), P(2), P(3));
    case 260: return var_billiard_sinai(p, w, P(0), P(1), P(2), P(3));
    case 261: return var_billiard_polygon(p, w, P(0), P(1), P(2), P(3));
    case 262: return var_lorentz_boost(p, w, P(0), P(1));
    case 263: return var_schwarzschild_lensing(p, w, P(0), P(1));
    case 264: return var_field_dipole(p, w, P(0), P(1), P(2), P(3));
    case 265: return var_magnetic_pendulum_pos(p, w, P(0), P(1), P(2), P(3));
    case 266: return var_jacobi_theta(p, w, P(0));
    case 267: return var_modular_lambda(p, w, P(0));
    case 268: return var_klein_j(p, w, P(0));
    case 269: return var_weierstrass_p(p, w, P(0), P(1), P(2), P(3));
    case 270: return var_gauss_map(p, w);
    case 271: return var_nbody_lensing(p, w, P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7));
    case 272: return var_curl_noise(p, w, P(0), P(1));
    case 273: return var_bessel_j0(p, w, P(0));
    case 274: return var_airy_radial(p, w, P(0), P(1));
    case 275: return var_cornu_spiral(p, w, P(0));
    case 276: return var_struve_h1(p, w, P(0));
    case 277: return var_radical_inverse(p, w, P(0), P(1));
    case 278: return var_gray_code(p, w, P(0), P(1));
    case 279: return var_morton_zorder(p, w, P(0), P(1));
    case 280: return var_chebyshev(p, w, P(0), P(1));
    case 281: return var_legendre(p, w, P(0), P(1));
    case 282: return var_spherical_harmonic(p, w, P(0), P(1), P(2));
    case 283: return var_fourier_warp(p, w, P(0), P(1), P(2));
    case 284: return var_snell_refraction(p, w, P(0), P(1));
    case 285: return var_grin_lens(p, w, P(0), P(1));
    case 286: return var_caustic_fold(p, w, P(0), P(1), P(2));
    case 287: return var_chladni(p, w, P(0), P(1), P(2));
    case 288: return var_standing_wave(p, w, P(0), P(1), P(2), P(3));
    case 289: return var_moire(p, w, P(0), P(1), P(2), P(3));
    case 290: return var_radial_shell(p, w, P(0), P(1), P(2));
    case 291: return var_hydrogen_orbital(p, w, P(0), P(1), P(2), P(3), P(4));
    case 292: return var_weibull_cdf(p, w, P(0), P(1));
    case 293: return var_logistic_cdf(p, w, P(0), P(1));
    case 294: return var_cauchy_cdf(p, w, P(0), P(1));
    case 295: return var_pareto_cdf(p, w, P(0), P(1), P(2));
    case 296: return var_morlet(p, w, P(0), P(1), P(2));
    case 297: return var_mexican_hat(p, w, P(0), P(1));
    case 298: return var_chirp(p, w, P(0), P(1), P(2));
    case 299: return var_kepler_orbit(p, w, P(0), P(1));
    case 300: return var_restricted_3body(p, w, P(0), P(1), P(2));
    case 301: return var_hill_epicyclic(p, w, P(0), P(1), P(2));
    case 302: return var_torus_knot(p, w, P(0), P(1), P(2), P(3));
    case 303: return var_braid_warp(p, w, P(0), P(1), P(2));
    case 304: return var_airy_caustic(p, w, P(0), P(1), P(2));
    case 305: return var_elliptic_E(p, w, P(0));
    case 306: return var_elliptic_K(p, w, P(0), P(1));
    case 307: return var_gaussian_cdf(p, w, P(0), P(1));
    case 308: return var_levy_cdf(p, w, P(0), P(1));
    case 309: return var_peano(p, w, P(0), P(1));
    case 310: return var_burning_ship(p, w, P(0), P(1));
    case 311: return var_magnet1(p, w, P(0), P(1));
    case 312: return var_nova(p, w, P(0), P(1), P(2));
    case 313: return var_halley(p, w, P(0), P(1));
    case 314: return var_lichtenberg(p, w, P(0), P(1), P(2), P(3), P(4));
    case 315: return var_copula_gaussian(p, w, P(0), P(1));
    case 316: return var_copula_clayton(p, w, P(0), P(1));
    case 317: return var_schwarz_christoffel(p, w, P(0));
    case 318: return var_doyle(p, w, P(0), P(1));
    case 319: return var_quasicrystal(p, w, P(0), P(1));
    case 320: return var_penrose(p, w, P(0), P(1));
    case 321: return var_collatz(p, w, P(0), P(1));
    case 322: return var_digamma(p, w, P(0), P(1));
    case 323: {
      const float c[10] = {P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7), P(8), P(9)};
      return var_sprott_poly(p, w, c);
    }
    case 324: return var_hopalong(p, w, P(0), P(1), P(2));
    case 325: return var_gumowski_mira(p, w, P(0), P(1));
    default: return make_float2(0.0f, 0.0f);
  }
#undef P
}

// chaos.wgsl:8192 — colour override of slot k of record xf (parameters p0, p1
// at vars[k].zw, p2 at vars_extra[k].x).
P3_INLINE bool dc_color_ext(int idx, float2 p, xrec_t xf, int k, float3& rgb) {
  const float p0 = var_param(xf, k, 0), p1 = var_param(xf, k, 1), p2 = var_param(xf, k, 2);
  switch (idx) {
    case 99: rgb = dc_linear_color(p); return true;
    case 100: rgb = dc_perlin_color(p, p0, p1, p2); return true;
    case 101: rgb = dc_gridout_color(p, p0); return true;
    case 102: rgb = dc_cylinder_color(p); return true;
    case 220: rgb = newton_color(p, p0); return true;
    case 265: rgb = magnet_color(p, p0, p1); return true;
    case 310: case 311: case 313: rgb = escape_color(p, idx, p0, p1, 0.0f); return true;
    case 312: rgb = escape_color(p, idx, p0, p1, p2); return true;
    default: return false;
  }
}

}  // namespace p3
