// Variation library, catalog entries V99..V325 — oracle restatement,
// templated on the float type like orc_variations.hpp (test infrastructure
// only: see oracle.hpp).
//
// Each function restates the WGSL function of the same name in
// src/shaders/chaos.wgsl (file:line in the comment above it) with the same
// operation order and the same RNG draw sequence: every rand01/irand call the
// WGSL makes is made here, in the same order, whether or not its value is
// used. Vector +/- of a literal (x, 0) keeps the `+ 0` / `- 0` on the
// zero lane (it changes the sign of a zero). WGSL semantics kept: f32 `%` is x - y*trunc(x/y) (wmod), i32(f) and
// u32(f) saturate (f2i / f2u), i32 arithmetic wraps (done in uint32_t),
// select(f, t, c) evaluates both sides, round() is half-to-even (rint),
// clamp(x, lo, hi) = min(max(x, lo), hi), mix(a, b, t) = a(1-t) + bt.
#pragma once

namespace orc {

// ---- shared helpers ----------------------------------------------------------

template <typename T> inline T fract(T x) { return x - std::floor(x); }
template <typename T> inline T clampv(T x, T lo, T hi) { return wmin(wmax(x, lo), hi); }
template <typename T> inline T mixv(T a, T b, T t) { return a * ((T)1 - t) + b * t; }
template <typename T> inline T rint_even(T x) { return std::nearbyint(x); }
template <typename T> inline T signv(T x) { return x > (T)0 ? (T)1 : (x < (T)0 ? (T)-1 : (T)0); }
template <typename T> inline uint32_t f2u(T f) {  // WGSL u32(f): saturating, NaN -> 0
  if (!(f == f) || f <= (T)0) return 0u;
  if (f >= (T)4294967295.0) return 4294967295u;
  return (uint32_t)f;
}
inline int32_t imin(int32_t a, int32_t b) { return b < a ? b : a; }
inline int32_t imax(int32_t a, int32_t b) { return a < b ? b : a; }
inline uint32_t umin(uint32_t a, uint32_t b) { return b < a ? b : a; }
inline int32_t imod(int32_t a, int32_t b) { return b == 0 ? a : a % b; }  // WGSL i32 %
inline int32_t idiv(int32_t a, int32_t b) { return b == 0 ? a : a / b; }  // WGSL i32 /
inline int32_t iwrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
inline int32_t iwrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// Complex helpers, z = (re, im) (chaos.wgsl:389-460).
template <typename T> inline V2<T> c_mul(V2<T> a, V2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> inline V2<T> c_sqr(V2<T> z) { return {z.x * z.x - z.y * z.y, (T)2 * z.x * z.y}; }
// |b|^2 floored at the f32 value of 1e-100 (zero: the floor is a no-op in f32).
template <typename T> inline V2<T> c_div(V2<T> a, V2<T> b) {
  T m2 = wmax(b.x * b.x + b.y * b.y, (T)(float)1e-100);
  return {(a.x * b.x + a.y * b.y) / m2, (a.y * b.x - a.x * b.y) / m2};
}
template <typename T> inline V2<T> c_recip(V2<T> z) {
  T m2 = wmax(z.x * z.x + z.y * z.y, (T)(float)1e-100);
  return {z.x / m2, -z.y / m2};
}
template <typename T> inline V2<T> c_sqrt(V2<T> z) {
  T rad = len2(z);
  T sb = sel((T)1, (T)-1, z.y < (T)0);
  return {std::sqrt(wmax((T)0.5 * (rad + z.x), (T)0)),
          sb * std::sqrt(wmax((T)0.5 * (rad - z.x), (T)0))};
}
template <typename T> inline V2<T> c_log(V2<T> z) {
  T m2 = (z.x * z.x + z.y * z.y) + (T)1e-20;
  return {(T)0.5 * std::log(m2), std::atan2(z.y, z.x)};
}
template <typename T> inline V2<T> c_exp(V2<T> z) {
  T e = std::exp(clampv(z.x, (T)-20, (T)20));
  return {e * safe_cos(z.y), e * safe_sin(z.y)};
}
template <typename T> inline V2<T> c_pow(V2<T> t, V2<T> p) { return c_exp(c_mul(p, c_log(t))); }
template <typename T> inline V2<T> c_sin(V2<T> z) {
  T y = clampv(z.y, (T)-20, (T)20);
  T ep = std::exp(y), en = std::exp(-y);
  T ch = (T)0.5 * (ep + en), sh = (T)0.5 * (ep - en);
  return {safe_sin(z.x) * ch, safe_cos(z.x) * sh};
}

// Improved-noise gradient lattice (src/shaders/noise_perlin.wgsl): Ken
// Perlin's published 256-entry permutation, indexed modulo 256.
static const uint8_t kPerlinPerm[256] = {
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

inline uint32_t perm(uint32_t i) { return kPerlinPerm[i & 255u]; }

template <typename T> inline T perlin_fade(T t) {
  return t * t * t * (t * (t * (T)6 - (T)15) + (T)10);
}
// Gradient pick: low 3 bits of the hash choose axis order, signs, and a
// doubled second component.
template <typename T> inline T perlin_grad2(uint32_t h, T x, T y) {
  h &= 7u;
  T u = h < 4u ? x : y;
  T v = h < 4u ? y : x;
  T a = (h & 1u) ? -u : u;
  T b = (h & 2u) ? (T)-2 * v : (T)2 * v;
  return a + b;
}
template <typename T> inline T perlin2d(T px, T py) {
  T fx = std::floor(px), fy = std::floor(py);
  uint32_t X = (uint32_t)(f2i(fx) & 255), Y = (uint32_t)(f2i(fy) & 255);
  T xf = px - fx, yf = py - fy;
  T u = perlin_fade(xf), v = perlin_fade(yf);
  uint32_t A = (perm(X) + Y) & 255u, B = (perm(X + 1u) + Y) & 255u;
  T g00 = perlin_grad2(perm(A), xf, yf);
  T g10 = perlin_grad2(perm(B), xf - (T)1, yf);
  T g01 = perlin_grad2(perm(A + 1u), xf, yf - (T)1);
  T g11 = perlin_grad2(perm(B + 1u), xf - (T)1, yf - (T)1);
  return mixv(mixv(g00, g10, u), mixv(g01, g11, u), v);
}
template <typename T> inline T perlin_fbm(T px, T py, T octaves, T scale) {
  T total = 0, amp = 1, freq = scale, norm = 0;
  const uint32_t n = f2u(clampv(octaves, (T)1, (T)8));
  for (uint32_t i = 0; i < n; ++i) {
    total = total + perlin2d(px * freq, py * freq) * amp;
    norm = norm + amp;
    amp = amp * (T)0.5;
    freq = freq * (T)2;
  }
  return total / norm;
}

// ---- direct-colour entries, position side (chaos.wgsl:1752) -------------------

template <typename T> V2<T> var_dc_cylinder_pos(V2<T> p, T w) { return {w * safe_sin(p.x), w * p.y}; }

// ---- post-flam3 plugin pack (chaos.wgsl:1784-1931) ---------------------------

// chaos.wgsl:1784 — three draws, always taken.
template <typename T, typename R>
V2<T> var_cpow2(V2<T> p, T w, T pr, T pa, T divisor, T range, R& rng) {
  T div = sel(divisor, (T)1, divisor == (T)0);
  T rmax = wmax(range, (T)1);
  T step = K<T>::TAU / div;
  T c = pr * safe_cos(K<T>::PI * (T)0.5 * pa) / div;
  T d = pr * safe_sin(K<T>::PI * (T)0.5 * pa) / div;
  T hc = c * (T)0.5;
  T inv_range = (T)0.5 / rmax;
  T full = K<T>::TAU * rmax;
  T u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01();
  T a = std::atan2(p.y, p.x);
  T n = std::floor(u0 * rmax) + sel((T)0, (T)1, a < (T)0);
  a = a + K<T>::TAU * n;
  if (safe_cos(a * inv_range) < (u1 * (T)2 - (T)1)) a = a - full;
  T lnr2 = std::log(dot2(p));
  T r = w * std::exp(hc * lnr2 - d * a);
  T th = c * a + (d * (T)0.5) * lnr2 + step * std::floor(div * u2);
  return {r * safe_cos(th), r * safe_sin(th)};
}

// chaos.wgsl:1814 — four draws, always taken.
template <typename T, typename R>
V2<T> var_cpow3(V2<T> p, T w, T pr, T pd, T divisor, T spread, R& rng) {
  T div = sel(divisor, (T)1, divisor == (T)0);
  T step = K<T>::TAU / div;
  T dabs = wmax(std::fabs(pd), (T)1e-30);
  T slog = sel(std::log(dabs), -std::log(dabs), pd < (T)0);
  T ang = std::atan2(slog * pr, K<T>::TAU);
  T ca = safe_cos(ang), sa = safe_sin(ang);
  T tc = ca * pr * ca / div;
  T td = ca * pr * sa / div;
  T hc = tc * (T)0.5, hd = td * (T)0.5;
  T coeff = sel((T)-0.095 * spread / td, (T)0, td == (T)0);
  T u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01(), u3 = rng.rand01();
  T a = std::atan2(p.y, p.x);
  if (a < (T)0) a = a + K<T>::TAU;
  if (safe_cos(a * (T)0.5) < (u0 * (T)2 - (T)1)) a = a - K<T>::TAU;
  T bsign = sel(-K<T>::TAU, K<T>::TAU, u1 < (T)0.5);
  a = a + bsign * rint_even(std::log(wmax(u2, (T)1e-30)) * coeff);
  T lnr2 = std::log(dot2(p));
  T r = w * std::exp(hc * lnr2 - td * a);
  T th = tc * a + hd * lnr2 + step * std::floor(div * u3);
  return {r * safe_cos(th), r * safe_sin(th)};
}

// chaos.wgsl:1856 — n-gon "loonie" with star and circle blends.
template <typename T> V2<T> var_loonie2(V2<T> p, T w, T sides_f, T star, T circle) {
  const int32_t sides = imin(imax(f2i(sides_f), 1), 16);
  T a = K<T>::TAU / (T)sides;
  T sa = safe_sin(a), ca = safe_cos(a);
  T sins = safe_sin(-star * K<T>::PI * (T)0.5), coss = safe_cos(star * K<T>::PI * (T)0.5);
  T sinc = safe_sin(circle * K<T>::PI * (T)0.5), cosc = safe_cos(circle * K<T>::PI * (T)0.5);
  T ww = w * w;
  T xr = p.x, yr = p.y;
  T r2 = xr * coss + std::fabs(yr) * sins;
  T cr = std::sqrt(xr * xr + yr * yr);
  for (int32_t i = 0; i < sides - 1; ++i) {
    T nx = xr * ca - yr * sa;
    yr = xr * sa + yr * ca;
    xr = nx;
    r2 = wmax(r2, xr * coss + std::fabs(yr) * sins);
  }
  r2 = r2 * cosc + cr * sinc;
  if (sides > 2) r2 = r2 * r2;
  else r2 = std::fabs(r2) * r2;
  if (r2 > (T)0 && r2 < ww) {
    T r = w * std::sqrt(std::fabs(ww / r2 - (T)1));
    return {r * p.x, r * p.y};
  }
  if (r2 < (T)0) {
    T r = w / std::sqrt(std::fabs(ww / r2) - (T)1);
    return {r * p.x, r * p.y};
  }
  return {w * p.x, w * p.y};
}

// chaos.wgsl:1908 — one draw, always taken.
template <typename T, typename R>
V2<T> var_epispiral(V2<T> p, T w, T n, T thickness, T holes, R& rng) {
  T th = std::atan2(p.y, p.x);
  T d = safe_cos(n * th);
  T u = rng.rand01();
  T inv = (T)1 / d;
  T t = sel(-holes + inv, -holes + (u * thickness) * inv, std::fabs(thickness) > (T)1e-30);
  return {w * t * safe_cos(th), w * t * safe_sin(th)};
}

// ---- cellular family (chaos.wgsl:1933-2147) ----------------------------------

// chaos.wgsl:1933 — integer mix of a cell coordinate to two uniforms.
template <typename T> V2<T> worley_hash2(int32_t cx, int32_t cy) {
  uint32_t s = (uint32_t)cx * 2654435769u;
  s = s ^ ((uint32_t)cy * 2246822519u);
  s = s ^ (s >> 16);
  s = s * 0x85ebca6bu;
  s = s ^ (s >> 13);
  s = s * 0xc2b2ae35u;
  s = s ^ (s >> 16);
  return {(T)(float)(s & 0xffffu) * (T)(float)(1.0 / 65535.0),
          (T)(float)((s >> 16) & 0xffffu) * (T)(float)(1.0 / 65535.0)};
}

// chaos.wgsl:1988 — bubble-wrap lattice.
template <typename T>
V2<T> var_bwraps(V2<T> p, T w, T cell, T space, T gain, T itwist, T otwist) {
  if (std::fabs(cell) < (T)1e-30) return {w * p.x, w * p.y};
  T radius = (T)0.5 * (cell / ((T)1 + space * space));
  T g2 = gain * gain / cell + (T)1e-6;
  T mb = g2 * radius;
  if (mb > (T)2) mb = (T)1;
  else mb = mb * ((T)1 / ((mb * mb) / (T)4 + (T)1));
  T rr2 = radius * radius;
  T rfac = radius / wmax(mb, (T)1e-30);
  T cx = (std::floor(p.x / cell) + (T)0.5) * cell;
  T cy = (std::floor(p.y / cell) + (T)0.5) * cell;
  T lx = p.x - cx, ly = p.y - cy;
  if ((lx * lx + ly * ly) > rr2) return {w * p.x, w * p.y};
  lx = lx * g2;
  ly = ly * g2;
  T rd = rfac / ((lx * lx + ly * ly) / (T)4 + (T)1);
  lx = lx * rd;
  ly = ly * rd;
  T rf = (lx * lx + ly * ly) / wmax(rr2, (T)1e-30);
  T th = itwist * ((T)1 - rf) + otwist * rf;
  T st = safe_sin(th), ct = safe_cos(th);
  return {w * (cx + ct * lx + st * ly), w * (cy - st * lx + ct * ly)};
}

// chaos.wgsl:2034 — noise-perturbed cell centre.
template <typename T> V2<T> crackle_centre(int32_t cx, int32_t cy, T s, T d) {
  T fx = (T)cx, fy = (T)cy;
  T ex = perlin2d(fx * (T)2.5, fy * (T)2.5);
  T ey = perlin2d(fy * (T)2.5 + (T)30.2, fx * (T)2.5 - (T)12.1);
  return {(fx + d * ex) * s, (fy + d * ey) * s};
}
// chaos.wgsl:2046 — Voronoi boundary ratio of u between cells p and q.
template <typename T> T crackle_vratio(V2<T> p, V2<T> q, V2<T> u) {
  T dx = p.x - q.x, dy = p.y - q.y;
  T den = dx * dx + dy * dy;
  if (den < (T)1e-30) return (T)1;
  T num = (u.x - q.x) * dx + (u.y - q.y) * dy;
  return (T)2 * num / den;
}
// chaos.wgsl:2073 — four draws (input blur), then a 3x3 Voronoi search.
template <typename T, typename R>
V2<T> var_crackle(V2<T> p, T w, T cell, T power, T distort, T scale, R& rng) {
  if (std::fabs(cell) < (T)1e-30) return {w * p.x, w * p.y};
  T s = cell * (T)0.5;
  T u1 = rng.rand01(), u2 = rng.rand01(), u3 = rng.rand01(), u4 = rng.rand01();
  T blur = (u1 + u2) * (T)0.5 + (u3 - (T)0.5) * (T)0.25;
  T th = (T)2 * K<T>::PI * u4;
  V2<T> u = {blur * std::sin(th), blur * std::cos(th)};
  const int32_t x0 = f2i(std::floor(u.x / s)), y0 = f2i(std::floor(u.y / s));
  V2<T> P[9];
  for (int i = 0; i < 9; ++i)
    P[i] = crackle_centre(iwrap_add(x0, i / 3 - 1), iwrap_add(y0, i % 3 - 1), s, distort);
  int q = 0;
  T best = (T)1e30;
  for (int i = 0; i < 9; ++i) {
    T dx = P[i].x - u.x, dy = P[i].y - u.y;
    T d2 = dx * dx + dy * dy;
    if (d2 < best) { best = d2; q = i; }
  }
  const int32_t xc = iwrap_add(x0, q / 3 - 1), yc = iwrap_add(y0, q % 3 - 1);
  for (int i = 0; i < 9; ++i)
    P[i] = crackle_centre(iwrap_add(xc, i / 3 - 1), iwrap_add(yc, i % 3 - 1), s, distort);
  T L = (T)-1e30;
  for (int i = 0; i < 9; ++i) {
    if (i == 4) continue;
    T ratio = crackle_vratio(P[i], P[4], u);
    if (ratio > L) L = ratio;
  }
  T dxo = u.x - P[4].x, dyo = u.y - P[4].y;
  T ls = L + (T)1e-30;
  T trg = std::pow(std::fabs(ls), power) * scale;
  T Rr = trg / ls;
  return {w * (P[4].x + dxo * Rr), w * (P[4].y + dyo * Rr)};
}

// ---- JWildfire S-tier, first half (chaos.wgsl:2149-2360) ---------------------

// chaos.wgsl:2149 — one draw.
template <typename T, typename R> V2<T> var_juliaq(V2<T> p, T w, T power, T divisor, R& rng) {
  T ps = sel(power, (T)1, std::fabs(power) < (T)1e-30);
  T ap = wmax((T)1, std::fabs(ps));
  T ip = divisor / ps;
  T hip = (T)0.5 * ip;
  T ip2 = K<T>::TAU / ps;
  T u = rng.rand01();
  T n = std::floor(u * ap);
  T a = std::atan2(p.y, p.x) * ip + n * ip2;
  T r = w * std::pow(p.x * p.x + p.y * p.y, hip);
  return {r * safe_cos(a), r * safe_sin(a)};
}

// chaos.wgsl:2169 — one draw picks one of two formulas per half-plane.
template <typename T, typename R> V2<T> var_glynnia(V2<T> p, T w, R& rng) {
  T v2 = w * (T)0.7071067811865476;
  T r = std::sqrt(p.x * p.x + p.y * p.y);
  T coin = rng.rand01();
  T sx = r >= (T)1 ? (T)1 : (T)-1;  // outside the unit disc keeps +x
  if (coin > (T)0.5) {
    T in = r + p.x;
    if (in <= (T)0) return {(T)0, (T)0};
    T d = std::sqrt(in);
    if (d == (T)0) return {(T)0, (T)0};
    return {sx * v2 * d, -(v2 / d) * p.y};
  }
  T d = r + p.x;
  T rad = r * (p.y * p.y + d * d);
  if (rad <= (T)0) return {(T)0, (T)0};
  T dx = std::sqrt(rad);
  if (dx == (T)0) return {(T)0, (T)0};
  T rr = w / dx;
  return {sx * rr * d, rr * p.y};
}

// chaos.wgsl:2213
template <typename T> V2<T> var_loonie3(V2<T> p, T w) {
  T ww = w * w;
  T r2 = (T)2 * ww;
  if (p.x > (T)1e-30) {
    T q = (p.x * p.x + p.y * p.y) / p.x;
    r2 = q * q;
  }
  if (r2 < ww) {
    T r = w * std::sqrt(ww / r2 - (T)1);
    return {r * p.x, r * p.y};
  }
  return {w * p.x, w * p.y};
}

// chaos.wgsl:2236 — two draws.
template <typename T, typename R>
V2<T> var_falloff(V2<T> p, T w, T scatter, T mindist, T mx, T my, T x0, T y0, R& rng) {
  T rmax = (T)0.04 * scatter;
  T dx = p.x - x0, dy = p.y - y0;
  T d = std::sqrt(dx * dx + dy * dy);
  if (d < (T)0) d = (T)0;
  d = (d - mindist) * rmax;
  if (d < (T)0) d = (T)0;
  T u0 = rng.rand01(), u1 = rng.rand01();
  return {w * (p.x + mx * u0 * d), w * (p.y + my * u1 * d)};
}

// chaos.wgsl:2263 — three draws; type 0 (linear), 1 (radial), 2 (gaussian).
template <typename T, typename R>
V2<T> var_falloff2(V2<T> p, T w, T scatter, T typ, T mx, T my, T x0, T y0, T mindist, R& rng) {
  T rmax = (T)0.04 * scatter;
  T dx = p.x - x0, dy = p.y - y0;
  T d = std::sqrt(dx * dx + dy * dy);
  if (d < (T)0) d = (T)0;
  d = (d - mindist) * rmax;
  if (d < (T)0) d = (T)0;
  T u0 = rng.rand01(), u1 = rng.rand01(), u2 = rng.rand01();
  const int32_t t = f2i(clampv(typ, (T)0, (T)2));
  if (t == 1) {
    T ra = std::sqrt(p.x * p.x + p.y * p.y);
    T phi = std::atan2(p.y, p.x) + my * d * u1;
    T rr = ra + mx * u0 * d;
    return {w * rr * safe_cos(phi), w * rr * safe_sin(phi)};
  }
  if (t == 2) {
    T sg = d * u1 * K<T>::TAU;
    T phi = d * u2 * K<T>::PI;
    T rad = d * u0;
    T sc = safe_cos(sg);
    return {w * (p.x + mx * rad * sc * safe_cos(phi)), w * (p.y + my * rad * sc * safe_sin(phi))};
  }
  return {w * (p.x + mx * u0 * d), w * (p.y + my * u1 * d)};
}

// chaos.wgsl:2318 — three draws, gaussian type with optional inversion.
template <typename T, typename R>
V2<T> var_falloff3(V2<T> p, T w, T scatter, T mx, T my, T x0, T y0, T mindist, T inv, R& rng) {
  T rmax = (T)0.04 * scatter;
  T dx = p.x - x0, dy = p.y - y0;
  T radius = std::sqrt(dx * dx + dy * dy);
  T base = sel(wmax(radius, (T)0), wmax((T)1 - radius, (T)0), inv > (T)0.5);
  T dist = wmax((base - mindist) * rmax, (T)0);
  T u0 = rng.rand01() - (T)0.5;
  T u1 = rng.rand01() - (T)0.5;
  T u2 = rng.rand01() - (T)0.5;
  T sg = dist * u1 * K<T>::TAU;
  T phi = dist * u2 * K<T>::PI;
  T rad = dist * u0;
  T sc = safe_cos(sg);
  return {w * (p.x + mx * rad * sc * safe_cos(phi)), w * (p.y + my * rad * sc * safe_sin(phi))};
}

// ---- kaleidoscope / circle family (chaos.wgsl:2362-2527) ---------------------

// chaos.wgsl:2362
template <typename T> V2<T> var_collideoscope(V2<T> p, T w, T ap, T nump) {
  const int32_t ni = imax(1, f2i(nump));
  T nf = (T)ni;
  T knpi = nf / K<T>::PI, pikn = K<T>::PI / nf;
  T kakn = K<T>::PI * ap / nf;
  T a = std::atan2(p.y, p.x);
  T r = w * std::sqrt(p.x * p.x + p.y * p.y);
  if (a >= (T)0) {
    const int32_t alt = f2i(a * knpi);
    a = (T)alt * pikn + wmod((imod(alt, 2) == 0 ? kakn : -kakn) + a, pikn);
  } else {
    const int32_t alt = f2i(-a * knpi);
    a = -((T)alt * pikn + wmod((imod(alt, 2) != 0 ? -kakn : kakn) - a, pikn));
  }
  return {r * safe_cos(a), r * safe_sin(a)};
}

// Square-to-circle perimeter walk shared by circlize / circlize2
// (chaos.wgsl:2400-2464): returns false when the point is the origin.
template <typename T> bool square_perimeter(V2<T> p, T& perim, T& side) {
  T ax = std::fabs(p.x), ay = std::fabs(p.y);
  if (ax >= ay) {
    perim = p.x >= ay ? ax + p.y : (T)5 * ax - p.y;
    side = ax;
  } else {
    perim = p.y >= ax ? (T)3 * ay - p.x : (T)7 * ay + p.x;
    side = ay;
  }
  return side != (T)0;
}
// chaos.wgsl:2400
template <typename T> V2<T> var_circlize(V2<T> p, T w, T hole) {
  T v4 = w / (K<T>::PI * (T)0.25);
  T perim, side;
  if (!square_perimeter(p, perim, side)) return {(T)0, (T)0};
  T r = v4 * side + hole;
  T a = (K<T>::PI * (T)0.25) * perim / side - (K<T>::PI * (T)0.25);
  return {r * safe_cos(a), r * safe_sin(a)};
}
// chaos.wgsl:2434
template <typename T> V2<T> var_circlize2(V2<T> p, T w, T hole) {
  T perim, side;
  if (!square_perimeter(p, perim, side)) return {(T)0, (T)0};
  T r = w * (side + hole);
  T a = (K<T>::PI * (T)0.25) * perim / side - (K<T>::PI * (T)0.25);
  return {r * safe_cos(a), r * safe_sin(a)};
}

// chaos.wgsl:2466 — elliptic-coordinate swirl.
template <typename T> V2<T> var_eswirl(V2<T> p, T w, T inp, T outp) {
  T tmp = p.y * p.y + p.x * p.x + (T)1;
  T tmp2 = (T)2 * p.x;
  T a1 = tmp + tmp2, a2 = tmp - tmp2;
  T s1 = sel((T)0, std::sqrt(wmax(a1, (T)0)), a1 > (T)0);
  T s2 = sel((T)0, std::sqrt(wmax(a2, (T)0)), a2 > (T)0);
  T xmax = (s1 + s2) * (T)0.5;
  if (xmax < (T)1) xmax = (T)1;
  T mu = std::log(xmax + std::sqrt(wmax(xmax * xmax - (T)1, (T)0)));
  T t = p.x / xmax;
  if (t > (T)1) t = (T)1;
  if (t < (T)-1) t = (T)-1;
  T nu = std::acos(t);
  if (p.y < (T)0) nu = -nu;
  T mus = sel(mu, (T)1e-30, mu == (T)0);
  T nw = nu + mu * outp + inp / mus;
  return {w * std::cosh(mu) * safe_cos(nw), w * std::sinh(mu) * safe_sin(nw)};
}

// chaos.wgsl:2503
template <typename T> V2<T> var_petal(V2<T> p, T w) {
  T cx = safe_cos(p.x), sx = safe_sin(p.x), cy = safe_cos(p.y);
  T a = cx * cy, b = sx * cy;
  T bx = a * a * a, by = b * b * b;
  return {w * cx * bx, w * cx * by};
}

// ---- Xyrus02 mid-tier, hexes (chaos.wgsl:2529-2760) --------------------------

// chaos.wgsl:2529 — the draw is taken only outside the unit circle.
template <typename T, typename R> V2<T> var_bcircle(V2<T> p, T w, T scale, T border, R& rng) {
  T bw = std::fabs(border);
  if (p.x == (T)0 && p.y == (T)0) return {(T)0, (T)0};
  T x = p.x * scale, y = p.y * scale;
  T r = std::sqrt(x * x + y * y);
  if (r <= (T)1) return {w * x, w * y};
  if (bw == (T)0) return {(T)0, (T)0};
  T ang = std::atan2(y, x);
  T u = rng.rand01();
  T om = (T)0.2 * bw * u + (T)1;
  return {w * om * safe_cos(ang), w * om * safe_sin(ang)};
}

// chaos.wgsl:2551 — z / (1 + c1 z + c2 z^2 + c3 z^3), expanded in reals.
template <typename T> V2<T> var_curl2(V2<T> p, T w, T c1, T c2, T c3) {
  T c2d = (T)2 * c2, c3t = (T)3 * c3;
  T x = p.x, y = p.y;
  T x2 = x * x, x3 = x2 * x, y2 = y * y, y3 = y2 * y;
  T re = c3 * x3 - c3t * x * y2 + c2 * x2 - c2 * y2 + c1 * x + (T)1;
  T im = c3t * x2 * y - c3 * y3 + c2d * x * y + c1 * y;
  T r = w / (re * re + im * im);
  return {(x * re + y * im) * r, (y * re - x * im) * r};
}

// chaos.wgsl:2579
template <typename T> V2<T> var_murl(V2<T> p, T w, T cin, T pin) {
  const int32_t pi = f2i(pin);
  T pf = (T)pi;
  T c = sel(cin / (pf - (T)1), cin, pi == 1);
  T p2 = pf / (T)2;
  T vp = w * (c + (T)1);
  T a = std::atan2(p.y, p.x) * pf;
  T sa = safe_sin(a), ca = safe_cos(a);
  T r = c * std::pow(p.x * p.x + p.y * p.y, p2);
  T re = r * ca + (T)1, im = r * sa;
  T r1 = vp / (re * re + im * im + (T)1e-29);
  return {r1 * (p.x * re + p.y * im), r1 * (p.y * re - p.x * im)};
}

// chaos.wgsl:2605
template <typename T> V2<T> var_stwins(V2<T> p, T w, T distort) {
  T x = p.x * w * (T)0.05, y = p.y * w * (T)0.05;
  T x2 = x * x, y2 = y * y;
  T num = (x2 - y2) * safe_sin(K<T>::TAU * distort * (x + y));
  T sum = x2 + y2;
  T res = num / sel(sum, (T)1, sum == (T)0);
  return {w * p.x + res, w * p.y + res};
}

// chaos.wgsl:2736 — running max of Voronoi ratios (P relative to centre Q).
template <typename T> T hex_ratio_step(V2<T> P, V2<T> Q, T ux, T uy, T prev) {
  T dx = P.x - Q.x, dy = P.y - Q.y;
  if (dx == (T)0 && dy == (T)0) return sel(prev, (T)1, (T)1 > prev);
  T num = (T)2 * ((ux - Q.x) * dx + (uy - Q.y) * dy);
  T ratio = num / (dx * dx + dy * dy);
  return sel(prev, ratio, ratio > prev);
}
// chaos.wgsl:2725 — ratios against the six neighbours H[1..6] of H[0].
template <typename T> T hex_ratio_max(const V2<T>* H, T ux, T uy) {
  T m = (T)-1e20;
  for (int i = 1; i < 7; ++i) m = hex_ratio_step(H[i], H[0], ux, uy, m);
  return m;
}
// Hex-lattice centre of cell (i, j) (axial coordinates, cell size s).
template <typename T> V2<T> hex_centre(int32_t i, int32_t j, T s) {
  const T sq3 = (T)1.7320508075688772935;
  T fi = (T)i, fj = (T)j;
  return {((T)1.5 * fi + (T)-1.5 * fj) * s, (sq3 / (T)2 * fi + sq3 / (T)2 * fj) * s};
}
// chaos.wgsl:2637 — hexagonal Voronoi cells with power falloff and rotation.
template <typename T> V2<T> var_hexes(V2<T> p, T w, T cell, T power, T rotate, T scale) {
  if (cell == (T)0) return {(T)0, (T)0};
  const T sq3 = (T)1.7320508075688772935;
  const T ah = (T)(1.0 / 3.0), bh = sq3 / (T)3, ch = (T)(-1.0 / 3.0), dh = sq3 / (T)3;
  T rs = safe_sin(rotate * K<T>::TAU), rc = safe_cos(rotate * K<T>::TAU);
  T ux = p.x, uy = p.y, s = cell;
  const int32_t h0 = f2i(std::floor((ah * ux + bh * uy) / s));
  const int32_t k0 = f2i(std::floor((ch * ux + dh * uy) / s));
  T best = (T)1e30;
  int q = 0;
  for (int di = -1; di < 2; ++di)
    for (int dj = -1; dj < 2; ++dj) {
      V2<T> c = hex_centre(iwrap_add(h0, di), iwrap_add(k0, dj), s);
      T dx = c.x - ux, dy = c.y - uy;
      T d2 = dx * dx + dy * dy;
      if (d2 < best) { best = d2; q = (di + 1) * 3 + (dj + 1); }
    }
  const int32_t hx = iwrap_add(h0, q / 3 - 1), hy = iwrap_add(k0, q % 3 - 1);
  // centre, then the six neighbours in the reference's order
  const int32_t oi[7] = {0, 0, 1, 1, 0, -1, -1}, oj[7] = {0, 1, 1, 0, -1, -1, 0};
  V2<T> H[7];
  for (int i = 0; i < 7; ++i) H[i] = hex_centre(iwrap_add(hx, oi[i]), iwrap_add(hy, oj[i]), s);
  T L1 = hex_ratio_max(H, ux, uy);
  T dxo = ux - H[0].x, dyo = uy - H[0].y;
  T trg = std::pow(L1 + (T)1e-30, power) * scale;
  T vx = dxo * rc + dyo * rs;
  T vy = -dxo * rs + dyo * rc;
  T L2 = hex_ratio_max(H, vx + H[0].x, vy + H[0].y);
  T L = wmax(L1, L2);
  T Rr;
  if (L < (T)0.5) Rr = trg / L1;
  else if (L > (T)0.8) Rr = trg / L2;
  else Rr = ((trg / L1) * ((T)0.8 - L) + (trg / L2) * (L - (T)0.5)) / (T)0.3;
  return {w * (vx * Rr + H[0].x), w * (vy * Rr + H[0].y)};
}

// ---- Xyrus02 X-family, blur_circle (chaos.wgsl:2762-3100) --------------------

// chaos.wgsl:2762
template <typename T> V2<T> var_xheart(V2<T> p, T w, T angle, T ratio) {
  T ang = K<T>::PI / (T)4 + ((T)0.5 * (K<T>::PI / (T)4) * angle);
  T ca = safe_cos(ang), sa = safe_sin(ang);
  T rat = (T)6 + (T)2 * ratio;
  T r24 = p.x * p.x + p.y * p.y + (T)4;
  if (r24 == (T)0) r24 = (T)1;
  T bx = (T)4 / r24, by = rat / r24;
  T xr = ca * (bx * p.x) - sa * (by * p.y);
  T yr = sa * (bx * p.x) + ca * (by * p.y);
  return {w * xr, w * sel(-yr, yr, xr > (T)0)};
}

// chaos.wgsl:2786
template <typename T>
V2<T> var_xhyperbol(V2<T> p, T w, T m00, T m01, T m10, T m11, T m20, T m21) {
  T r = (T)1 / (p.x * p.x + p.y * p.y + K<T>::EPS);
  T x = p.x * r, y = p.y * r;
  T re = m00 * x + m01 * y + m20;
  T im = m10 * x + m11 * y + m21;
  T al = std::atan2(im, re) + K<T>::TAU;
  T sa = safe_sin(al), ca = safe_cos(al);
  T rsq = re * re + im * im;
  T xo = rsq * ca, yo = rsq * sa;
  T ri = w / (xo * xo + yo * yo + K<T>::EPS);
  return {xo * ri, yo * ri};
}

// chaos.wgsl:2815
template <typename T> V2<T> var_xcurl2(V2<T> p, T w, T c1, T c2, T c3) {
  T x = p.x, y = p.y;
  T x2 = x * x, y2 = y * y, x3 = x2 * x;
  T re = (T)1 + c1 * x + c2 * (x2 - y2) + c3 * (x3 - (T)3 * x);
  T im = c1 * y + c2 * ((T)2 * x * y) + c3 * ((T)3 * x * y - (T)1);
  T r = w / (re * re + im * im);
  return {(x * re + y * im) * r, (y * re + x * im) * r};
}

// chaos.wgsl:2838 — triangular (trilinear) lattice tiling; two draws, the
// second always taken.
template <typename T, typename R>
V2<T> var_xtrb(V2<T> p, T w, T pin, T dist, T radius, T width, T ap, T bp, R& rng) {
  const int32_t pi = imax(1, f2i(pin));
  T pf = (T)pi;
  T angB = (T)0.047 + ap, angC = (T)0.047 + bp;
  T angA = K<T>::PI - angB - angC;
  T sA2 = safe_sin((T)0.5 * angA), cA2 = safe_cos((T)0.5 * angA);
  T sB2 = safe_sin((T)0.5 * angB), cB2 = safe_cos((T)0.5 * angB);
  T sC2 = safe_sin((T)0.5 * angC), cC2 = safe_cos((T)0.5 * angC);
  T sC = safe_sin(angC), cC = safe_cos(angC);
  T as = radius * (sC2 / cC2 + sB2 / cB2);
  T bs = radius * (sC2 / cC2 + sA2 / cA2);
  T cs = radius * (sB2 / cB2 + sA2 / cA2);
  T w1 = (T)1 - width, w2 = (T)2 * width, w3 = (T)1 - width * width;
  T S2 = radius * (as + bs + cs);
  T Ha = S2 / as / (T)6, Hb = S2 / bs / (T)6, Hc = S2 / cs / (T)6;
  T ab = as / bs, ac = as / cs, ba = bs / as, bc = bs / cs, ca = cs / as, cb = cs / bs;
  T S2a = (T)6 * Ha, S2b = (T)6 * Hb, S2c = (T)6 * Hc;
  T S2bc = S2 / (bs + cs) / (T)6, S2ab = S2 / (as + bs) / (T)6, S2ac = S2 / (as + cs) / (T)6;
  const uint32_t absN = (uint32_t)(pi < 0 ? -pi : pi);
  T cN = dist / pf / (T)2;
  T U = p.y + radius;
  T V = p.x * sC - p.y * cC + radius;
  T M = std::floor(U / S2a);
  T oal = U - M * S2a;
  T N = std::floor(V / S2b);
  T obe = V - N * S2b;
  T oga = S2c - ac * oal - bc * obe;
  T Rn = rng.rand01();
  const bool neg = oga <= (T)0;
  if (neg) { oal = S2a - oal; obe = S2b - obe; oga = -oga; }
  T Al = oal, Be = obe, Ga = oga;
  T Al1 = 0, Be1 = 0, Ga1 = 0, De1 = 0;
  const bool edge = Rn >= w3;
  if (Be < Al) {
    if (Ga < Be) {
      if (edge) { De1 = width * Be; Ga1 = width * Ga; }
      else { Ga1 = w1 * Ga + w2 * Hc * Ga / Be; De1 = w1 * Be + w2 * S2ab * ((T)3 - Ga / Be); }
      Al1 = S2a - ba * De1 - ca * Ga1;
      Be1 = De1;
    } else if (Ga < Al) {
      if (edge) { Ga1 = width * Ga; De1 = width * Be; }
      else { De1 = w1 * Be + w2 * Hb * Be / Ga; Ga1 = w1 * Ga + w2 * S2ac * ((T)3 - Be / Ga); }
      Al1 = S2a - ba * De1 - ca * Ga1;
      Be1 = De1;
    } else {
      if (edge) { Al1 = width * Al; Be1 = width * Be; }
      else { Be1 = w1 * Be + w2 * Hb * Be / Al; Al1 = w1 * Al + w2 * S2ac * ((T)3 - Be / Al); }
    }
  } else {
    if (Ga < Al) {
      if (edge) { De1 = width * Al; Ga1 = width * Ga; }
      else { Ga1 = w1 * Ga + w2 * Hc * Ga / Al; De1 = w1 * Al + w2 * S2ab * ((T)3 - Ga / Al); }
      Be1 = S2b - ab * De1 - cb * Ga1;
      Al1 = De1;
    } else if (Ga < Be) {
      if (edge) { Ga1 = width * Ga; De1 = width * Al; }
      else { De1 = w1 * Al + w2 * Ha * Al / Ga; Ga1 = w1 * Ga + w2 * S2bc * ((T)3 - Al / Ga); }
      Be1 = S2b - ab * De1 - cb * Ga1;
      Al1 = De1;
    } else {
      if (edge) { Be1 = width * Be; Al1 = width * Al; }
      else { Al1 = w1 * Al + w2 * Ha * Al / Be; Be1 = w1 * Be + w2 * S2bc * ((T)3 - Al / Be); }
    }
  }
  (void)Ga1;
  T alpha = Al1, beta = Be1;
  if (neg) { alpha = S2a - alpha; beta = S2b - beta; }
  alpha = alpha + M * S2a;
  beta = beta + N * S2b;
  T inx = (beta - radius + (alpha - radius) * cC) / sC;
  T iny = alpha - radius;
  T u2 = rng.rand01();
  const uint32_t br = absN > 0u ? f2u(u2 * (T)absN) : 0u;
  T ang = (std::atan2(iny, inx) + K<T>::TAU * (T)br) / pf;
  T rp = std::pow(inx * inx + iny * iny, cN);
  return {w * w * rp * safe_cos(ang), w * w * rp * safe_sin(ang)};
}

// chaos.wgsl:3019 — step to a neighbouring cell along the square spiral.
template <typename T> V2<T> var_xyrus_gridout(V2<T> p, T w) {
  T x = p.x, y = p.y;
  T rx = sel(std::ceil(x - (T)0.5), std::floor(x + (T)0.5), x >= (T)0);
  T ry = sel(std::ceil(y - (T)0.5), std::floor(y + (T)0.5), y >= (T)0);
  T dx = 0, dy = 0;
  if (ry <= (T)0) {
    if (rx > (T)0) { if (-ry >= rx) dx = (T)1; else dy = (T)1; }
    else { if (ry <= rx) dx = (T)1; else dy = (T)-1; }
  } else {
    if (rx > (T)0) { if (ry >= rx) dx = (T)-1; else dy = (T)1; }
    else { if (ry > -rx) dx = (T)-1; else dy = (T)-1; }
  }
  return {w * (x + dx), w * (y + dy)};
}

// chaos.wgsl:3053 — two draws; a uniform square mapped onto a disc.
template <typename T, typename R> V2<T> var_blur_circle(V2<T> p, T w, T hole, R& rng) {
  (void)p;
  T u0 = rng.rand01(), u1 = rng.rand01();
  T x = (T)2 * u0 - (T)1, y = (T)2 * u1 - (T)1;
  T ax = std::fabs(x), ay = std::fabs(y);
  T s = sel(ay, ax, ax > ay);
  T a = std::atan2(y, x);
  T q3 = (T)3 * K<T>::PI / (T)4, q1 = K<T>::PI / (T)4;
  T ps;
  if (a < -q3) ps = ay;
  else if (a < -q1) ps = (T)2 * s + x;
  else if (a < q1) ps = (T)4 * s + y;
  else if (a < q3) ps = (T)6 * s - x;
  else ps = (T)8 * s - y;
  T r = (w * (T)4 / K<T>::PI) * s + hole;
  T phi = q1 * ps / s - K<T>::PI;
  return {r * safe_cos(phi), r * safe_sin(phi)};
}

// ---- bipolar2 .. waffle (chaos.wgsl:3102-3451) ----------------------------------

// chaos.wgsl:3102
template <typename T>
V2<T> var_bipolar2(V2<T> p, T w, T shift, T a, T b, T c, T d, T e, T f1, T g1, T h) {
  T hp = K<T>::PI * (T)0.5, tp = (T)2 / K<T>::PI;
  T x2y2 = (p.x * p.x + p.y * p.y) * g1;
  T t = x2y2 + a;
  T x2 = b * p.x;
  T ps = -hp * shift;
  T y = c * std::atan2(e * p.y, x2y2 - d) + ps;
  if (y > hp) y = -hp + wmod(y + hp, K<T>::PI);
  else if (y < -hp) y = hp - wmod(hp - y, K<T>::PI);
  T fn = t + x2, gn = t - x2;
  if (gn == (T)0 || (fn / gn) <= (T)0) return {(T)0, (T)0};
  return {w * f1 * tp * std::log(fn / gn), w * tp * y * h};
}

// chaos.wgsl:3136
template <typename T> V2<T> var_bubble2(V2<T> p, T w, T xs, T ys) {
  T r = w / ((T)0.25 * dot2(p) + (T)1);
  return {p.x * r * xs, p.y * r * ys};
}

// Inverse hyperbolic family (chaos.wgsl:3155-3237); outputs scaled by 2w/pi.
// chaos.wgsl:3155 — one draw picks the sign.
template <typename T, typename R> V2<T> var_acosh(V2<T> p, T w, R& rng) {
  V2<T> s = c_sqr(p);
  V2<T> d = c_sqrt(V2<T>{s.x - (T)1, s.y - (T)0});
  V2<T> z = c_log(V2<T>{p.x + d.x, p.y + d.y});
  T k = w * (T)2 / K<T>::PI;
  T sg = sel((T)1, (T)-1, rng.rand01() >= (T)0.5);
  return {sg * (z.x * k), sg * (z.y * k)};
}
// chaos.wgsl:3163
template <typename T> V2<T> var_arcsinh(V2<T> p, T w) {
  V2<T> s = c_sqr(p);
  V2<T> d = c_sqrt(V2<T>{s.x + (T)1, s.y + (T)0});
  V2<T> z = c_log(V2<T>{p.x + d.x, p.y + d.y});
  T k = w * (T)2 / K<T>::PI;
  return {z.x * k, z.y * k};
}
// chaos.wgsl:3172
template <typename T> V2<T> var_arctanh(V2<T> p, T w) {
  V2<T> z = c_log(c_div(V2<T>{p.x + (T)1, p.y + (T)0}, V2<T>{(T)1 - p.x, (T)0 - p.y}));
  T k = w * (T)2 / K<T>::PI;
  return {z.x * k, z.y * k};
}
// chaos.wgsl:3180
template <typename T> V2<T> var_acoth(V2<T> p, T w) {
  V2<T> rz = c_recip(p);
  V2<T> l = c_log(c_div(V2<T>{rz.x + (T)1, rz.y + (T)0}, V2<T>{(T)1 - rz.x, (T)0 - rz.y}));
  T k = w * (T)2 / K<T>::PI;
  return {((T)0.5 * l.y) * k, ((T)0.5 * l.x) * k};
}
// chaos.wgsl:3192 — one draw picks the sign.
template <typename T, typename R> V2<T> var_acosech(V2<T> p, T w, R& rng) {
  V2<T> rz = c_recip(p);
  V2<T> s = c_sqr(rz);
  V2<T> d = c_sqrt(V2<T>{s.x - (T)1, s.y - (T)0});
  V2<T> l = c_log(V2<T>{rz.x + d.x, rz.y + d.y});
  T k = w * (T)2 / K<T>::PI;
  T fx = l.y * k, fy = l.x * k;
  T sg = sel((T)1, (T)-1, rng.rand01() >= (T)0.5);
  return {sg * fx, sg * fy};
}
// chaos.wgsl:3206
template <typename T> V2<T> var_arcsech2(V2<T> p, T w) {
  V2<T> z = c_recip(p);
  V2<T> zs = c_sqrt(V2<T>{z.x - (T)1, z.y - (T)0});
  V2<T> za = c_sqrt(V2<T>{z.x + (T)1, z.y + (T)0});
  V2<T> m = c_mul(za, zs);
  V2<T> l = c_log(V2<T>{z.x + m.x, z.y + m.y});
  T k = w * (T)2 / K<T>::PI;
  T lx = l.x * k, ly = l.y * k;
  const bool neg = ly < (T)0;
  return {sel(-lx, lx, neg), sel(ly - (T)1, ly + (T)1, neg)};
}

// chaos.wgsl:3239
template <typename T>
V2<T> var_cell2(V2<T> p, T w, T size, T a, T snx, T sny, T ssx, T ssy) {
  T ss = sel(size, (T)1e-30, std::fabs(size) < (T)1e-30);
  T inv = a / ss;
  T cx = std::floor(p.x * inv), cy = std::floor(p.y * inv);
  T dx = p.x - cx * ss, dy = p.y - cy * ss;
  T sx = cx, sy = cy;
  if (sy >= (T)0) { sy = sy * sny; sx = sx * snx; }
  else { sy = -ssy * sy; sx = sx * ssx; }
  return {w * (dx + sx * ss), -w * (dy + sy * ss)};
}

// chaos.wgsl:3279
template <typename T> V2<T> var_curl_sp(V2<T> p, T w, T powp, T c1, T c2, T sx, T sy) {
  T pw = sel(powp, (T)1e-30, powp == (T)0);
  T pinv = (T)1 / pw;
  T c2x2 = (T)2 * c2;
  T x = std::pow(std::fabs(p.x), pw) * signv(p.x);
  T y = std::pow(std::fabs(p.y), pw) * signv(p.y);
  T d = x * x - y * y;
  T aa = c1 * x + c2 * d;
  T re = std::sqrt(aa * aa + sx * sx) * sel((T)-1, (T)1, aa > (T)0) + (T)1;
  T bb = c1 * y + c2x2 * x * y;
  T im = std::sqrt(bb * bb + sy * sy) * sel((T)-1, (T)1, bb > (T)0);
  T c = std::pow(std::fabs(re * re + im * im), pinv);
  T r = w / wmax(c, (T)1e-30);
  return {(x * re + y * im) * r, (y * re - x * im) * r};
}

// chaos.wgsl:3314
template <typename T> V2<T> var_murl2(V2<T> p, T w, T c, T pf) {
  T p2 = pf * (T)0.5;
  T sp = sel(pf, (T)1e-30, pf == (T)0);
  T invp = (T)1 / sp;
  T cp1 = c + (T)1;
  T vp = w * std::pow(std::fabs(cp1), (T)2 * invp) * sel((T)-1, (T)1, cp1 >= (T)0);
  T a1 = std::atan2(p.y, p.x) * sp;
  T r0 = c * std::pow(std::fabs(dot2(p)), p2);
  T re0 = r0 * safe_cos(a1) + (T)1, im0 = r0 * safe_sin(a1);
  T r1 = std::pow(std::fabs(re0 * re0 + im0 * im0), invp);
  T a2 = std::atan2(im0, re0) * (T)2 * invp;
  T re1 = r1 * safe_cos(a2), im1 = r1 * safe_sin(a2);
  T rl = vp / wmax(r1 * r1, (T)1e-30);
  return {rl * (p.x * re1 + p.y * im1), rl * (p.y * re1 - p.x * im1)};
}

// chaos.wgsl:3348 — two draws; a point on a Lissajous curve.
template <typename T, typename R>
V2<T> var_lissajous(V2<T> p, T w, T tmin, T tmax, T a, T b, T c, T d, T e, R& rng) {
  (void)p;
  T t = (tmax - tmin) * rng.rand01() + tmin;
  T yj = rng.rand01() - (T)0.5;
  T drift = c * t + e * yj;
  return {w * (safe_sin(a * t + d) + drift), w * (safe_sin(b * t) + drift)};
}

// chaos.wgsl:3369 — two draws; a point on a hypotrochoid.
template <typename T, typename R>
V2<T> var_spirograph(V2<T> p, T w, T a, T b, T d, T tmin, T tmax, T ymin, T ymax, T c1, T c2,
                     R& rng) {
  (void)p;
  T t = (tmax - tmin) * rng.rand01() + tmin;
  T yj = (ymax - ymin) * rng.rand01() + ymin;
  T ab = a + b;
  T sb = sel(b, (T)1e-30, std::fabs(b) < (T)1e-30);
  T ratio = ab / sb;
  T x1 = ab * safe_cos(t) - c1 * safe_cos(ratio * t);
  T y1 = ab * safe_sin(t) - c2 * safe_sin(ratio * t);
  return {w * (x1 + d * safe_cos(t) + yj), w * (y1 + d * safe_sin(t) + yj)};
}

// chaos.wgsl:3399 — one mode draw, then two or three more per mode
// (operands drawn left to right).
template <typename T, typename R>
V2<T> var_waffle(V2<T> p, T w, T slices_f, T xth, T yth, T rotation, R& rng) {
  (void)p;
  T slices = wmax((T)1, std::floor(std::fabs(slices_f)));
  T inv = (T)1 / slices;
  T vc = w * safe_cos(rotation), vs = w * safe_sin(rotation);
  const uint32_t mode = f2u(wmin((T)4, std::floor(rng.rand01() * (T)5)));
  T a, r;
  if (mode == 0u) {
    T k0 = std::floor(rng.rand01() * slices);
    T u0 = rng.rand01();
    a = (k0 + u0 * xth) * inv;
    T k1 = std::floor(rng.rand01() * slices);
    T u1 = rng.rand01();
    r = (k1 + u1 * yth) * inv;
  } else if (mode == 1u) {
    T k0 = std::floor(rng.rand01() * slices);
    T u0 = rng.rand01();
    a = (k0 + u0) * inv;
    r = (std::floor(rng.rand01() * slices) + yth) * inv;
  } else if (mode == 2u) {
    a = (std::floor(rng.rand01() * slices) + xth) * inv;
    T k1 = std::floor(rng.rand01() * slices);
    T u1 = rng.rand01();
    r = (k1 + u1) * inv;
  } else if (mode == 3u) {
    a = rng.rand01();
    T k1 = std::floor(rng.rand01() * slices);
    T u1 = rng.rand01();
    r = (k1 + yth + u1 * ((T)1 - yth)) * inv;
  } else {
    T k0 = std::floor(rng.rand01() * slices);
    T u0 = rng.rand01();
    a = (k0 + xth + u0 * ((T)1 - xth)) * inv;
    r = rng.rand01();
  }
  return {vc * a + vs * r, -vs * a + vc * r};
}

// ---- glynnSim family .. corners (chaos.wgsl:3453-4111) ------------------------

// chaos.wgsl:3453 — a disc of radius `radius` at angle phi1 filled by two
// draws; outside it, points are inverted into it with probability
// contrast * alpha^|pow| (one draw).
template <typename T, typename R>
V2<T> var_glynnSim1(V2<T> p, T w, T radius, T radius1, T phi1, T thick, T powp, T contrast,
                    R& rng) {
  T a = phi1 * (K<T>::PI / (T)180);
  T x1 = radius * safe_cos(a), y1 = radius * safe_sin(a);
  T ap = std::fabs(powp);
  T r = std::sqrt(dot2(p));
  if (r < radius) {
    T ri = radius1 * (thick + ((T)1 - thick) * rng.rand01());
    T phi = K<T>::TAU * rng.rand01();
    return {w * (ri * safe_cos(phi) + x1), w * (ri * safe_sin(phi) + y1)};
  }
  T alpha = radius / wmax(r, (T)1e-30);
  T x, y;
  if (rng.rand01() > contrast * std::pow(std::fabs(alpha), ap)) { x = p.x; y = p.y; }
  else { T a2 = alpha * alpha; x = a2 * p.x; y = a2 * p.y; }
  T dx = x - x1, dy = y - y1;
  if (dx * dx + dy * dy < radius1 * radius1) {
    T ri = radius1 * (thick + ((T)1 - thick) * rng.rand01());
    T phi = K<T>::TAU * rng.rand01();
    return {w * (ri * safe_cos(phi) + x1), w * (ri * safe_sin(phi) + y1)};
  }
  return {w * x, w * y};
}

// chaos.wgsl:3496 — an annular sector between phi1 and phi2.
template <typename T, typename R>
V2<T> var_glynnSim2(V2<T> p, T w, T radius, T thick, T contrast, T powp, T phi1, T phi2, R& rng) {
  T d2r = K<T>::PI / (T)180;
  T f1 = phi1 * d2r, f2 = phi2 * d2r;
  T rt = radius + thick;
  T gam = thick * ((T)2 * radius + thick) / wmax(rt, (T)1e-30);
  T delta = f2 - f1;
  T ap = std::fabs(powp);
  T r = std::sqrt(dot2(p));
  if (r < radius) {
    T ri = rt - gam * rng.rand01();
    T ph = f1 + delta * rng.rand01();
    return {w * (ri * safe_cos(ph)), w * (ri * safe_sin(ph))};
  }
  T alpha = radius / wmax(r, (T)1e-30);
  if (rng.rand01() > contrast * std::pow(std::fabs(alpha), ap)) return {w * p.x, w * p.y};
  T a2 = alpha * alpha;
  return {w * (a2 * p.x), w * (a2 * p.y)};
}

// chaos.wgsl:3527 — two concentric circles picked with probability gamma.
template <typename T, typename R>
V2<T> var_glynnSim3(V2<T> p, T w, T radius, T thick, T contrast, T powp, R& rng) {
  T r1 = radius + thick;
  T r2 = (radius * radius) / wmax(r1, (T)1e-30);
  T gam = r1 / wmax(r1 + r2, (T)1e-30);
  T ap = std::fabs(powp);
  T r = std::sqrt(dot2(p));
  if (r < r1) {
    T phi = K<T>::TAU * rng.rand01();
    T ri = sel(r2, r1, rng.rand01() < gam);
    return {w * (ri * safe_cos(phi)), w * (ri * safe_sin(phi))};
  }
  T alpha = radius / wmax(r, (T)1e-30);
  if (rng.rand01() > contrast * std::pow(std::fabs(alpha), ap)) return {w * p.x, w * p.y};
  T a2 = alpha * alpha;
  return {w * (a2 * p.x), w * (a2 * p.y)};
}

// chaos.wgsl:3562
template <typename T> V2<T> var_flipy(V2<T> p, T w) {
  return {w * p.x, w * p.y * sel((T)1, (T)-1, p.x > (T)0)};
}

// chaos.wgsl:3571
template <typename T> V2<T> var_eclipse(V2<T> p, T w, T shift) {
  if (!(std::fabs(p.y) <= w)) return {w * p.x, w * p.y};
  T c2 = std::sqrt(wmax(w * w - p.y * p.y, (T)0));
  T ox;
  if (std::fabs(p.x) <= c2) {
    T xs = p.x + shift * w;
    ox = std::fabs(xs) >= c2 ? -w * p.x : w * xs;
  } else {
    ox = w * p.x;
  }
  return {ox, w * p.y};
}

// chaos.wgsl:3596 — barycentric coordinates against basis (a,b), (c,d).
template <typename T> V2<T> var_barycentroid(V2<T> p, T w, T a, T b, T c, T d) {
  T d00 = a * a + b * b, d01 = a * c + b * d, d02 = a * p.x + b * p.y;
  T d11 = c * c + d * d, d12 = c * p.x + d * p.y;
  T den = d00 * d11 - d01 * d01;
  if (std::fabs(den) < (T)1e-30) return {w * p.x, w * p.y};
  T inv = (T)1 / den;
  T u = (d11 * d02 - d01 * d12) * inv;
  T v = (d00 * d12 - d01 * d02) * inv;
  return {w * (std::sqrt(u * u + p.x * p.x) * signv(u)), w * (std::sqrt(v * v + p.y * p.y) * signv(v))};
}

// chaos.wgsl:3624 — keep the point on one side of a conic, else (0, 0).
template <typename T> V2<T> var_chunk(V2<T> p, T w, T a, T b, T c, T d, T e, T f, T modep) {
  T aa = w * a, bb = w * b, cc = w * c, dd = w * d, ee = w * e, ff = w * f;
  T r = aa * p.x * p.x + bb * p.x * p.y + cc * p.y * p.y + dd * p.x + ee * p.y + ff;
  const int32_t mode = f2i(modep);
  if (mode == 0 && r <= (T)0) return {p.x, p.y};
  if (mode == 1 && r > (T)0) return {p.x, p.y};
  return {(T)0, (T)0};
}

// chaos.wgsl:3668
template <typename T> V2<T> var_ennepers(V2<T> p, T w) {
  T x = p.x, y = p.y;
  return {w * (x - (x * x * x) / (T)3) + x * y * y, w * (y - (y * y * y) / (T)3) + y * x * x};
}

// chaos.wgsl:3681 — Abramowitz-Stegun 7.1.26 rational approximation.
template <typename T> T erf_as(T x) {
  const T a1 = (T)0.254829592, a2 = (T)-0.284496736, a3 = (T)1.421413741;
  const T a4 = (T)-1.453152027, a5 = (T)1.061405429, pp = (T)0.3275911;
  T s = sel((T)-1, (T)1, x >= (T)0);
  T ax = std::fabs(x);
  T t = (T)1 / ((T)1 + pp * ax);
  T poly = (((((a5 * t) + a4) * t + a3) * t + a2) * t + a1) * t;
  return s * ((T)1 - poly * std::exp(-ax * ax));
}
// chaos.wgsl:3695
template <typename T> V2<T> var_erf(V2<T> p, T w) { return {w * erf_as(p.x), w * erf_as(p.y)}; }

// chaos.wgsl:3703
template <typename T> V2<T> var_circus(V2<T> p, T w, T scale) {
  T s1 = sel((T)1 / scale, (T)1e30, scale == (T)0);
  T r = std::sqrt(p.x * p.x + p.y * p.y);
  T a = std::atan2(p.y, p.x);
  T sa = std::sin(a), ca = std::cos(a);
  T rs = sel(r * s1, r * scale, r <= (T)1);
  return {w * rs * ca, w * rs * sa};
}

// chaos.wgsl:3722 — one draw where the disc and the corner region overlap.
template <typename T, typename R> V2<T> var_asteria(V2<T> p, T w, T alpha, R& rng) {
  T sa = safe_sin(K<T>::PI * alpha), ca = safe_cos(K<T>::PI * alpha);
  T x0 = w * p.x, y0 = w * p.y;
  T r = x0 * x0 + y0 * y0;
  T xx = (std::fabs(x0) - (T)1) * (std::fabs(x0) - (T)1);
  T yy = (std::fabs(y0) - (T)1) * (std::fabs(y0) - (T)1);
  T r2 = std::sqrt(yy + xx);
  const bool inside = r < (T)1;
  bool keep;
  if (inside && r2 < (T)1) keep = rng.rand01() > (T)0.35;
  else keep = !inside;
  if (keep) return {x0, y0};
  T rx = x0 * ca - y0 * sa, ry = x0 * sa + y0 * ca;
  T den = std::sqrt(wmax((T)1 - ry * ry, (T)1e-30));
  T in = wmax((T)1 - (-std::fabs(ry) + (T)1) * (-std::fabs(ry) + (T)1), (T)0);
  T nx = rx / den * ((T)1 - std::sqrt(in));
  return {nx * ca + ry * sa, -nx * sa + ry * ca};
}

// chaos.wgsl:3769
template <typename T> V2<T> var_clifford_js(V2<T> p, T w, T a, T b, T c, T d) {
  return {w * (safe_sin(a * p.y) + c * safe_cos(a * p.x)),
          w * (safe_sin(b * p.x) + d * safe_cos(b * p.y))};
}

// chaos.wgsl:3785
template <typename T>
V2<T> var_devil_warp(V2<T> p, T w, T a, T b, T effect, T warp, T rmin, T rmax) {
  (void)w;
  T x = p.x, y = p.y;
  T r2 = (T)1 / wmax(x * x + y * y, (T)1e-30);
  T ba = x * x + r2 * b * y * y;
  T bb = y * y + r2 * a * x * x;
  T pa = sel((T)0, std::pow(ba, warp), ba > (T)0);
  T pb = sel((T)0, std::pow(bb, warp), bb > (T)0);
  T r = clampv(pa - pb, rmin, rmax);
  r = effect * r;
  return {x * ((T)1 + r), y * ((T)1 + r)};
}

// chaos.wgsl:3808 — integer lattice noise in [0, 1] (i32 arithmetic wraps).
template <typename T> T voron_noise(int32_t x) {
  uint32_t s = ((uint32_t)x << 13) ^ (uint32_t)x;
  uint32_t r = (s * (s * s * 15731u + 789221u) + 1376312589u) & 0x7fffffffu;
  return (T)(float)(int32_t)r * (T)(float)(1.0 / 2147483647.0);
}
// chaos.wgsl:3815 — nearest of up to 26 hashed feature points per cell.
template <typename T> V2<T> var_voron(V2<T> p, T w, T k, T step, T num_f, T xs_f, T ys_f) {
  const int32_t xs = f2i(xs_f), ys = f2i(ys_f);
  T ss = sel(step, (T)1e-30, step == (T)0);
  const int32_t M = f2i(std::floor(p.x / ss)), N = f2i(std::floor(p.y / ss));
  T rmin = (T)20, X0 = 0, Y0 = 0;
  for (int32_t i = -1; i < 2; ++i) {
    const int32_t M1 = iwrap_add(M, i);
    for (int32_t j = -1; j < 2; ++j) {
      const int32_t N1 = iwrap_add(N, j);
      const int32_t h = iwrap_add(iwrap_add(iwrap_mul(19, M1), iwrap_mul(257, N1)), xs);
      const int32_t kx = iwrap_add(1, f2i(std::floor(num_f * voron_noise<T>(h))));
      const int32_t cnt = imin(imax(kx, 1), 26);
      for (int32_t l = 0; l < cnt; ++l) {
        const int32_t hx = iwrap_add(iwrap_add(iwrap_add(l, iwrap_mul(64, M1)), iwrap_mul(15, N1)), xs);
        const int32_t hy = iwrap_add(iwrap_add(iwrap_add(l, iwrap_mul(21, M1)), iwrap_mul(33, N1)), ys);
        T X = (voron_noise<T>(hx) + (T)M1) * step;
        T Y = (voron_noise<T>(hy) + (T)N1) * step;
        T ox = p.x - X, oy = p.y - Y;
        T r = std::sqrt(ox * ox + oy * oy);
        if (r < rmin) { rmin = r; X0 = X; Y0 = Y; }
      }
    }
  }
  return {w * (k * (p.x - X0) + X0), w * (k * (p.y - Y0) + Y0)};
}

// chaos.wgsl:3865
template <typename T> V2<T> var_henon(V2<T> p, T w, T a, T b, T c) {
  return {w * (c - a * p.x * p.x + p.y), w * b * p.x};
}

// chaos.wgsl:3874
template <typename T> V2<T> var_atan(V2<T> p, T w, T modep, T stretch) {
  T nrm = w * (T)2 / K<T>::PI;
  const int32_t mode = f2i(modep);
  if (mode == 0) return {w * p.x, nrm * std::atan(stretch * p.y)};
  if (mode == 1) return {nrm * std::atan(stretch * p.x), w * p.y};
  if (mode == 2) return {nrm * std::atan(stretch * p.x), nrm * std::atan(stretch * p.y)};
  return {(T)0, (T)0};
}

// chaos.wgsl:3896
template <typename T> V2<T> var_cardioid(V2<T> p, T w, T a) {
  T th = std::atan2(p.y, p.x);
  T rs = p.x * p.x + p.y * p.y + safe_sin(a * th) + (T)1;
  T r = w * std::sqrt(wmax(rs, (T)0));
  return {r * std::cos(th), r * std::sin(th)};
}

// chaos.wgsl:3910 — one draw: a point on Fay's chrysanthemum curve.
template <typename T, typename R> V2<T> var_chrysanthemum(V2<T> p, T w, R& rng) {
  (void)p;
  T u = (T)21 * K<T>::PI * rng.rand01();
  T p4 = std::sin((T)17 * u / (T)3);
  T p8 = std::sin((T)2 * std::cos((T)3 * u) - (T)28 * u);
  T p44 = p4 * p4 * p4 * p4;
  T p82 = p8 * p8;
  T p88 = p82 * p82 * p82 * p82;
  T rr = (T)5 * ((T)1 + std::sin((T)11 * u / (T)5)) - (T)4 * p44 * p88;
  T r = w * (T)0.1 * rr;
  return {r * std::cos(u), r * std::sin(u)};
}

// chaos.wgsl:3928 — bipolar coordinates with angular folding.
template <typename T> V2<T> var_bcollide(V2<T> p, T w, T nump, T a) {
  T num = wmax((T)1, nump);
  T bnpi = num / K<T>::PI, pibn = K<T>::PI / num, bab = K<T>::PI * a / num;
  T xp = p.x + (T)1, xm = p.x - (T)1, y2 = p.y * p.y;
  T tau = (T)0.5 * (std::log(wmax(xp * xp + y2, (T)1e-30)) - std::log(wmax(xm * xm + y2, (T)1e-30)));
  T sr = K<T>::PI - std::atan2(p.y, xp) - std::atan2(p.y, (T)1 - p.x);
  const int32_t alt = f2i(sr * bnpi);
  T off = sel(-bab, bab, (alt & 1) == 0);
  T folded = sr + off - std::floor((sr + off) / pibn) * pibn;
  T sg = (T)alt * pibn + folded;
  T tmp = std::cosh(tau) - safe_cos(sg);
  T ts = sel(tmp, (T)1e-30, std::fabs(tmp) < (T)1e-30);
  return {w * std::sinh(tau) / ts, w * safe_sin(sg) / ts};
}

// chaos.wgsl:3969
template <typename T> V2<T> var_bsplit(V2<T> p, T w, T sx, T sy) {
  T ax = p.x + sx;
  T sn = std::sin(ax);
  if (std::fabs(sn) < (T)1e-6) return {(T)0, (T)0};
  T cs = std::cos(ax);
  T tn = sn / cs;
  T tsafe = sel(tn, signv(tn) * (T)1e6, std::fabs(cs) < (T)1e-6);
  return {w / tsafe * safe_cos(p.y + sy), w / sn * ((T)-1 * p.y + sy)};
}

// chaos.wgsl:3990
template <typename T> V2<T> var_bulge(V2<T> p, T w, T n) {
  T rs = wmax(std::sqrt(p.x * p.x + p.y * p.y), (T)1e-30);
  T sc = std::pow(rs, n) / rs;
  return {w * p.x * sc, w * p.y * sc};
}

// chaos.wgsl:4012 — two draws, always taken.
template <typename T, typename R> V2<T> var_checks(V2<T> p, T w, T cx, T cy, T size, T rnd, R& rng) {
  T cs = (T)1 / (size + (T)1e-6);
  const int32_t par = iwrap_add(f2i(rint_even(p.x * cs)), f2i(rint_even(p.y * cs)));
  T rnx = rnd * rng.rand01();
  T rny = rnd * rng.rand01();
  T dx, dy;
  if ((par & 1) == 0) { dx = -cx + rnx; dy = -cy; }
  else { dx = cx; dy = cy + rny; }
  return {w * (p.x + dx), w * (p.y + dy)};
}

// chaos.wgsl:4046 — one draw plus a sin-hash of the position.
template <typename T, typename R> V2<T> var_circular(V2<T> p, T w, T angle, T seed, R& rng) {
  T ca = angle * K<T>::PI / (T)180;
  T ar = std::sin(p.x * (T)12.9898 + p.y * (T)78.233 + seed) * (T)43758.5453;
  T aux = ar - std::trunc(ar);
  T rnd = ((T)2 * (rng.rand01() + aux) - (T)2) * ca;
  T rad = std::sqrt(p.x * p.x + p.y * p.y);
  T ang = std::atan2(p.y, p.x);
  return {w * std::cos(ang + rnd) * rad, w * std::sin(ang + rnd) * rad};
}
// chaos.wgsl:4066
template <typename T, typename R>
V2<T> var_circular2(V2<T> p, T w, T angle, T seed, T xx, T yy, R& rng) {
  T ca = angle * K<T>::PI / (T)180;
  T ar = std::sin(p.x * xx + p.y * yy + seed) * (T)43758.5453;
  T aux = ar - std::trunc(ar);
  T rnd = ((T)2 * (rng.rand01() + aux) - (T)2) * ca;
  T rad = std::sqrt(p.x * p.x + p.y * p.y);
  T ang = std::atan2(p.y, p.x);
  return {w * std::cos(ang + rnd) * rad, w * std::sin(ang + rnd) * rad};
}

// chaos.wgsl:4081
template <typename T>
V2<T> var_corners(V2<T> p, T w, T cx, T cy, T mx, T my, T xp, T yp, T xyp, T logm, T logb) {
  T xs = p.x * p.x, ys = p.y * p.y;
  T ex, ey;
  if (logm == (T)0) {
    ex = std::pow(wmax(xs, (T)0), xp + xyp) * mx;
    ey = std::pow(wmax(ys, (T)0), yp + xyp) * my;
  } else {
    T lb = std::log(wmax(std::fabs(logb), (T)1.000001));
    ex = std::pow(std::log((xs * mx) + (T)3) / lb, xp + (T)2.25 + xyp) - (T)1.33;
    ey = std::pow(std::log((ys * my) + (T)3) / lb, yp + (T)2.25 + xyp) - (T)1.33;
  }
  return {sel(-w * ex - cx, w * ex + cx, p.x > (T)0), sel(-w * ey - cy, w * ey + cy, p.y > (T)0)};
}

// ---- circleblur .. twoface (chaos.wgsl:4113-4812) -----------------------------

// chaos.wgsl:4113 — two draws: a uniform point in the disc.
template <typename T, typename R> V2<T> var_circleblur(V2<T> p, T w, R& rng) {
  (void)p;
  T rad = std::sqrt(rng.rand01());
  T a = rng.rand01() * (T)2 * K<T>::PI;
  return {w * std::cos(a) * rad, w * std::sin(a) * rad};
}

// chaos.wgsl:4130 — Binet-form Fibonacci spiral.
template <typename T> V2<T> var_fibonacci2(V2<T> p, T w, T sc, T sc2) {
  const T ffive = (T)0.447213595, lnphi = (T)0.481211825;
  T a = p.y * lnphi;
  T s1 = safe_sin(a), c1 = safe_cos(a);
  T b = (p.x * K<T>::PI + p.y * lnphi) * (T)-1;
  T s2 = safe_sin(b), c2 = safe_cos(b);
  T e1 = sc * std::exp(sc2 * (p.x * lnphi));
  T e2 = sc * std::exp(sc2 * ((p.x * lnphi - p.y * K<T>::PI) * (T)-1));
  return {w * (e1 * c1 - e2 * c2) * ffive, w * (e1 * s1 - e2 * s2) * ffive};
}

// Hyperbolic {p, q} tiling radius (chaos.wgsl:4152-4220); `form` picks the
// reference's two algebraically equal spellings of r^2.
template <typename T> T hypertile_r(T pp, T q, bool form1, T& pa) {
  T pi_ = wmax((T)3, pp), qs = wmax((T)3, q);
  pa = (T)2 * K<T>::PI / pi_;
  T r2;
  if (!form1) {
    T qa = (T)2 * K<T>::PI / qs;
    T den = std::cos(pa) + std::cos(qa);
    r2 = sel((T)1, ((T)1 - std::cos(pa)) / den + (T)1, std::fabs(den) > (T)1e-30);
  } else {
    T cp = std::cos(pa);
    T den = cp + std::cos((T)2 * K<T>::PI / qs);
    r2 = sel((T)1, (T)1 - (cp - (T)1) / den, std::fabs(den) > (T)1e-30);
  }
  return sel((T)1, (T)1 / std::sqrt(wmax(r2, (T)1e-30)), r2 > (T)0);
}
// chaos.wgsl:4152 — Moebius step of the tiling group by the nth rotation.
template <typename T> V2<T> var_hypertile(V2<T> p, T w, T pp, T q, T n) {
  T pa;
  T r = hypertile_r(pp, q, false, pa);
  T an = n * pa;
  T re = r * safe_cos(an), im = r * safe_sin(an);
  T a = p.x + re, b = p.y - im;
  T c = re * p.x - im * p.y + (T)1, d = re * p.y + im * p.x;
  T vr = w / wmax(c * c + d * d, (T)1e-30);
  return {vr * (a * c + b * d), vr * (b * c - a * d)};
}
// chaos.wgsl:4176 — one draw picks one of ten rotations.
template <typename T, typename R> V2<T> var_hypertile1(V2<T> p, T w, T pp, T q, R& rng) {
  T pa;
  T r = hypertile_r(pp, q, true, pa);
  T rpa = std::floor(rng.rand01() * (T)10) * pa;
  T re = r * std::cos(rpa), im = r * std::sin(rpa);
  T a = p.x + re, b = p.y - im;
  T c = re * p.x - im * p.y + (T)1, d = re * p.y + im * p.x;
  T vr = w / wmax(c * c + d * d, (T)1e-30);
  return {vr * (a * c + b * d), vr * (b * c - a * d)};
}
// chaos.wgsl:4200 — one draw picks the rotation applied after the step.
template <typename T, typename R> V2<T> var_hypertile2(V2<T> p, T w, T pp, T q, R& rng) {
  T pa;
  T r = hypertile_r(pp, q, true, pa);
  T a = p.x + r, b = p.y;
  T c = r * p.x + (T)1, d = r * p.y;
  T xx = a * c + b * d, yy = b * c - a * d;
  T vr = w / wmax(c * c + d * d, (T)1e-30);
  T rpa = std::floor(rng.rand01() * (T)32767) * pa;
  T ca = safe_cos(rpa), sa = safe_sin(rpa);
  return {vr * (xx * ca + yy * sa), vr * (yy * ca - xx * sa)};
}

// chaos.wgsl:4226
template <typename T> V2<T> var_idisc(V2<T> p, T w) {
  T r = std::sqrt(p.x * p.x + p.y * p.y);
  T a = K<T>::PI / (r + (T)1);
  T v = std::atan2(p.y, p.x) * w / K<T>::PI;
  return {v * std::cos(a), v * std::sin(a)};
}

// chaos.wgsl:4243
template <typename T> V2<T> var_hole(V2<T> p, T w, T a, T insidep) {
  T al = std::atan2(p.y, p.x);
  T delta = std::pow(wmax(al / K<T>::PI + (T)1, (T)1e-30), a);
  T ss = p.x * p.x + p.y * p.y;
  T r = sel(w * std::sqrt(wmax(ss + delta, (T)0)), w * delta / wmax(ss + delta, (T)1e-30),
            f2i(insidep) != 0);
  return {r * std::cos(al), r * std::sin(al)};
}

// chaos.wgsl:4260 — cos/sin of 45 radians, folded to f32 constants.
template <typename T>
V2<T> var_kaleidoscope(V2<T> p, T w, T pull, T rotate, T lineup, T offx, T offy) {
  const T c45 = (T)0.5253219888177297, s45 = (T)0.8509035245341184;
  T ox = w * (((rotate * p.x) * c45 - p.y * s45 + lineup) + offx);
  T oy = sel(w * ((rotate * p.y) * c45 + p.x * s45 - pull - lineup),
             w * (((rotate * p.y) * c45 + p.x * s45 + pull + lineup) + offy), p.y > (T)0);
  return {ox, oy};
}

// chaos.wgsl:4275
template <typename T> V2<T> var_layered_spiral(V2<T> p, T w, T radius) {
  T a = p.x * radius;
  T t = p.x * p.x + p.y * p.y + (T)1e-30;
  return {w * a * safe_cos(t), w * a * safe_sin(t)};
}

// chaos.wgsl:4283
template <typename T> V2<T> var_linear_t(V2<T> p, T w, T px, T py) {
  T sx = sel((T)-1, (T)1, p.x >= (T)0), sy = sel((T)-1, (T)1, p.y >= (T)0);
  return {w * sx * std::pow(wmax(std::fabs(p.x), (T)1e-30), px),
          w * sy * std::pow(wmax(std::fabs(p.y), (T)1e-30), py)};
}

// chaos.wgsl:4296 — one draw along a fixed direction.
template <typename T, typename R> V2<T> var_line(V2<T> p, T w, T delta, T phi, R& rng) {
  (void)p;
  T cd = safe_cos(delta * K<T>::PI), sd = safe_sin(delta * K<T>::PI);
  T cp = safe_cos(phi * K<T>::PI), sp = safe_sin(phi * K<T>::PI);
  T ux = cd * cp, uy = sd * cp, uz = sp;
  T r = std::sqrt(wmax(ux * ux + uy * uy + uz * uz, (T)1e-30));
  ux = ux / r;
  uy = uy / r;
  T t = rng.rand01() * w;
  return {ux * t, uy * t};
}

// chaos.wgsl:4323
template <typename T> V2<T> var_ovoid(V2<T> p, T w, T px, T py) {
  T r = w / (p.x * p.x + p.y * p.y + (T)1e-6);
  return {p.x * r * px, p.y * r * py};
}

// chaos.wgsl:4331 — one draw picks the root.
template <typename T, typename R>
V2<T> var_phoenix_julia(V2<T> p, T w, T power, T dist, T xd, T yd, R& rng) {
  T ps = sel(power, (T)1, power == (T)0);
  T inv_n = dist / ps, inv2 = (T)2 * K<T>::PI / ps;
  T cn = dist / ((T)2 * ps);
  T px = p.x * (xd + (T)1), py = p.y * (yd + (T)1);
  T k = std::floor(rng.rand01() * (T)32768);
  T a = std::atan2(py, px) * inv_n + k * inv2;
  T r = w * std::pow(wmax(p.x * p.x + p.y * p.y, (T)1e-30), cn);
  return {r * safe_cos(a), r * safe_sin(a)};
}

// chaos.wgsl:4359
template <typename T> V2<T> var_unpolar(V2<T> p, T w) {
  T v = (w / K<T>::PI) * (T)0.5;
  T r = std::exp(p.y);
  return {v * r * safe_sin(p.x), v * r * safe_cos(p.x)};
}

// chaos.wgsl:4368
template <typename T> V2<T> var_shredrad(V2<T> p, T w, T n, T widthp) {
  T width = clampv(widthp, (T)-1, (T)1);
  T alpha = (T)2 * K<T>::PI / wmax((T)0.001, n);
  T ang = std::atan2(p.y, p.x);
  T rad = std::sqrt(p.x * p.x + p.y * p.y);
  T xa = (ang + (T)3 * K<T>::PI + alpha * (T)0.5) / alpha;
  T za = ((xa - std::floor(xa)) * width + std::floor(xa)) * alpha - K<T>::PI - alpha * (T)0.5 * width;
  return {w * rad * safe_cos(za), w * rad * safe_sin(za)};
}

// chaos.wgsl:4395 — one draw: a seed of Vogel's sunflower model.
template <typename T, typename R> V2<T> var_vogel(V2<T> p, T w, T n, T scale, R& rng) {
  const T phi = (T)1.61803398874989;
  T k2 = (T)2 * K<T>::PI / (phi * phi);
  T i = std::floor(rng.rand01() * wmax((T)1, n)) + (T)1;
  T a = i * k2;
  T r = w * (std::sqrt(p.x * p.x + p.y * p.y) + std::sqrt(i));
  return {r * (safe_cos(a) + scale * p.x), r * (safe_sin(a) + scale * p.y)};
}

// chaos.wgsl:4414 — one draw only in dual mode inside the unit disc.
template <typename T, typename R>
V2<T> var_yin_yang(V2<T> p, T w, T radius, T ang1, T ang2, T dualp, T outp, R& rng) {
  T sa = safe_sin(K<T>::PI * ang1), ca = safe_cos(K<T>::PI * ang1);
  T sb = safe_sin(K<T>::PI * ang2), cb = safe_cos(K<T>::PI * ang2);
  const int32_t dual = f2i(dualp), outside = f2i(outp);
  T x = p.x, y = p.y;
  T inv = (T)1, RR = radius;
  if (x * x + y * y < (T)1) {
    T nx = x * ca - y * sa, ny = x * sa + y * ca;
    if (dual == 1 && rng.rand01() > (T)0.5) {
      inv = (T)-1;
      RR = (T)1 - radius;
      nx = x * cb - y * sb;
      ny = x * sb + y * cb;
    }
    x = nx;
    y = ny;
    if (y > (T)0) {
      T t = std::sqrt(wmax((T)1 - y * y, (T)0));
      T k = x / wmax(std::fabs(t), (T)1e-30) * signv(t);
      T t1 = (t - (T)0.5) * (T)2;
      T al = ((T)1 - k) * (T)0.5;
      T be = (T)1 - al;
      T dx = al * (RR - (T)1);
      T k1 = al * RR + be * (T)1;
      return {w * (t1 * k1 + dx) * inv, w * std::sqrt(wmax((T)1 - t1 * t1, (T)0)) * k1 * inv};
    }
    return {w * (x * ((T)1 - RR) + RR) * inv, w * (y * ((T)1 - RR)) * inv};
  }
  if (outside == 1) return {w * p.x, w * p.y};
  return {(T)0, (T)0};
}

// chaos.wgsl:4466 — one draw picks the winding.
template <typename T, typename R> V2<T> var_squish(V2<T> p, T w, T powp, R& rng) {
  T power = wmax((T)2, powp);
  T ip = (T)1 / power;
  T ax = std::fabs(p.x), ay = std::fabs(p.y);
  T s, pp;
  if (ax > ay) { s = ax; pp = sel((T)4 * s - p.y, p.y, p.x > (T)0); }
  else { s = ay; pp = sel((T)6 * s + p.x, (T)2 * s - p.x, p.y > (T)0); }
  T rot = std::floor(power * rng.rand01());
  pp = ip * (pp + (T)8 * s * rot);
  if (pp <= (T)1 * s) return {w * s, w * pp};
  if (pp <= (T)3 * s) return {w * ((T)2 * s - pp), w * s};
  if (pp <= (T)5 * s) return {-w * s, w * ((T)4 * s - pp)};
  if (pp <= (T)7 * s) return {-w * ((T)6 * s - pp), -w * s};
  return {w * s, -w * ((T)8 * s - pp)};
}

// chaos.wgsl:4501
template <typename T> V2<T> var_target(V2<T> p, T w, T even, T odd, T size) {
  T h = (T)0.5 * size;
  T a = std::atan2(p.y, p.x);
  T r = std::sqrt(p.x * p.x + p.y * p.y);
  T t = std::log(wmax(r, (T)1e-30));
  if (t < (T)0) t = t - h;
  T at = std::fabs(t);
  T ss = wmax(std::fabs(size), (T)1e-30);
  t = at - std::floor(at / ss) * ss;
  a = a + sel(odd, even, t < h);
  return {r * safe_cos(a) * w, r * safe_sin(a) * w};
}

// chaos.wgsl:4530
template <typename T> V2<T> var_funnel(V2<T> p, T w, T effect) {
  T cx = safe_cos(p.x), cy = safe_cos(p.y);
  T secx = (T)1 / sel(cx, (T)1e-30, std::fabs(cx) < (T)1e-30);
  T secy = (T)1 / sel(cy, (T)1e-30, std::fabs(cy) < (T)1e-30);
  T off = effect * K<T>::PI;
  return {w * safe_tanh(p.x) * (secx + off), w * safe_tanh(p.y) * (secy + off)};
}

// chaos.wgsl:4550
template <typename T> V2<T> var_holesq(V2<T> p, T w) {
  T x = w * p.x, y = w * p.y;
  T fx = std::fabs(x), fy = std::fabs(y);
  if (fx + fy > (T)1) return {x, y};
  if (fx > fy) return {sel((x + fy - (T)1) * (T)0.5, (x - fy + (T)1) * (T)0.5, x >= (T)0), y};
  return {x, sel((y + fx - (T)1) * (T)0.5, (y - fx + (T)1) * (T)0.5, y >= (T)0)};
}

// chaos.wgsl:4568 — ten hole shapes.
template <typename T>
V2<T> var_hole2(V2<T> p, T w, T a, T b, T c, T d, T insidep, T shapep) {
  T rho = p.x * p.x + p.y * p.y;
  T th = std::atan2(p.y, p.x) * d;
  T delta = std::pow(wmax(th / K<T>::PI + (T)1, (T)1e-30), a) * c;
  T r1;
  switch (f2i(shapep)) {
    case 0: r1 = std::sqrt(wmax(rho, (T)0)) + delta; break;
    case 1: r1 = std::sqrt(wmax(rho + delta, (T)0)); break;
    case 2: r1 = std::sqrt(wmax(rho + safe_sin(b * th) + delta, (T)0)); break;
    case 3: r1 = std::sqrt(wmax(rho + safe_sin(th) + delta, (T)0)); break;
    case 4: r1 = std::sqrt(wmax(rho + th * th - delta + (T)1, (T)0)); break;
    case 5: r1 = std::sqrt(wmax(rho + std::fabs(safe_tan(th)) + delta, (T)0)); break;
    case 6: r1 = std::sqrt(wmax(rho * ((T)1 + safe_sin(b * th)) + delta, (T)0)); break;
    case 7: r1 = std::sqrt(wmax(rho + std::fabs(safe_sin((T)0.5 * b * th)) + delta, (T)0)); break;
    case 8: r1 = std::sqrt(wmax(rho + safe_sin(K<T>::PI * safe_sin(b * th)) + delta, (T)0)); break;
    case 9:
      r1 = std::sqrt(wmax(
          rho + (safe_sin(b * th) + safe_sin((T)2 * b * th + K<T>::PI * (T)0.5)) * (T)0.5 + delta, (T)0));
      break;
    default: r1 = (T)1; break;
  }
  T rf = sel(w * r1, w / wmax(r1, (T)1e-30), f2i(insidep) != 0);
  return {rf * safe_cos(th), rf * safe_sin(th)};
}

// chaos.wgsl:4604 — one draw picks one of four inversion centres.
template <typename T, typename R> V2<T> var_lace_js(V2<T> p, T w, R& rng) {
  const T r = (T)2, h3 = (T)0.8660254037844386;
  T r0 = std::sqrt(p.x * p.x + p.y * p.y);
  T u = rng.rand01();
  T x, y;
  if (u > (T)0.75) {
    T th = std::atan2(p.y, p.x - (T)1);
    y = -r0 * std::cos(th) / r + (T)1;
    x = -r0 * std::sin(th) / r;
  } else if (u > (T)0.5) {
    T th = std::atan2(p.y - h3, p.x + (T)0.5);
    y = -r0 * std::cos(th) / r - (T)0.5;
    x = -r0 * std::sin(th) / r + h3;
  } else if (u > (T)0.25) {
    T th = std::atan2(p.y + h3, p.x + (T)0.5);
    y = -r0 * std::cos(th) / r - (T)0.5;
    x = -r0 * std::sin(th) / r - h3;
  } else {
    T th = std::atan2(p.y, p.x);
    y = -r0 * std::cos(th) / r;
    x = -r0 * std::sin(th) / r;
  }
  return {w * x, w * y};
}

// chaos.wgsl:4633 — modes 0/1 draw once for the sign of the root.
template <typename T, typename R>
V2<T> var_julia_outside(V2<T> p, T w, T rediv, T imdiv, T modep, R& rng) {
  const int32_t mode = f2i(modep);
  const bool sq = mode == 0 || mode == 2, rt = mode == 0 || mode == 1;
  V2<T> z = p, z2 = p;
  if (sq) z = c_sqrt(z);
  z = V2<T>{z.x + (T)1, z.y};
  if (sq) z = c_sqr(z);
  if (sq) z2 = c_sqrt(z2);
  z2 = V2<T>{z2.x - (T)1, z2.y};
  if (sq) z2 = c_sqr(z2);
  z = c_div(z, z2);
  if (rt) z = c_sqrt(z);
  z = c_div(z, V2<T>{rediv, imdiv});
  if (rt) {
    T sg = sel((T)-1, (T)1, rng.rand01() < (T)0.5);
    return {w * sg * z.x, w * sg * z.y};
  }
  return {w * z.x, w * z.y};
}

// chaos.wgsl:4680 — a different map per quadrant.
template <typename T>
V2<T> var_fourth(V2<T> p, T w, T spin, T space, T twist, T offx, T offy) {
  T ww = w * w;
  if (p.x > (T)0 && p.y > (T)0) {
    T th = std::atan2(p.y, p.x);
    T r = (T)1 / wmax(std::sqrt(p.x * p.x + p.y * p.y), (T)1e-30);
    return {w * r * std::cos(th), w * r * std::sin(th)};
  }
  if (p.x > (T)0 && p.y < (T)0) {
    T r2 = p.x * p.x + p.y * p.y;
    if (r2 < ww) {
      T r = w * std::sqrt(wmax(ww / wmax(r2, (T)1e-30) - (T)1, (T)0));
      return {r * p.x, r * p.y};
    }
    return {w * p.x, w * p.y};
  }
  if (p.x < (T)0 && p.y > (T)0) {
    T xx = p.x - offx, yy = p.y + offy;
    T r0 = std::sqrt(xx * xx + yy * yy);
    if (r0 < w) {
      T th = std::atan2(yy, xx) + spin + twist * (w - r0);
      T r = w * r0;
      return {r * safe_cos(th) + offx, r * safe_sin(th) - offy};
    }
    T r = w * ((T)1 + space / wmax(r0, (T)1e-30));
    return {r * xx + offx, r * yy - offy};
  }
  return {w * p.x, w * p.y};
}

// chaos.wgsl:4728
template <typename T> V2<T> var_pulse(V2<T> p, T w, T fx, T fy, T sx, T sy) {
  return {w * (p.x + sx * safe_sin(p.x * fx)), w * (p.y + sy * safe_sin(p.y * fy))};
}

// chaos.wgsl:4738
template <typename T> V2<T> var_rays1(V2<T> p, T w) {
  T t = p.x * p.x + p.y * p.y;
  T tv = safe_tan(std::sqrt(wmax(t, (T)1e-30)));
  T it = (T)1 / sel(tv, (T)1e-30, std::fabs(tv) < (T)1e-30);
  T u = it + w * ((T)2 / K<T>::PI) * ((T)2 / K<T>::PI);
  return {w * u * t / sel(p.x, (T)1e-30, p.x == (T)0), w * u * t / sel(p.y, (T)1e-30, p.y == (T)0)};
}
// chaos.wgsl:4751
template <typename T> V2<T> var_rays2(V2<T> p, T w) {
  T t = p.x * p.x + p.y * p.y;
  T ts = wmax(t, (T)1e-30);
  T in = (ts + (T)1e-6) * safe_tan((T)1 / ts + (T)1e-6);
  T ci = safe_cos(in);
  T u = (T)1 / sel(ci, (T)1e-30, std::fabs(ci) < (T)1e-30);
  T co = w / (T)10;
  return {co * u * t / sel(p.x, (T)1e-30, p.x == (T)0), co * u * t / sel(p.y, (T)1e-30, p.y == (T)0)};
}
// chaos.wgsl:4766
template <typename T> V2<T> var_rays3(V2<T> p, T w) {
  T t = p.x * p.x + p.y * p.y;
  T ts = wmax(t, (T)1e-30);
  T in = safe_sin(t * t + (T)1e-6) * safe_sin((T)1 / (ts * ts) + (T)1e-6);
  T den = std::sqrt(wmax(safe_cos(in), (T)1e-30));
  T u = (T)1 / wmax(den, (T)1e-30);
  T co = w / (T)10;
  return {co * u * safe_cos(t) * t / sel(p.x, (T)1e-30, p.x == (T)0),
          co * u * safe_tan(t) * t / sel(p.y, (T)1e-30, p.y == (T)0)};
}

// chaos.wgsl:4786
template <typename T> V2<T> var_tancos(V2<T> p, T w) {
  T d1 = (T)1e-6 + p.x * p.x + p.y * p.y;
  T d2 = w / d1;
  return {d2 * safe_tanh(d1) * (T)2 * p.x, d2 * safe_cos(d1) * (T)2 * p.y};
}

// chaos.wgsl:4802
template <typename T> V2<T> var_twoface(V2<T> p, T w) {
  T v = w;
  if (p.x > (T)0) v = v / wmax(p.x * p.x + p.y * p.y, (T)1e-30);
  return {v * p.x, v * p.y};
}

// ---- elliptic / bipolar family .. circlecrop (chaos.wgsl:4814-5363) -----------

// Elliptic coordinates (mu, nu) of p shared by e_julia / e_collide / e_mod
// (chaos.wgsl:4814-4936): returns xmax = cosh(mu) (>= 1) and nu in [0, pi].
template <typename T> T ell_xmax(T x, T r2) {
  T tmp = r2 + (T)1, tmp2 = (T)2 * x;
  T sa = std::sqrt(wmax(tmp + tmp2, (T)0)), sb = std::sqrt(wmax(tmp - tmp2, (T)0));
  T xmax = (sa + sb) * (T)0.5;
  return xmax < (T)1 ? (T)1 : xmax;
}
template <typename T> T ell_nu(T x, T xmax) {
  T t = x / xmax;
  if (t > (T)1) t = (T)1;
  if (t < (T)-1) t = (T)-1;
  return std::acos(t);
}

// chaos.wgsl:4814 — one draw picks the root.
template <typename T, typename R> V2<T> var_e_julia(V2<T> p, T w, T powp, R& rng) {
  T ps = sel(powp, (T)1, powp == (T)0);
  T pa = std::fabs(ps);
  T r2 = p.y * p.y + p.x * p.x;
  T x = p.x;
  if (!(ps > (T)0)) {
    r2 = (T)1 / wmax(r2, (T)1e-30);
    x = p.x * r2;
  }
  T xmax = ell_xmax(x, r2);
  T mu0 = std::acosh(xmax);
  T nu = ell_nu(x, xmax);
  if (p.y < (T)0) nu = -nu;
  T k = std::floor(rng.rand01() * pa);
  nu = nu / ps + ((T)2 * K<T>::PI / ps) * k;
  T mu = mu0 / ps;
  return {w * std::cosh(mu) * safe_cos(nu), w * std::sinh(mu) * safe_sin(nu)};
}

// chaos.wgsl:4857 — one draw when filled > 0, a second when it fills.
template <typename T, typename R> V2<T> var_cannabis_curve_wf(V2<T> p, T w, T filled, R& rng) {
  T a = std::atan2(p.y, p.x);
  T r = ((T)1 + (T)0.9 * std::cos((T)8 * a)) * ((T)1 + (T)0.1 * std::cos((T)24 * a)) *
        ((T)0.9 + (T)0.1 * std::cos((T)200 * a)) * ((T)1 + std::sin(a));
  a = a + K<T>::PI * (T)0.5;
  if (filled > (T)0 && filled > rng.rand01()) r = r * rng.rand01();
  return {w * std::sin(a) * r, w * std::cos(a) * r};
}

// chaos.wgsl:4868
template <typename T> V2<T> var_e_collide(V2<T> p, T w, T nump, T a) {
  T num = wmax((T)1, nump);
  T knpi = num / K<T>::PI, pikn = K<T>::PI / num, ka = K<T>::PI * a / num;
  T xmax = ell_xmax(p.x, p.y * p.y + p.x * p.x);
  T nu = ell_nu(p.x, xmax);
  const int32_t alt = f2i(nu * knpi);
  T off = sel(-ka, ka, (alt & 1) == 0);
  T arg = nu + off;
  nu = (T)alt * pikn + (arg - std::floor(arg / pikn) * pikn);
  if (p.y <= (T)0) nu = -nu;
  return {w * xmax * std::cos(nu),
          w * std::sqrt(wmax(xmax - (T)1, (T)0)) * std::sqrt(xmax + (T)1) * std::sin(nu)};
}

// chaos.wgsl:4897
template <typename T> V2<T> var_e_mod(V2<T> p, T w, T radius, T distance) {
  T xmax = ell_xmax(p.x, p.y * p.y + p.x * p.x);
  T mu = std::acosh(xmax);
  T nu = ell_nu(p.x, xmax);
  if (p.y < (T)0) nu = -nu;
  if (mu < radius && -mu < radius) {
    T rs = wmax(radius, (T)1e-30);
    T two = (T)2 * rs;
    if (nu > (T)0) mu = wmod(mu + rs + distance * rs, two) - rs;
    else mu = wmod(mu - rs - distance * rs, two) + rs;
  }
  return {w * std::cosh(mu) * std::cos(nu), w * std::sinh(mu) * std::sin(nu)};
}

// Tiled band shared by intersection / parallel: the coordinate folded into
// [-m1, m1) with period m2*m1 (chaos.wgsl:4938-5204, `%` spelling).
template <typename T> T band_fold_mod(T v, T m1, T period) {
  if (v > m1) return -m1 + wmod(v + m1, wmax(period, (T)1e-30));
  if (v < -m1) return m1 - wmod(m1 - v, wmax(period, (T)1e-30));
  return v;
}
// The same fold written with floor (parallel).
template <typename T> T band_fold_floor(T v, T m1, T period) {
  T rs = wmax(period, (T)1e-30);
  if (v > m1) { T a = v + m1; return -m1 + (a - std::floor(a / rs) * rs); }
  if (v < -m1) { T a = m1 - v; return m1 - (a - std::floor(a / rs) * rs); }
  return v;
}
// Random integer step along one axis: three draws (branch already taken).
template <typename T, typename R> T band_jump(T width, R& rng) {
  T x = sel(-width, width, rng.rand01() < (T)0.5);
  return rint_even(x * std::log(wmax(rng.rand01(), (T)1e-30)));
}

// chaos.wgsl:4938 — one branch draw, then two more.
template <typename T, typename R>
V2<T> var_intersection(V2<T> p, T w, T xwidth, T xtile, T xm1, T xm2, T xheight, T yheight,
                       T ytile, T ym1, T ym2, T ywidth, R& rng) {
  if (w == (T)0) return {(T)0, (T)0};
  T xr1 = xm2 * xm1, yr1 = ym2 * ym1;
  if (rng.rand01() < (T)0.5) {
    T ox = xtile * (p.x + band_jump(xwidth, rng));
    return {ox, xheight * band_fold_mod(p.y, xm1, xr1)};
  }
  T oy = ytile * (p.y + band_jump(yheight, rng));
  return {ywidth * band_fold_mod(p.x, ym1, yr1), oy};
}

// chaos.wgsl:4985
template <typename T> V2<T> var_inv_squircular(V2<T> p, T w) {
  if (w == (T)0) return {(T)0, (T)0};
  const T sq2 = (T)1.41421356237;
  T u = p.x, v = p.y;
  T ri = u * u + v * v;
  T r2 = std::sqrt(wmax(ri * (w * w * ri - (T)4 * u * u * v * v) / w, (T)0));
  T r = std::sqrt(wmax(ri - r2, (T)0)) / sq2;
  return {r / sel(u, (T)1e-30, u == (T)0), r / sel(v, (T)1e-30, v == (T)0)};
}

// chaos.wgsl:5008
template <typename T> V2<T> var_lozi(V2<T> p, T w, T a, T b, T c) {
  return {w * (c - a * std::fabs(p.x) + p.y), w * b * p.x};
}

// chaos.wgsl:5014
template <typename T> V2<T> var_hypershift(V2<T> p, T w, T shift, T stretch) {
  T scale = (T)1 - shift * shift;
  T r1 = (T)1 / wmax(p.x * p.x + p.y * p.y, (T)1e-30);
  T x = r1 * p.x + shift, y = r1 * p.y;
  T rad = w * scale / wmax(x * x + y * y, (T)1e-30);
  return {rad * x + shift, rad * y * stretch};
}

// chaos.wgsl:5026 — offset from the nearest hexagon centre (cube rounding).
template <typename T> V2<T> var_hex_modulus(V2<T> p, T w, T size) {
  const T h32 = (T)0.8660254037844386, sq3 = (T)1.7320508075688772;
  T hs = h32 / wmax(size, (T)1e-30);
  T wt = w / h32;
  T X = p.x * hs, Y = p.y * hs;
  T x = (T)0.5773502691896258 * X - Y / (T)3;
  T z = (T)2 * Y / (T)3;
  T y = -x - z;
  T rx = rint_even(x), ry = rint_even(y), rz = rint_even(z);
  T dx = std::fabs(rx - x), dy = std::fabs(ry - y), dz = std::fabs(rz - z);
  if (dx > dy && dx > dz) rx = -ry - rz;
  else if (dy > dz) ry = -rx - rz;
  else rz = -rx - ry;
  (void)ry;
  T fx = sq3 * rx + h32 * rz, fy = (T)1.5 * rz;
  return {(X - fx) * wt, (Y - fy) * wt};
}

// chaos.wgsl:5067 — one draw decides whether the point is pushed to a border.
template <typename T, typename R> V2<T> var_boarders2(V2<T> p, T w, T c, T left, T right, R& rng) {
  T ac = wmax(std::fabs(c), (T)1e-6);
  T al = wmax(std::fabs(left), (T)1e-6), ar = wmax(std::fabs(right), (T)1e-6);
  T cl = ac * al, cr = ac + ac * ar;
  T rx = rint_even(p.x), ry = rint_even(p.y);
  T ox = p.x - rx, oy = p.y - ry;
  if (rng.rand01() >= cr) return {w * (ox * ac + rx), w * (oy * ac + ry)};
  T sx = sel(ox, (T)1e-30, ox == (T)0), sy = sel(oy, (T)1e-30, oy == (T)0);
  if (std::fabs(ox) >= std::fabs(oy)) {
    if (ox >= (T)0) return {w * (ox * ac + rx + cl), w * (oy * ac + ry + cl * oy / sx)};
    return {w * (ox * ac + rx - cl), w * (oy * ac + ry - cl * oy / sx)};
  }
  if (oy >= (T)0) return {w * (ox * ac + rx + ox / sy * cl), w * (oy * ac + ry + cl)};
  return {w * (ox * ac + rx - ox / sy * cl), w * (oy * ac + ry - cl)};
}

// Bipolar coordinates (tau, sigma) of p (chaos.wgsl:5107-5145).
template <typename T> T bip_tau(V2<T> p) {
  T xp = p.x + (T)1, xm = p.x - (T)1, y2 = p.y * p.y;
  return (T)0.5 * (std::log(wmax(xp * xp + y2, (T)1e-30)) - std::log(wmax(xm * xm + y2, (T)1e-30)));
}
template <typename T> T bip_sigma(V2<T> p) {
  return K<T>::PI - std::atan2(p.y, p.x + (T)1) - std::atan2(p.y, (T)1 - p.x);
}
// chaos.wgsl:5107
template <typename T> V2<T> var_b_mod(V2<T> p, T w, T radius, T distance) {
  T tau = bip_tau(p);
  T sg = bip_sigma(p);
  if (tau < radius && -tau < radius) {
    T rs = wmax(radius, (T)1e-30);
    T two = (T)2 * rs;
    T arg = tau + rs + distance * rs;
    tau = arg - std::floor(arg / two) * two - rs;
  }
  T tmp = std::cosh(tau) - std::cos(sg);
  T ts = sel(tmp, (T)1e-30, std::fabs(tmp) < (T)1e-30);
  return {w * std::sinh(tau) / ts, w * std::sin(sg) / ts};
}
// chaos.wgsl:5127 — one draw picks the root.
template <typename T, typename R>
V2<T> var_b_transform(V2<T> p, T w, T rotate, T powp, T move, T split, R& rng) {
  T power = wmax((T)1, powp);
  T tau = bip_tau(p) / power + move;
  T sg = bip_sigma(p) + rotate;
  T k = std::floor(rng.rand01() * power);
  sg = sg / power + ((T)2 * K<T>::PI / power) * k;
  tau = tau + sel(-split, split, p.x >= (T)0);
  T tmp = std::cosh(tau) - safe_cos(sg);
  T ts = sel(tmp, (T)1e-30, std::fabs(tmp) < (T)1e-30);
  return {w * std::sinh(tau) / ts, w * safe_sin(sg) / ts};
}

// chaos.wgsl:5147 — like intersection, with two horizontal band sets.
template <typename T, typename R>
V2<T> var_parallel(V2<T> p, T w, T x1w, T x1t, T x1m1, T x1m2, T x1h, T x1mv, T x2w, T x2t, T x2m1,
                   T x2m2, R& rng) {
  const T x2h = (T)0.5, x2mv = (T)1;
  if (w == (T)0) return {(T)0, (T)0};
  T xr1 = x1m2 * x1m1, xr2 = x2m2 * x2m1;
  if (rng.rand01() < (T)0.5) {
    T ox = x1t * (p.x + band_jump(x1w, rng));
    return {ox, x1h * band_fold_floor(p.y, x1m1, xr1) + w * x1mv};
  }
  T ox = x2t * (p.x + band_jump(x2w, rng));
  return {ox, x2h * band_fold_floor(p.y, x2m1, xr2) - w * x2mv};
}

// chaos.wgsl:5206
template <typename T> V2<T> var_waves3(V2<T> p, T w, T sx, T sy, T fx, T fy, T sxf, T syf) {
  T sxx = (T)0.5 * sx * ((T)1 + safe_sin(p.y * sxf));
  T syy = (T)0.5 * sy * ((T)1 + safe_sin(p.x * syf));
  return {w * (p.x + safe_sin(p.y * fx) * sxx), w * (p.y + safe_sin(p.x * fy) * syy)};
}

// chaos.wgsl:5219 — per-row sin-hash amplitude.
template <typename T> V2<T> var_waves4(V2<T> p, T w, T sx, T sy, T fx, T fy, T contp, T yfact) {
  T cell = std::floor(p.y * fx / K<T>::TAU);
  T ax = std::sin(cell * (T)12.9898 + cell * (T)78.233 + (T)1 + p.y * (T)0.001 * yfact) * (T)43758.5453;
  ax = ax - std::trunc(ax);
  if (contp > (T)0.5) ax = sel((T)0, (T)1, ax > (T)0.5);
  return {w * (p.x + safe_sin(p.y * fx) * ax * ax * sx), w * (p.y + safe_sin(p.x * fy) * sy)};
}

// chaos.wgsl:5237
template <typename T> V2<T> var_scry2(V2<T> p, T w, T sides_f, T star, T circle) {
  const int32_t sides = imin(imax(f2i(sides_f), 1), 50);
  T a = K<T>::TAU / (T)sides;
  T sa = safe_sin(a), ca = safe_cos(a);
  T sins = safe_sin(-star * K<T>::PI * (T)0.5), coss = safe_cos(star * K<T>::PI * (T)0.5);
  T sinc = safe_sin(circle * K<T>::PI * (T)0.5), cosc = safe_cos(circle * K<T>::PI * (T)0.5);
  T xr = p.x, yr = p.y;
  T r2 = xr * coss + std::fabs(yr) * sins;
  T cr = std::sqrt(xr * xr + yr * yr);
  int32_t iters = 0;
  for (int32_t i = 0; i < sides - 1; ++i) {
    T nx = xr * ca - yr * sa;
    yr = xr * sa + yr * ca;
    xr = nx;
    r2 = wmax(r2, xr * coss + std::fabs(yr) * sins);
    iters = i + 1;
  }
  r2 = r2 * cosc + cr * sinc;
  T r1 = r2;
  if (iters > 1) r2 = r2 * r2;
  else r2 = std::fabs(r2) * r2;
  if (w == (T)0) return {(T)0, (T)0};
  T d = r1 * (r2 + (T)1 / w);
  if (d == (T)0) return {(T)0, (T)0};
  T r = (T)1 / d;
  return {p.x * r, p.y * r};
}

// chaos.wgsl:5281
template <typename T> V2<T> var_ennepers2(V2<T> p, T w, T ap, T bp, T cp) {
  T x = p.x, y = p.y;
  T r2 = (T)1 / wmax(x * x + y * y, (T)1e-30);
  T dxy = (ap * x) * (ap * x) - (bp * y) * (bp * y);
  return {w * x * (ap * ap - dxy * r2 - cp * std::sqrt(std::fabs(x))),
          w * y * (bp * bp - dxy * r2 - cp * std::sqrt(std::fabs(y)))};
}

// chaos.wgsl:5294 — one draw picks one of three Moebius maps.
template <typename T, typename R> V2<T> var_apollony(V2<T> p, T w, R& rng) {
  const T sq3 = (T)1.7320508075688772;
  T dx = (T)1 + sq3 - p.x;
  T den = wmax(dx * dx + p.y * p.y, (T)1e-30);
  T a0 = (T)3 * dx / den - ((T)1 + sq3) / ((T)2 + sq3);
  T b0 = (T)3 * p.y / den;
  T ab = wmax(a0 * a0 + b0 * b0, (T)1e-30);
  T fx = a0 / ab, fy = -b0 / ab;
  const int32_t br = imod(f2i(std::floor((T)4 * rng.rand01())), 3);
  T x, y;
  if (br == 0) { x = a0; y = b0; }
  else if (br == 1) { x = -fx * (T)0.5 - fy * sq3 * (T)0.5; y = fx * sq3 * (T)0.5 - fy * (T)0.5; }
  else { x = -fx * (T)0.5 + fy * sq3 * (T)0.5; y = -fx * sq3 * (T)0.5 - fy * (T)0.5; }
  return {w * x, w * y};
}

// chaos.wgsl:5326 — one draw (scatter band), always taken.
template <typename T, typename R>
V2<T> var_circlecrop(V2<T> p, T w, T radius, T cx, T cy, T scatter, T zerop, R& rng) {
  T ca = clampv(scatter, (T)-1, (T)1);
  T dx = p.x - cx, dy = p.y - cy;
  T rad = std::sqrt(dx * dx + dy * dy);
  T ang = std::atan2(dy, dx);
  T rdc = radius + (rng.rand01() * (T)0.5 * ca);
  const bool esc = rad > radius, zero = zerop > (T)0.5;
  if (zero && esc) return {(T)0, (T)0};
  if (!zero && esc) return {w * rdc * std::cos(ang) + cx, w * rdc * std::sin(ang) + cy};
  return {w * dx + cx, w * dy + cy};
}

// ---- conformal / escape-time maps (chaos.wgsl:5365-5520) ----------------------

// chaos.wgsl:5365 — z^k by binary exponentiation.
template <typename T> V2<T> c_pow_int(V2<T> z, int32_t k) {
  V2<T> res{(T)1, (T)0}, base = z;
  for (int32_t e = k; e > 0; e >>= 1) {
    if ((e & 1) == 1) res = c_mul(res, base);
    base = c_sqr(base);
  }
  return res;
}
// chaos.wgsl:5383 — one Newton step for z^n = 1.
template <typename T> V2<T> var_newton(V2<T> p, T w, T nin) {
  const int32_t n = imin(imax(f2i(nin + (T)0.5), 2), 8);
  V2<T> zn = c_pow_int(p, n), zm = c_pow_int(p, n - 1);
  T nf = (T)n, nm1 = (T)(n - 1);
  V2<T> num{nm1 * zn.x + (T)1, nm1 * zn.y};
  V2<T> den{nf * zm.x, nf * zm.y};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_div(num, den);
  return {w * q.x, w * q.y};
}
// chaos.wgsl:5438
template <typename T> V2<T> var_burning_ship(V2<T> p, T w, T cx, T cy) {
  V2<T> s = c_sqr(V2<T>{std::fabs(p.x), std::fabs(p.y)});
  return {w * (s.x + cx), w * (s.y + cy)};
}
// chaos.wgsl:5444
template <typename T> V2<T> var_magnet1(V2<T> p, T w, T cx, T cy) {
  V2<T> s = c_sqr(p);
  V2<T> num{s.x + cx - (T)1, s.y + cy - (T)0};
  V2<T> den{(T)2 * p.x + cx - (T)2, (T)2 * p.y + cy - (T)0};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_sqr(c_div(num, den));
  return {w * q.x, w * q.y};
}
// chaos.wgsl:5455 — relaxed Newton step for z^3 = 1 plus c.
template <typename T> V2<T> var_nova(V2<T> p, T w, T cx, T cy, T relax) {
  V2<T> z3 = c_pow_int(p, 3), z2 = c_sqr(p);
  V2<T> num{z3.x - (T)1, z3.y - (T)0};
  V2<T> den{(T)3 * z2.x, (T)3 * z2.y};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_div(num, den);
  T sx = relax * q.x, sy = relax * q.y;
  return {w * (p.x - sx + cx), w * (p.y - sy + cy)};
}
// chaos.wgsl:5467 — Halley step for z^3 = 1 plus c.
template <typename T> V2<T> var_halley(V2<T> p, T w, T cx, T cy) {
  V2<T> z3 = c_pow_int(p, 3), z2 = c_sqr(p);
  V2<T> f{z3.x - (T)1, z3.y - (T)0};
  V2<T> fp{(T)3 * z2.x, (T)3 * z2.y};
  V2<T> fpp{(T)6 * p.x, (T)6 * p.y};
  V2<T> m = c_mul(f, fp);
  V2<T> num{(T)2 * m.x, (T)2 * m.y};
  V2<T> s = c_sqr(fp), m2 = c_mul(f, fpp);
  V2<T> den{(T)2 * s.x - m2.x, (T)2 * s.y - m2.y};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_div(num, den);
  return {w * (p.x - q.x + cx), w * (p.y - q.y + cy)};
}
// chaos.wgsl:5506 — Blaschke factor z (z - a) / (1 - conj(a) z).
template <typename T> V2<T> var_blaschke(V2<T> p, T w, T ax, T ay) {
  V2<T> num = c_mul(p, V2<T>{p.x - ax, p.y - ay});
  V2<T> m = c_mul(V2<T>{ax, -ay}, p);
  V2<T> den{(T)1 - m.x, (T)0 - m.y};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_div(num, den);
  return {w * q.x, w * q.y};
}

// chaos.wgsl:5518 — Cayley map (z - is) / (z + is).
template <typename T> V2<T> var_cayley(V2<T> p, T w, T s) {
  V2<T> den{p.x + (T)0, p.y + s};
  if (den.x * den.x + den.y * den.y < (T)1e-20) return {w * p.x, w * p.y};
  V2<T> q = c_div(V2<T>{p.x - (T)0, p.y - s}, den);
  return {w * q.x, w * q.y};
}

// chaos.wgsl:5532 — Lanczos (g = 7, 9 terms) complex Gamma with reflection.
template <typename T> V2<T> var_complex_gamma(V2<T> p, T w, T scale) {
  const T sq2pi = (T)2.5066282746310002;
  const T kC[8] = {(T)(float)676.5203681218851,     (T)(float)-1259.1392167224028,
                   (T)(float)771.32342877765313,    (T)(float)-176.61502916214059,
                   (T)(float)12.507343278686905,    (T)(float)-0.13857109526572012,
                   (T)(float)9.9843695780195716e-6, (T)(float)1.5056327351493116e-7};
  V2<T> z = p;
  const bool refl = z.x < (T)0.5;
  if (refl) z = V2<T>{(T)1 - z.x, -z.y};
  V2<T> x{z.x - (T)1, z.y - (T)0};
  V2<T> A{(T)0.99999999999980993, (T)0};
  for (int i = 0; i < 8; ++i) {
    V2<T> d = c_div(V2<T>{kC[i], (T)0}, V2<T>{x.x + (T)(i + 1), x.y + (T)0});
    A = V2<T>{A.x + d.x, A.y + d.y};
  }
  V2<T> t{x.x + ((T)7 + (T)0.5), x.y + (T)0};
  V2<T> tp = c_pow(t, V2<T>{x.x + (T)0.5, x.y + (T)0});
  V2<T> en = c_exp(V2<T>{-t.x, -t.y});
  V2<T> m = c_mul(tp, c_mul(en, A));
  V2<T> res{sq2pi * m.x, sq2pi * m.y};
  if (refl) {
    V2<T> sp = c_sin(V2<T>{(T)3.14159265 * p.x, (T)3.14159265 * p.y});
    res = c_div(c_div(V2<T>{(T)3.14159265, (T)0}, sp), res);
  }
  T k = w * scale;
  return {k * res.x, k * res.y};
}

// chaos.wgsl:5581 — Halley iterations for W e^W = z (1..4 steps).
template <typename T> V2<T> var_lambert_w(V2<T> p, T w, T itin) {
  const int32_t iters = imin(imax(f2i(itin + (T)0.5), 1), 4);
  V2<T> wn;
  if (len2(p) < (T)2) {
    wn = c_log(V2<T>{(T)1 + p.x, (T)0 + p.y});
  } else {
    V2<T> lz = c_log(p), llz = c_log(lz);
    wn = V2<T>{lz.x - llz.x, lz.y - llz.y};
  }
  for (int32_t i = 0; i < iters; ++i) {
    V2<T> ew = c_exp(wn);
    V2<T> we = c_mul(wn, ew);
    V2<T> f{we.x - p.x, we.y - p.y};
    V2<T> w1{wn.x + (T)1, wn.y + (T)0}, w2{wn.x + (T)2, wn.y + (T)0};
    V2<T> t1{(T)2 * w1.x, (T)2 * w1.y};
    if (t1.x * t1.x + t1.y * t1.y < (T)1e-20) break;
    V2<T> in = c_div(c_mul(w2, f), t1);
    V2<T> fp = c_mul(ew, w1);
    V2<T> den{fp.x - in.x, fp.y - in.y};
    if (den.x * den.x + den.y * den.y < (T)1e-20) break;
    V2<T> q = c_div(f, den);
    wn = V2<T>{wn.x - q.x, wn.y - q.y};
  }
  return {w * wn.x, w * wn.y};
}

// ---- map projections (chaos.wgsl:5621-5669) -----------------------------------

// chaos.wgsl:5621
template <typename T> V2<T> var_mercator(V2<T> p, T w) {
  T lat = clampv(p.y, (T)-1.5, (T)1.5);
  return {w * p.x, w * std::log(std::fabs(safe_tan((T)0.78539816 + lat * (T)0.5)) + (T)1e-6)};
}
// chaos.wgsl:5628
template <typename T> V2<T> var_lambert(V2<T> p, T w) {
  T k = std::sqrt((T)2 / ((T)1 + safe_cos(p.y) * safe_cos(p.x) + (T)1e-6));
  return {w * (k * safe_cos(p.y) * safe_sin(p.x)), w * (k * safe_sin(p.y))};
}
// chaos.wgsl:5635 — four Newton steps for the auxiliary angle.
template <typename T> V2<T> var_mollweide(V2<T> p, T w) {
  T t = p.y;
  T tgt = (T)3.14159265 * safe_sin(p.y);
  for (int i = 0; i < 4; ++i) {
    T s2 = safe_sin((T)2 * t), c2 = safe_cos((T)2 * t);
    T f = (T)2 * t + s2 - tgt;
    T df = (T)2 + (T)2 * c2;
    if (std::fabs(df) < (T)1e-6) break;
    t = t - f / df;
  }
  return {w * ((T)0.9003163 * p.x * safe_cos(t)), w * ((T)1.4142135 * safe_sin(t))};
}
// chaos.wgsl:5652
template <typename T> V2<T> var_hammer(V2<T> p, T w) {
  T z = std::sqrt((T)1 + safe_cos(p.y) * safe_cos(p.x * (T)0.5) + (T)1e-6);
  return {w * (((T)2.828427 * safe_cos(p.y) * safe_sin(p.x * (T)0.5)) / z),
          w * (((T)1.4142135 * safe_sin(p.y)) / z)};
}
// chaos.wgsl:5660
template <typename T> V2<T> var_stereographic(V2<T> p, T w) {
  T k = (T)2 / ((T)1 + safe_cos(p.y) * safe_cos(p.x) + (T)1e-6);
  return {w * (k * safe_cos(p.y) * safe_sin(p.x)), w * (k * safe_sin(p.y))};
}

// ---- one-step attractor maps, folds, toral maps (chaos.wgsl:5671-5808) ----------

// chaos.wgsl:5671 — Chirikov standard map.
template <typename T> V2<T> var_standard_map(V2<T> p, T w, T k) {
  T xp = p.x + k * safe_sin(p.y);
  return {w * xp, w * (p.y + xp)};
}
// chaos.wgsl:5678
template <typename T> V2<T> var_de_jong(V2<T> p, T w, T a, T b, T c, T d) {
  return {w * (safe_sin(a * p.y) - safe_cos(b * p.x)), w * (safe_sin(c * p.x) - safe_cos(d * p.y))};
}
// chaos.wgsl:5685
template <typename T> V2<T> var_ikeda(V2<T> p, T w, T u) {
  T t = (T)0.4 - (T)6 / ((T)1 + p.x * p.x + p.y * p.y);
  T st = safe_sin(t), ct = safe_cos(t);
  return {w * ((T)1 + u * (p.x * ct - p.y * st)), w * (u * (p.x * st + p.y * ct))};
}
// chaos.wgsl:5699 — reflect each coordinate into [-limit, limit].
template <typename T> T box_fold1(T v, T lim) {
  if (v > lim) return (T)2 * lim - v;
  if (v < -lim) return (T)-2 * lim - v;
  return v;
}
template <typename T> V2<T> var_box_fold(V2<T> p, T w, T limit) {
  return {w * box_fold1(p.x, limit), w * box_fold1(p.y, limit)};
}
// Mandelbox sphere-fold scale (chaos.wgsl:5710-5745).
template <typename T> T sphere_scale(T r2, T rmin, T rmax) {
  T a = rmin * rmin, b = rmax * rmax;
  if (r2 < a) return b / a;
  if (r2 < b) return b / wmax(r2, (T)1e-6);
  return (T)1;
}
// chaos.wgsl:5710
template <typename T> V2<T> var_sphere_fold(V2<T> p, T w, T rmin, T rmax) {
  T sc = sphere_scale(p.x * p.x + p.y * p.y, rmin, rmax);
  return {w * (p.x * sc), w * (p.y * sc)};
}
// chaos.wgsl:5724
template <typename T> V2<T> var_mandelbox_step(V2<T> p, T w, T scale, T rmin, T rmax, T cx, T cy) {
  T x = box_fold1(p.x, (T)1), y = box_fold1(p.y, (T)1);
  T sf = sphere_scale(x * x + y * y, rmin, rmax);
  return {w * (x * sf * scale + cx), w * (y * sf * scale + cy)};
}
// chaos.wgsl:5747 — mirror the angle into one wedge of n.
template <typename T> V2<T> var_kifs_fold(V2<T> p, T w, T n, T offset) {
  T r = len2(p);
  if (r < (T)1e-6) return {(T)0, (T)0};
  T a = std::atan2(p.y, p.x) - offset;
  T th = (T)6.283185307179586 / wmax((T)1, std::fabs(n));
  a = a - th * std::floor(a / th);
  if (a > th * (T)0.5) a = th - a;
  a = a + offset;
  return {w * (r * safe_cos(a)), w * (r * safe_sin(a))};
}
// chaos.wgsl:5765
template <typename T> V2<T> var_arnold_cat(V2<T> p, T w) {
  T xp = p.x * (T)2 + p.y, yp = p.x + p.y;
  return {w * (fract(xp + (T)0.5) - (T)0.5), w * (fract(yp + (T)0.5) - (T)0.5)};
}
// chaos.wgsl:5772
template <typename T> V2<T> var_bakers_map(V2<T> p, T w) {
  T x = fract(p.x + (T)0.5), y = fract(p.y + (T)0.5);
  T xp, yp;
  if (x < (T)0.5) { xp = (T)2 * x; yp = y * (T)0.5; }
  else { xp = (T)2 * x - (T)1; yp = y * (T)0.5 + (T)0.5; }
  return {w * (xp - (T)0.5), w * (yp - (T)0.5)};
}
// chaos.wgsl:5788
template <typename T> V2<T> var_tent_map(V2<T> p, T w) {
  T x = fract(p.x + (T)0.5), y = fract(p.y + (T)0.5);
  T xp = (T)1 - std::fabs((T)1 - (T)2 * x), yp = (T)1 - std::fabs((T)1 - (T)2 * y);
  return {w * (xp - (T)0.5), w * (yp - (T)0.5)};
}
// chaos.wgsl:5797
template <typename T> V2<T> var_logistic_map(V2<T> p, T w, T r) {
  T x = fract(p.x + (T)0.5), y = fract(p.y + (T)0.5);
  return {w * (r * x * ((T)1 - x) - (T)0.5), w * (r * y * ((T)1 - y) - (T)0.5)};
}

// ---- plane curves, flows, spirals, series (chaos.wgsl:5810-5969) ----------------

// Polar point (r, theta) scaled by w.
template <typename T> V2<T> polar_w(T w, T r, T th) { return {w * (r * safe_cos(th)), w * (r * safe_sin(th))}; }

// chaos.wgsl:5810
template <typename T> V2<T> var_superellipse(V2<T> p, T w, T a, T b, T n) {
  T th = std::atan2(p.y, p.x);
  T sa = wmax(std::fabs(a), (T)1e-4), sb = wmax(std::fabs(b), (T)1e-4);
  T sn = wmax(n, (T)0.01);
  T c = std::fabs(safe_cos(th) / sa), s = std::fabs(safe_sin(th) / sb);
  T r = std::pow(wmax(std::pow(c, sn) + std::pow(s, sn), (T)1e-6), (T)-1 / sn);
  return polar_w(w, r, th);
}
// chaos.wgsl:5822
template <typename T> V2<T> var_limacon(V2<T> p, T w, T a, T b) {
  T th = std::atan2(p.y, p.x);
  return polar_w(w, b + a * safe_cos(th), th);
}
// chaos.wgsl:5829
template <typename T> V2<T> var_epicycloid(V2<T> p, T w, T k) {
  T t = std::atan2(p.y, p.x);
  T k1 = k + (T)1;
  return {w * (k1 * safe_cos(t) - safe_cos(k1 * t)), w * (k1 * safe_sin(t) - safe_sin(k1 * t))};
}
// chaos.wgsl:5838
template <typename T> V2<T> var_catenary(V2<T> p, T w, T a) {
  T sa = wmax(std::fabs(a), (T)1e-4) * signv(a + (T)1e-8);
  return {w * p.x, w * (sa * std::cosh(p.x / sa))};
}
// chaos.wgsl:5845
template <typename T> V2<T> var_tractrix(V2<T> p, T w) {
  return {w * (p.x - safe_tanh(p.x)), w * ((T)1 / std::cosh(p.x))};
}
// chaos.wgsl:5852
template <typename T> V2<T> var_tinkerbell(V2<T> p, T w, T a, T b, T c, T d) {
  return {w * (p.x * p.x - p.y * p.y + a * p.x + b * p.y), w * ((T)2 * p.x * p.y + c * p.x + d * p.y)};
}
// chaos.wgsl:5859 — one Euler step of the forced Duffing oscillator.
template <typename T> V2<T> var_duffing(V2<T> p, T w, T h, T delta, T gamma, T omega) {
  T xp = p.x + h * p.y;
  T yp = p.y + h * (p.x - p.x * p.x * p.x - delta * p.y + gamma * safe_cos(omega * p.x));
  return {w * xp, w * yp};
}
// chaos.wgsl:5868
template <typename T> V2<T> var_vanderpol(V2<T> p, T w, T h, T mu) {
  return {w * (p.x + h * p.y), w * (p.y + h * (mu * ((T)1 - p.x * p.x) * p.y - p.x))};
}
// chaos.wgsl:5875 — Euler step with |p| as the third coordinate.
template <typename T> V2<T> var_rossler(V2<T> p, T w, T h, T a) {
  T z = len2(p);
  return {w * (p.x + h * (-p.y - z)), w * (p.y + h * (p.x + a * p.y))};
}
// chaos.wgsl:5883
template <typename T> V2<T> var_droste(V2<T> p, T w, T s) {
  T r = wmax(len2(p), (T)1e-6);
  T th = std::atan2(p.y, p.x);
  T k = std::log(wmax(s, (T)1e-4)) / (T)6.283185307179586;
  T re = std::log(r) - k * th;
  T im = th + k * std::log(r);
  T er = std::exp(re);
  return {w * (er * safe_cos(im)), w * (er * safe_sin(im))};
}
// chaos.wgsl:5896
template <typename T> V2<T> var_logspiral(V2<T> p, T w, T a, T k) {
  T th = std::atan2(p.y, p.x);
  return polar_w(w, a * std::exp(k * th), th);
}
// Angle in (0, 2pi], floored at 1e-6 (chaos.wgsl:5903-5925).
template <typename T> T pos_angle(V2<T> p) {
  T t = std::atan2(p.y, p.x);
  return wmax(sel(t, t + (T)6.283185307179586, t < (T)0), (T)1e-6);
}
// chaos.wgsl:5903
template <typename T> V2<T> var_fermat_spiral(V2<T> p, T w, T a) {
  T th = pos_angle(p);
  return polar_w(w, a * std::sqrt(th), th);
}
// chaos.wgsl:5911
template <typename T> V2<T> var_lituus(V2<T> p, T w, T a) {
  T th = pos_angle(p);
  return polar_w(w, a / std::sqrt(th), th);
}
// chaos.wgsl:5919
template <typename T> V2<T> var_hyperbolic_spiral(V2<T> p, T w, T a) {
  T th = pos_angle(p);
  return polar_w(w, a / th, th);
}
// chaos.wgsl:5927 — truncated Weierstrass cosine series.
template <typename T> V2<T> var_weierstrass(V2<T> p, T w, T a, T b, T terms, T amp) {
  T wx = 0, wy = 0, ap = 1, bp = 1;
  const uint32_t n = f2u(clampv(terms, (T)1, (T)16));
  for (uint32_t i = 0; i < n; ++i) {
    wx += ap * safe_cos(bp * (T)3.1415926535 * p.x);
    wy += ap * safe_cos(bp * (T)3.1415926535 * p.y);
    ap *= a;
    bp *= b;
  }
  return {w * (p.x + amp * wx), w * (p.y + amp * wy)};
}
// chaos.wgsl:5943 — truncated Takagi (blancmange) series.
template <typename T> V2<T> var_takagi(V2<T> p, T w, T terms, T amp) {
  T tx = 0, ty = 0, p2 = 1;
  const uint32_t n = f2u(clampv(terms, (T)1, (T)16));
  for (uint32_t i = 0; i < n; ++i) {
    T xs = p2 * p.x, ys = p2 * p.y;
    tx += std::fabs(xs - std::floor(xs + (T)0.5)) / p2;
    ty += std::fabs(ys - std::floor(ys + (T)0.5)) / p2;
    p2 *= (T)2;
  }
  return {w * (p.x + amp * tx), w * (p.y + amp * ty)};
}
// chaos.wgsl:5959
template <typename T> V2<T> var_cantor_stairs(V2<T> p, T w, T terms, T amp) {
  T cx = p.x, cy = p.y;
  const uint32_t n = f2u(clampv(terms, (T)1, (T)8));
  for (uint32_t i = 0; i < n; ++i) {
    cx = (cx + safe_sin(cx * (T)6.2831853)) * (T)0.5;
    cy = (cy + safe_sin(cy * (T)6.2831853)) * (T)0.5;
  }
  return {w * (p.x + amp * cx), w * (p.y + amp * cy)};
}

// ---- billiards (chaos.wgsl:5971-6327) -------------------------------------------

// Specular reflection of v about unit normal n.
template <typename T> V2<T> reflect2(T vx, T vy, T nx, T ny) {
  T d = vx * nx + vy * ny;
  return {vx - (T)2 * d * nx, vy - (T)2 * d * ny};
}

// chaos.wgsl:5971 — one step of length `step` inside a circular table.
template <typename T> V2<T> var_billiard_circle(V2<T> p, T w, T radius, T step, T angle) {
  T rc = wmax(radius, (T)1.0e-5);
  T px = p.x, py = p.y;
  T r0 = len2(p);
  if (r0 > rc) { px = (px / r0) * rc; py = (py / r0) * rc; r0 = rc; }
  T bx = (T)1, by = (T)0;
  if (r0 > (T)1.0e-6) { bx = px / r0; by = py / r0; }
  T ca = safe_cos(angle), sa = safe_sin(angle);
  T vx = bx * ca - by * sa, vy = bx * sa + by * ca;
  T dpv = px * vx + py * vy;
  T disc = dpv * dpv + rc * rc - (px * px + py * py);
  T th = -dpv + std::sqrt(wmax(disc, (T)0));
  T fx = px + step * vx, fy = py + step * vy;
  if (th > (T)0 && th < step) {
    T hx = px + th * vx, hy = py + th * vy;
    V2<T> r = reflect2(vx, vy, -hx / rc, -hy / rc);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return {w * fx, w * fy};
}

// Exit time of the ray p + t v through the circle of radius R centred at
// (cx, 0); negative when the discriminant is (chaos.wgsl:6069-6100).
template <typename T> bool circle_exit(T px, T py, T vx, T vy, T cx, T R, T& t) {
  T dx = px - cx;
  T dpv = dx * vx + py * vy;
  T disc = dpv * dpv + R * R - (dx * dx + py * py);
  if (!(disc >= (T)0)) return false;
  t = -dpv + std::sqrt(disc);
  return true;
}

// chaos.wgsl:6014 — stadium: two flat walls and two semicircular caps.
template <typename T> V2<T> var_billiard_stadium(V2<T> p, T w, T width, T height, T step, T angle) {
  T R = wmax(height / (T)2, (T)1.0e-5), w2 = wmax(width / (T)2, (T)1.0e-5);
  T px = p.x, py = p.y;
  if (std::fabs(px) > w2) {
    T cx = sel(-w2, w2, px > (T)0);
    T dx = px - cx;
    T dist = std::sqrt(dx * dx + py * py);
    if (dist > R) { px = cx + (dx / dist) * R; py = (py / dist) * R; }
  } else if (std::fabs(py) > R) {
    py = sel(-R, R, py > (T)0);
  }
  T vx = safe_cos(angle), vy = safe_sin(angle);
  T th = (T)1.0e10;
  int wall = 0;  // 1 top, 2 bottom, 3 right cap, 4 left cap
  if (vy > (T)0) {
    T t = (R - py) / vy, x = px + t * vx;
    if (t >= (T)0 && x >= -w2 && x <= w2 && t < th) { th = t; wall = 1; }
  }
  if (vy < (T)0) {
    T t = (-R - py) / vy, x = px + t * vx;
    if (t >= (T)0 && x >= -w2 && x <= w2 && t < th) { th = t; wall = 2; }
  }
  T t;
  if (circle_exit(px, py, vx, vy, w2, R, t)) {
    T x = px + t * vx;
    if (t >= (T)0 && x >= w2 - (T)1.0e-5 && t < th) { th = t; wall = 3; }
  }
  if (circle_exit(px, py, vx, vy, -w2, R, t)) {
    T x = px + t * vx;
    if (t >= (T)0 && x <= -w2 + (T)1.0e-5 && t < th) { th = t; wall = 4; }
  }
  T fx = px + step * vx, fy = py + step * vy;
  if (th > (T)0 && th < step) {
    T hx = px + th * vx, hy = py + th * vy;
    T nx = 0, ny = 0;
    if (wall == 1) ny = (T)-1;
    else if (wall == 2) ny = (T)1;
    else if (wall == 3) {
      T rl = wmax(len2(V2<T>{hx - w2, hy}), (T)1.0e-5);
      nx = -(hx - w2) / rl;
      ny = -hy / rl;
    } else if (wall == 4) {
      T rl = wmax(len2(V2<T>{hx + w2, hy}), (T)1.0e-5);
      nx = -(hx + w2) / rl;
      ny = -hy / rl;
    }
    V2<T> r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return {w * fx, w * fy};
}

// chaos.wgsl:6121 — square table with a central circular scatterer.
template <typename T> V2<T> var_billiard_sinai(V2<T> p, T w, T length, T radius, T step, T angle) {
  T L2 = wmax(length / (T)2, (T)1.0e-5), R = wmax(radius, (T)1.0e-5);
  T px = p.x, py = p.y;
  if (std::fabs(px) > L2) px = sel(-L2, L2, px > (T)0);
  if (std::fabs(py) > L2) py = sel(-L2, L2, py > (T)0);
  T r0 = std::sqrt(px * px + py * py);
  if (r0 < R) {
    if (r0 > (T)1.0e-6) { px = (px / r0) * R; py = (py / r0) * R; }
    else {
      T th = (T)hash01(fbits((float)p.x) ^ fbits((float)p.y)) * (T)6.283185307179586;
      px = R * std::cos(th);
      py = R * std::sin(th);
    }
  }
  T vx = safe_cos(angle), vy = safe_sin(angle);
  T th = (T)1.0e10;
  int wall = 0;  // 1 left, 2 right, 3 bottom, 4 top, 5 scatterer
  if (vx < (T)0) {
    T t = (-L2 - px) / vx, y = py + t * vy;
    if (t >= (T)0 && y >= -L2 && y <= L2 && t < th) { th = t; wall = 1; }
  }
  if (vx > (T)0) {
    T t = (L2 - px) / vx, y = py + t * vy;
    if (t >= (T)0 && y >= -L2 && y <= L2 && t < th) { th = t; wall = 2; }
  }
  if (vy < (T)0) {
    T t = (-L2 - py) / vy, x = px + t * vx;
    if (t >= (T)0 && x >= -L2 && x <= L2 && t < th) { th = t; wall = 3; }
  }
  if (vy > (T)0) {
    T t = (L2 - py) / vy, x = px + t * vx;
    if (t >= (T)0 && x >= -L2 && x <= L2 && t < th) { th = t; wall = 4; }
  }
  {
    T dpv = px * vx + py * vy;
    T disc = dpv * dpv - ((px * px + py * py) - R * R);
    if (disc >= (T)0) {
      T t = -dpv - std::sqrt(disc);
      if (t >= (T)0 && t < th) { th = t; wall = 5; }
    }
  }
  T fx = px + step * vx, fy = py + step * vy;
  if (th > (T)0 && th < step) {
    T hx = px + th * vx, hy = py + th * vy;
    T nx = 0, ny = 0;
    if (wall == 1) nx = (T)1;
    else if (wall == 2) nx = (T)-1;
    else if (wall == 3) ny = (T)1;
    else if (wall == 4) ny = (T)-1;
    else if (wall == 5) { nx = hx / R; ny = hy / R; }
    V2<T> r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return {w * fx, w * fy};
}

// chaos.wgsl:6227 — regular polygon table (3..12 sides).
template <typename T> V2<T> var_billiard_polygon(V2<T> p, T w, T sides_f, T radius, T step, T angle) {
  const uint32_t n = f2u(clampv(sides_f + (T)0.5, (T)3, (T)12));
  const T two_pi = (T)(2.0 * 3.141592653589793);
  T R = wmax(radius, (T)1.0e-5);
  T inr = R * std::cos((T)3.141592653589793 / (T)n);
  T px = p.x, py = p.y;
  T r0 = std::sqrt(px * px + py * py);
  if (r0 > inr) {
    T sc = inr / wmax(r0, (T)1.0e-6);
    px = px * sc;
    py = py * sc;
  }
  T vx = safe_cos(angle), vy = safe_sin(angle);
  T th = (T)1.0e10;
  int32_t seg = -1;
  T ax = R * safe_cos((T)0), ay = R * safe_sin((T)0);
  for (uint32_t k = 0; k < n; ++k) {
    T ph = (two_pi * (T)(k + 1u)) / (T)n;
    T bx = R * safe_cos(ph), by = R * safe_sin(ph);
    T dx = bx - ax, dy = by - ay;
    T det = -vx * dy + vy * dx;
    if (std::fabs(det) > (T)1.0e-6) {
      T t = (-(ax - px) * dy + (ay - py) * dx) / det;
      T u = (vx * (ay - py) - vy * (ax - px)) / det;
      if (t >= (T)0 && u >= (T)0 && u <= (T)1 && t < th) { th = t; seg = (int32_t)k; }
    }
    ax = bx;
    ay = by;
  }
  T fx = px + step * vx, fy = py + step * vy;
  if (th > (T)0 && th < step && seg != -1) {
    T hx = px + th * vx, hy = py + th * vy;
    T p0 = (two_pi * (T)seg) / (T)n, p1 = (two_pi * (T)(seg + 1)) / (T)n;
    T sx = R * safe_cos(p0), sy = R * safe_sin(p0);
    T ex_ = R * safe_cos(p1), ey_ = R * safe_sin(p1);
    T ex = ex_ - sx, ey = ey_ - sy;
    T el = wmax(std::sqrt(ex * ex + ey * ey), (T)1.0e-5);
    T nx = -ey / el, ny = ex / el;
    T mx = (T)0.5 * (sx + ex_), my = (T)0.5 * (sy + ey_);
    if (nx * mx + ny * my > (T)0) { nx = -nx; ny = -ny; }
    V2<T> r = reflect2(vx, vy, nx, ny);
    fx = hx + (step - th) * r.x;
    fy = hy + (step - th) * r.y;
  }
  return {w * fx, w * fy};
}

// ---- physics warps (chaos.wgsl:6329-6433) ---------------------------------------

// chaos.wgsl:6329 — hyperbolic rotation along a rotated axis.
template <typename T> V2<T> var_lorentz_boost(V2<T> p, T w, T rapidity, T angle) {
  T c = safe_cos(angle), s = safe_sin(angle);
  T rx = p.x * c + p.y * s, ry = -p.x * s + p.y * c;
  T ch = std::cosh(rapidity), sh = std::sinh(rapidity);
  T xp = rx * ch + ry * sh, yp = rx * sh + ry * ch;
  return {w * (xp * c - yp * s), w * (xp * s + yp * c)};
}
// chaos.wgsl:6355
template <typename T> V2<T> var_schwarzschild_lensing(V2<T> p, T w, T mass, T eps) {
  T b = len2(p) + wmax(eps, (T)1e-6);
  T al = mass / b;
  T ca = safe_cos(al), sa = safe_sin(al);
  return {w * (p.x * ca - p.y * sa), w * (p.x * sa + p.y * ca)};
}
// chaos.wgsl:6372 — Euler step along the field of a +/- charge pair.
template <typename T> V2<T> var_field_dipole(V2<T> p, T w, T charge, T sep, T step, T angle) {
  T d = sep * (T)0.5;
  T dx = safe_cos(angle), dy = safe_sin(angle);
  T ppx = p.x - d * dx, ppy = p.y - d * dy;
  T npx = p.x - -d * dx, npy = p.y - -d * dy;
  T d3p = std::pow(ppx * ppx + ppy * ppy + (T)1e-4, (T)1.5);
  T d3n = std::pow(npx * npx + npy * npy + (T)1e-4, (T)1.5);
  T ex = charge * (ppx / d3p - npx / d3n), ey = charge * (ppy / d3p - npy / d3n);
  return {w * (p.x + step * ex), w * (p.y + step * ey)};
}
// chaos.wgsl:6391 — pull toward N magnets on a circle, with damping.
template <typename T>
V2<T> var_magnetic_pendulum_pos(V2<T> p, T w, T magnets, T radius, T strength, T damping) {
  const uint32_t n = f2u(clampv(magnets, (T)3, (T)6));
  T inv = (T)1 / (T)n;
  T fx = 0, fy = 0;
  for (uint32_t i = 0; i < n; ++i) {
    T th = (T)6.283185307179586 * (T)i * inv;
    T rx = radius * safe_cos(th) - p.x, ry = radius * safe_sin(th) - p.y;
    T pw = std::pow(rx * rx + ry * ry + (T)1e-4, (T)1.5);
    fx = fx + rx / pw;
    fy = fy + ry / pw;
  }
  return {w * (p.x + strength * fx - damping * p.x), w * (p.y + strength * fy - damping * p.y)};
}

// ---- modular forms (chaos.wgsl:462-522, 6435-6498) --------------------------------

template <typename T> V2<T> upper_half(V2<T> p, T im_floor) {
  T im = std::fabs(p.y) + wmax(im_floor, (T)0.02);
  return {p.x, wmax(im, (T)0.006435)};
}
template <typename T> V2<T> nome(V2<T> tau) { return c_exp(V2<T>{-K<T>::PI * tau.y, K<T>::PI * tau.x}); }
// Jacobi theta series in the nome (8 terms; powers by recurrence).
template <typename T> V2<T> theta3(V2<T> q) {
  V2<T> q2 = c_sqr(q), qn = q, odd = c_mul(q2, q), sum = q;
  for (int n = 2; n <= 8; ++n) {
    qn = c_mul(qn, odd);
    sum = V2<T>{sum.x + qn.x, sum.y + qn.y};
    odd = c_mul(odd, q2);
  }
  return {(T)1 + (T)2 * sum.x, (T)0 + (T)2 * sum.y};
}
template <typename T> V2<T> theta2(V2<T> q) {
  V2<T> q2 = c_sqr(q), term{(T)1, (T)0}, ev = q2, sum = term;
  for (int n = 1; n <= 7; ++n) {
    term = c_mul(term, ev);
    sum = V2<T>{sum.x + term.x, sum.y + term.y};
    ev = c_mul(ev, q2);
  }
  V2<T> m = c_mul(c_pow(q, V2<T>{(T)0.25, (T)0}), sum);
  return {(T)2 * m.x, (T)2 * m.y};
}
// log(1 + |z|) z / |z| of z clamped to +/-1e18 per component.
template <typename T> V2<T> log_compress(V2<T> z, T w) {
  V2<T> c{clampv(z.x, (T)-1e18, (T)1e18), clampv(z.y, (T)-1e18, (T)1e18)};
  T mag = len2(c);
  T sc = std::log((T)1 + mag) / wmax(mag, (T)1e-12);
  return {w * sc * c.x, w * sc * c.y};
}
// chaos.wgsl:6435
template <typename T> V2<T> var_jacobi_theta(V2<T> p, T w, T imf) {
  V2<T> t = theta3(nome(upper_half(p, imf)));
  return {w * t.x, w * t.y};
}
// chaos.wgsl:6442 — modular lambda (theta2 / theta3)^4.
template <typename T> V2<T> lambda_of(V2<T> p, T imf) {
  V2<T> q = nome(upper_half(p, imf));
  V2<T> r = c_div(theta2(q), theta3(q));
  return c_sqr(c_sqr(r));
}
template <typename T> V2<T> var_modular_lambda(V2<T> p, T w, T imf) {
  V2<T> l = lambda_of(p, imf);
  return {w * l.x, w * l.y};
}
// chaos.wgsl:6453 — Klein j from lambda, log-compressed.
template <typename T> V2<T> var_klein_j(V2<T> p, T w, T imf) {
  V2<T> lam = lambda_of(p, imf);
  V2<T> l2 = c_sqr(lam);
  V2<T> a{(T)1 - lam.x + l2.x, (T)0 - lam.y + l2.y};
  V2<T> a3 = c_pow_int(a, 3);
  V2<T> num{(T)256 * a3.x, (T)256 * a3.y};
  V2<T> oml{(T)1 - lam.x, (T)0 - lam.y};
  V2<T> den = c_mul(l2, c_sqr(oml));
  return log_compress(c_div(num, den), w);
}
// chaos.wgsl:6479 — Weierstrass P over a 5x5 lattice patch.
template <typename T> V2<T> var_weierstrass_p(V2<T> p, T w, T o1r, T o1i, T o2r, T o2i) {
  V2<T> acc = c_recip(c_sqr(p));
  for (int m = -2; m <= 2; ++m)
    for (int n = -2; n <= 2; ++n) {
      if (m == 0 && n == 0) continue;
      T ox = (T)m * o1r + (T)n * o2r, oy = (T)m * o1i + (T)n * o2i;
      V2<T> a = c_recip(c_sqr(V2<T>{p.x - ox, p.y - oy}));
      V2<T> b = c_recip(c_sqr(V2<T>{ox, oy}));
      acc = V2<T>{acc.x + a.x - b.x, acc.y + a.y - b.y};
    }
  return log_compress(acc, w);
}
// chaos.wgsl:6500 — fractional part of 1/t, |t| floored at 1e-4.
template <typename T> T gauss_frac(T t) {
  T at = wmax(std::fabs(t), (T)1.0e-4);
  T r = sel((T)1, (T)-1, t < (T)0) / at;
  return r - std::floor(r);
}
template <typename T> V2<T> var_gauss_map(V2<T> p, T w) { return {w * gauss_frac(p.x), w * gauss_frac(p.y)}; }

// ---- lattice value noise, special functions, digit codes (chaos.wgsl:6518-6870) ---

template <typename T> T vnoise_corner(int32_t ix, int32_t iy) {
  uint32_t ux = (uint32_t)ix * 0x9e3779b1u, uy = (uint32_t)iy * 0x85ebca77u;
  return (T)hash01(ux ^ uy) * (T)2 - (T)1;
}
template <typename T> T vnoise_smooth(T t) { return t * t * ((T)3 - (T)2 * t); }
// Value noise with x cells wrapped modulo `period` (period 0: no wrap).
template <typename T> T value_noise2(T px, T py, int32_t period) {
  T fx = std::floor(px), fy = std::floor(py);
  const int32_t ix = f2i(fx), iy = f2i(fy);
  T ux = vnoise_smooth(px - fx), uy = vnoise_smooth(py - fy);
  int32_t i0 = ix, i1 = iwrap_add(ix, 1);
  if (period > 0) {
    i0 = imod(i0, period); if (i0 < 0) i0 += period;
    i1 = imod(i1, period); if (i1 < 0) i1 += period;
  }
  const int32_t j1 = iwrap_add(iy, 1);
  T bottom = mixv(vnoise_corner<T>(i0, iy), vnoise_corner<T>(i1, iy), ux);
  T top = mixv(vnoise_corner<T>(i0, j1), vnoise_corner<T>(i1, j1), ux);
  return mixv(bottom, top, uy);
}
// chaos.wgsl:6575 — fBm of angle-periodic value noise, polar/iso blend.
template <typename T> T lich_field(V2<T> p, T freq, T branches, T radial, int32_t octaves) {
  T r = len2(p);
  T ang = std::atan2(p.y, p.x) / K<T>::TAU;
  T u = mixv(p.x * freq, (ang + (T)0.5) * branches, radial);
  T v = mixv(p.y * freq, r * freq, radial);
  int32_t period = imax(f2i(rint_even(branches)), 1);
  T sum = 0, amp = 1, tot = 0;
  for (int32_t k = 0; k < octaves; ++k) {
    sum = sum + amp * value_noise2(u, v, period);
    tot = tot + amp;
    amp = amp * (T)0.5;
    u = u * (T)2;
    v = v * (T)2;
    period = iwrap_mul(period, 2);
  }
  return sum / tot;
}
// chaos.wgsl:6600 — Abramowitz-Stegun 9.4.1 / 9.4.3 polynomial J0.
template <typename T> T bessel_j0(T x) {
  T ax = std::fabs(x);
  if (ax < (T)3) {
    T y = (x * x) / (T)9;
    return (T)1 + y * ((T)-2.2499997 + y * ((T)1.2656208 + y * ((T)-0.3163866 +
                  y * ((T)0.0444479 + y * ((T)-0.0039444 + y * (T)0.0002100)))));
  }
  T z = (T)3 / ax, y = z * z;
  T amp = (T)0.79788456 + y * ((T)-0.00000077 + y * ((T)-0.00552740 + y * ((T)-0.00009512 +
          y * ((T)0.00137237 + y * ((T)-0.00072805 + y * (T)0.00014476)))));
  T ph = ax - (T)0.78539816 + y * ((T)-0.04166397 + y * ((T)-0.00003954 +
         y * ((T)0.00262573 + y * ((T)-0.00054125 + y * ((T)-0.00029333 + y * (T)0.00013558)))));
  return (amp / std::sqrt(ax)) * safe_cos(ph);
}
// chaos.wgsl:6617 — Airy Ai: asymptotic tails, Maclaurin series between.
template <typename T> T airy_ai(T x) {
  if (x > (T)4) {
    T xi = (T)(2.0 / 3.0) * std::pow(x, (T)1.5);
    return std::exp(-xi) / ((T)(2.0 * 1.7724539) * std::pow(x, (T)0.25));
  }
  if (x < (T)-5) {
    T ax = -x;
    T xi = (T)(2.0 / 3.0) * std::pow(ax, (T)1.5);
    return safe_sin(xi + (T)0.78539816) / ((T)1.7724539 * std::pow(ax, (T)0.25));
  }
  const T c1 = (T)0.355028053887817, c2 = (T)0.258819403792807;
  T x3 = x * x * x;
  T f = 1, tf = 1, g = x, tg = x;
  for (int k = 1; k < 12; ++k) {
    T kf = (T)k;
    tf = tf * (x3 / (((T)3 * kf - (T)1) * ((T)3 * kf)));
    f = f + tf;
    tg = tg * (x3 / (((T)3 * kf) * ((T)3 * kf + (T)1)));
    g = g + tg;
  }
  return c1 * f - c2 * g;
}
// chaos.wgsl:6643 — Struve H1 power series (16 terms), x clamped to +/-10.
template <typename T> T struve_h1(T xin) {
  T x = clampv(xin, (T)-10, (T)10);
  T h = (T)0.5 * x, h2 = h * h;
  T term = h2 / (T)(0.8862269 * 1.3293404);
  T sum = term;
  for (int m = 1; m < 16; ++m) {
    T mf = (T)m;
    term = term * (-(h2)) / ((mf + (T)0.5) * (mf + (T)1.5));
    sum = sum + term;
  }
  return sum;
}
// chaos.wgsl:6658-6675 — fixed-point cell index of a coordinate in
// [-extent, extent) (wrapped) and the cell centre of an index.
inline uint32_t bits_mask(uint32_t bits) { return (1u << (bits & 31u)) - 1u; }
template <typename T> uint32_t fp_encode(T c, T extent, uint32_t bits) {
  T s = wmax(extent, (T)1.0e-4);
  T levels = (T)(float)(1u << (bits & 31u));
  T norm = (c + s) / ((T)2 * s);
  T folded = norm - std::floor(norm);
  return umin(f2u(folded * levels), bits_mask(bits));
}
template <typename T> T fp_decode(uint32_t i, T extent, uint32_t bits) {
  T s = wmax(extent, (T)1.0e-4);
  T levels = (T)(float)(1u << (bits & 31u));
  T unit = ((T)(float)i + (T)0.5) / levels;
  return unit * (T)2 * s - s;
}
// chaos.wgsl:6677 — Chebyshev T_n by the three-term recurrence.
template <typename T> T cheb_T(T x, int32_t n) {
  if (n <= 0) return (T)1;
  if (n == 1) return x;
  T a = 1, b = x, t = x;
  for (int32_t k = 2; k <= n; ++k) { t = (T)2 * x * b - a; a = b; b = t; }
  return t;
}
// chaos.wgsl:6690 — Legendre P_n (Bonnet recurrence).
template <typename T> T legendre_P(T x, int32_t n) {
  if (n <= 0) return (T)1;
  if (n == 1) return x;
  T a = 1, b = x, t = x;
  for (int32_t k = 1; k < n; ++k) {
    T fk = (T)k;
    t = (((T)2 * fk + (T)1) * x * b - fk * a) / (fk + (T)1);
    a = b;
    b = t;
  }
  return t;
}
// chaos.wgsl:6706 — associated Legendre P_l^m (Condon-Shortley phase).
template <typename T> T assoc_legendre(int32_t l, int32_t m, T u) {
  T pmm = 1;
  if (m > 0) {
    T s = std::sqrt(wmax((T)1 - u * u, (T)0));
    T fact = 1;
    for (int32_t i = 1; i <= m; ++i) { pmm = pmm * (-fact) * s; fact = fact + (T)2; }
  }
  if (l == m) return pmm;
  T pmm1 = u * ((T)2 * (T)m + (T)1) * pmm;
  if (l == m + 1) return pmm1;
  T pll = 0;
  for (int32_t ll = m + 2; ll <= l; ++ll) {
    T fl = (T)ll;
    pll = (((T)2 * fl - (T)1) * u * pmm1 - (fl + (T)m - (T)1) * pmm) / (fl - (T)m);
    pmm = pmm1;
    pmm1 = pll;
  }
  return pll;
}
// chaos.wgsl:6728 — |Y_l^m(theta)| without normalisation, capped at 4.
template <typename T> T sph_harmonic(T th, int32_t l, int32_t m) {
  const int32_t mm = m < 0 ? -m : m;
  T y = assoc_legendre(l, mm, safe_cos(th)) * safe_cos((T)mm * th);
  return clampv(std::fabs(y), (T)0, (T)4);
}
// chaos.wgsl:6742 — Chladni plate field F and its gradient.
template <typename T> void chladni(T x, T y, T n, T m, T& F, T& Fx, T& Fy) {
  T an = K<T>::PI * n, am = K<T>::PI * m;
  T cnx = safe_cos(an * x), snx = safe_sin(an * x);
  T cmy = safe_cos(am * y), smy = safe_sin(am * y);
  T cmx = safe_cos(am * x), smx = safe_sin(am * x);
  T cny = safe_cos(an * y), sny = safe_sin(an * y);
  F = cnx * cmy - cmx * cny;
  Fx = -an * snx * cmy + am * smx * cny;
  Fy = -am * cnx * smy + an * cmx * sny;
}
// chaos.wgsl:6760 — generalised Laguerre L_k^alpha.
template <typename T> T assoc_laguerre(int32_t k, T alpha, T x) {
  if (k <= 0) return (T)1;
  T a = 1, b = (T)1 + alpha - x;
  if (k == 1) return b;
  for (int32_t j = 1; j < k; ++j) {
    T fj = (T)j;
    T c = (((T)2 * fj + (T)1 + alpha - x) * b - (fj + alpha) * a) / (fj + (T)1);
    a = b;
    b = c;
  }
  return b;
}
// chaos.wgsl:6773 — hydrogen radial density rho^(2l) e^-rho (L^(2l+1)_(n-l-1))^2.
template <typename T> T radial_psi2(int32_t n, int32_t l, T rho) {
  T rf = wmax(rho, (T)0);
  T lag = assoc_laguerre(n - l - 1, (T)2 * (T)l + (T)1, rf);
  T rp = 1;
  if (l > 0)
    for (int32_t j = 0; j < 2 * l; ++j) rp = rp * rf;
  return rp * std::exp(-rf) * (lag * lag);
}
// chaos.wgsl:6796 — Hastings polynomial approximations of E(m), K(m).
template <typename T> T elliptic_E(T m) {
  T m1 = clampv((T)1 - m, (T)0, (T)1);
  T l = std::log((T)1 / wmax(m1, (T)1.0e-6));
  T a = (T)1 + m1 * ((T)0.4630151 + m1 * (T)0.1077812);
  T b = m1 * ((T)0.2452727 + m1 * (T)0.0412496);
  return a + b * l;
}
template <typename T> T elliptic_K(T m) {
  T m1 = clampv((T)1 - m, (T)1.0e-3, (T)1);
  T l = std::log((T)1 / m1);
  T a = (T)1.3862944 + m1 * ((T)0.1119723 + m1 * (T)0.0725296);
  T b = (T)0.5 + m1 * ((T)0.1213478 + m1 * (T)0.0288729);
  return a + b * l;
}
// chaos.wgsl:6813 — Winitzki's closed-form inverse error function.
template <typename T> T erfinv_w(T xin) {
  T x = clampv(xin, (T)-0.999999, (T)0.999999);
  const T a = (T)0.147;
  T l1 = std::log((T)1 - x * x);
  T t1 = (T)2 / (K<T>::PI * a) + (T)0.5 * l1;
  T in = std::sqrt(wmax(t1 * t1 - l1 / a, (T)0)) - t1;
  return signv(x) * std::sqrt(wmax(in, (T)0));
}
// chaos.wgsl:6825-6870 — base-3 digit codes (Peano curve).
inline uint32_t pow3u(uint32_t n) {
  uint32_t p = 1u;
  for (uint32_t i = 0; i < n; ++i) p = p * 3u;
  return p;
}
template <typename T> uint32_t tri_encode(T c, T extent, uint32_t trits) {
  T s = wmax(extent, (T)1.0e-4);
  const uint32_t cells = pow3u(trits);
  T norm = (c + s) / ((T)2 * s);
  T folded = norm - std::floor(norm);
  return umin(f2u(folded * (T)(float)cells), cells - 1u);
}
template <typename T> T tri_decode(uint32_t i, T extent, uint32_t trits) {
  T s = wmax(extent, (T)1.0e-4);
  T levels = (T)(float)pow3u(trits);
  return (((T)(float)i + (T)0.5) / levels) * (T)2 * s - s;
}
inline uint32_t peano_scramble(uint32_t idx, uint32_t trits) {
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
template <typename T> V2<T> var_bessel_j0(V2<T> p, T w, T freq) {
  T j = bessel_j0(freq * len2(p));
  return {w * (j * p.x), w * (j * p.y)};
}
// chaos.wgsl:6879
template <typename T> V2<T> var_airy_radial(V2<T> p, T w, T scale, T shift) {
  T a = airy_ai(scale * (len2(p) - shift));
  T k = (T)3 * a;
  return {w * (k * p.x), w * (k * p.y)};
}
// chaos.wgsl:6886 — rational approximation of the Fresnel integrals.
template <typename T> V2<T> var_cornu_spiral(V2<T> p, T w, T freq) {
  T t = freq * p.x;
  T s = std::fabs(t);
  T sg = sel((T)1, (T)-1, t < (T)0);
  T f = ((T)1 + (T)0.926 * s) / ((T)2 + (T)1.792 * s + (T)3.104 * (s * s));
  T g = (T)1 / ((T)2 + (T)4.142 * s + (T)3.492 * (s * s) + (T)6.670 * (s * s * s));
  T a = (K<T>::PI * (s * s)) / (T)2;
  T sa = safe_sin(a), ca = safe_cos(a);
  T ci = (T)0.5 + f * sa - g * ca;
  T si = (T)0.5 - f * ca - g * sa;
  return {w * (sg * ci), w * (sg * si + (T)0.25 * p.y)};
}
// chaos.wgsl:6902
template <typename T> V2<T> var_struve_h1(V2<T> p, T w, T freq) {
  T k = (T)0.6 * struve_h1(freq * len2(p));
  return {w * (k * p.x), w * (k * p.y)};
}
// chaos.wgsl:6911 — softened inverse-square pull toward mass m at c.
template <typename T> V2<T> nbody_pull(V2<T> p, T cx, T cy, T m, T g, T eps) {
  T dx = cx - p.x, dy = cy - p.y;
  T r2 = (dx * dx + dy * dy) + eps;
  T inv = g * m / (r2 * std::sqrt(r2));
  return {inv * dx, inv * dy};
}
// chaos.wgsl:6917
template <typename T>
V2<T> var_nbody_lensing(V2<T> p, T w, T c1x, T c1y, T c2x, T c2y, T m1, T m2, T g, T eps) {
  T e = wmax(eps, (T)5.0e-3);
  V2<T> a = nbody_pull(p, c1x, c1y, m1, g, e), b = nbody_pull(p, c2x, c2y, m2, g, e);
  T dx = a.x + b.x, dy = a.y + b.y;
  return {w * (p.x + dx), w * (p.y + dy)};
}
// chaos.wgsl:6926 — displacement along the curl of a value-noise stream function.
template <typename T> V2<T> var_curl_noise(V2<T> p, T w, T freq, T amp) {
  T f = wmax(freq, (T)1.0e-3);
  T sx = p.x * f, sy = p.y * f;
  const T h = (T)1.0e-2;
  T dx = (value_noise2(sx + h, sy, 0) - value_noise2(sx - h, sy, 0)) / ((T)2 * h);
  T dy = (value_noise2(sx, sy + h, 0) - value_noise2(sx, sy - h, 0)) / ((T)2 * h);
  T ux = clampv(dy, (T)-8, (T)8), uy = clampv(-dx, (T)-8, (T)8);
  return {w * (p.x + amp * ux), w * (p.y + amp * uy)};
}
// chaos.wgsl:6944 — one Newton step toward the zero set of a noise field.
template <typename T>
V2<T> var_lichtenberg(V2<T> p, T w, T freqin, T brin, T radin, T detin, T strin) {
  const T H = (T)1.0e-2, R_EPS = (T)1.0e-4, G_EPS = (T)1.0e-3, MAX_STEP = (T)2;
  T freq = wmax(freqin, (T)1.0e-3);
  T br = clampv(rint_even(brin), (T)1, (T)16);
  T rad = clampv(radin, (T)0, (T)1);
  const int32_t oct = f2i(clampv(rint_even(detin), (T)1, (T)4));
  T str = wmax(strin, (T)0);
  if (len2(p) < R_EPS) return {w * p.x, w * p.y};
  T f0 = lich_field(p, freq, br, rad, oct);
  T fxp = lich_field(V2<T>{p.x + H, p.y + (T)0}, freq, br, rad, oct);
  T fxm = lich_field(V2<T>{p.x - H, p.y - (T)0}, freq, br, rad, oct);
  T fyp = lich_field(V2<T>{p.x + (T)0, p.y + H}, freq, br, rad, oct);
  T fym = lich_field(V2<T>{p.x - (T)0, p.y - H}, freq, br, rad, oct);
  T gx = (fxp - fxm) / ((T)2 * H), gy = (fyp - fym) / ((T)2 * H);
  T g2 = wmax(gx * gx + gy * gy, G_EPS);
  T k = -(f0 / g2);
  T stx = k * gx, sty = k * gy;
  T sl = std::sqrt(stx * stx + sty * sty);
  if (sl > MAX_STEP) { T c = MAX_STEP / sl; stx = stx * c; sty = sty * c; }
  return {w * (p.x + str * stx), w * (p.y + str * sty)};
}
template <typename T> T cop_sigmoid(T t) { return (T)1 / ((T)1 + std::exp(-t)); }
// chaos.wgsl:6981 — Gaussian copula: y replaced by its rho-correlated normal.
template <typename T> V2<T> var_copula_gaussian(V2<T> p, T w, T strength, T rhoin) {
  T s = wmax(strength, (T)1.0e-3);
  const T hi = (T)(1.0 - 1.0e-4);
  T u = clampv(cop_sigmoid(s * p.x), (T)1.0e-4, hi);
  T v = clampv(cop_sigmoid(s * p.y), (T)1.0e-4, hi);
  T zx = (T)1.4142135 * erfinv_w((T)2 * u - (T)1);
  T zy = (T)1.4142135 * erfinv_w((T)2 * v - (T)1);
  T rho = clampv(rhoin, (T)-0.95, (T)0.95);
  T zyp = rho * zx + std::sqrt(wmax((T)1 - rho * rho, (T)0)) * zy;
  return {w * p.x, w * (zyp / s)};
}
// chaos.wgsl:6998 — Clayton copula conditional quantile.
template <typename T> V2<T> var_copula_clayton(V2<T> p, T w, T strength, T thin) {
  T s = wmax(strength, (T)1.0e-3);
  const T hi = (T)(1.0 - 1.0e-4);
  T u = clampv(cop_sigmoid(s * p.x), (T)1.0e-4, hi);
  T v = clampv(cop_sigmoid(s * p.y), (T)1.0e-4, hi);
  T th = clampv(thin, (T)0.05, (T)8);
  T ua = std::pow(u, -th), va = std::pow(v, -th);
  T base = wmax(ua + va - (T)1, (T)1.0e-6);
  T h = std::pow(u, -th - (T)1) * std::pow(base, (T)-1 - (T)1 / th);
  T up = clampv(h, (T)1.0e-4, hi);
  T yp = clampv(std::log(up / ((T)1 - up)), (T)-12, (T)12) / s;
  return {w * p.x, w * yp};
}
// chaos.wgsl:7021 — truncated series of the disc-to-polygon map.
template <typename T> V2<T> var_schwarz_christoffel(V2<T> p, T w, T sides) {
  T n = wmax(rint_even(sides), (T)3);
  const int32_t ni = f2i(n);
  T r = len2(p);
  V2<T> z = p;
  if (r > (T)0.999) { T k = (T)0.999 / wmax(r, (T)1.0e-12); z = V2<T>{p.x * k, p.y * k}; }
  T a = (T)2 / n;
  T ax = 0, ay = 0, coef = 1;
  for (int32_t k = 0; k < 10; ++k) {
    const int32_t pi = iwrap_add(iwrap_mul(ni, k), 1);
    V2<T> zp = c_pow_int(z, pi);
    T c = coef / (T)pi;
    ax = ax + c * zp.x;
    ay = ay + c * zp.y;
    coef = coef * (a + (T)k) / (T)(k + 1);
  }
  return {w * ax, w * ay};
}
// chaos.wgsl:7047 — Doyle-style log-spiral shear with a ripple.
template <typename T> V2<T> var_doyle(V2<T> p, T w, T dp, T dq) {
  T pp = wmax(std::fabs(dp), (T)1);
  V2<T> L = c_log(p);
  T u = L.x, v = L.y;
  T pitch = dq / pp;
  T us = u + pitch * v;
  T rip = (T)0.12 * (safe_sin(pp * v) + safe_cos(pp * v - K<T>::TAU * us));
  T vs = v + pitch * rip;
  V2<T> e = c_exp(V2<T>{us + rip, vs});
  return {w * e.x, w * e.y};
}
// chaos.wgsl:7070 — gradient of an n-fold sum of plane waves.
template <typename T> V2<T> var_quasicrystal(V2<T> p, T w, T symin, T freqin) {
  const int32_t n = f2i(clampv(rint_even(symin), (T)2, (T)12));
  T freq = wmax(freqin, (T)1.0e-3);
  T gx = 0, gy = 0;
  for (int32_t k = 0; k < n; ++k) {
    T th = K<T>::PI * (T)k / (T)n;
    T dx = std::cos(th), dy = std::sin(th);
    T d = freq * (p.x * dx + p.y * dy);
    T s = safe_sin(d);
    gx = gx - s * dx;
    gy = gy - s * dy;
  }
  gx = gx / (T)n;
  gy = gy / (T)n;
  return {w * (p.x + (T)0.4 * gx), w * (p.y + (T)0.4 * gy)};
}
// chaos.wgsl:7093 — de Bruijn pentagrid cell vertex.
template <typename T> V2<T> var_penrose(V2<T> p, T w, T scalein, T offset) {
  T scale = wmax(scalein, (T)1.0e-3);
  T vx = 0, vy = 0;
  for (int j = 0; j < 5; ++j) {
    T th = K<T>::TAU * (T)j / (T)5;
    T dx = std::cos(th), dy = std::sin(th);
    T kk = std::floor(scale * (p.x * dx + p.y * dy) + offset);
    vx = vx + kk * dx;
    vy = vy + kk * dy;
  }
  return {w * vx / scale, w * vy / scale};
}
// chaos.wgsl:7115 — analytic interpolation of the Collatz map.
template <typename T> V2<T> var_collatz(V2<T> p, T w, T scalein, T shift) {
  T scale = wmax(std::fabs(scalein), (T)1.0e-3);
  V2<T> z{p.x * scale + shift, p.y * scale};
  const T hp = K<T>::PI * (T)0.5;
  V2<T> arg{z.x * hp, z.y * hp};
  V2<T> sn = c_sin(arg), cs = c_sin(V2<T>{arg.x + hp, arg.y + (T)0});
  V2<T> s2 = c_mul(sn, sn), c2 = c_mul(cs, cs);
  V2<T> ev = c_mul(V2<T>{(T)0.5 * z.x, (T)0.5 * z.y}, c2);
  V2<T> od = c_mul(V2<T>{(T)0.5 * ((T)3 * z.x + (T)1), (T)0.5 * ((T)3 * z.y + (T)0)}, s2);
  return {w * (ev.x + od.x), w * (ev.y + od.y)};
}
// chaos.wgsl:7137 — digamma by recurrence to z + 6 and the asymptotic series.
template <typename T> V2<T> var_digamma(V2<T> p, T w, T scalein, T shift) {
  T scale = wmax(std::fabs(scalein), (T)1.0e-3);
  V2<T> z{p.x * scale + shift, p.y * scale};
  V2<T> sum{(T)0, (T)0};
  for (int k = 0; k < 6; ++k) {
    V2<T> r = c_recip(z);
    sum = V2<T>{sum.x + r.x, sum.y + r.y};
    z = V2<T>{z.x + (T)1, z.y + (T)0};
  }
  V2<T> lz = c_log(z), zi = c_recip(z);
  V2<T> zi2 = c_mul(zi, zi), zi4 = c_mul(zi2, zi2);
  const T k12 = (T)(1.0 / 12.0), k120 = (T)(1.0 / 120.0);
  T px = lz.x - (T)0.5 * zi.x - k12 * zi2.x + k120 * zi4.x;
  T py = lz.y - (T)0.5 * zi.y - k12 * zi2.y + k120 * zi4.y;
  return {w * (px - sum.x), w * (py - sum.y)};
}
// chaos.wgsl:7160 — Sprott quadratic map (ten coefficients).
template <typename T> V2<T> var_sprott_poly(V2<T> p, T w, const T* c) {
  T x = p.x, y = p.y;
  T x2 = x * x, xy = x * y, y2 = y * y;
  return {w * (c[0] * x + c[1] * x2 + c[2] * xy + c[3] * y + c[4] * y2),
          w * (c[5] * x + c[6] * x2 + c[7] * xy + c[8] * y + c[9] * y2)};
}
// chaos.wgsl:7175 — Martin's hopalong map.
template <typename T> V2<T> var_hopalong(V2<T> p, T w, T a, T b, T c) {
  return {w * (p.y - signv(p.x) * std::sqrt(std::fabs(b * p.x - c))), w * (a - p.x)};
}
// chaos.wgsl:7184
template <typename T> T gm_G(T x, T a) { return a * x + (T)2 * ((T)1 - a) * x * x / ((T)1 + x * x); }
template <typename T> V2<T> var_gumowski_mira(V2<T> p, T w, T a, T b) {
  T xp = b * p.y + gm_G(p.x, a);
  T yp = -p.x + gm_G(xp, a);
  return {w * xp, w * yp};
}

// ---- polynomials, digit codes, optics, waves, orbitals (chaos.wgsl:7196-7466) ----

// Integer order round(v) clamped to [lo, hi].
template <typename T> int32_t order_of(T v, T lo, T hi) { return f2i(clampv(rint_even(v), lo, hi)); }
// chaos.wgsl:7196
template <typename T> V2<T> var_chebyshev(V2<T> p, T w, T ox, T oy) {
  const int32_t nx = order_of(ox, (T)0, (T)12), ny = order_of(oy, (T)0, (T)12);
  return {w * cheb_T(clampv(p.x, (T)-1, (T)1), nx), w * cheb_T(clampv(p.y, (T)-1, (T)1), ny)};
}
// chaos.wgsl:7204
template <typename T> V2<T> var_legendre(V2<T> p, T w, T ox, T oy) {
  const int32_t nx = order_of(ox, (T)0, (T)12), ny = order_of(oy, (T)0, (T)12);
  return {w * legendre_P(clampv(p.x, (T)-1, (T)1), nx), w * legendre_P(clampv(p.y, (T)-1, (T)1), ny)};
}
// Radial rescale of p to length rp (r > 0).
template <typename T> V2<T> radial_to(V2<T> p, T w, T r, T rp) {
  T ir = (T)1 / r;
  return {w * (p.x * ir * rp), w * (p.y * ir * rp)};
}
// chaos.wgsl:7214
template <typename T> V2<T> var_spherical_harmonic(V2<T> p, T w, T dl, T om, T amount) {
  T r = len2(p);
  if (r < (T)1.0e-6) return {(T)0, (T)0};
  T th = std::atan2(p.y, p.x);
  const int32_t l = order_of(dl, (T)0, (T)6);
  const int32_t m = f2i(clampv(rint_even(om), (T)0, (T)l));
  T y = sph_harmonic(th, l, m);
  T amt = clampv(amount, (T)0, (T)2);
  T sc = ((T)1 - amt) + amt * y;
  return radial_to(p, w, r, r * clampv(sc, (T)-3, (T)3));
}
// chaos.wgsl:7230 — random-phase angular Fourier envelope.
template <typename T> V2<T> var_fourier_warp(V2<T> p, T w, T harmonics, T amp, T seed) {
  T r = len2(p);
  if (r < (T)1.0e-6) return {(T)0, (T)0};
  T th = std::atan2(p.y, p.x);
  const int32_t n = order_of(harmonics, (T)1, (T)8);
  T amt = clampv(amp, (T)0, (T)0.9);
  const uint32_t sb = f2u(clampv(seed, (T)0, (T)1024));
  T acc = 0;
  for (int32_t k = 1; k <= n; ++k) {
    T fk = (T)k;
    T phi = (T)hash01(sb * 2654435761u + (uint32_t)k * 40503u) * K<T>::TAU;
    acc = acc + safe_cos(fk * th + phi) / fk;
  }
  T env = (T)1 + amt * acc;
  return radial_to(p, w, r, r * clampv(env, (T)0.05, (T)3));
}
inline uint32_t reverse_bits_n(uint32_t x, uint32_t bits) {
  uint32_t r = 0u;
  for (uint32_t k = 0; k < bits; ++k) { r = (r << 1) | (x & 1u); x >>= 1; }
  return r;
}
// chaos.wgsl:7264 — van der Corput digit reversal of the fixed-point cell.
template <typename T> V2<T> var_radical_inverse(V2<T> p, T w, T extent, T bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (T)2, (T)24));
  uint32_t rx = reverse_bits_n(fp_encode(p.x, extent, bits), bits) & bits_mask(bits);
  uint32_t ry = reverse_bits_n(fp_encode(p.y, extent, bits), bits) & bits_mask(bits);
  return {w * fp_decode(rx, extent, bits), w * fp_decode(ry, extent, bits)};
}
// chaos.wgsl:7276
template <typename T> V2<T> var_gray_code(V2<T> p, T w, T extent, T bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (T)2, (T)24));
  const uint32_t m = bits_mask(bits);
  uint32_t ix = fp_encode(p.x, extent, bits), iy = fp_encode(p.y, extent, bits);
  uint32_t gx = (ix ^ (ix >> 1)) & m, gy = (iy ^ (iy >> 1)) & m;
  return {w * fp_decode(gx, extent, bits), w * fp_decode(gy, extent, bits)};
}
inline uint32_t part1by1(uint32_t x, uint32_t bits) {
  uint32_t r = 0u;
  for (uint32_t i = 0; i < bits; ++i) r = r | (((x >> (i & 31u)) & 1u) << ((i << 1) & 31u));
  return r;
}
// chaos.wgsl:7301 — interleave, then split back into two coordinates.
template <typename T> V2<T> var_morton_zorder(V2<T> p, T w, T extent, T bits_f) {
  const uint32_t bits = f2u(clampv(bits_f, (T)2, (T)16));
  const uint32_t m = bits_mask(bits);
  uint32_t ix = fp_encode(p.x, extent, bits), iy = fp_encode(p.y, extent, bits);
  uint32_t code = part1by1(ix, bits) | (part1by1(iy, bits) << 1);
  uint32_t nx = code & m, ny = (code >> (bits & 31u)) & m;
  return {w * fp_decode(nx, extent, bits), w * fp_decode(ny, extent, bits)};
}
// chaos.wgsl:7316 — refract the radial direction at a horizontal interface.
template <typename T> V2<T> var_snell_refraction(V2<T> p, T w, T nratio, T strength) {
  T r = len2(p);
  if (r < (T)1.0e-6) return {w * p.x, w * p.y};
  T dx = p.x / r, dy = p.y / r;
  T nr = clampv(nratio, (T)0.05, (T)5);
  T s2 = nr * dx;
  T ox, oy;
  if (std::fabs(s2) >= (T)1) { ox = dx; oy = -dy; }
  else {
    T c2 = std::sqrt(wmax((T)1 - s2 * s2, (T)0));
    ox = s2;
    oy = sel(-c2, c2, dy >= (T)0);
  }
  T s = clampv(strength, (T)0, (T)1);
  T bx = mixv(dx, ox, s), by = mixv(dy, oy, s);
  T bl = wmax(std::sqrt(bx * bx + by * by), (T)1.0e-6);
  bx = bx / bl;
  by = by / bl;
  return {w * r * bx, w * r * by};
}
// chaos.wgsl:7345
template <typename T> V2<T> var_grin_lens(V2<T> p, T w, T focal, T eps) {
  T f = wmax(std::fabs(focal), (T)0.05), e = wmax(std::fabs(eps), (T)1.0e-3);
  T r = len2(p);
  if (r < (T)1.0e-6) return {w * p.x, w * p.y};
  T soft = (e * e) / (r * r + e * e);
  T pull = (r / f) * soft;
  T dx = p.x / r, dy = p.y / r;
  T nr = wmax(r - pull, -r);
  return {w * nr * dx, w * nr * dy};
}
// chaos.wgsl:7359
template <typename T> V2<T> var_caustic_fold(V2<T> p, T w, T freq, T amp, T phase) {
  T k = clampv(freq, (T)0, (T)16), a = clampv(amp, (T)0, (T)2);
  T gx = -safe_sin(k * p.x + phase), gy = -safe_sin(k * p.y + phase);
  return {w * clampv(p.x + a * gx, (T)-8, (T)8), w * clampv(p.y + a * gy, (T)-8, (T)8)};
}
// chaos.wgsl:7377 — unit step down the Chladni field toward its nodal lines.
template <typename T> V2<T> var_chladni(V2<T> p, T w, T nin, T min_, T step) {
  T n = clampv(std::floor(nin + (T)0.5), (T)1, (T)12), m = clampv(std::floor(min_ + (T)0.5), (T)1, (T)12);
  T F, gx, gy;
  chladni(p.x, p.y, n, m, F, gx, gy);
  T s = sel((T)1, (T)-1, F < (T)0);
  gx = s * gx;
  gy = s * gy;
  T inv = (T)1 / (std::sqrt(gx * gx + gy * gy) + (T)1.0e-4);
  return {w * (p.x - step * gx * inv), w * (p.y - step * gy * inv)};
}
// chaos.wgsl:7395 — up to four standing-wave modes (inactive ones gated by 0).
template <typename T> V2<T> var_standing_wave(V2<T> p, T w, T modes, T freq, T amp, T decay) {
  T kmax = clampv(std::floor(modes + (T)0.5), (T)1, (T)4);
  T d = clampv(decay, (T)0, (T)0.999);
  T dx = 0, dy = 0, ak = amp;
  for (int k = 1; k <= 4; ++k) {
    T kf = (T)k * K<T>::PI * freq;
    T gate = sel((T)0, (T)1, (T)k <= kmax);
    dx = dx + gate * ak * safe_sin(kf * p.x) * safe_sin(kf * p.y);
    dy = dy + gate * ak * safe_sin(kf * p.y) * safe_cos(kf * p.x);
    ak = ak * d;
  }
  return {w * (p.x + dx), w * (p.y + dy)};
}
// chaos.wgsl:7413
template <typename T> V2<T> var_moire(V2<T> p, T w, T freq, T beat, T angle, T amp) {
  T ca = safe_cos(angle), sa = safe_sin(angle);
  T u = p.x * ca + p.y * sa, v = -p.x * sa + p.y * ca;
  T f2 = freq + beat;
  T gx = safe_sin(freq * K<T>::PI * p.x) * safe_sin(f2 * K<T>::PI * u);
  T gy = safe_sin(freq * K<T>::PI * p.y) * safe_sin(f2 * K<T>::PI * v);
  return {w * (p.x + amp * gx), w * (p.y + amp * gy)};
}
// Radius 1.5 a / (1 + a) along p's direction ((1, 0) at the origin).
template <typename T> V2<T> shell_out(V2<T> p, T w, T r, T amp) {
  T ro = (T)1.5 * (amp / ((T)1 + amp));
  T dx = (T)1, dy = (T)0;
  if (r > (T)1.0e-6) { dx = p.x / r; dy = p.y / r; }
  return {w * ro * dx, w * ro * dy};
}
// chaos.wgsl:7428
template <typename T> V2<T> var_radial_shell(V2<T> p, T w, T nin, T lin, T shell) {
  const int32_t n = imin(imax(f2i(rint_even(nin)), 1), 4);
  const int32_t l = imin(imax(f2i(rint_even(lin)), 0), n - 1);
  T r = len2(p);
  T rho = ((T)2 * r) / ((T)n * wmax(shell, (T)1.0e-3));
  return shell_out(p, w, r, radial_psi2(n, l, rho));
}
// chaos.wgsl:7444
template <typename T> V2<T> var_hydrogen_orbital(V2<T> p, T w, T nin, T lin, T min_, T shell, T lobe) {
  const int32_t n = imin(imax(f2i(rint_even(nin)), 1), 4);
  const int32_t l = imin(imax(f2i(rint_even(lin)), 0), n - 1);
  const int32_t m = imin(imax(f2i(rint_even(min_)), -l), l);
  T r = len2(p);
  T rho = ((T)2 * r) / ((T)n * wmax(shell, (T)1.0e-3));
  T rad = radial_psi2(n, l, rho);
  T yl = sph_harmonic(std::atan2(p.y, p.x), l, m);
  T amp = mixv(rad, rad * (yl * yl), clampv(lobe, (T)0, (T)1));
  return shell_out(p, w, r, amp);
}

// ---- distributions, wavelets, celestial, knots, follow-ons (chaos.wgsl:7467-7729) --

// u = r^2 / (1 + r^2) clamped to [1e-4, 1 - 1e-4].
template <typename T> T cdf_u(T r2) {
  return clampv(r2 / ((T)1 + r2), (T)1.0e-4, (T)(1.0 - 1.0e-4));
}
// w * rp along p / |p| ((0, 0) when |p| < 1e-9).
template <typename T> V2<T> along(V2<T> p, T w, T r, T rp) {
  if (r < (T)1.0e-9) return {w * rp * (T)0, w * rp * (T)0};
  T m = wmax(r, (T)1.0e-12);
  return {w * rp * (p.x / m), w * rp * (p.y / m)};
}
// chaos.wgsl:7467
template <typename T> V2<T> var_weibull_cdf(V2<T> p, T w, T lambda, T k) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T u = cdf_u(r2);
  T kk = wmax(std::fabs(k), (T)0.05);
  return along(p, w, r, lambda * std::pow(-std::log((T)1 - u), (T)1 / kk));
}
// chaos.wgsl:7477
template <typename T> V2<T> var_logistic_cdf(V2<T> p, T w, T mu, T s) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T u = cdf_u(r2);
  T lg = clampv(std::log(u / ((T)1 - u)), (T)-12, (T)12);
  return along(p, w, r, wmax(mu + s * lg, (T)0));
}
// chaos.wgsl:7490
template <typename T> V2<T> var_cauchy_cdf(V2<T> p, T w, T gamma, T tail) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T u = cdf_u(r2);
  T raw = gamma * safe_tan(K<T>::PI * (u - (T)0.5));
  T cap = wmax(tail, (T)0.01);
  return along(p, w, r, clampv(raw, -cap, cap));
}
// chaos.wgsl:7502
template <typename T> V2<T> var_pareto_cdf(V2<T> p, T w, T xm, T alpha, T tail) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T u = cdf_u(r2);
  T a = wmax(std::fabs(alpha), (T)0.05);
  T raw = xm * std::pow((T)1 - u, (T)-1 / a);
  return along(p, w, r, wmin(raw, wmax(tail, (T)0.01)));
}
// Radial scale (1 + amp psi) applied to p.
template <typename T> V2<T> scaled(V2<T> p, T w, T sc) { return {w * sc * p.x, w * sc * p.y}; }
// chaos.wgsl:7517
template <typename T> V2<T> var_morlet(V2<T> p, T w, T freq, T sigma, T amp) {
  T r = len2(p);
  T t = r / wmax(sigma, (T)0.05);
  T psi = std::exp((T)-0.5 * wmin(t * t, (T)80)) * safe_cos(freq * r);
  return scaled(p, w, (T)1 + amp * psi);
}
// chaos.wgsl:7527
template <typename T> V2<T> var_mexican_hat(V2<T> p, T w, T sigma, T amp) {
  T r = len2(p);
  T t = r / wmax(sigma, (T)0.05);
  T t2 = wmin(t * t, (T)80);
  T psi = ((T)1 - t2) * std::exp((T)-0.5 * t2);
  return scaled(p, w, (T)1 + amp * psi);
}
// chaos.wgsl:7537
template <typename T> V2<T> var_chirp(V2<T> p, T w, T alpha, T amp, T decay) {
  T r2 = dot2(p);
  T psi = std::exp(-decay * wmin(r2, (T)160)) * safe_sin(alpha * r2);
  return scaled(p, w, (T)1 + amp * psi);
}
// chaos.wgsl:7548 — two Newton steps of Kepler's equation.
template <typename T> V2<T> var_kepler_orbit(V2<T> p, T w, T ein, T scale) {
  T e = clampv(ein, (T)0, (T)0.95);
  T M = std::atan2(p.y, p.x);
  T a = scale * (len2(p) + (T)1.0e-6);
  T E = M;
  for (int i = 0; i < 2; ++i) {
    T s = safe_sin(E), c = safe_cos(E);
    E = E - (E - e * s - M) / wmax((T)1 - e * c, (T)0.05);
  }
  T b = std::sqrt(wmax((T)1 - e * e, (T)0));
  return {w * (a * (safe_cos(E) - e)), w * (a * b * safe_sin(E))};
}
// chaos.wgsl:7567 — Euler step in the rotating-frame restricted 3-body field.
template <typename T> V2<T> var_restricted_3body(V2<T> p, T w, T muin, T step, T coriolis) {
  T mu = clampv(muin, (T)0.01, (T)0.5);
  T r1x = p.x - -mu, r1y = p.y - (T)0;
  T r2x = p.x - ((T)1 - mu), r2y = p.y - (T)0;
  T d1 = std::pow(r1x * r1x + r1y * r1y + (T)1.0e-3, (T)1.5);
  T d2 = std::pow(r2x * r2x + r2y * r2y + (T)1.0e-3, (T)1.5);
  T gx = p.x - ((T)1 - mu) * (r1x / d1) - mu * (r2x / d2);
  T gy = p.y - ((T)1 - mu) * (r1y / d1) - mu * (r2y / d2);
  T dx = step * (gx + coriolis * -gy), dy = step * (gy + coriolis * gx);
  T dl = std::sqrt(dx * dx + dy * dy);
  if (dl > (T)2) { T c = (T)2 / dl; dx = dx * c; dy = dy * c; }
  return {w * (p.x + dx), w * (p.y + dy)};
}
// chaos.wgsl:7585 — Hill's epicyclic solution advanced by `phase`.
template <typename T> V2<T> var_hill_epicyclic(V2<T> p, T w, T kappa, T phase, T shear) {
  T k = clampv(kappa, (T)0.2, (T)2);
  T phi = phase * k;
  T cf = safe_cos(phi), sf = safe_sin(phi);
  T xo = p.x * cf - (sf / k) * p.y;
  T yo = ((T)2 * k) * sf * p.x + p.y * cf;
  return {w * xo, w * (yo + shear * p.x)};
}
// chaos.wgsl:7598
template <typename T> V2<T> var_torus_knot(V2<T> p, T w, T pp, T qq, T radius, T tube) {
  T phi = std::atan2(p.y, p.x);
  T pf = wmax(rint_even(pp), (T)1), qf = wmax(rint_even(qq), (T)1);
  T amp = tube * clampv(len2(p), (T)0, (T)1);
  T rr = radius + amp * safe_cos(qf * phi);
  return {w * (rr * safe_cos(pf * phi)), w * (rr * safe_sin(pf * phi))};
}
// chaos.wgsl:7612 — alternate angular lanes twisted in opposite senses.
template <typename T> V2<T> var_braid_warp(V2<T> p, T w, T strands, T twist, T crossings) {
  T n = wmax(rint_even(strands), (T)2), cr = wmax(rint_even(crossings), (T)1);
  T r = len2(p);
  T th = std::atan2(p.y, p.x);
  T a = (th + K<T>::PI) / K<T>::TAU;
  T lane = std::floor(a * n);
  T frac = a * n - lane;
  T sg = sel((T)1, (T)-1, (f2i(lane) & 1) == 1);
  T weave = safe_sin(K<T>::PI * frac);
  T dth = twist * sg * weave * safe_sin(cr * K<T>::PI * frac) * clampv(r, (T)0, (T)1);
  T nt = th + dth;
  return {w * r * safe_cos(nt), w * r * safe_sin(nt)};
}
// chaos.wgsl:7638
template <typename T> V2<T> var_airy_caustic(V2<T> p, T w, T scale, T r0, T amp) {
  T a = airy_ai(scale * (len2(p) - r0));
  T env = clampv((T)1 + amp * a, (T)0.05, (T)3);
  return {w * (env * p.x), w * (env * p.y)};
}
// chaos.wgsl:7650
template <typename T> V2<T> var_elliptic_E(V2<T> p, T w, T scale) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T e = elliptic_E(r2 / ((T)1 + r2));
  T k = scale * e;
  if (r < (T)1.0e-9) return {w * k * (T)0, w * k * (T)0};
  T m = wmax(r, (T)1.0e-12);
  return {w * k * (p.x / m), w * k * (p.y / m)};
}
// chaos.wgsl:7661
template <typename T> V2<T> var_elliptic_K(V2<T> p, T w, T scale, T tail) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T k = elliptic_K(r2 / ((T)1 + r2));
  return along(p, w, r, wmin(scale * k, wmax(tail, (T)0.01)));
}
// chaos.wgsl:7676
template <typename T> V2<T> var_gaussian_cdf(V2<T> p, T w, T mu, T sigma) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T q = erfinv_w((T)2 * cdf_u(r2) - (T)1);
  return along(p, w, r, wmax(mu + sigma * (T)1.4142135 * q, (T)0));
}
// chaos.wgsl:7689
template <typename T> V2<T> var_levy_cdf(V2<T> p, T w, T c, T tail) {
  T r2 = dot2(p), r = std::sqrt(r2);
  T e = erfinv_w((T)1 - cdf_u(r2));
  T raw = wmax(c, (T)1.0e-4) / wmax((T)2 * e * e, (T)1.0e-6);
  return along(p, w, r, wmin(raw, wmax(tail, (T)0.01)));
}
// chaos.wgsl:7707
template <typename T> V2<T> var_peano(V2<T> p, T w, T extent, T trits_f) {
  const uint32_t t = f2u(clampv(trits_f, (T)2, (T)15));
  uint32_t sx = peano_scramble(tri_encode(p.x, extent, t), t);
  uint32_t sy = peano_scramble(tri_encode(p.y, extent, t), t);
  return {w * tri_decode(sx, extent, t), w * tri_decode(sy, extent, t)};
}

// ---- dispatch, catalog indices 99..325 (chaos.wgsl:7848-8106) --------------------
// Entries 99..101 are colour-only (no position contribution).
template <typename T, typename R>
V2<T> apply_variation_ext(int idx, V2<T> p, const VarArgs<T>& va, R& rng) {
  const T w = va.w;
  const T* q = va.p;
  switch (idx) {
    case 99: case 100: case 101: return {(T)0, (T)0};
    case 102: return var_dc_cylinder_pos(p, w);
    case 103: return var_cpow2(p, w, q[0], q[1], q[2], q[3], rng);
    case 104: return var_cpow3(p, w, q[0], q[1], q[2], q[3], rng);
    case 105: return var_loonie2(p, w, q[0], q[1], q[2]);
    case 106: return var_epispiral(p, w, q[0], q[1], q[2], rng);
    case 107: return var_bwraps(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 108: return var_crackle(p, w, q[0], q[1], q[2], q[3], rng);
    case 109: return var_juliaq(p, w, q[0], q[1], rng);
    case 110: return var_glynnia(p, w, rng);
    case 111: return var_loonie3(p, w);
    case 112: return var_falloff(p, w, q[0], q[1], q[2], q[3], q[4], q[5], rng);
    case 113: return var_falloff2(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], rng);
    case 114: return var_falloff3(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], rng);
    case 115: return var_collideoscope(p, w, q[0], q[1]);
    case 116: return var_circlize(p, w, q[0]);
    case 117: return var_circlize2(p, w, q[0]);
    case 118: return var_eswirl(p, w, q[0], q[1]);
    case 119: return var_petal(p, w);
    case 120: return var_bcircle(p, w, q[0], q[1], rng);
    case 121: return var_curl2(p, w, q[0], q[1], q[2]);
    case 122: return var_murl(p, w, q[0], q[1]);
    case 123: return var_stwins(p, w, q[0]);
    case 124: return var_hexes(p, w, q[0], q[1], q[2], q[3]);
    case 125: return var_xheart(p, w, q[0], q[1]);
    case 126: return var_xhyperbol(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 127: return var_xcurl2(p, w, q[0], q[1], q[2]);
    case 128: return var_xtrb(p, w, q[0], q[1], q[2], q[3], q[4], q[5], rng);
    case 129: return var_xyrus_gridout(p, w);
    case 130: return var_blur_circle(p, w, q[0], rng);
    case 131: return var_bipolar2(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8]);
    case 132: return var_bubble2(p, w, q[0], q[1]);
    case 133: return var_acosh(p, w, rng);
    case 134: return var_arcsinh(p, w);
    case 135: return var_arctanh(p, w);
    case 136: return var_acoth(p, w);
    case 137: return var_acosech(p, w, rng);
    case 138: return var_arcsech2(p, w);
    case 139: return var_cell2(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 140: return var_curl_sp(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 141: return var_murl2(p, w, q[0], q[1]);
    case 142: return var_lissajous(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], rng);
    case 143: return var_spirograph(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], rng);
    case 144: return var_waffle(p, w, q[0], q[1], q[2], q[3], rng);
    case 145: return var_glynnSim1(p, w, q[0], q[1], q[2], q[3], q[4], q[5], rng);
    case 146: return var_glynnSim2(p, w, q[0], q[1], q[2], q[3], q[4], q[5], rng);
    case 147: return var_glynnSim3(p, w, q[0], q[1], q[2], q[3], rng);
    case 148: return var_flipy(p, w);
    case 149: return var_eclipse(p, w, q[0]);
    case 150: return var_barycentroid(p, w, q[0], q[1], q[2], q[3]);
    case 151: return var_chunk(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6]);
    case 152: return var_ennepers(p, w);
    case 153: return var_erf(p, w);
    case 154: return var_circus(p, w, q[0]);
    case 155: return var_asteria(p, w, q[0], rng);
    case 156: return var_clifford_js(p, w, q[0], q[1], q[2], q[3]);
    case 157: return var_devil_warp(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 158: return var_voron(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 159: return var_henon(p, w, q[0], q[1], q[2]);
    case 160: return var_atan(p, w, q[0], q[1]);
    case 161: return var_cardioid(p, w, q[0]);
    case 162: return var_chrysanthemum(p, w, rng);
    case 163: return var_bcollide(p, w, q[0], q[1]);
    case 164: return var_bsplit(p, w, q[0], q[1]);
    case 165: return var_bulge(p, w, q[0]);
    case 166: return var_checks(p, w, q[0], q[1], q[2], q[3], rng);
    case 167: return var_circular(p, w, q[0], q[1], rng);
    case 168: return var_circular2(p, w, q[0], q[1], q[2], q[3], rng);
    case 169: return var_corners(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8]);
    case 170: return var_circleblur(p, w, rng);
    case 171: return var_fibonacci2(p, w, q[0], q[1]);
    case 172: return var_hypertile(p, w, q[0], q[1], q[2]);
    case 173: return var_hypertile1(p, w, q[0], q[1], rng);
    case 174: return var_hypertile2(p, w, q[0], q[1], rng);
    case 175: return var_idisc(p, w);
    case 176: return var_hole(p, w, q[0], q[1]);
    case 177: return var_kaleidoscope(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 178: return var_layered_spiral(p, w, q[0]);
    case 179: return var_linear_t(p, w, q[0], q[1]);
    case 180: return var_line(p, w, q[0], q[1], rng);
    case 181: return var_ovoid(p, w, q[0], q[1]);
    case 182: return var_phoenix_julia(p, w, q[0], q[1], q[2], q[3], rng);
    case 183: return var_unpolar(p, w);
    case 184: return var_shredrad(p, w, q[0], q[1]);
    case 185: return var_vogel(p, w, q[0], q[1], rng);
    case 186: return var_yin_yang(p, w, q[0], q[1], q[2], q[3], q[4], rng);
    case 187: return var_squish(p, w, q[0], rng);
    case 188: return var_target(p, w, q[0], q[1], q[2]);
    case 189: return var_funnel(p, w, q[0]);
    case 190: return var_holesq(p, w);
    case 191: return var_hole2(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 192: return var_lace_js(p, w, rng);
    case 193: return var_julia_outside(p, w, q[0], q[1], q[2], rng);
    case 194: return var_fourth(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 195: return var_pulse(p, w, q[0], q[1], q[2], q[3]);
    case 196: return var_rays1(p, w);
    case 197: return var_rays2(p, w);
    case 198: return var_rays3(p, w);
    case 199: return var_tancos(p, w);
    case 200: return var_twoface(p, w);
    case 201: return var_e_julia(p, w, q[0], rng);
    case 202: return var_cannabis_curve_wf(p, w, q[0], rng);
    case 203: return var_e_collide(p, w, q[0], q[1]);
    case 204: return var_e_mod(p, w, q[0], q[1]);
    case 205: return var_intersection(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], rng);
    case 206: return var_inv_squircular(p, w);
    case 207: return var_lozi(p, w, q[0], q[1], q[2]);
    case 208: return var_hypershift(p, w, q[0], q[1]);
    case 209: return var_hex_modulus(p, w, q[0]);
    case 210: return var_boarders2(p, w, q[0], q[1], q[2], rng);
    case 211: return var_b_mod(p, w, q[0], q[1]);
    case 212: return var_b_transform(p, w, q[0], q[1], q[2], q[3], rng);
    case 213: return var_parallel(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], rng);
    case 214: return var_waves3(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 215: return var_waves4(p, w, q[0], q[1], q[2], q[3], q[4], q[5]);
    case 216: return var_scry2(p, w, q[0], q[1], q[2]);
    case 217: return var_ennepers2(p, w, q[0], q[1], q[2]);
    case 218: return var_apollony(p, w, rng);
    case 219: return var_circlecrop(p, w, q[0], q[1], q[2], q[3], q[4], rng);
    case 220: return var_newton(p, w, q[0]);
    case 221: return var_blaschke(p, w, q[0], q[1]);
    case 222: return var_cayley(p, w, q[0]);
    case 223: return var_complex_gamma(p, w, q[0]);
    case 224: return var_lambert_w(p, w, q[0]);
    case 225: return var_mercator(p, w);
    case 226: return var_lambert(p, w);
    case 227: return var_mollweide(p, w);
    case 228: return var_hammer(p, w);
    case 229: return var_stereographic(p, w);
    case 230: return var_standard_map(p, w, q[0]);
    case 231: return var_de_jong(p, w, q[0], q[1], q[2], q[3]);
    case 232: return var_ikeda(p, w, q[0]);
    case 233: return var_box_fold(p, w, q[0]);
    case 234: return var_sphere_fold(p, w, q[0], q[1]);
    case 235: return var_mandelbox_step(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 236: return var_kifs_fold(p, w, q[0], q[1]);
    case 237: return var_arnold_cat(p, w);
    case 238: return var_bakers_map(p, w);
    case 239: return var_tent_map(p, w);
    case 240: return var_logistic_map(p, w, q[0]);
    case 241: return var_superellipse(p, w, q[0], q[1], q[2]);
    case 242: return var_limacon(p, w, q[0], q[1]);
    case 243: return var_epicycloid(p, w, q[0]);
    case 244: return var_catenary(p, w, q[0]);
    case 245: return var_tractrix(p, w);
    case 246: return var_tinkerbell(p, w, q[0], q[1], q[2], q[3]);
    case 247: return var_duffing(p, w, q[0], q[1], q[2], q[3]);
    case 248: return var_vanderpol(p, w, q[0], q[1]);
    case 249: return var_rossler(p, w, q[0], q[1]);
    case 250: return var_droste(p, w, q[0]);
    case 251: return var_logspiral(p, w, q[0], q[1]);
    case 252: return var_fermat_spiral(p, w, q[0]);
    case 253: return var_lituus(p, w, q[0]);
    case 254: return var_hyperbolic_spiral(p, w, q[0]);
    case 255: return var_weierstrass(p, w, q[0], q[1], q[2], q[3]);
    case 256: return var_takagi(p, w, q[0], q[1]);
    case 257: return var_cantor_stairs(p, w, q[0], q[1]);
    case 258: return var_billiard_circle(p, w, q[0], q[1], q[2]);
    case 259: return var_billiard_stadium(p, w, q[0], q[1], q[2], q[3]);
    case 260: return var_billiard_sinai(p, w, q[0], q[1], q[2], q[3]);
    case 261: return var_billiard_polygon(p, w, q[0], q[1], q[2], q[3]);
    case 262: return var_lorentz_boost(p, w, q[0], q[1]);
    case 263: return var_schwarzschild_lensing(p, w, q[0], q[1]);
    case 264: return var_field_dipole(p, w, q[0], q[1], q[2], q[3]);
    case 265: return var_magnetic_pendulum_pos(p, w, q[0], q[1], q[2], q[3]);
    case 266: return var_jacobi_theta(p, w, q[0]);
    case 267: return var_modular_lambda(p, w, q[0]);
    case 268: return var_klein_j(p, w, q[0]);
    case 269: return var_weierstrass_p(p, w, q[0], q[1], q[2], q[3]);
    case 270: return var_gauss_map(p, w);
    case 271: return var_nbody_lensing(p, w, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]);
    case 272: return var_curl_noise(p, w, q[0], q[1]);
    case 273: return var_bessel_j0(p, w, q[0]);
    case 274: return var_airy_radial(p, w, q[0], q[1]);
    case 275: return var_cornu_spiral(p, w, q[0]);
    case 276: return var_struve_h1(p, w, q[0]);
    case 277: return var_radical_inverse(p, w, q[0], q[1]);
    case 278: return var_gray_code(p, w, q[0], q[1]);
    case 279: return var_morton_zorder(p, w, q[0], q[1]);
    case 280: return var_chebyshev(p, w, q[0], q[1]);
    case 281: return var_legendre(p, w, q[0], q[1]);
    case 282: return var_spherical_harmonic(p, w, q[0], q[1], q[2]);
    case 283: return var_fourier_warp(p, w, q[0], q[1], q[2]);
    case 284: return var_snell_refraction(p, w, q[0], q[1]);
    case 285: return var_grin_lens(p, w, q[0], q[1]);
    case 286: return var_caustic_fold(p, w, q[0], q[1], q[2]);
    case 287: return var_chladni(p, w, q[0], q[1], q[2]);
    case 288: return var_standing_wave(p, w, q[0], q[1], q[2], q[3]);
    case 289: return var_moire(p, w, q[0], q[1], q[2], q[3]);
    case 290: return var_radial_shell(p, w, q[0], q[1], q[2]);
    case 291: return var_hydrogen_orbital(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 292: return var_weibull_cdf(p, w, q[0], q[1]);
    case 293: return var_logistic_cdf(p, w, q[0], q[1]);
    case 294: return var_cauchy_cdf(p, w, q[0], q[1]);
    case 295: return var_pareto_cdf(p, w, q[0], q[1], q[2]);
    case 296: return var_morlet(p, w, q[0], q[1], q[2]);
    case 297: return var_mexican_hat(p, w, q[0], q[1]);
    case 298: return var_chirp(p, w, q[0], q[1], q[2]);
    case 299: return var_kepler_orbit(p, w, q[0], q[1]);
    case 300: return var_restricted_3body(p, w, q[0], q[1], q[2]);
    case 301: return var_hill_epicyclic(p, w, q[0], q[1], q[2]);
    case 302: return var_torus_knot(p, w, q[0], q[1], q[2], q[3]);
    case 303: return var_braid_warp(p, w, q[0], q[1], q[2]);
    case 304: return var_airy_caustic(p, w, q[0], q[1], q[2]);
    case 305: return var_elliptic_E(p, w, q[0]);
    case 306: return var_elliptic_K(p, w, q[0], q[1]);
    case 307: return var_gaussian_cdf(p, w, q[0], q[1]);
    case 308: return var_levy_cdf(p, w, q[0], q[1]);
    case 309: return var_peano(p, w, q[0], q[1]);
    case 310: return var_burning_ship(p, w, q[0], q[1]);
    case 311: return var_magnet1(p, w, q[0], q[1]);
    case 312: return var_nova(p, w, q[0], q[1], q[2]);
    case 313: return var_halley(p, w, q[0], q[1]);
    case 314: return var_lichtenberg(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 315: return var_copula_gaussian(p, w, q[0], q[1]);
    case 316: return var_copula_clayton(p, w, q[0], q[1]);
    case 317: return var_schwarz_christoffel(p, w, q[0]);
    case 318: return var_doyle(p, w, q[0], q[1]);
    case 319: return var_quasicrystal(p, w, q[0], q[1]);
    case 320: return var_penrose(p, w, q[0], q[1]);
    case 321: return var_collatz(p, w, q[0], q[1]);
    case 322: return var_digamma(p, w, q[0], q[1]);
    case 323: return var_sprott_poly(p, w, q);
    case 324: return var_hopalong(p, w, q[0], q[1], q[2]);
    case 325: return var_gumowski_mira(p, w, q[0], q[1]);
    default: return {(T)0, (T)0};
  }
}

// ---- direct-colour side (chaos.wgsl:1682-1780, 5408, 5484, 6410, 8192-8210) ------
// Colours are f32 only: they override the palette sample at the splat.

struct RGB { float r, g, b; };

// chaos.wgsl:1696 — HSL to RGB (h wrapped, s and l clamped).
inline RGB hsl_rgb(float h_in, float s_in, float l_in) {
  float h = fract(h_in);
  float s = clampv(s_in, 0.0f, 1.0f), l = clampv(l_in, 0.0f, 1.0f);
  float c = (1.0f - std::fabs(2.0f * l - 1.0f)) * s;
  float h6 = h * 6.0f;
  float x = c * (1.0f - std::fabs((h6 - 2.0f * std::floor(h6 * 0.5f)) - 1.0f));
  float m = l - c * 0.5f;
  RGB o;
  if (h6 < 1.0f) o = {c, x, 0.0f};
  else if (h6 < 2.0f) o = {x, c, 0.0f};
  else if (h6 < 3.0f) o = {0.0f, c, x};
  else if (h6 < 4.0f) o = {0.0f, x, c};
  else if (h6 < 5.0f) o = {x, 0.0f, c};
  else o = {c, 0.0f, x};
  return {o.r + m, o.g + m, o.b + m};
}
// chaos.wgsl:1682
inline RGB dc_linear_color(V2<float> p) {
  return {clampv(0.5f + 0.5f * p.x, 0.0f, 1.0f), clampv(0.5f + 0.5f * p.y, 0.0f, 1.0f),
          clampv(0.5f - 0.25f * (p.x + p.y), 0.0f, 1.0f)};
}
// chaos.wgsl:1721
inline RGB dc_perlin_color(V2<float> p, float scale, float octaves, float seed) {
  float n = perlin_fbm(p.x, p.y, octaves, wmax(scale, 1e-6f));
  return hsl_rgb(fract(0.5f + 0.5f * n + seed), 1.0f, 0.55f);
}
// chaos.wgsl:1732 — hashed colour per grid cell.
inline RGB dc_gridout_color(V2<float> p, float cells) {
  float n = wmax(cells, 1.0f);
  const int32_t cx = f2i(std::floor(p.x * n)), cy = f2i(std::floor(p.y * n));
  const uint32_t mixed = (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u)) + 0x9e3779b9u;
  return {hash01(mixed), hash01(mixed ^ 0xdeadbeefu), hash01(mixed ^ 0x13579bdfu)};
}
// chaos.wgsl:1756
inline RGB dc_cylinder_color(V2<float> p) {
  float hue = fract(0.5f + 0.5f * safe_sin(p.x));
  float lit = clampv(0.5f + 0.25f * safe_tanh(p.y * 0.5f), 0.2f, 0.85f);
  return hsl_rgb(hue, 0.9f, lit);
}
// chaos.wgsl:5408 — hue of the root of unity nearest the Newton image.
inline RGB newton_color(V2<float> p, float nin) {
  const int32_t n = imin(imax(f2i(nin + 0.5f), 2), 8);
  V2<float> z = var_newton(p, 1.0f, (float)n);
  int32_t best = 0;
  float bd = 1.0e30f;
  const float step = 6.2831853f / (float)n;
  for (int32_t k = 0; k < n; ++k) {
    float ang = step * (float)k;
    float dx = z.x - std::cos(ang), dy = z.y - std::sin(ang);
    float d2 = dx * dx + dy * dy;
    if (d2 < bd) { bd = d2; best = k; }
  }
  return hsl_rgb((float)best / (float)n, 1.0f, 0.55f);
}
// chaos.wgsl:6410 — hue of the nearest magnet.
inline RGB magnet_color(V2<float> p, float magnets, float radius) {
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
inline RGB escape_color(V2<float> p, int idx, float cx, float cy, float relax) {
  V2<float> z = p;
  int32_t it = 0;
  for (int32_t i = 0; i < 12; ++i) {
    if (idx == 310) z = var_burning_ship(z, 1.0f, cx, cy);
    else if (idx == 311) z = var_magnet1(z, 1.0f, cx, cy);
    else if (idx == 312) z = var_nova(z, 1.0f, cx, cy, relax);
    else z = var_halley(z, 1.0f, cx, cy);
    it = i + 1;
    if (z.x * z.x + z.y * z.y > 256.0f) break;
  }
  float r = wmax(len2(z), 1.0001f);
  float mu = (float)it + 1.0f - std::log2(wmax(std::log(r), 1.0e-6f));
  return hsl_rgb(fract(mu * 0.12f), 1.0f, 0.55f);
}
// chaos.wgsl:8192 — colour override of one slot with parameters p0, p1, p2.
inline bool dc_color_ext(int idx, V2<float> p, float p0, float p1, float p2, float* rgb) {
  RGB c;
  switch (idx) {
    case 99: c = dc_linear_color(p); break;
    case 100: c = dc_perlin_color(p, p0, p1, p2); break;
    case 101: c = dc_gridout_color(p, p0); break;
    case 102: c = dc_cylinder_color(p); break;
    case 220: c = newton_color(p, p0); break;
    case 265: c = magnet_color(p, p0, p1); break;
    case 310: case 311: case 313: c = escape_color(p, idx, p0, p1, 0.0f); break;
    case 312: c = escape_color(p, idx, p0, p1, p2); break;
    default: return false;
  }
  rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b;
  return true;
}

}  // namespace orc
