// Variation library — oracle restatement, templated on the float type.
// Test infrastructure only: see oracle.hpp.
//
// T = float  follows the FP32 WGSL kernels (src/shaders/chaos.wgsl:524-7715,
//            dispatch apply_variation :7731-8107) and is what the chaos oracle runs.
// T = double is the same expression tree in f64; it is checked against the
//            instrumented-flam3 f64 fixtures (tests/golden/variations/*.json),
//            the way src/variations.test.ts:145-175 checks its f64 oracles.
//
// The RNG is a template parameter R with irand()/rand01()/rand11(); the chaos
// oracle passes the walker's ISAAC stream, the fixture tests pass a scripted
// stream (branch injection, as src/variations.test.ts:193-236 does).
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <type_traits>
#include <cstring>

namespace orc {

// WGSL max / min (and clamp = min(max(x, lo), hi)): a NaN operand yields the
// other operand; otherwise max(a, b) = a < b ? b : a and min(a, b) =
// b < a ? b : a, which also fixes which zero a tie of +0 and -0 returns.
template <typename T> inline T wmax(T a, T b) {
  if (std::is_floating_point<T>::value && a != a) return b;
  return a < b ? b : a;
}
template <typename T> inline T wmin(T a, T b) {
  if (std::is_floating_point<T>::value && a != a) return b;
  return b < a ? b : a;
}

template <typename T> struct V2 { T x, y; };

template <typename T> struct K;
template <> struct K<float> {
  static constexpr float PI = 3.14159265358979323846f;
  static constexpr float TAU = 6.28318530717958647692f;
  static constexpr float EPS = 1e-10f;
};
template <> struct K<double> {
  static constexpr double PI = 3.14159265358979323846;
  static constexpr double TAU = 6.28318530717958647692;
  static constexpr double EPS = 1e-10;
};

// Arguments of one variation slot: weight, ten params and the owning xform's
// affine rows (waves/popcorn/rings/fan read c and f from them).
template <typename T> struct VarArgs {
  T w;
  T p[10];
  T a0[4];  // a, b, c, weight
  T a1[4];  // d, e, f, nvars
};

// chaos.wgsl:356-362 — integer hash used when a trig argument is out of range.
inline float hash01(uint32_t h) {
  h ^= h >> 17; h *= 0xed5ad4bbu;
  h ^= h >> 11; h *= 0xac4c1b51u;
  h ^= h >> 15;
  return (float)h / 4294967296.0f;
}
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// chaos.wgsl:363-384 — trig with a range guard at |a| > 1e6 and tanh clamp.
template <typename T> inline T safe_sin(T a) {
  if (std::fabs(a) <= (T)1.0e6) return std::sin(a);
  return std::sin((T)(hash01(fbits((float)a)) * K<float>::TAU));
}
template <typename T> inline T safe_cos(T a) {
  if (std::fabs(a) <= (T)1.0e6) return std::cos(a);
  return std::cos((T)(hash01(fbits((float)a)) * K<float>::TAU));
}
template <typename T> inline T safe_tan(T a) { return safe_sin(a) / safe_cos(a); }
template <typename T> inline T safe_tanh(T a) {
  T c = a < (T)-20 ? (T)-20 : (a > (T)20 ? (T)20 : a);
  return std::tanh(c);
}

template <typename T> inline T dot2(V2<T> p) { return p.x * p.x + p.y * p.y; }
template <typename T> inline T len2(V2<T> p) { return std::sqrt(p.x * p.x + p.y * p.y); }

template <typename T> inline T wmod(T x, T y) { return x - y * std::trunc(x / y); }  // WGSL f32 %
template <typename T> inline int32_t f2i(T f) {  // WGSL i32(f): saturating, NaN -> 0
  if (!(f == f)) return 0;
  if (f >= (T)2147483647.0) return 2147483647;
  if (f <= (T)-2147483648.0) return (-2147483647 - 1);
  return (int32_t)f;
}
template <typename T> inline T sel(T f, T t, bool c) { return c ? t : f; }  // WGSL select

// ---- flam3 core set, V0..V98 (chaos.wgsl:524-1661) -------------------------
// Angle conventions follow the reference: atan2(x, y) for polar, handkerchief,
// heart, disc, ex, julia, fan, blob, fan2, disc2; atan2(y, x) elsewhere.

template <typename T> V2<T> var_linear(V2<T> p, T w) { return {p.x * w, p.y * w}; }
template <typename T> V2<T> var_sinusoidal(V2<T> p, T w) {
  return {w * safe_sin(p.x), w * safe_sin(p.y)};
}
template <typename T> V2<T> var_spherical(V2<T> p, T w) {
  T r2 = dot2(p) + K<T>::EPS;
  T s = w / r2;
  return {p.x * s, p.y * s};
}
template <typename T> V2<T> var_swirl(V2<T> p, T w) {
  T r2 = dot2(p);
  T s = safe_sin(r2), c = safe_cos(r2);
  return {w * (s * p.x - c * p.y), w * (c * p.x + s * p.y)};
}
template <typename T> V2<T> var_horseshoe(V2<T> p, T w) {
  T r = len2(p) + K<T>::EPS;
  T k = w / r;
  return {k * ((p.x - p.y) * (p.x + p.y)), k * ((T)2 * p.x * p.y)};
}
template <typename T> V2<T> var_polar(V2<T> p, T w) {
  T phi = std::atan2(p.x, p.y), r = len2(p);
  return {w * (phi / K<T>::PI), w * (r - (T)1)};
}
template <typename T> V2<T> var_handkerchief(V2<T> p, T w) {
  T phi = std::atan2(p.x, p.y), r = len2(p);
  T k = w * r;
  return {k * safe_sin(phi + r), k * safe_cos(phi - r)};
}
template <typename T> V2<T> var_heart(V2<T> p, T w) {
  T phi = std::atan2(p.x, p.y), r = len2(p);
  T k = w * r;
  return {k * safe_sin(phi * r), k * -safe_cos(phi * r)};
}
template <typename T> V2<T> var_disc(V2<T> p, T w) {
  T phi = std::atan2(p.x, p.y), r = len2(p);
  T k = w * (phi / K<T>::PI);
  return {k * safe_sin(K<T>::PI * r), k * safe_cos(K<T>::PI * r)};
}
template <typename T> V2<T> var_spiral(V2<T> p, T w) {
  T r = len2(p) + K<T>::EPS;
  T sina = p.x / r, cosa = p.y / r;
  T k = w / r;
  return {k * (cosa + safe_sin(r)), k * (sina - safe_cos(r))};
}
template <typename T> V2<T> var_hyperbolic(V2<T> p, T w) {
  T r = len2(p) + K<T>::EPS;
  T sina = p.x / r, cosa = p.y / r;
  return {w * (sina / r), w * (r * cosa)};
}
template <typename T> V2<T> var_diamond(V2<T> p, T w) {
  T r = len2(p) + K<T>::EPS;
  T sina = p.x / r, cosa = p.y / r;
  T ro = len2(p);
  return {w * (sina * safe_cos(ro)), w * (cosa * safe_sin(ro))};
}
template <typename T> V2<T> var_ex(V2<T> p, T w) {
  T phi = std::atan2(p.x, p.y), r = len2(p);
  T n0 = safe_sin(phi + r), n1 = safe_cos(phi - r);
  T m0 = n0 * n0 * n0, m1 = n1 * n1 * n1;
  T k = w * r;
  return {k * (m0 + m1), k * (m0 - m1)};
}
template <typename T, typename R> V2<T> var_julia(V2<T> p, T w, R& rng) {
  T phi = std::atan2(p.x, p.y);
  T theta = phi * (T)0.5 + sel((T)0, K<T>::PI, (rng.irand() & 1u) == 1u);
  T r = std::sqrt(len2(p));
  T k = w * r;
  return {k * safe_cos(theta), k * safe_sin(theta)};
}
template <typename T, typename R> V2<T> var_julian(V2<T> p, T w, T power, T dist, R& rng) {
  T r = len2(p), phi = std::atan2(p.y, p.x);
  T pa = std::fabs(power);
  T n = std::floor(rng.rand01() * pa);
  T theta = (phi + K<T>::TAU * n) / power;
  T nr = w * std::pow(r, dist / power);
  return {nr * safe_cos(theta), nr * safe_sin(theta)};
}
template <typename T> V2<T> var_bent(V2<T> p, T w) {
  T x = sel(p.x, p.x * (T)2, p.x < 0), y = sel(p.y, p.y * (T)0.5, p.y < 0);
  return {w * x, w * y};
}
template <typename T> V2<T> var_waves(V2<T> p, T w, const T* a0, const T* a1) {
  T b = a0[1], c = a0[2], e = a1[1], f = a1[2];
  return {w * (p.x + b * safe_sin(p.y / (c * c + K<T>::EPS))),
          w * (p.y + e * safe_sin(p.x / (f * f + K<T>::EPS)))};
}
template <typename T> V2<T> var_fisheye(V2<T> p, T w) {
  T r = (T)2 / (len2(p) + (T)1);
  T k = w * r;
  return {k * p.y, k * p.x};
}
template <typename T> V2<T> var_popcorn(V2<T> p, T w, const T* a0, const T* a1) {
  T c = a0[2], f = a1[2];
  return {w * (p.x + c * safe_sin(safe_tan((T)3 * p.y))),
          w * (p.y + f * safe_sin(safe_tan((T)3 * p.x)))};
}
template <typename T> V2<T> var_eyefish(V2<T> p, T w) {
  T r = (T)2 / (len2(p) + (T)1);
  T k = w * r;
  return {k * p.x, k * p.y};
}
template <typename T> V2<T> var_bubble(V2<T> p, T w) {
  T r = w / ((T)0.25 * dot2(p) + (T)1);
  return {r * p.x, r * p.y};
}
template <typename T> V2<T> var_cylinder(V2<T> p, T w) { return {w * safe_sin(p.x), w * p.y}; }
template <typename T> V2<T> var_pdj(V2<T> p, T w, T pa, T pb, T pc, T pd) {
  T nx1 = safe_cos(pb * p.x), nx2 = safe_sin(pc * p.x);
  T ny1 = safe_sin(pa * p.y), ny2 = safe_cos(pd * p.y);
  return {w * (ny1 - nx1), w * (nx2 - ny2)};
}
template <typename T> V2<T> var_disc2(V2<T> p, T w, T rot, T twist) {
  T timespi = rot * K<T>::PI;
  T cosadd = safe_cos(twist) - (T)1;
  T sinadd = safe_sin(twist);
  if (twist > K<T>::TAU) {
    T k = (T)1 + twist - K<T>::TAU;
    cosadd = cosadd * k; sinadd = sinadd * k;
  } else if (twist < -K<T>::TAU) {
    T k = (T)1 + twist + K<T>::TAU;
    cosadd = cosadd * k; sinadd = sinadd * k;
  }
  T t = timespi * (p.x + p.y);
  T sinr = safe_sin(t), cosr = safe_cos(t);
  T r = (w * std::atan2(p.x, p.y)) / K<T>::PI;
  return {(sinr + cosadd) * r, (cosr + sinadd) * r};
}
template <typename T> V2<T> var_exponential(V2<T> p, T w) {
  T dx = w * std::exp(p.x - (T)1);
  T dy = K<T>::PI * p.y;
  return {dx * safe_cos(dy), dx * safe_sin(dy)};
}
template <typename T> V2<T> var_power(V2<T> p, T w) {
  T r = len2(p) + K<T>::EPS;
  T sina = p.x / r, cosa = p.y / r;
  T k = w * std::pow(r, sina);
  return {k * cosa, k * sina};
}
template <typename T> V2<T> var_cosine(V2<T> p, T w) {
  T a = p.x * K<T>::PI;
  return {w * (safe_cos(a) * std::cosh(p.y)), w * (-safe_sin(a) * std::sinh(p.y))};
}
template <typename T> V2<T> var_tangent(V2<T> p, T w) {
  return {w * (safe_sin(p.x) / safe_cos(p.y)), w * safe_tan(p.y)};
}
template <typename T> V2<T> var_secant2(V2<T> p, T w) {
  T r = w * len2(p);
  T cr = safe_cos(r);
  T icr = (T)1 / cr;
  T y = sel(w * (icr - (T)1), w * (icr + (T)1), cr < 0);
  return {w * p.x, y};
}
template <typename T> V2<T> var_cross(V2<T> p, T w) {
  T s = p.x * p.x - p.y * p.y;
  T r = w * std::sqrt((T)1 / (s * s + K<T>::EPS));
  return {p.x * r, p.y * r};
}
template <typename T> V2<T> var_rings(V2<T> p, T w, const T* a0) {
  T c = a0[2];
  T r0 = len2(p), re = r0 + K<T>::EPS;
  T sina = p.x / re, cosa = p.y / re;
  T dx = c * c + K<T>::EPS;
  T r = w * ((wmod(r0 + dx, (T)2 * dx) - dx) + r0 * ((T)1 - dx));
  return {r * cosa, r * sina};
}
template <typename T> V2<T> var_fan(V2<T> p, T w, const T* a0, const T* a1) {
  T c = a0[2], f = a1[2];
  T dx = K<T>::PI * (c * c + K<T>::EPS);
  T dy = f, dx2 = (T)0.5 * dx;
  T phi = std::atan2(p.x, p.y);
  T r = w * len2(p);
  T t = (phi + dy) - dx * std::trunc((phi + dy) / dx);
  T a = sel(phi + dx2, phi - dx2, t > dx2);
  return {r * safe_cos(a), r * safe_sin(a)};
}
template <typename T> V2<T> var_rings2(V2<T> p, T w, T val) {
  T r0 = len2(p), re = r0 + K<T>::EPS;
  T sina = p.x / re, cosa = p.y / re;
  T dx = val * val + K<T>::EPS;
  T r = (r0 - (T)2 * dx * std::trunc((r0 + dx) / ((T)2 * dx))) + r0 * ((T)1 - dx);
  T k = w * r;
  return {k * sina, k * cosa};
}
template <typename T> V2<T> var_fan2(V2<T> p, T w, T fx, T fy) {
  T phi = std::atan2(p.x, p.y);
  T r = w * len2(p);
  T dy = fy, dx = K<T>::PI * (fx * fx + K<T>::EPS), dx2 = (T)0.5 * dx;
  T t = (phi + dy) - dx * std::trunc((phi + dy) / dx);
  T a = sel(phi + dx2, phi - dx2, t > dx2);
  return {r * safe_sin(a), r * safe_cos(a)};
}
template <typename T> V2<T> var_perspective(V2<T> p, T w, T angle, T dist) {
  T hpa = angle * (K<T>::PI * (T)0.5);
  T vsin = safe_sin(hpa), vfcos = dist * safe_cos(hpa);
  T t = (T)1 / (dist - p.y * vsin);
  return {w * (dist * p.x * t), w * (vfcos * p.y * t)};
}
template <typename T> V2<T> var_bipolar(V2<T> p, T w, T shift) {
  const T HP = K<T>::PI * (T)0.5;
  const T TOP = (T)2 / K<T>::PI;
  T x2y2 = dot2(p);
  T t = x2y2 + (T)1;
  T x2 = (T)2 * p.x;
  T ps = -HP * shift;
  T y = (T)0.5 * std::atan2((T)2 * p.y, x2y2 - (T)1) + ps;
  if (y > HP) y = -HP + wmod(y + HP, K<T>::PI);
  else if (y < -HP) y = HP - wmod(HP - y, K<T>::PI);
  return {w * (T)0.25 * TOP * std::log((t + x2) / (t - x2)), w * TOP * y};
}
template <typename T> V2<T> var_curl(V2<T> p, T w, T c1, T c2) {
  T re = (T)1 + c1 * p.x + c2 * (p.x * p.x - p.y * p.y);
  T im = c1 * p.y + (T)2 * c2 * p.x * p.y;
  T r = w / (re * re + im * im);
  return {(p.x * re + p.y * im) * r, (p.y * re - p.x * im) * r};
}
template <typename T> V2<T> var_rectangles(V2<T> p, T w, T rx, T ry) {
  T rxs = sel(rx, (T)1, rx == 0), rys = sel(ry, (T)1, ry == 0);
  T ox = sel(w * (((T)2 * std::floor(p.x / rxs) + (T)1) * rxs - p.x), w * p.x, rx == 0);
  T oy = sel(w * (((T)2 * std::floor(p.y / rys) + (T)1) * rys - p.y), w * p.y, ry == 0);
  return {ox, oy};
}
template <typename T> V2<T> var_blob(V2<T> p, T w, T low, T high, T waves) {
  T r0 = len2(p), re = r0 + K<T>::EPS;
  T sina = p.x / re, cosa = p.y / re;
  T a = std::atan2(p.x, p.y);
  T r = r0 * (low + (high - low) * ((T)0.5 + (T)0.5 * safe_sin(waves * a)));
  T k = w * r;
  return {k * sina, k * cosa};
}
template <typename T> V2<T> var_ngon(V2<T> p, T w, T sides, T power, T circle, T corners) {
  T sumsq = dot2(p);
  T rf = std::pow(sumsq, power * (T)0.5);
  T theta = std::atan2(p.y, p.x);
  T b = K<T>::TAU / sides;
  T phi = theta - b * std::floor(theta / b);
  if (phi > b * (T)0.5) phi = phi - b;
  T amp = (corners * ((T)1 / (safe_cos(phi) + K<T>::EPS) - (T)1) + circle) / (rf + K<T>::EPS);
  T k = w * amp;
  return {k * p.x, k * p.y};
}
template <typename T> V2<T> var_wedge(V2<T> p, T w, T angle, T hole, T count, T swirl) {
  T r0 = len2(p);
  T a = std::atan2(p.y, p.x) + swirl * r0;
  const T OOP = (T)1 / K<T>::PI;
  T c = std::floor((count * a + K<T>::PI) * OOP * (T)0.5);
  T cf = (T)1 - angle * count * OOP * (T)0.5;
  a = a * cf + c * angle;
  T r = w * (r0 + hole);
  return {r * safe_cos(a), r * safe_sin(a)};
}
template <typename T, typename R> V2<T> var_cpow(V2<T> p, T w, T pr, T pi, T pw, R& rng) {
  T a = std::atan2(p.y, p.x);
  T lnr = (T)0.5 * std::log(dot2(p));
  T va = K<T>::TAU / pw, vc = pr / pw, vd = pi / pw;
  T n = std::floor(rng.rand01() * std::fabs(pw));
  T ang = vc * a + vd * lnr + va * n;
  T m = w * std::exp(vc * lnr - vd * a);
  return {m * safe_cos(ang), m * safe_sin(ang)};
}
template <typename T> V2<T> var_curve(V2<T> p, T w, T xamp, T yamp, T xlen, T ylen) {
  T pcx = wmax(xlen * xlen, (T)1e-20), pcy = wmax(ylen * ylen, (T)1e-20);
  return {w * (p.x + xamp * std::exp((-p.y * p.y) / pcx)),
          w * (p.y + yamp * std::exp((-p.x * p.x) / pcy))};
}
template <typename T, typename R> V2<T> var_noise(V2<T> p, T w, R& rng) {
  T tmpr = rng.rand01() * K<T>::TAU;
  T r = w * rng.rand01();
  return {p.x * r * safe_cos(tmpr), p.y * r * safe_sin(tmpr)};
}
template <typename T, typename R> V2<T> var_blur(V2<T> p, T w, R& rng) {
  (void)p;
  T tmpr = rng.rand01() * K<T>::TAU;
  T r = w * rng.rand01();
  return {r * safe_cos(tmpr), r * safe_sin(tmpr)};
}
template <typename T, typename R> V2<T> var_gaussian_blur(V2<T> p, T w, R& rng) {
  (void)p;
  T ang = rng.rand01() * K<T>::TAU;
  T r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
  T r = w * (r0 + r1 + r2 + r3 - (T)2);
  return {r * safe_cos(ang), r * safe_sin(ang)};
}
template <typename T, typename R> V2<T> var_arch(V2<T> p, T w, R& rng) {
  (void)p;
  T ang = rng.rand01() * w * K<T>::PI;
  T s = safe_sin(ang), c = safe_cos(ang);
  return {w * s, w * (s * s) / c};
}
template <typename T, typename R> V2<T> var_radial_blur(V2<T> p, T w, T angle, R& rng) {
  T hpa = angle * (K<T>::PI * (T)0.5);
  T spin = safe_sin(hpa), zoom = safe_cos(hpa);
  T r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
  T g = w * (r0 + r1 + r2 + r3 - (T)2);
  T ra = len2(p);
  T tmpa = std::atan2(p.y, p.x) + spin * g;
  T sa = safe_sin(tmpa), ca = safe_cos(tmpa);
  T rz = zoom * g - (T)1;
  return {ra * ca + rz * p.x, ra * sa + rz * p.y};
}
template <typename T, typename R> V2<T> var_juliascope(V2<T> p, T w, T power, T dist, R& rng) {
  T phi = std::atan2(p.y, p.x);
  T sumsq = dot2(p);
  T pa = std::fabs(power);
  int32_t t = f2i(std::floor(rng.rand01() * pa));
  T tmpr = sel((K<T>::TAU * (T)t - phi) / power, (K<T>::TAU * (T)t + phi) / power, (t & 1) == 0);
  T r = w * std::pow(sumsq, dist / power / (T)2);
  return {r * safe_cos(tmpr), r * safe_sin(tmpr)};
}
template <typename T, typename R> V2<T> var_square(V2<T> p, T w, R& rng) {
  (void)p;
  T r0 = rng.rand01(), r1 = rng.rand01();
  return {w * (r0 - (T)0.5), w * (r1 - (T)0.5)};
}
template <typename T, typename R> V2<T> var_rays(V2<T> p, T w, R& rng) {
  T ang = w * rng.rand01() * K<T>::PI;
  T r = w / (dot2(p) + K<T>::EPS);
  T tanr = w * safe_tan(ang) * r;
  return {tanr * safe_cos(p.x), tanr * safe_sin(p.y)};
}
template <typename T, typename R> V2<T> var_blade(V2<T> p, T w, R& rng) {
  T r = rng.rand01() * w * len2(p);
  T s = safe_sin(r), c = safe_cos(r);
  return {w * p.x * (c + s), w * p.x * (c - s)};
}
template <typename T, typename R> V2<T> var_twintrian(V2<T> p, T w, R& rng) {
  const T LN10 = (T)2.302585092994046;
  T r = rng.rand01() * w * len2(p);
  T s = safe_sin(r), c = safe_cos(r);
  T diff = std::log(s * s) / LN10 + c;
  if (diff != diff || diff > (T)1e10 || diff < (T)-1e10) diff = (T)-30;
  return {w * p.x * diff, w * p.x * (diff - s * K<T>::PI)};
}
template <typename T> V2<T> var_exp(V2<T> p, T w) {
  T e = std::exp(p.x);
  T k = w * e;
  return {k * safe_cos(p.y), k * safe_sin(p.y)};
}
template <typename T> V2<T> var_log(V2<T> p, T w) {
  return {w * ((T)0.5 * std::log(dot2(p))), w * std::atan2(p.y, p.x)};
}
template <typename T> V2<T> var_sin(V2<T> p, T w) {
  return {w * (safe_sin(p.x) * std::cosh(p.y)), w * (safe_cos(p.x) * std::sinh(p.y))};
}
template <typename T> V2<T> var_cos(V2<T> p, T w) {
  return {w * (safe_cos(p.x) * std::cosh(p.y)), w * (-safe_sin(p.x) * std::sinh(p.y))};
}
template <typename T> V2<T> var_tan(V2<T> p, T w) {
  T den = (T)1 / (safe_cos((T)2 * p.x) + std::cosh((T)2 * p.y));
  T k = w * den;
  return {k * safe_sin((T)2 * p.x), k * std::sinh((T)2 * p.y)};
}
template <typename T> V2<T> var_sec(V2<T> p, T w) {
  T den = (T)2 / (safe_cos((T)2 * p.x) + std::cosh((T)2 * p.y));
  T k = w * den;
  return {k * (safe_cos(p.x) * std::cosh(p.y)), k * (safe_sin(p.x) * std::sinh(p.y))};
}
template <typename T> V2<T> var_csc(V2<T> p, T w) {
  T den = (T)2 / (std::cosh((T)2 * p.y) - safe_cos((T)2 * p.x));
  T k = w * den;
  return {k * (safe_sin(p.x) * std::cosh(p.y)), k * (-safe_cos(p.x) * std::sinh(p.y))};
}
template <typename T> V2<T> var_cot(V2<T> p, T w) {
  T den = (T)1 / (std::cosh((T)2 * p.y) - safe_cos((T)2 * p.x));
  T k = w * den;
  return {k * safe_sin((T)2 * p.x), k * -std::sinh((T)2 * p.y)};
}
template <typename T> V2<T> var_sinh(V2<T> p, T w) {
  return {w * (std::sinh(p.x) * safe_cos(p.y)), w * (std::cosh(p.x) * safe_sin(p.y))};
}
template <typename T> V2<T> var_cosh(V2<T> p, T w) {
  return {w * (std::cosh(p.x) * safe_cos(p.y)), w * (std::sinh(p.x) * safe_sin(p.y))};
}
template <typename T> V2<T> var_tanh(V2<T> p, T w) {
  T den = (T)1 / (safe_cos((T)2 * p.y) + std::cosh((T)2 * p.x));
  T k = w * den;
  return {k * std::sinh((T)2 * p.x), k * safe_sin((T)2 * p.y)};
}
template <typename T> V2<T> var_sech(V2<T> p, T w) {
  T den = (T)2 / (safe_cos((T)2 * p.y) + std::cosh((T)2 * p.x));
  T k = w * den;
  return {k * (safe_cos(p.y) * std::cosh(p.x)), k * (-safe_sin(p.y) * std::sinh(p.x))};
}
template <typename T> V2<T> var_csch(V2<T> p, T w) {
  T den = (T)2 / (std::cosh((T)2 * p.x) - safe_cos((T)2 * p.y));
  T k = w * den;
  return {k * (std::sinh(p.x) * safe_cos(p.y)), k * (-std::cosh(p.x) * safe_sin(p.y))};
}
template <typename T> V2<T> var_coth(V2<T> p, T w) {
  T den = (T)1 / (std::cosh((T)2 * p.x) - safe_cos((T)2 * p.y));
  T k = w * den;
  return {k * std::sinh((T)2 * p.x), k * safe_sin((T)2 * p.y)};
}
template <typename T> V2<T> var_butterfly(V2<T> p, T w) {
  T wx = w * (T)1.3029400317411197908970256609023;
  T y2 = p.y * (T)2;
  T r = wx * std::sqrt(std::fabs(p.y * p.x) / (K<T>::EPS + p.x * p.x + y2 * y2));
  return {r * p.x, r * y2};
}
template <typename T> V2<T> var_edisc(V2<T> p, T w) {
  T tmp = dot2(p) + (T)1;
  T tmp2 = (T)2 * p.x;
  T r1 = std::sqrt(tmp + tmp2), r2 = std::sqrt(tmp - tmp2);
  T xmax = (r1 + r2) * (T)0.5;
  T a1 = std::log(xmax + std::sqrt(wmax(xmax - (T)1, (T)0)));
  T a2 = -std::acos(wmin(wmax(p.x / xmax, (T)-1), (T)1));
  T wn = w / (T)11.57034632;
  T csv = safe_cos(a1);
  T snv = sel(safe_sin(a1), -safe_sin(a1), p.y > 0);
  T snhu = std::sinh(a2), cshu = std::cosh(a2);
  return {wn * cshu * csv, wn * snhu * snv};
}
template <typename T> V2<T> var_elliptic(V2<T> p, T w) {
  T tmp = dot2(p) + (T)1;
  T x2 = (T)2 * p.x;
  T xmax = (T)0.5 * (std::sqrt(tmp + x2) + std::sqrt(tmp - x2));
  T a = p.x / xmax;
  T b = std::sqrt(wmax((T)1 - a * a, (T)0));
  T ssx = std::sqrt(wmax(xmax - (T)1, (T)0));
  T wn = w / (K<T>::PI * (T)0.5);
  T ylog = wn * std::log(xmax + ssx);
  return {wn * std::atan2(a, b), sel(-ylog, ylog, p.y > 0)};
}
template <typename T> V2<T> var_foci(V2<T> p, T w) {
  T expx = std::exp(p.x) * (T)0.5;
  T expnx = (T)0.25 / expx;
  T tmp = w / (expx + expnx - safe_cos(p.y));
  return {tmp * (expx - expnx), tmp * safe_sin(p.y)};
}
template <typename T> V2<T> var_loonie(V2<T> p, T w) {
  T r2 = dot2(p), w2 = w * w;
  if (r2 < w2) {
    T r = w * std::sqrt(w2 / r2 - (T)1);
    return {r * p.x, r * p.y};
  }
  return {w * p.x, w * p.y};
}
template <typename T> V2<T> var_polar2(V2<T> p, T w) {
  T p2v = w / K<T>::PI;
  return {p2v * std::atan2(p.x, p.y), (p2v * (T)0.5) * std::log(dot2(p) + K<T>::EPS)};
}
template <typename T> V2<T> var_scry(V2<T> p, T w) {
  T s = dot2(p);
  T r = (T)1 / (std::sqrt(s) * (s + (T)1 / (w + K<T>::EPS)));
  return {p.x * r, p.y * r};
}
template <typename T> V2<T> var_bent2(V2<T> p, T w, T bx, T by) {
  return {w * sel(p.x, p.x * bx, p.x < 0), w * sel(p.y, p.y * by, p.y < 0)};
}
template <typename T> V2<T> var_cell(V2<T> p, T w, T size) {
  T inv = (T)1 / size;
  int32_t x = f2i(std::floor(p.x * inv)), y = f2i(std::floor(p.y * inv));
  T dx = p.x - (T)x * size, dy = p.y - (T)y * size;
  // WGSL i32 arithmetic wraps: do it in u32.
  auto twice = [](int32_t v) { return (int32_t)((uint32_t)v * 2u); };
  auto neg_odd = [](int32_t v) { return (int32_t)(0u - ((uint32_t)v * 2u + 1u)); };
  if (y >= 0) {
    if (x >= 0) { y = twice(y); x = twice(x); }
    else { y = twice(y); x = neg_odd(x); }
  } else {
    if (x >= 0) { y = neg_odd(y); x = twice(x); }
    else { y = neg_odd(y); x = neg_odd(x); }
  }
  return {w * (dx + (T)x * size), -w * (dy + (T)y * size)};
}
template <typename T> V2<T> var_escher(V2<T> p, T w, T beta) {
  T a = std::atan2(p.y, p.x);
  T lnr = (T)0.5 * std::log(dot2(p));
  T seb = safe_sin(beta), ceb = safe_cos(beta);
  T vc = (T)0.5 * ((T)1 + ceb), vd = (T)0.5 * seb;
  T m = w * std::exp(vc * lnr - vd * a);
  T n = vc * a + vd * lnr;
  return {m * safe_cos(n), m * safe_sin(n)};
}
template <typename T> V2<T> var_modulus(V2<T> p, T w, T mx, T my) {
  T xr = (T)2 * mx, yr = (T)2 * my;
  T ox, oy;
  if (p.x > mx) ox = w * (-mx + wmod(p.x + mx, xr));
  else if (p.x < -mx) ox = w * (mx - wmod(mx - p.x, xr));
  else ox = w * p.x;
  if (p.y > my) oy = w * (-my + wmod(p.y + my, yr));
  else if (p.y < -my) oy = w * (my - wmod(my - p.y, yr));
  else oy = w * p.y;
  return {ox, oy};
}
template <typename T> V2<T> var_split(V2<T> p, T w, T xs, T ys) {
  T oy = sel(-w * p.y, w * p.y, safe_cos(p.x * xs * K<T>::PI) >= 0);
  T ox = sel(-w * p.x, w * p.x, safe_cos(p.y * ys * K<T>::PI) >= 0);
  return {ox, oy};
}
template <typename T> V2<T> var_splits(V2<T> p, T w, T sx, T sy) {
  return {sel(w * (p.x - sx), w * (p.x + sx), p.x >= 0), sel(w * (p.y - sy), w * (p.y + sy), p.y >= 0)};
}
template <typename T> V2<T> var_stripes(V2<T> p, T w, T space, T warp) {
  T rx = std::floor(p.x + (T)0.5);
  T ox = p.x - rx;
  return {w * (ox * ((T)1 - space) + rx), w * (p.y + ox * ox * warp)};
}
template <typename T> V2<T> var_whorl(V2<T> p, T w, T inside, T outside) {
  T r = len2(p);
  T base = std::atan2(p.y, p.x);
  T a = sel(base + outside / (w - r), base + inside / (w - r), r < w);
  return {w * r * safe_cos(a), w * r * safe_sin(a)};
}
template <typename T> V2<T> var_flux(V2<T> p, T w, T spread) {
  T xpw = p.x + w, xmw = p.x - w;
  T ty = p.y * p.y;
  T avgr = w * ((T)2 + spread) * std::sqrt(std::sqrt(ty + xpw * xpw) / std::sqrt(ty + xmw * xmw));
  T avga = (std::atan2(p.y, xmw) - std::atan2(p.y, xpw)) * (T)0.5;
  return {avgr * safe_cos(avga), avgr * safe_sin(avga)};
}
template <typename T> V2<T> var_popcorn2(V2<T> p, T w, T px, T py, T pc) {
  return {w * (p.x + px * safe_sin(safe_tan(p.y * pc))), w * (p.y + py * safe_sin(safe_tan(p.x * pc)))};
}
template <typename T> V2<T> var_lazysusan(V2<T> p, T w, T lx, T ly, T spin, T twist, T space) {
  T x = p.x - lx, y = p.y + ly;
  T r0 = std::sqrt(x * x + y * y);
  if (r0 < w) {
    T a = std::atan2(y, x) + spin + twist * (w - r0);
    T r = w * r0;
    return {r * safe_cos(a) + lx, r * safe_sin(a) - ly};
  }
  T r = w * ((T)1 + space / r0);
  return {r * x + lx, r * y - ly};
}
template <typename T> V2<T> var_waves2(V2<T> p, T w, T sx, T fx, T sy, T fy) {
  return {w * (p.x + sx * safe_sin(p.y * fx)), w * (p.y + sy * safe_sin(p.x * fy))};
}
template <typename T> V2<T> var_oscope(V2<T> p, T w, T freq, T amp, T damping, T sep) {
  T tpf = K<T>::TAU * freq;
  T t = sel(amp * std::exp(-std::fabs(p.x) * damping) * safe_cos(tpf * p.x) + sep,
            amp * safe_cos(tpf * p.x) + sep, damping == 0);
  if (std::fabs(p.y) <= t) return {w * p.x, -w * p.y};
  return {w * p.x, w * p.y};
}
template <typename T> V2<T> var_separation(V2<T> p, T w, T sx, T sxi, T sy, T syi) {
  T sx2 = sx * sx, sy2 = sy * sy;
  T ox = sel(-w * (std::sqrt(p.x * p.x + sx2) + p.x * sxi), w * (std::sqrt(p.x * p.x + sx2) - p.x * sxi), p.x > 0);
  T oy = sel(-w * (std::sqrt(p.y * p.y + sy2) + p.y * syi), w * (std::sqrt(p.y * p.y + sy2) - p.y * syi), p.y > 0);
  return {ox, oy};
}
template <typename T> V2<T> var_auger(V2<T> p, T w, T freq, T ww, T scale, T sym) {
  T s = safe_sin(freq * p.x), t = safe_sin(freq * p.y);
  T dy = p.y + ww * (scale * s * (T)0.5 + std::fabs(p.y) * s);
  T dx = p.x + ww * (scale * t * (T)0.5 + std::fabs(p.x) * t);
  return {w * (p.x + sym * (dx - p.x)), w * dy};
}
template <typename T> V2<T> var_wedge_sph(V2<T> p, T w, T angle, T hole, T count, T swirl) {
  T r0 = len2(p);
  T ri = (T)1 / (r0 + K<T>::EPS);
  T a = std::atan2(p.y, p.x) + swirl * ri;
  const T OOP = (T)1 / K<T>::PI;
  T c = std::floor((count * a + K<T>::PI) * OOP * (T)0.5);
  T cf = (T)1 - angle * count * OOP * (T)0.5;
  a = a * cf + c * angle;
  T r = w * (ri + hole);
  return {r * safe_cos(a), r * safe_sin(a)};
}
template <typename T, typename R>
V2<T> var_super_shape(V2<T> p, T w, T rnd, T m, T n1, T n2, T n3, T holes, R& rng) {
  T pm4 = m * (T)0.25, pn1 = (T)-1 / n1;
  T r0 = len2(p);
  T theta = pm4 * std::atan2(p.y, p.x) + K<T>::PI * (T)0.25;
  T st = safe_sin(theta), ct = safe_cos(theta);
  T t1 = std::pow(std::fabs(ct), n2), t2 = std::pow(std::fabs(st), n3);
  T r = w * ((rnd * rng.rand01() + ((T)1 - rnd) * r0) - holes) * std::pow(t1 + t2, pn1) / r0;
  return {r * p.x, r * p.y};
}
template <typename T, typename R> V2<T> var_flower(V2<T> p, T w, T petals, T holes, R& rng) {
  T theta = std::atan2(p.y, p.x);
  T r0 = len2(p);
  T r = w * (rng.rand01() - holes) * safe_cos(petals * theta) / r0;
  return {r * p.x, r * p.y};
}
template <typename T, typename R> V2<T> var_conic(V2<T> p, T w, T ecc, T holes, R& rng) {
  T r0 = len2(p);
  T ct = p.x / r0;
  T r = w * (rng.rand01() - holes) * ecc / ((T)1 + ecc * ct) / r0;
  return {r * p.x, r * p.y};
}
template <typename T, typename R> V2<T> var_parabola(V2<T> p, T w, T height, T width, R& rng) {
  T r = len2(p);
  T sr = safe_sin(r), cr = safe_cos(r);
  T r0 = rng.rand01(), r1 = rng.rand01();
  return {height * w * sr * sr * r0, width * w * cr * r1};
}
template <typename T, typename R> V2<T> var_pie(V2<T> p, T w, T slices, T rotation, T thick, R& rng) {
  (void)p;
  T r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01();
  T sl = std::trunc(r0 * slices + (T)0.5);
  T a = rotation + K<T>::TAU * (sl + r1 * thick) / slices;
  T r = w * r2;
  return {r * safe_cos(a), r * safe_sin(a)};
}
template <typename T, typename R> V2<T> var_boarders(V2<T> p, T w, R& rng) {
  T r0 = rng.rand01();
  T rx = std::floor(p.x + (T)0.5), ry = std::floor(p.y + (T)0.5);
  T ox = p.x - rx, oy = p.y - ry;
  if (r0 >= (T)0.75) return {w * (ox * (T)0.5 + rx), w * (oy * (T)0.5 + ry)};
  if (std::fabs(ox) >= std::fabs(oy)) {
    if (ox >= 0) return {w * (ox * (T)0.5 + rx + (T)0.25), w * (oy * (T)0.5 + ry + (T)0.25 * oy / ox)};
    return {w * (ox * (T)0.5 + rx - (T)0.25), w * (oy * (T)0.5 + ry - (T)0.25 * oy / ox)};
  }
  if (oy >= 0) return {w * (ox * (T)0.5 + rx + ox / oy * (T)0.25), w * (oy * (T)0.5 + ry + (T)0.25)};
  return {w * (ox * (T)0.5 + rx - ox / oy * (T)0.25), w * (oy * (T)0.5 + ry - (T)0.25)};
}
template <typename T> V2<T> var_mobius(V2<T> p, T w, const T* q) {
  T reu = q[0] * p.x - q[1] * p.y + q[2];
  T imu = q[0] * p.y + q[1] * p.x + q[3];
  T rev = q[4] * p.x - q[5] * p.y + q[6];
  T imv = q[4] * p.y + q[5] * p.x + q[7];
  T rad = w / (rev * rev + imv * imv);
  return {rad * (reu * rev + imu * imv), rad * (imu * rev - reu * imv)};
}
template <typename T, typename R>
V2<T> var_wedge_julia(V2<T> p, T w, T angle, T count, T power, T dist, R& rng) {
  T sumsq = dot2(p);
  T cn = dist / power * (T)0.5;
  T r = w * std::pow(sumsq, cn);
  T cf = (T)1 - angle * count / K<T>::TAU;
  T n = std::floor(rng.rand01() * std::fabs(power));
  T a = (std::atan2(p.y, p.x) + K<T>::TAU * n) / power;
  T c = std::floor((count * a + K<T>::PI) / K<T>::PI * (T)0.5);
  a = a * cf + c * angle;
  return {r * safe_cos(a), r * safe_sin(a)};
}

}  // namespace orc

#include "orc_variations_ext.hpp"

namespace orc {

// Dispatch by catalog index (src/variations.ts:24 V table). Index 97
// (pre_blur) never reaches here: the chain applies it before the sum.
// Unknown indices contribute (0, 0), like the WGSL default arm.
template <typename T, typename R>
V2<T> apply_variation(int idx, V2<T> p, const VarArgs<T>& a, R& rng) {
  const T w = a.w;
  const T* q = a.p;
  switch (idx) {
    case 0: return var_linear(p, w);
    case 1: return var_sinusoidal(p, w);
    case 2: return var_spherical(p, w);
    case 3: return var_swirl(p, w);
    case 4: return var_horseshoe(p, w);
    case 5: return var_polar(p, w);
    case 6: return var_handkerchief(p, w);
    case 7: return var_heart(p, w);
    case 8: return var_disc(p, w);
    case 9: return var_spiral(p, w);
    case 10: return var_hyperbolic(p, w);
    case 11: return var_diamond(p, w);
    case 12: return var_ex(p, w);
    case 13: return var_julia(p, w, rng);
    case 14: return var_julian(p, w, q[0], q[1], rng);
    case 15: return var_bent(p, w);
    case 16: return var_waves(p, w, a.a0, a.a1);
    case 17: return var_fisheye(p, w);
    case 18: return var_popcorn(p, w, a.a0, a.a1);
    case 19: return var_eyefish(p, w);
    case 20: return var_bubble(p, w);
    case 21: return var_cylinder(p, w);
    case 22: return var_disc2(p, w, q[0], q[1]);
    case 23: return var_pdj(p, w, q[0], q[1], q[2], q[3]);
    case 24: return var_exponential(p, w);
    case 25: return var_power(p, w);
    case 26: return var_cosine(p, w);
    case 27: return var_tangent(p, w);
    case 28: return var_secant2(p, w);
    case 29: return var_cross(p, w);
    case 30: return var_rings(p, w, a.a0);
    case 31: return var_fan(p, w, a.a0, a.a1);
    case 32: return var_rings2(p, w, q[0]);
    case 33: return var_fan2(p, w, q[0], q[1]);
    case 34: return var_perspective(p, w, q[0], q[1]);
    case 35: return var_bipolar(p, w, q[0]);
    case 36: return var_curl(p, w, q[0], q[1]);
    case 37: return var_rectangles(p, w, q[0], q[1]);
    case 38: return var_blob(p, w, q[0], q[1], q[2]);
    case 39: return var_ngon(p, w, q[0], q[1], q[2], q[3]);
    case 40: return var_wedge(p, w, q[0], q[1], q[2], q[3]);
    case 41: return var_cpow(p, w, q[0], q[1], q[2], rng);
    case 42: return var_curve(p, w, q[0], q[1], q[2], q[3]);
    case 43: return var_noise(p, w, rng);
    case 44: return var_blur(p, w, rng);
    case 45: return var_gaussian_blur(p, w, rng);
    case 46: return var_arch(p, w, rng);
    case 47: return var_radial_blur(p, w, q[0], rng);
    case 48: return var_juliascope(p, w, q[0], q[1], rng);
    case 49: return var_square(p, w, rng);
    case 50: return var_rays(p, w, rng);
    case 51: return var_blade(p, w, rng);
    case 52: return var_twintrian(p, w, rng);
    case 53: return var_exp(p, w);
    case 54: return var_log(p, w);
    case 55: return var_sin(p, w);
    case 56: return var_cos(p, w);
    case 57: return var_tan(p, w);
    case 58: return var_sec(p, w);
    case 59: return var_csc(p, w);
    case 60: return var_cot(p, w);
    case 61: return var_sinh(p, w);
    case 62: return var_cosh(p, w);
    case 63: return var_tanh(p, w);
    case 64: return var_sech(p, w);
    case 65: return var_csch(p, w);
    case 66: return var_coth(p, w);
    case 67: return var_butterfly(p, w);
    case 68: return var_edisc(p, w);
    case 69: return var_elliptic(p, w);
    case 70: return var_foci(p, w);
    case 71: return var_loonie(p, w);
    case 72: return var_polar2(p, w);
    case 73: return var_scry(p, w);
    case 74: return var_bent2(p, w, q[0], q[1]);
    case 75: return var_cell(p, w, q[0]);
    case 76: return var_escher(p, w, q[0]);
    case 77: return var_modulus(p, w, q[0], q[1]);
    case 78: return var_split(p, w, q[0], q[1]);
    case 79: return var_splits(p, w, q[0], q[1]);
    case 80: return var_stripes(p, w, q[0], q[1]);
    case 81: return var_whorl(p, w, q[0], q[1]);
    case 82: return var_flux(p, w, q[0]);
    case 83: return var_popcorn2(p, w, q[0], q[1], q[2]);
    case 84: return var_lazysusan(p, w, q[0], q[1], q[2], q[3], q[4]);
    case 85: return var_waves2(p, w, q[0], q[1], q[2], q[3]);
    case 86: return var_oscope(p, w, q[0], q[1], q[2], q[3]);
    case 87: return var_separation(p, w, q[0], q[1], q[2], q[3]);
    case 88: return var_auger(p, w, q[0], q[1], q[2], q[3]);
    case 89: return var_wedge_sph(p, w, q[0], q[1], q[2], q[3]);
    case 90: return var_super_shape(p, w, q[0], q[1], q[2], q[3], q[4], q[5], rng);
    case 91: return var_flower(p, w, q[0], q[1], rng);
    case 92: return var_conic(p, w, q[0], q[1], rng);
    case 93: return var_parabola(p, w, q[0], q[1], rng);
    case 94: return var_pie(p, w, q[0], q[1], q[2], rng);
    case 95: return var_boarders(p, w, rng);
    case 96: return var_wedge_julia(p, w, q[0], q[1], q[2], q[3], rng);
    case 98: return var_mobius(p, w, q);
    default: return apply_variation_ext(idx, p, a, rng);
  }
}

// Direct-colour override of slot k of record xf (chaos.wgsl:8192-8210):
// parameters p0, p1 at vars[k].zw and p2 at vars_extra[k].x.
inline bool dc_color_for(int idx, V2<float> p, const float* xf, int k, float* rgb) {
  return dc_color_ext(idx, p, xf[20 + 4 * k + 2], xf[20 + 4 * k + 3], xf[52 + 4 * k], rgb);
}

// Catalog indices with a position or colour side (97, pre_blur, is applied
// by the chain itself).
inline bool variation_supported(int idx) { return idx >= 0 && idx <= 325 && idx != 97; }

}  // namespace orc
