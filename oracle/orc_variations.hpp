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
#include <cstring>

namespace orc {

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

// ---- V0..V2 (config-0 set) -------------------------------------------------
// chaos.wgsl:524-527
template <typename T> V2<T> var_linear(V2<T> p, T w) { return {p.x * w, p.y * w}; }
// chaos.wgsl:528-531
template <typename T> V2<T> var_sinusoidal(V2<T> p, T w) {
  return {w * safe_sin(p.x), w * safe_sin(p.y)};
}
// chaos.wgsl:532-536
template <typename T> V2<T> var_spherical(V2<T> p, T w) {
  T r2 = dot2(p) + K<T>::EPS;
  T s = w / r2;
  return {p.x * s, p.y * s};
}

// Dispatch by catalog index (src/variations.ts:24 V table). Index 97
// (pre_blur) never reaches here: the chain applies it before the sum.
// Unknown indices contribute (0, 0), like the WGSL default arm.
template <typename T, typename R>
V2<T> apply_variation(int idx, V2<T> p, const VarArgs<T>& a, R& rng) {
  (void)rng;
  const T w = a.w;
  switch (idx) {
    case 0: return var_linear(p, w);
    case 1: return var_sinusoidal(p, w);
    case 2: return var_spherical(p, w);
    default: return {(T)0, (T)0};
  }
}

// Direct-colour override of one slot (chaos.wgsl:8192-8210); the dc_*
// catalog entries are not restated yet, so no slot overrides the palette.
inline bool dc_color_for(int idx, V2<float> p, const float* xf, int k, float* rgb) {
  (void)idx; (void)p; (void)xf; (void)k; (void)rgb;
  return false;
}

// Variation indices this restatement covers (used by tests to skip the rest).
inline bool variation_supported(int idx) { return idx >= 0 && idx <= 2; }

}  // namespace orc
