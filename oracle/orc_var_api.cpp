// Single-variation entry points for the fixture tests (test infrastructure).
// The RNG is a scripted stream: draws return the given words in order, and
// the return value is the number of draws consumed (or -1 for an index the
// restatement does not cover).
#include "oracle.hpp"
#include "orc_variations.hpp"

namespace {

template <typename T>
struct Scripted {
  const uint32_t* words;
  int n;
  int used = 0;
  uint32_t irand() { uint32_t v = used < n ? words[used] : 0u; ++used; return v; }
  T rand01() { return (T)orc::rand01_of(irand()); }
  T rand11() { return (T)orc::rand11_of(irand()); }
};

template <typename T>
int eval(int idx, T x, T y, const T* w, const T* params, const T* aff, const uint32_t* words,
         int n_words, T* out) {
  if (!orc::variation_supported(idx)) return -1;
  orc::VarArgs<T> a;
  a.w = *w;
  for (int i = 0; i < 10; ++i) a.p[i] = params[i];
  for (int i = 0; i < 4; ++i) { a.a0[i] = aff[i]; a.a1[i] = aff[4 + i]; }
  Scripted<T> rng{words, n_words};
  orc::V2<T> o = orc::apply_variation<T>(idx, orc::V2<T>{x, y}, a, rng);
  out[0] = o.x;
  out[1] = o.y;
  return rng.used;
}

}  // namespace

extern "C" {

int orc_variation_f64(int idx, double x, double y, const double* w, const double* params,
                      const double* aff, const uint32_t* words, int n_words, double* out) {
  return eval<double>(idx, x, y, w, params, aff, words, n_words, out);
}

int orc_variation_f32(int idx, float x, float y, const float* w, const float* params,
                      const float* aff, const uint32_t* words, int n_words, float* out) {
  return eval<float>(idx, x, y, w, params, aff, words, n_words, out);
}

int orc_variation_supported(int idx) { return orc::variation_supported(idx) ? 1 : 0; }

// Direct-colour side of entry idx at (x, y) with parameters p0..p2: writes
// RGB and returns 1, or returns 0 for an entry without one.
int orc_dc_color(int idx, float x, float y, float p0, float p1, float p2, float* rgb) {
  return orc::dc_color_ext(idx, orc::V2<float>{x, y}, p0, p1, p2, rgb) ? 1 : 0;
}

}  // extern "C"
