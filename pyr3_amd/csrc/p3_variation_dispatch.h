// Catalog dispatch of the gfx950 chaos kernel: index -> variation function
// (V0..V98 in p3_variations.h, V99..V325 in p3_variations_ext.h) and the
// direct-colour lookup.
#pragma once
#include "p3_variations_ext.h"

namespace p3 {

// Dispatch by catalog index (src/variations.ts:24 V table). Index 97
// (pre_blur) is applied by the chain before the sum; unknown indices
// contribute (0, 0) like the WGSL default arm (the host rejects them).
P3_INLINE float2 apply_variation(int idx, float2 p, xrec_t xf, int k, float w,
                                         Rng& rng) {
#define P(i) var_param(xf, k, i)
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
    case 14: return var_julian(p, w, P(0), P(1), rng);
    case 15: return var_bent(p, w);
    case 16: return var_waves(p, w, xf, xf + 4);
    case 17: return var_fisheye(p, w);
    case 18: return var_popcorn(p, w, xf, xf + 4);
    case 19: return var_eyefish(p, w);
    case 20: return var_bubble(p, w);
    case 21: return var_cylinder(p, w);
    case 22: return var_disc2(p, w, P(0), P(1));
    case 23: return var_pdj(p, w, P(0), P(1), P(2), P(3));
    case 24: return var_exponential(p, w);
    case 25: return var_power(p, w);
    case 26: return var_cosine(p, w);
    case 27: return var_tangent(p, w);
    case 28: return var_secant2(p, w);
    case 29: return var_cross(p, w);
    case 30: return var_rings(p, w, xf);
    case 31: return var_fan(p, w, xf, xf + 4);
    case 32: return var_rings2(p, w, P(0));
    case 33: return var_fan2(p, w, P(0), P(1));
    case 34: return var_perspective(p, w, P(0), P(1));
    case 35: return var_bipolar(p, w, P(0));
    case 36: return var_curl(p, w, P(0), P(1));
    case 37: return var_rectangles(p, w, P(0), P(1));
    case 38: return var_blob(p, w, P(0), P(1), P(2));
    case 39: return var_ngon(p, w, P(0), P(1), P(2), P(3));
    case 40: return var_wedge(p, w, P(0), P(1), P(2), P(3));
    case 41: return var_cpow(p, w, P(0), P(1), P(2), rng);
    case 42: return var_curve(p, w, P(0), P(1), P(2), P(3));
    case 43: return var_noise(p, w, rng);
    case 44: return var_blur(p, w, rng);
    case 45: return var_gaussian_blur(p, w, rng);
    case 46: return var_arch(p, w, rng);
    case 47: return var_radial_blur(p, w, P(0), rng);
    case 48: return var_juliascope(p, w, P(0), P(1), rng);
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
    case 74: return var_bent2(p, w, P(0), P(1));
    case 75: return var_cell(p, w, P(0));
    case 76: return var_escher(p, w, P(0));
    case 77: return var_modulus(p, w, P(0), P(1));
    case 78: return var_split(p, w, P(0), P(1));
    case 79: return var_splits(p, w, P(0), P(1));
    case 80: return var_stripes(p, w, P(0), P(1));
    case 81: return var_whorl(p, w, P(0), P(1));
    case 82: return var_flux(p, w, P(0));
    case 83: return var_popcorn2(p, w, P(0), P(1), P(2));
    case 84: return var_lazysusan(p, w, P(0), P(1), P(2), P(3), P(4));
    case 85: return var_waves2(p, w, P(0), P(1), P(2), P(3));
    case 86: return var_oscope(p, w, P(0), P(1), P(2), P(3));
    case 87: return var_separation(p, w, P(0), P(1), P(2), P(3));
    case 88: return var_auger(p, w, P(0), P(1), P(2), P(3));
    case 89: return var_wedge_sph(p, w, P(0), P(1), P(2), P(3));
    case 90: return var_super_shape(p, w, P(0), P(1), P(2), P(3), P(4), P(5), rng);
    case 91: return var_flower(p, w, P(0), P(1), rng);
    case 92: return var_conic(p, w, P(0), P(1), rng);
    case 93: return var_parabola(p, w, P(0), P(1), rng);
    case 94: return var_pie(p, w, P(0), P(1), P(2), rng);
    case 95: return var_boarders(p, w, rng);
    case 96: return var_wedge_julia(p, w, P(0), P(1), P(2), P(3), rng);
    case 98: {
      const float m[8] = {P(0), P(1), P(2), P(3), P(4), P(5), P(6), P(7)};
      return var_mobius(p, w, m);
    }
    default: return apply_variation_ext(idx, p, xf, k, w, rng);
  }

#undef P
}

// Direct-colour override of slot k (chaos.wgsl:8192-8210); false when the
// entry has no colour side.
P3_INLINE bool dc_color_for(int idx, float2 p, xrec_t xf, int k, float3& rgb) {
  return dc_color_ext(idx, p, xf, k, rgb);
}

}  // namespace p3
