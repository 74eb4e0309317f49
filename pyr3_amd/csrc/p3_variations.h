// Variation catalog for the gfx950 chaos kernel.
//
// Each entry is the FP32 expression of the WGSL function of the same name in
// src/shaders/chaos.wgsl (V0..V98 at :524-1661, dispatch :7731-8107), with
// the same operation order (the library is built with -ffp-contract=off, so
// a*b+c keeps two roundings). The expression trees are the same ones the CPU
// oracle restates (its f64 instantiation is pinned by 4,607 instrumented-flam3
// rows). Parameters are read from the 464-byte xform record on demand:
// param0/1 at vars[k].zw, param2..5 at vars_extra[k], param6..9 at
// vars_extra2[k].
#pragma once
#include "p3_device.h"

namespace p3 {

P3_INLINE float dot2(float2 p) { return p.x * p.x + p.y * p.y; }
P3_INLINE float len2(float2 p) { return sqrtf(p.x * p.x + p.y * p.y); }
// WGSL max / min (clamp = min(max(x, lo), hi)): a NaN operand yields the other
// operand; a tie of +0 and -0 returns the first (max) or first (min) argument.
P3_INLINE float smax(float a, float b) { return isnan(a) ? b : (a < b ? b : a); }
P3_INLINE float smin(float a, float b) { return isnan(a) ? b : (b < a ? b : a); }
P3_INLINE float sel(float f, float t, bool c) { return c ? t : f; }  // WGSL select
P3_INLINE float wmod(float x, float y) { return x - y * truncf(x / y); }  // WGSL f32 %
P3_INLINE int32_t f2i(float f) {  // WGSL i32(f): saturating, NaN -> 0
  if (!(f == f)) return 0;
  if (f >= 2147483647.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)f;
}

// Parameter i of slot k of the xform record xf.
P3_INLINE float var_param(xrec_t xf, int k, int i) {
  if (i < 2) return xf[20 + 4 * k + 2 + i];
  if (i < 6) return xf[52 + 4 * k + (i - 2)];
  return xf[84 + 4 * k + (i - 6)];
}

// ---- flam3 core set, V0..V98 (chaos.wgsl:524-1661) -------------------------
// Angle conventions follow the reference: atan2(x, y) for polar, handkerchief,
// heart, disc, ex, julia, fan, blob, fan2, disc2; atan2(y, x) elsewhere.

P3_INLINE float2 var_linear(float2 p, float w) { return make_float2(p.x * w, p.y * w); }
P3_INLINE float2 var_sinusoidal(float2 p, float w) {
  return make_float2(w * safe_sin(p.x), w * safe_sin(p.y));
}
P3_INLINE float2 var_spherical(float2 p, float w) {
  float r2 = dot2(p) + kEps;
  float s = w / r2;
  return make_float2(p.x * s, p.y * s);
}
P3_INLINE float2 var_swirl(float2 p, float w) {
  float r2 = dot2(p);
  float s = safe_sin(r2), c = safe_cos(r2);
  return make_float2(w * (s * p.x - c * p.y), w * (c * p.x + s * p.y));
}
P3_INLINE float2 var_horseshoe(float2 p, float w) {
  float r = len2(p) + kEps;
  float k = w / r;
  return make_float2(k * ((p.x - p.y) * (p.x + p.y)), k * ((float)2 * p.x * p.y));
}
P3_INLINE float2 var_polar(float2 p, float w) {
  float phi = atan2f(p.x, p.y), r = len2(p);
  return make_float2(w * (phi / kPi), w * (r - (float)1));
}
P3_INLINE float2 var_handkerchief(float2 p, float w) {
  float phi = atan2f(p.x, p.y), r = len2(p);
  float k = w * r;
  return make_float2(k * safe_sin(phi + r), k * safe_cos(phi - r));
}
P3_INLINE float2 var_heart(float2 p, float w) {
  float phi = atan2f(p.x, p.y), r = len2(p);
  float k = w * r;
  return make_float2(k * safe_sin(phi * r), k * -safe_cos(phi * r));
}
P3_INLINE float2 var_disc(float2 p, float w) {
  float phi = atan2f(p.x, p.y), r = len2(p);
  float k = w * (phi / kPi);
  return make_float2(k * safe_sin(kPi * r), k * safe_cos(kPi * r));
}
P3_INLINE float2 var_spiral(float2 p, float w) {
  float r = len2(p) + kEps;
  float sina = p.x / r, cosa = p.y / r;
  float k = w / r;
  return make_float2(k * (cosa + safe_sin(r)), k * (sina - safe_cos(r)));
}
P3_INLINE float2 var_hyperbolic(float2 p, float w) {
  float r = len2(p) + kEps;
  float sina = p.x / r, cosa = p.y / r;
  return make_float2(w * (sina / r), w * (r * cosa));
}
P3_INLINE float2 var_diamond(float2 p, float w) {
  float r = len2(p) + kEps;
  float sina = p.x / r, cosa = p.y / r;
  float ro = len2(p);
  return make_float2(w * (sina * safe_cos(ro)), w * (cosa * safe_sin(ro)));
}
P3_INLINE float2 var_ex(float2 p, float w) {
  float phi = atan2f(p.x, p.y), r = len2(p);
  float n0 = safe_sin(phi + r), n1 = safe_cos(phi - r);
  float m0 = n0 * n0 * n0, m1 = n1 * n1 * n1;
  float k = w * r;
  return make_float2(k * (m0 + m1), k * (m0 - m1));
}
P3_INLINE float2 var_julia(float2 p, float w, Rng& rng) {
  float phi = atan2f(p.x, p.y);
  float theta = phi * (float)0.5 + sel((float)0, kPi, (rng.irand() & 1u) == 1u);
  float r = sqrtf(len2(p));
  float k = w * r;
  return make_float2(k * safe_cos(theta), k * safe_sin(theta));
}
P3_INLINE float2 var_julian(float2 p, float w, float power, float dist, Rng& rng) {
  float r = len2(p), phi = atan2f(p.y, p.x);
  float pa = fabsf(power);
  float n = floorf(rng.rand01() * pa);
  float theta = (phi + kTau * n) / power;
  float nr = w * powf(r, dist / power);
  return make_float2(nr * safe_cos(theta), nr * safe_sin(theta));
}
P3_INLINE float2 var_bent(float2 p, float w) {
  float x = sel(p.x, p.x * (float)2, p.x < 0), y = sel(p.y, p.y * (float)0.5, p.y < 0);
  return make_float2(w * x, w * y);
}
P3_INLINE float2 var_waves(float2 p, float w, xrec_t a0, xrec_t a1) {
  float b = a0[1], c = a0[2], e = a1[1], f = a1[2];
  return make_float2(w * (p.x + b * safe_sin(p.y / (c * c + kEps))),
          w * (p.y + e * safe_sin(p.x / (f * f + kEps))));
}
P3_INLINE float2 var_fisheye(float2 p, float w) {
  float r = (float)2 / (len2(p) + (float)1);
  float k = w * r;
  return make_float2(k * p.y, k * p.x);
}
P3_INLINE float2 var_popcorn(float2 p, float w, xrec_t a0, xrec_t a1) {
  float c = a0[2], f = a1[2];
  return make_float2(w * (p.x + c * safe_sin(safe_tan((float)3 * p.y))),
          w * (p.y + f * safe_sin(safe_tan((float)3 * p.x))));
}
P3_INLINE float2 var_eyefish(float2 p, float w) {
  float r = (float)2 / (len2(p) + (float)1);
  float k = w * r;
  return make_float2(k * p.x, k * p.y);
}
P3_INLINE float2 var_bubble(float2 p, float w) {
  float r = w / ((float)0.25 * dot2(p) + (float)1);
  return make_float2(r * p.x, r * p.y);
}
P3_INLINE float2 var_cylinder(float2 p, float w) { return make_float2(w * safe_sin(p.x), w * p.y); }
P3_INLINE float2 var_pdj(float2 p, float w, float pa, float pb, float pc, float pd) {
  float nx1 = safe_cos(pb * p.x), nx2 = safe_sin(pc * p.x);
  float ny1 = safe_sin(pa * p.y), ny2 = safe_cos(pd * p.y);
  return make_float2(w * (ny1 - nx1), w * (nx2 - ny2));
}
P3_INLINE float2 var_disc2(float2 p, float w, float rot, float twist) {
  float timespi = rot * kPi;
  float cosadd = safe_cos(twist) - (float)1;
  float sinadd = safe_sin(twist);
  if (twist > kTau) {
    float k = (float)1 + twist - kTau;
    cosadd = cosadd * k; sinadd = sinadd * k;
  } else if (twist < -kTau) {
    float k = (float)1 + twist + kTau;
    cosadd = cosadd * k; sinadd = sinadd * k;
  }
  float t = timespi * (p.x + p.y);
  float sinr = safe_sin(t), cosr = safe_cos(t);
  float r = (w * atan2f(p.x, p.y)) / kPi;
  return make_float2((sinr + cosadd) * r, (cosr + sinadd) * r);
}
P3_INLINE float2 var_exponential(float2 p, float w) {
  float dx = w * expf(p.x - (float)1);
  float dy = kPi * p.y;
  return make_float2(dx * safe_cos(dy), dx * safe_sin(dy));
}
P3_INLINE float2 var_power(float2 p, float w) {
  float r = len2(p) + kEps;
  float sina = p.x / r, cosa = p.y / r;
  float k = w * powf(r, sina);
  return make_float2(k * cosa, k * sina);
}
P3_INLINE float2 var_cosine(float2 p, float w) {
  float a = p.x * kPi;
  return make_float2(w * (safe_cos(a) * coshf(p.y)), w * (-safe_sin(a) * sinhf(p.y)));
}
P3_INLINE float2 var_tangent(float2 p, float w) {
  return make_float2(w * (safe_sin(p.x) / safe_cos(p.y)), w * safe_tan(p.y));
}
P3_INLINE float2 var_secant2(float2 p, float w) {
  float r = w * len2(p);
  float cr = safe_cos(r);
  float icr = (float)1 / cr;
  float y = sel(w * (icr - (float)1), w * (icr + (float)1), cr < 0);
  return make_float2(w * p.x, y);
}
P3_INLINE float2 var_cross(float2 p, float w) {
  float s = p.x * p.x - p.y * p.y;
  float r = w * sqrtf((float)1 / (s * s + kEps));
  return make_float2(p.x * r, p.y * r);
}
P3_INLINE float2 var_rings(float2 p, float w, xrec_t a0) {
  float c = a0[2];
  float r0 = len2(p), re = r0 + kEps;
  float sina = p.x / re, cosa = p.y / re;
  float dx = c * c + kEps;
  float r = w * ((wmod(r0 + dx, (float)2 * dx) - dx) + r0 * ((float)1 - dx));
  return make_float2(r * cosa, r * sina);
}
P3_INLINE float2 var_fan(float2 p, float w, xrec_t a0, xrec_t a1) {
  float c = a0[2], f = a1[2];
  float dx = kPi * (c * c + kEps);
  float dy = f, dx2 = (float)0.5 * dx;
  float phi = atan2f(p.x, p.y);
  float r = w * len2(p);
  float t = (phi + dy) - dx * truncf((phi + dy) / dx);
  float a = sel(phi + dx2, phi - dx2, t > dx2);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}
P3_INLINE float2 var_rings2(float2 p, float w, float val) {
  float r0 = len2(p), re = r0 + kEps;
  float sina = p.x / re, cosa = p.y / re;
  float dx = val * val + kEps;
  float r = (r0 - (float)2 * dx * truncf((r0 + dx) / ((float)2 * dx))) + r0 * ((float)1 - dx);
  float k = w * r;
  return make_float2(k * sina, k * cosa);
}
P3_INLINE float2 var_fan2(float2 p, float w, float fx, float fy) {
  float phi = atan2f(p.x, p.y);
  float r = w * len2(p);
  float dy = fy, dx = kPi * (fx * fx + kEps), dx2 = (float)0.5 * dx;
  float t = (phi + dy) - dx * truncf((phi + dy) / dx);
  float a = sel(phi + dx2, phi - dx2, t > dx2);
  return make_float2(r * safe_sin(a), r * safe_cos(a));
}
P3_INLINE float2 var_perspective(float2 p, float w, float angle, float dist) {
  float hpa = angle * (kPi * (float)0.5);
  float vsin = safe_sin(hpa), vfcos = dist * safe_cos(hpa);
  float t = (float)1 / (dist - p.y * vsin);
  return make_float2(w * (dist * p.x * t), w * (vfcos * p.y * t));
}
P3_INLINE float2 var_bipolar(float2 p, float w, float shift) {
  const float HP = kPi * (float)0.5;
  const float TOP = (float)2 / kPi;
  float x2y2 = dot2(p);
  float t = x2y2 + (float)1;
  float x2 = (float)2 * p.x;
  float ps = -HP * shift;
  float y = (float)0.5 * atan2f((float)2 * p.y, x2y2 - (float)1) + ps;
  if (y > HP) y = -HP + wmod(y + HP, kPi);
  else if (y < -HP) y = HP - wmod(HP - y, kPi);
  return make_float2(w * (float)0.25 * TOP * logf((t + x2) / (t - x2)), w * TOP * y);
}
P3_INLINE float2 var_curl(float2 p, float w, float c1, float c2) {
  float re = (float)1 + c1 * p.x + c2 * (p.x * p.x - p.y * p.y);
  float im = c1 * p.y + (float)2 * c2 * p.x * p.y;
  float r = w / (re * re + im * im);
  return make_float2((p.x * re + p.y * im) * r, (p.y * re - p.x * im) * r);
}
P3_INLINE float2 var_rectangles(float2 p, float w, float rx, float ry) {
  float rxs = sel(rx, (float)1, rx == 0), rys = sel(ry, (float)1, ry == 0);
  float ox = sel(w * (((float)2 * floorf(p.x / rxs) + (float)1) * rxs - p.x), w * p.x, rx == 0);
  float oy = sel(w * (((float)2 * floorf(p.y / rys) + (float)1) * rys - p.y), w * p.y, ry == 0);
  return make_float2(ox, oy);
}
P3_INLINE float2 var_blob(float2 p, float w, float low, float high, float waves) {
  float r0 = len2(p), re = r0 + kEps;
  float sina = p.x / re, cosa = p.y / re;
  float a = atan2f(p.x, p.y);
  float r = r0 * (low + (high - low) * ((float)0.5 + (float)0.5 * safe_sin(waves * a)));
  float k = w * r;
  return make_float2(k * sina, k * cosa);
}
P3_INLINE float2 var_ngon(float2 p, float w, float sides, float power, float circle, float corners) {
  float sumsq = dot2(p);
  float rf = powf(sumsq, power * (float)0.5);
  float theta = atan2f(p.y, p.x);
  float b = kTau / sides;
  float phi = theta - b * floorf(theta / b);
  if (phi > b * (float)0.5) phi = phi - b;
  float amp = (corners * ((float)1 / (safe_cos(phi) + kEps) - (float)1) + circle) / (rf + kEps);
  float k = w * amp;
  return make_float2(k * p.x, k * p.y);
}
P3_INLINE float2 var_wedge(float2 p, float w, float angle, float hole, float count, float swirl) {
  float r0 = len2(p);
  float a = atan2f(p.y, p.x) + swirl * r0;
  const float OOP = (float)1 / kPi;
  float c = floorf((count * a + kPi) * OOP * (float)0.5);
  float cf = (float)1 - angle * count * OOP * (float)0.5;
  a = a * cf + c * angle;
  float r = w * (r0 + hole);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}
P3_INLINE float2 var_cpow(float2 p, float w, float pr, float pi, float pw, Rng& rng) {
  float a = atan2f(p.y, p.x);
  float lnr = (float)0.5 * logf(dot2(p));
  float va = kTau / pw, vc = pr / pw, vd = pi / pw;
  float n = floorf(rng.rand01() * fabsf(pw));
  float ang = vc * a + vd * lnr + va * n;
  float m = w * expf(vc * lnr - vd * a);
  return make_float2(m * safe_cos(ang), m * safe_sin(ang));
}
P3_INLINE float2 var_curve(float2 p, float w, float xamp, float yamp, float xlen, float ylen) {
  float pcx = smax(xlen * xlen, (float)1e-20), pcy = smax(ylen * ylen, (float)1e-20);
  return make_float2(w * (p.x + xamp * expf((-p.y * p.y) / pcx)),
          w * (p.y + yamp * expf((-p.x * p.x) / pcy)));
}
P3_INLINE float2 var_noise(float2 p, float w, Rng& rng) {
  float tmpr = rng.rand01() * kTau;
  float r = w * rng.rand01();
  return make_float2(p.x * r * safe_cos(tmpr), p.y * r * safe_sin(tmpr));
}
P3_INLINE float2 var_blur(float2 p, float w, Rng& rng) {
  (void)p;
  float tmpr = rng.rand01() * kTau;
  float r = w * rng.rand01();
  return make_float2(r * safe_cos(tmpr), r * safe_sin(tmpr));
}
P3_INLINE float2 var_gaussian_blur(float2 p, float w, Rng& rng) {
  (void)p;
  float ang = rng.rand01() * kTau;
  float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
  float r = w * (r0 + r1 + r2 + r3 - (float)2);
  return make_float2(r * safe_cos(ang), r * safe_sin(ang));
}
P3_INLINE float2 var_arch(float2 p, float w, Rng& rng) {
  (void)p;
  float ang = rng.rand01() * w * kPi;
  float s = safe_sin(ang), c = safe_cos(ang);
  return make_float2(w * s, w * (s * s) / c);
}
P3_INLINE float2 var_radial_blur(float2 p, float w, float angle, Rng& rng) {
  float hpa = angle * (kPi * (float)0.5);
  float spin = safe_sin(hpa), zoom = safe_cos(hpa);
  float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
  float g = w * (r0 + r1 + r2 + r3 - (float)2);
  float ra = len2(p);
  float tmpa = atan2f(p.y, p.x) + spin * g;
  float sa = safe_sin(tmpa), ca = safe_cos(tmpa);
  float rz = zoom * g - (float)1;
  return make_float2(ra * ca + rz * p.x, ra * sa + rz * p.y);
}
P3_INLINE float2 var_juliascope(float2 p, float w, float power, float dist, Rng& rng) {
  float phi = atan2f(p.y, p.x);
  float sumsq = dot2(p);
  float pa = fabsf(power);
  int32_t t = f2i(floorf(rng.rand01() * pa));
  float tmpr = sel((kTau * (float)t - phi) / power, (kTau * (float)t + phi) / power, (t & 1) == 0);
  float r = w * powf(sumsq, dist / power / (float)2);
  return make_float2(r * safe_cos(tmpr), r * safe_sin(tmpr));
}
P3_INLINE float2 var_square(float2 p, float w, Rng& rng) {
  (void)p;
  float r0 = rng.rand01(), r1 = rng.rand01();
  return make_float2(w * (r0 - (float)0.5), w * (r1 - (float)0.5));
}
P3_INLINE float2 var_rays(float2 p, float w, Rng& rng) {
  float ang = w * rng.rand01() * kPi;
  float r = w / (dot2(p) + kEps);
  float tanr = w * safe_tan(ang) * r;
  return make_float2(tanr * safe_cos(p.x), tanr * safe_sin(p.y));
}
P3_INLINE float2 var_blade(float2 p, float w, Rng& rng) {
  float r = rng.rand01() * w * len2(p);
  float s = safe_sin(r), c = safe_cos(r);
  return make_float2(w * p.x * (c + s), w * p.x * (c - s));
}
P3_INLINE float2 var_twintrian(float2 p, float w, Rng& rng) {
  const float LN10 = (float)2.302585092994046;
  float r = rng.rand01() * w * len2(p);
  float s = safe_sin(r), c = safe_cos(r);
  float diff = logf(s * s) / LN10 + c;
  if (diff != diff || diff > (float)1e10 || diff < (float)-1e10) diff = (float)-30;
  return make_float2(w * p.x * diff, w * p.x * (diff - s * kPi));
}
P3_INLINE float2 var_exp(float2 p, float w) {
  float e = expf(p.x);
  float k = w * e;
  return make_float2(k * safe_cos(p.y), k * safe_sin(p.y));
}
P3_INLINE float2 var_log(float2 p, float w) {
  return make_float2(w * ((float)0.5 * logf(dot2(p))), w * atan2f(p.y, p.x));
}
P3_INLINE float2 var_sin(float2 p, float w) {
  return make_float2(w * (safe_sin(p.x) * coshf(p.y)), w * (safe_cos(p.x) * sinhf(p.y)));
}
P3_INLINE float2 var_cos(float2 p, float w) {
  return make_float2(w * (safe_cos(p.x) * coshf(p.y)), w * (-safe_sin(p.x) * sinhf(p.y)));
}
P3_INLINE float2 var_tan(float2 p, float w) {
  float den = (float)1 / (safe_cos((float)2 * p.x) + coshf((float)2 * p.y));
  float k = w * den;
  return make_float2(k * safe_sin((float)2 * p.x), k * sinhf((float)2 * p.y));
}
P3_INLINE float2 var_sec(float2 p, float w) {
  float den = (float)2 / (safe_cos((float)2 * p.x) + coshf((float)2 * p.y));
  float k = w * den;
  return make_float2(k * (safe_cos(p.x) * coshf(p.y)), k * (safe_sin(p.x) * sinhf(p.y)));
}
P3_INLINE float2 var_csc(float2 p, float w) {
  float den = (float)2 / (coshf((float)2 * p.y) - safe_cos((float)2 * p.x));
  float k = w * den;
  return make_float2(k * (safe_sin(p.x) * coshf(p.y)), k * (-safe_cos(p.x) * sinhf(p.y)));
}
P3_INLINE float2 var_cot(float2 p, float w) {
  float den = (float)1 / (coshf((float)2 * p.y) - safe_cos((float)2 * p.x));
  float k = w * den;
  return make_float2(k * safe_sin((float)2 * p.x), k * -sinhf((float)2 * p.y));
}
P3_INLINE float2 var_sinh(float2 p, float w) {
  return make_float2(w * (sinhf(p.x) * safe_cos(p.y)), w * (coshf(p.x) * safe_sin(p.y)));
}
P3_INLINE float2 var_cosh(float2 p, float w) {
  return make_float2(w * (coshf(p.x) * safe_cos(p.y)), w * (sinhf(p.x) * safe_sin(p.y)));
}
P3_INLINE float2 var_tanh(float2 p, float w) {
  float den = (float)1 / (safe_cos((float)2 * p.y) + coshf((float)2 * p.x));
  float k = w * den;
  return make_float2(k * sinhf((float)2 * p.x), k * safe_sin((float)2 * p.y));
}
P3_INLINE float2 var_sech(float2 p, float w) {
  float den = (float)2 / (safe_cos((float)2 * p.y) + coshf((float)2 * p.x));
  float k = w * den;
  return make_float2(k * (safe_cos(p.y) * coshf(p.x)), k * (-safe_sin(p.y) * sinhf(p.x)));
}
P3_INLINE float2 var_csch(float2 p, float w) {
  float den = (float)2 / (coshf((float)2 * p.x) - safe_cos((float)2 * p.y));
  float k = w * den;
  return make_float2(k * (sinhf(p.x) * safe_cos(p.y)), k * (-coshf(p.x) * safe_sin(p.y)));
}
P3_INLINE float2 var_coth(float2 p, float w) {
  float den = (float)1 / (coshf((float)2 * p.x) - safe_cos((float)2 * p.y));
  float k = w * den;
  return make_float2(k * sinhf((float)2 * p.x), k * safe_sin((float)2 * p.y));
}
P3_INLINE float2 var_butterfly(float2 p, float w) {
  float wx = w * (float)1.3029400317411197908970256609023;
  float y2 = p.y * (float)2;
  float r = wx * sqrtf(fabsf(p.y * p.x) / (kEps + p.x * p.x + y2 * y2));
  return make_float2(r * p.x, r * y2);
}
P3_INLINE float2 var_edisc(float2 p, float w) {
  float tmp = dot2(p) + (float)1;
  float tmp2 = (float)2 * p.x;
  float r1 = sqrtf(tmp + tmp2), r2 = sqrtf(tmp - tmp2);
  float xmax = (r1 + r2) * (float)0.5;
  float a1 = logf(xmax + sqrtf(smax(xmax - (float)1, (float)0)));
  float a2 = -acosf(smin(smax(p.x / xmax, (float)-1), (float)1));
  float wn = w / (float)11.57034632;
  float csv = safe_cos(a1);
  float snv = sel(safe_sin(a1), -safe_sin(a1), p.y > 0);
  float snhu = sinhf(a2), cshu = coshf(a2);
  return make_float2(wn * cshu * csv, wn * snhu * snv);
}
P3_INLINE float2 var_elliptic(float2 p, float w) {
  float tmp = dot2(p) + (float)1;
  float x2 = (float)2 * p.x;
  float xmax = (float)0.5 * (sqrtf(tmp + x2) + sqrtf(tmp - x2));
  float a = p.x / xmax;
  float b = sqrtf(smax((float)1 - a * a, (float)0));
  float ssx = sqrtf(smax(xmax - (float)1, (float)0));
  float wn = w / (kPi * (float)0.5);
  float ylog = wn * logf(xmax + ssx);
  return make_float2(wn * atan2f(a, b), sel(-ylog, ylog, p.y > 0));
}
P3_INLINE float2 var_foci(float2 p, float w) {
  float expx = expf(p.x) * (float)0.5;
  float expnx = (float)0.25 / expx;
  float tmp = w / (expx + expnx - safe_cos(p.y));
  return make_float2(tmp * (expx - expnx), tmp * safe_sin(p.y));
}
P3_INLINE float2 var_loonie(float2 p, float w) {
  float r2 = dot2(p), w2 = w * w;
  if (r2 < w2) {
    float r = w * sqrtf(w2 / r2 - (float)1);
    return make_float2(r * p.x, r * p.y);
  }
  return make_float2(w * p.x, w * p.y);
}
P3_INLINE float2 var_polar2(float2 p, float w) {
  float p2v = w / kPi;
  return make_float2(p2v * atan2f(p.x, p.y), (p2v * (float)0.5) * logf(dot2(p) + kEps));
}
P3_INLINE float2 var_scry(float2 p, float w) {
  float s = dot2(p);
  float r = (float)1 / (sqrtf(s) * (s + (float)1 / (w + kEps)));
  return make_float2(p.x * r, p.y * r);
}
P3_INLINE float2 var_bent2(float2 p, float w, float bx, float by) {
  return make_float2(w * sel(p.x, p.x * bx, p.x < 0), w * sel(p.y, p.y * by, p.y < 0));
}
P3_INLINE float2 var_cell(float2 p, float w, float size) {
  float inv = (float)1 / size;
  int32_t x = f2i(floorf(p.x * inv)), y = f2i(floorf(p.y * inv));
  float dx = p.x - (float)x * size, dy = p.y - (float)y * size;
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
  return make_float2(w * (dx + (float)x * size), -w * (dy + (float)y * size));
}
P3_INLINE float2 var_escher(float2 p, float w, float beta) {
  float a = atan2f(p.y, p.x);
  float lnr = (float)0.5 * logf(dot2(p));
  float seb = safe_sin(beta), ceb = safe_cos(beta);
  float vc = (float)0.5 * ((float)1 + ceb), vd = (float)0.5 * seb;
  float m = w * expf(vc * lnr - vd * a);
  float n = vc * a + vd * lnr;
  return make_float2(m * safe_cos(n), m * safe_sin(n));
}
P3_INLINE float2 var_modulus(float2 p, float w, float mx, float my) {
  float xr = (float)2 * mx, yr = (float)2 * my;
  float ox, oy;
  if (p.x > mx) ox = w * (-mx + wmod(p.x + mx, xr));
  else if (p.x < -mx) ox = w * (mx - wmod(mx - p.x, xr));
  else ox = w * p.x;
  if (p.y > my) oy = w * (-my + wmod(p.y + my, yr));
  else if (p.y < -my) oy = w * (my - wmod(my - p.y, yr));
  else oy = w * p.y;
  return make_float2(ox, oy);
}
P3_INLINE float2 var_split(float2 p, float w, float xs, float ys) {
  float oy = sel(-w * p.y, w * p.y, safe_cos(p.x * xs * kPi) >= 0);
  float ox = sel(-w * p.x, w * p.x, safe_cos(p.y * ys * kPi) >= 0);
  return make_float2(ox, oy);
}
P3_INLINE float2 var_splits(float2 p, float w, float sx, float sy) {
  return make_float2(sel(w * (p.x - sx), w * (p.x + sx), p.x >= 0), sel(w * (p.y - sy), w * (p.y + sy), p.y >= 0));
}
P3_INLINE float2 var_stripes(float2 p, float w, float space, float warp) {
  float rx = floorf(p.x + (float)0.5);
  float ox = p.x - rx;
  return make_float2(w * (ox * ((float)1 - space) + rx), w * (p.y + ox * ox * warp));
}
P3_INLINE float2 var_whorl(float2 p, float w, float inside, float outside) {
  float r = len2(p);
  float base = atan2f(p.y, p.x);
  float a = sel(base + outside / (w - r), base + inside / (w - r), r < w);
  return make_float2(w * r * safe_cos(a), w * r * safe_sin(a));
}
P3_INLINE float2 var_flux(float2 p, float w, float spread) {
  float xpw = p.x + w, xmw = p.x - w;
  float ty = p.y * p.y;
  float avgr = w * ((float)2 + spread) * sqrtf(sqrtf(ty + xpw * xpw) / sqrtf(ty + xmw * xmw));
  float avga = (atan2f(p.y, xmw) - atan2f(p.y, xpw)) * (float)0.5;
  return make_float2(avgr * safe_cos(avga), avgr * safe_sin(avga));
}
P3_INLINE float2 var_popcorn2(float2 p, float w, float px, float py, float pc) {
  return make_float2(w * (p.x + px * safe_sin(safe_tan(p.y * pc))), w * (p.y + py * safe_sin(safe_tan(p.x * pc))));
}
P3_INLINE float2 var_lazysusan(float2 p, float w, float lx, float ly, float spin, float twist, float space) {
  float x = p.x - lx, y = p.y + ly;
  float r0 = sqrtf(x * x + y * y);
  if (r0 < w) {
    float a = atan2f(y, x) + spin + twist * (w - r0);
    float r = w * r0;
    return make_float2(r * safe_cos(a) + lx, r * safe_sin(a) - ly);
  }
  float r = w * ((float)1 + space / r0);
  return make_float2(r * x + lx, r * y - ly);
}
P3_INLINE float2 var_waves2(float2 p, float w, float sx, float fx, float sy, float fy) {
  return make_float2(w * (p.x + sx * safe_sin(p.y * fx)), w * (p.y + sy * safe_sin(p.x * fy)));
}
P3_INLINE float2 var_oscope(float2 p, float w, float freq, float amp, float damping, float sep) {
  float tpf = kTau * freq;
  float t = sel(amp * expf(-fabsf(p.x) * damping) * safe_cos(tpf * p.x) + sep,
            amp * safe_cos(tpf * p.x) + sep, damping == 0);
  if (fabsf(p.y) <= t) return make_float2(w * p.x, -w * p.y);
  return make_float2(w * p.x, w * p.y);
}
P3_INLINE float2 var_separation(float2 p, float w, float sx, float sxi, float sy, float syi) {
  float sx2 = sx * sx, sy2 = sy * sy;
  float ox = sel(-w * (sqrtf(p.x * p.x + sx2) + p.x * sxi), w * (sqrtf(p.x * p.x + sx2) - p.x * sxi), p.x > 0);
  float oy = sel(-w * (sqrtf(p.y * p.y + sy2) + p.y * syi), w * (sqrtf(p.y * p.y + sy2) - p.y * syi), p.y > 0);
  return make_float2(ox, oy);
}
P3_INLINE float2 var_auger(float2 p, float w, float freq, float ww, float scale, float sym) {
  float s = safe_sin(freq * p.x), t = safe_sin(freq * p.y);
  float dy = p.y + ww * (scale * s * (float)0.5 + fabsf(p.y) * s);
  float dx = p.x + ww * (scale * t * (float)0.5 + fabsf(p.x) * t);
  return make_float2(w * (p.x + sym * (dx - p.x)), w * dy);
}
P3_INLINE float2 var_wedge_sph(float2 p, float w, float angle, float hole, float count, float swirl) {
  float r0 = len2(p);
  float ri = (float)1 / (r0 + kEps);
  float a = atan2f(p.y, p.x) + swirl * ri;
  const float OOP = (float)1 / kPi;
  float c = floorf((count * a + kPi) * OOP * (float)0.5);
  float cf = (float)1 - angle * count * OOP * (float)0.5;
  a = a * cf + c * angle;
  float r = w * (ri + hole);
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}
P3_INLINE float2 var_super_shape(float2 p, float w, float rnd, float m, float n1, float n2, float n3, float holes, Rng& rng) {
  float pm4 = m * (float)0.25, pn1 = (float)-1 / n1;
  float r0 = len2(p);
  float theta = pm4 * atan2f(p.y, p.x) + kPi * (float)0.25;
  float st = safe_sin(theta), ct = safe_cos(theta);
  float t1 = powf(fabsf(ct), n2), t2 = powf(fabsf(st), n3);
  float r = w * ((rnd * rng.rand01() + ((float)1 - rnd) * r0) - holes) * powf(t1 + t2, pn1) / r0;
  return make_float2(r * p.x, r * p.y);
}
P3_INLINE float2 var_flower(float2 p, float w, float petals, float holes, Rng& rng) {
  float theta = atan2f(p.y, p.x);
  float r0 = len2(p);
  float r = w * (rng.rand01() - holes) * safe_cos(petals * theta) / r0;
  return make_float2(r * p.x, r * p.y);
}
P3_INLINE float2 var_conic(float2 p, float w, float ecc, float holes, Rng& rng) {
  float r0 = len2(p);
  float ct = p.x / r0;
  float r = w * (rng.rand01() - holes) * ecc / ((float)1 + ecc * ct) / r0;
  return make_float2(r * p.x, r * p.y);
}
P3_INLINE float2 var_parabola(float2 p, float w, float height, float width, Rng& rng) {
  float r = len2(p);
  float sr = safe_sin(r), cr = safe_cos(r);
  float r0 = rng.rand01(), r1 = rng.rand01();
  return make_float2(height * w * sr * sr * r0, width * w * cr * r1);
}
P3_INLINE float2 var_pie(float2 p, float w, float slices, float rotation, float thick, Rng& rng) {
  (void)p;
  float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01();
  float sl = truncf(r0 * slices + (float)0.5);
  float a = rotation + kTau * (sl + r1 * thick) / slices;
  float r = w * r2;
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}
P3_INLINE float2 var_boarders(float2 p, float w, Rng& rng) {
  float r0 = rng.rand01();
  float rx = floorf(p.x + (float)0.5), ry = floorf(p.y + (float)0.5);
  float ox = p.x - rx, oy = p.y - ry;
  if (r0 >= (float)0.75) return make_float2(w * (ox * (float)0.5 + rx), w * (oy * (float)0.5 + ry));
  if (fabsf(ox) >= fabsf(oy)) {
    if (ox >= 0) return make_float2(w * (ox * (float)0.5 + rx + (float)0.25), w * (oy * (float)0.5 + ry + (float)0.25 * oy / ox));
    return make_float2(w * (ox * (float)0.5 + rx - (float)0.25), w * (oy * (float)0.5 + ry - (float)0.25 * oy / ox));
  }
  if (oy >= 0) return make_float2(w * (ox * (float)0.5 + rx + ox / oy * (float)0.25), w * (oy * (float)0.5 + ry + (float)0.25));
  return make_float2(w * (ox * (float)0.5 + rx - ox / oy * (float)0.25), w * (oy * (float)0.5 + ry - (float)0.25));
}
P3_INLINE float2 var_mobius(float2 p, float w, const float q[8]) {
  float reu = q[0] * p.x - q[1] * p.y + q[2];
  float imu = q[0] * p.y + q[1] * p.x + q[3];
  float rev = q[4] * p.x - q[5] * p.y + q[6];
  float imv = q[4] * p.y + q[5] * p.x + q[7];
  float rad = w / (rev * rev + imv * imv);
  return make_float2(rad * (reu * rev + imu * imv), rad * (imu * rev - reu * imv));
}
P3_INLINE float2 var_wedge_julia(float2 p, float w, float angle, float count, float power, float dist, Rng& rng) {
  float sumsq = dot2(p);
  float cn = dist / power * (float)0.5;
  float r = w * powf(sumsq, cn);
  float cf = (float)1 - angle * count / kTau;
  float n = floorf(rng.rand01() * fabsf(power));
  float a = (atan2f(p.y, p.x) + kTau * n) / power;
  float c = floorf((count * a + kPi) / kPi * (float)0.5);
  a = a * cf + c * angle;
  return make_float2(r * safe_cos(a), r * safe_sin(a));
}


}  // namespace p3
