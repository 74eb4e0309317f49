// Chaos-game kernel body for gfx950, compiled at run time per genome shape.
//
// Behaviour follows chaos_main (src/shaders/chaos.wgsl:8269-8732):
//   init draws rand_11, rand_11, rand01 -> (x, y, z)
//   per turn: pick = table[row * 16384 + (irand & 16383)], row = previous
//   accepted xform or the fallback row; chain A (affine, pre_blur, variation
//   sum, direct colour, post-affine, colour mix); optional xform blend; bad
//   check (NaN or |v| > 1e10) with rand_11 reseed and up to four retried
//   turns that do not advance i (except at i = 0); final-xform lens with an
//   opacity gate; scale-relative jitter (two rand01); splat after the fuse.
//
// The xform chains come from a generated `Spec` (see p3_rtc.cpp): one
// straight-line case per xform of the genome's shape (its variation indices
// per slot, pre_blur slots, post-affine presence), so no variation switch
// runs per lane and each case reads its xform record at a fixed offset from
// a kernel-argument pointer (scalar loads).
//
// The splat adds u32(pal * opacity * 255) and u32(opacity * 255) to the four
// channels of bin (y * W + x). The reference's adds saturate at 2^32 - 1; a
// saturating add of non-negative values is order-independent (the final
// value is min(start + sum, 2^32 - 1)). The device's atomic rate is per
// request, whatever its width, so a splat is ONE 64-bit add into a packed
// 16-bit accumulator with exact carry counts (add_packed), merged into the
// u32 histogram before it is read. A splat with a word of 2^16 or more (a
// palette or opacity above 1, which the reference's packers never produce)
// goes straight into the histogram as four u32 adds whose returned old
// values detect a wrap, followed by an exchange to 2^32 - 1: after the last
// wrap of a word the last operation on it is always that exchange.
#pragma once
#include "p3_device.h"
#include "p3_variation_dispatch.h"

namespace p3 {

struct Chain {
  float2 pv, pv_pre, pa;
  float z;
};

// chaos.wgsl:8164-8181 for one pre_blur slot k of record xf.
P3_INLINE float2 pre_blur_slot(xrec_t xf, int k, float2 c, Rng& rng) {
  const float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
  const float g = xf[20 + 4 * k + 1] * (r0 + r1 + r2 + r3 - 2.0f);
  const float r4 = rng.rand01();
  const float ang = r4 * kTau;
  c.x = c.x + cosf(ang) * g;
  c.y = c.y + sinf(ang) * g;
  return c;
}

P3_INLINE float2 post_affine(xrec_t xf, float2 pv) {
  return make_float2(xf[12] * pv.x + xf[13] * pv.y + xf[14],
                     xf[16] * pv.x + xf[17] * pv.y + xf[18]);
}

P3_INLINE bool is_bad(float2 v) {
  return isnan(v.x) || isnan(v.y) || fabsf(v.x) > 1e10f || fabsf(v.y) > 1e10f;
}

// Device-scope relaxed atomics: only the final sums matter.
P3_INLINE uint32_t add_ret(P3_GLOBAL uint32_t* p, uint32_t d) {
  return __hip_atomic_fetch_add(p, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
P3_INLINE void pin_max(P3_GLOBAL uint32_t* p) {
  (void)__hip_atomic_exchange(p, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Packed add of v, whose four 16-bit fields are each an increment below
// 2^16 (one splat's (d0, d1, d2, d3), or the sum of a group's splats), as ONE
// 64-bit atomic on the bin's u64 accumulator. The returned old value o and
// the sum n = o + v give the carry into every bit, o ^ v ^ n; a field's
// increment (its word plus the carry in) is at most 2^16, so it carries out
// at most once per add, and counting each carry out (bits 16, 32, 48 and the
// 64-bit wrap) in the bin's carry[0..3] makes the field sums exact:
// S_k = field_k + 2^16 carry[k] - carry[k-1]. hist_merge (p3_image.hip) adds
// them to the u32 histogram, saturating.
P3_INLINE void count_carries(P3_GLOBAL uint64_t* carry, uint64_t o, uint64_t v) {
  const uint64_t n = o + v;
  const uint64_t c = o ^ v ^ n;
  const bool c0 = (c >> 16) & 1u, c1 = (c >> 32) & 1u, c2 = (c >> 48) & 1u, c3 = n < o;
  if (c0 | c1 | c2 | c3) {
    if (c0) (void)__hip_atomic_fetch_add(carry + 0, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c1) (void)__hip_atomic_fetch_add(carry + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c2) (void)__hip_atomic_fetch_add(carry + 2, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c3) (void)__hip_atomic_fetch_add(carry + 3, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
P3_INLINE uint64_t add_packed_issue(P3_GLOBAL uint64_t* acc, uint64_t v) {
  return __hip_atomic_fetch_add(acc, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
P3_INLINE void add_packed(P3_GLOBAL uint64_t* acc, P3_GLOBAL uint64_t* carry, uint64_t v) {
  count_carries(carry, add_packed_issue(acc, v), v);
}
P3_INLINE uint64_t pack_words(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3) {
  return (uint64_t)d0 | ((uint64_t)d1 << 16) | ((uint64_t)d2 << 32) | ((uint64_t)d3 << 48);
}
// Four saturating channel adds (R, G, B, count) at b: the returned old
// values are checked after all four are issued; a word that wrapped is
// pinned to 2^32 - 1.
P3_INLINE void hist_add4_sat(P3_GLOBAL uint32_t* b, uint32_t d0, uint32_t d1, uint32_t d2,
                             uint32_t d3) {
  const uint32_t o0 = d0 ? add_ret(b + 0, d0) : 0u;
  const uint32_t o1 = d1 ? add_ret(b + 1, d1) : 0u;
  const uint32_t o2 = d2 ? add_ret(b + 2, d2) : 0u;
  const uint32_t o3 = d3 ? add_ret(b + 3, d3) : 0u;
  if (d0 && o0 > 0xffffffffu - d0) pin_max(b + 0);
  if (d1 && o1 > 0xffffffffu - d1) pin_max(b + 1);
  if (d2 && o2 > 0xffffffffu - d2) pin_max(b + 2);
  if (d3 && o3 > 0xffffffffu - d3) pin_max(b + 3);
}

// The block's LDS copy of the 256-entry palette.
typedef const P3_LDS float4* pal_t;
P3_INLINE float4 palette_at(pal_t pal, uint32_t i) { return pal[i]; }

// chaos.wgsl:1662-1666: k = (1, 2/3, 1/3), p = |fract(h + k) * 6 - 3|,
// rgb = v * mix(1, clamp(p - 1, 0, 1), s); fract(x) = x - floor(x).
P3_INLINE float3 hsv_to_rgb(float h, float s, float v) {
  const float k[3] = {1.0f, 2.0f / 3.0f, 1.0f / 3.0f};
  float o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float f = h + k[c];
    const float p = fabsf((f - floorf(f)) * 6.0f - 3.0f);
    o[c] = v * mixf(1.0f, fminf(fmaxf(p - 1.0f, 0.0f), 1.0f), s);
  }
  return make_float3(o[0], o[1], o[2]);
}

// Colour-mode overrides of the splat colour (chaos.wgsl:8625-8697): flow
// (direction of this turn's displacement from (p0x, p0y) -> hue, saturated
// magnitude -> value), trap distance (palette at a glow / ring coordinate of
// the distance to a point, circle or line) and phase (arg z -> hue,
// log-modulus rings -> value), blended over the palette colour by strength.
P3_INLINE void colour_mode(const ChaosArgs& A, pal_t palette, float sx, float sy, float p0x,
                           float p0y, float4& pal) {
  float3 rgb;
  float t;
  if (A.color_mode == 1u) {
    const float dx = sx - p0x, dy = sy - p0y;
    const float mag = sqrtf(dx * dx + dy * dy);
    if (!(mag > 1e-20f)) return;
    const float hue = (atan2f(dy, dx) + kPi) / (2.0f * kPi);
    const float val = fminf(fmaxf(1.0f - expf(-fmaxf(A.flow_scale, 0.0f) * mag), 0.0f), 1.0f);
    rgb = hsv_to_rgb(hue, 1.0f, val);
    t = fminf(fmaxf(A.flow_strength, 0.0f), 1.0f);
  } else if (A.color_mode == 2u) {
    const float qx = sx - A.trap_cx, qy = sy - A.trap_cy;
    float d;
    if (A.trap_kind == 0u) d = sqrtf(qx * qx + qy * qy);
    else if (A.trap_kind == 1u) d = fabsf(sqrtf(qx * qx + qy * qy) - A.trap_radius);
    else d = fabsf(qx * A.trap_nx + qy * A.trap_ny);
    float tt;
    if (A.trap_mode == 0u) {
      tt = fminf(fmaxf(expf(-fmaxf(A.trap_falloff, 0.0f) * d), 0.0f), 1.0f);
    } else {
      const float f = d * fmaxf(A.trap_freq, 0.0f);
      tt = f - floorf(f);
    }
    const uint32_t idx = min(f2u_sat(tt * 255.0f), 255u);
    const float4 c = palette_at(palette, idx);
    rgb = make_float3(c.x, c.y, c.z);
    t = fminf(fmaxf(A.trap_strength, 0.0f), 1.0f);
  } else if (A.color_mode == 3u) {
    const float mag = sqrtf(sx * sx + sy * sy);
    if (!(mag > 1e-20f)) return;
    const float hue = (atan2f(sy, sx) + kPi) / (2.0f * kPi);
    const float f = logf(mag) * fmaxf(A.phase_freq, 0.0f);
    rgb = hsv_to_rgb(hue, 1.0f, mixf(1.0f, 0.5f, f - floorf(f)));
    t = fminf(fmaxf(A.phase_strength, 0.0f), 1.0f);
  } else {
    return;
  }
  pal.x = mixf(pal.x, rgb.x, t);
  pal.y = mixf(pal.y, rgb.y, t);
  pal.z = mixf(pal.z, rgb.z, t);
}

// Per-walker loop state that stays with the walker between turns.
struct Walker {
  float x, y, z;
  int prev;
  uint32_t consec, iter;
};

// A turn's splat request for the wave-level aggregation: the bin and its
// packed words (each at most kAggWord), or bin = kNoSplat.
constexpr uint32_t kNoSplat = 0xffffffffu;
constexpr uint32_t kAggWord = 1023u;   // 64 lanes x 1023 < 2^16: group sums stay in their fields
struct Splat {
  uint32_t bin;
  uint64_t v;
};

// Trace entry (walker wid, turn i), or null outside the traced window.
P3_INLINE P3_GLOBAL float* trace_entry(const ChaosArgs& A, uint32_t wid, uint32_t i) {
  if (A.trace == nullptr || wid >= A.trace_walkers || i >= A.trace_iters) return nullptr;
  return (P3_GLOBAL float*)A.trace + ((size_t)wid * A.trace_iters + i) * kTraceFloats;
}

// Values walker wid has drawn since its ISAAC state was seeded or loaded
// (for the trace only).
P3_INLINE uint32_t drawn_since_start(const ChaosArgs& A, const Rng& rng, uint32_t wid) {
  if (A.isaac_in) {   // the loaded state: its c and unread count
    const P3_CONST uint32_t* st = (const P3_CONST uint32_t*)A.isaac_in + (size_t)wid * 36u;
    return isaac_drawn(rng.s, st[3], st[0]);
  }
  return isaac_drawn(rng.s, 1u, 16u);   // device seeding ends after one round
}

// Loop state word of a walker: ISAAC unread count and flags (bits 0-6),
// previous xform + 1 (bits 8-15), retry count (bits 16-18).
P3_INLINE uint32_t pack_state(uint32_t cnt, int prev, uint32_t consec) {
  return (cnt & 127u) | ((uint32_t)(prev + 1) << 8) | (consec << 16);
}

// One turn of walker `wid` (global id) after its pick `fn` is known: chain,
// blend, bad check and retry window, final lens, jitter, splat. Advances
// `wk.iter` unless the turn is retried. A splat whose words are all at most
// kAggWord (every splat of a palette and opacity in [0, 1]: the words are at
// most 255) is returned in `sp` for splat_wave; a wider one is added here
// (packed below 2^16, else four saturating u32 adds). The trace entry's
// pick draw (field 11) was written when the pick was drawn.
template <class Spec>
P3_INLINE void run_turn(const ChaosArgs& A, pal_t palette, Walker& wk, Rng& rng, uint32_t fn,
                        uint32_t wid, Splat& sp) {
  const uint32_t i = wk.iter;
  const float p0x = wk.x, p0y = wk.y;   // position before this turn's chain
  const xrec_t X = (xrec_t)A.xforms;
  const P3_CONST uint32_t* D = (const P3_CONST uint32_t*)A.distrib;
  const xrec_t xf = X + (size_t)fn * kXformFloats;
  bool dc_on = false;
  float3 dc_rgb = make_float3(0.0f, 0.0f, 0.0f);
  // One call site of the chain, so its straight-line xform cases are
  // inlined once: pass 0 runs the picked xform; with xform blend, a second
  // pass runs the blended-in xform bi (whose direct colour is dropped).
  // Draw order as chaos.wgsl:8372-8380: chain(fn), blend test, bi's draw,
  // chain(bi), then the blend factor.
  Chain ca{}, cb{};
  uint32_t cur = fn;
  bool blended = false;
#pragma clang loop unroll(disable)
  for (int pass = 0; pass < 2; ++pass) {
    bool dcp = false;
    float3 rgbp = make_float3(0.0f, 0.0f, 0.0f);
    const Chain c = Spec::chain(A, cur, X, wk.x, wk.y, wk.z, rng, dcp, rgbp);
    if (pass != 0) {
      cb = c;
      break;
    }
    ca = c;
    dc_on = dcp;
    dc_rgb = rgbp;
    if (!(A.xform_blend > 0.0f) || !(rng.rand01() < A.xform_blend)) break;
    const uint32_t rb = rng.irand();
    cur = D[(size_t)A.fallback_row * kGrain + (rb & (kGrain - 1))];
    blended = true;
  }
  float2 pv = ca.pv;
  float nz = ca.z;
  if (blended) {
    const float t = rng.rand01();
    pv.x = mixf(pv.x, cb.pv.x, t);
    pv.y = mixf(pv.y, cb.pv.y, t);
    nz = mixf(nz, cb.z, t);
  }
  const bool bad = is_bad(pv);
  if (bad) {
    const float rx = rng.rand11();
    const float ry = rng.rand11();
    pv = make_float2(rx, ry);
    wk.consec += 1u;
  } else {
    wk.consec = 0u;
  }
  P3_GLOBAL float* tr = trace_entry(A, wid, i);
  if (tr) {
    tr[0] = (float)i; tr[1] = (float)fn;
    tr[2] = ca.pa.x; tr[3] = ca.pa.y; tr[4] = ca.pv_pre.x; tr[5] = ca.pv_pre.y;
    tr[6] = pv.x; tr[7] = pv.y; tr[8] = bad ? 1.0f : 0.0f; tr[9] = nz;
    tr[12] = pv.x; tr[13] = pv.y; tr[14] = 0.0f; tr[15] = 0.0f;
  }
  const float ppx = pv.x, ppy = pv.y, ppz = nz;
  if (wk.consec > 0u && wk.consec < 5u) {
    // Retried turn: the loop index only advances when it is 0.
    wk.x = ppx; wk.y = ppy; wk.z = ppz;
    if (tr) tr[10] = __uint_as_float(drawn_since_start(A, rng, wid));
    if (i == 0u) wk.iter = 1u;
    return;
  }
  if (wk.consec >= 5u) wk.consec = 0u;
  wk.prev = (int)fn;
  float sx = ppx, sy = ppy, sz = ppz;
  if (Spec::has_final(A)) {
    const int fslot = Spec::final_slot(A);
    const xrec_t fx = X + (size_t)fslot * kXformFloats;
    const bool apply = fx[10] == 1.0f ? true : (rng.rand01() < fx[10]);
    // During the fuse the final's output feeds neither the walker nor a
    // splat, so a final that draws nothing is left out there (its opacity
    // gate draw above is kept).
    const bool needed = !Spec::final_draw_free(A) || i >= A.fuse || tr != nullptr;
    if (apply && needed) {
      const float2 fpa = make_float2(fx[0] * pv.x + fx[1] * pv.y + fx[2],
                                     fx[4] * pv.x + fx[5] * pv.y + fx[6]);
      const float2 fpv = Spec::final_sum(A, X, fslot, fpa, rng, dc_on, dc_rgb);
      const float fz = mixf(nz, fx[8], fx[9]);
      if (!is_bad(fpv)) { sx = fpv.x; sy = fpv.y; sz = fz; }
    }
  }
  const float mag = fmaxf(fmaxf(fabsf(ppx), fabsf(ppy)), 1e-30f);
  const float amp = mag * A.jitter;
  const float jx = (rng.rand01() - 0.5f) * amp;
  const float jy = (rng.rand01() - 0.5f) * amp;
  wk.x = ppx + jx; wk.y = ppy + jy; wk.z = ppz;
  if (tr) {
    tr[12] = wk.x; tr[13] = wk.y; tr[14] = sx; tr[15] = sy;
    tr[10] = __uint_as_float(drawn_since_start(A, rng, wid));
  }
  wk.iter = i + 1u;
  if (i >= A.fuse) {
    const float opacity = xf[10];
    const float dx = sx - A.cx, dy = sy - A.cy;
    const float rx = dx * A.cos_r - dy * A.sin_r;
    const float ry = dx * A.sin_r + dy * A.cos_r;
    const float px = rx * A.scale + (float)A.width * 0.5f;
    const float py = ry * A.scale + (float)A.height * 0.5f;
    // floor(p) in [0, W) <=> p in [0, W) for finite p; NaN fails both tests.
    if (px >= 0.0f && px < (float)A.width && py >= 0.0f && py < (float)A.height) {
      const uint32_t xi = (uint32_t)floorf(px), yi = (uint32_t)floorf(py);
      const float cxf = fminf(fmaxf(sz, 0.0f), 1.0f) * 255.0f;
      float4 pal;
      if (A.palette_mode == 1u) {
        uint32_t i0 = f2u_sat(cxf);
        float frac = cxf - (float)i0;
        if (i0 >= 255u) { i0 = 254u; frac = 1.0f; }
        const float4 p0 = palette_at(palette, i0), p1 = palette_at(palette, i0 + 1u);
        pal = make_float4(mixf(p0.x, p1.x, frac), mixf(p0.y, p1.y, frac),
                          mixf(p0.z, p1.z, frac), mixf(p0.w, p1.w, frac));
      } else {
        const uint32_t pi = min(f2u_sat(cxf), 255u);
        pal = palette_at(palette, pi);
      }
      if (dc_on) { pal.x = dc_rgb.x; pal.y = dc_rgb.y; pal.z = dc_rgb.z; }
      if (A.color_mode != 0u) colour_mode(A, palette, sx, sy, p0x, p0y, pal);
      const float wgt = opacity * 255.0f;
      const size_t bin = ((size_t)yi * A.width + xi) * 4u;
      P3_GLOBAL uint32_t* b = (P3_GLOBAL uint32_t*)A.hist + bin;
      const uint32_t d0 = f2u_sat(pal.x * wgt), d1 = f2u_sat(pal.y * wgt),
                     d2 = f2u_sat(pal.z * wgt), d3 = f2u_sat(wgt);
      const uint32_t dmax = d0 | d1 | d2 | d3;
      if (dmax <= kAggWord) {
        if (dmax) { sp.bin = (uint32_t)(bin >> 2); sp.v = pack_words(d0, d1, d2, d3); }
      } else if (dmax <= 0xffffu) {
        add_packed((P3_GLOBAL uint64_t*)A.acc + (bin >> 2), (P3_GLOBAL uint64_t*)A.carry + bin,
                   pack_words(d0, d1, d2, d3));
      } else {
        hist_add4_sat(b, d0, d1, d2, d3);
      }
      if (A.idx_sum != nullptr) {   // capture of the colour index (chaos.wgsl:8720-8728)
        const uint32_t d = f2u_sat(fminf(fmaxf(sz, 0.0f), 1.0f) * wgt);
        P3_GLOBAL uint32_t* q = (P3_GLOBAL uint32_t*)A.idx_sum + (size_t)yi * A.width + xi;
        if (d) {
          const uint32_t o = add_ret(q, d);
          if (o > 0xffffffffu - d) pin_max(q);
        }
      }
    }
  }
}

// The wave's splats of one turn, aggregated by bin: lanes whose bins agree
// add their packed words into one sum (in LDS) and one lane issues the
// global add for the group. Every lane of the wave calls it (the wave is
// converged). Groups are found through a 64-slot hash of the bin: the last
// lane to write a slot leads the lanes of the same bin that hashed there;
// lanes of another bin in a taken slot add on their own. Each splat is
// added exactly once, and a group's field sums (at most 64 x kAggWord) stay
// below 2^16, so add_packed stays exact. Hot bins (a compact attractor
// concentrating its samples on a few bins) then cost one atomic per wave
// instead of one per lane.
// The wave's hash table is its 64 entries of `order`, which the wave's lanes
// read at the start of the turn and the next write of which follows the
// block's barrier.
//
// A lane's packed add is issued here and its carries are counted at the
// lane's next call (or by settle_add after the block's last turn), so the
// wave does not wait for the atomic's returned value within the turn.
struct PendingAdd {
  uint64_t o, v;   // the add's returned old value and its increment
  uint32_t bin;    // kNoSplat: none pending
};
P3_INLINE void settle_add(const ChaosArgs& A, PendingAdd& q) {
  if (q.bin != kNoSplat) count_carries((P3_GLOBAL uint64_t*)A.carry + (size_t)4 * q.bin, q.o, q.v);
  q.bin = kNoSplat;
}

template <int kB>
P3_INLINE void splat_wave(const ChaosArgs& A, BlockLds<kB>& L, const Splat& sp, int t,
                          PendingAdd& pend) {
  settle_add(A, pend);
  const int lane = t & 63;
  P3_LDS uint16_t* tab = (P3_LDS uint16_t*)L.order + (t & ~63);
  P3_LDS uint64_t* sum = (P3_LDS uint64_t*)L.agg_sum + (t & ~63);
  const bool on = sp.bin != kNoSplat;
  sum[lane] = 0ull;
  const uint32_t slot = (sp.bin * 2654435761u) >> 26;
  if (on) tab[slot] = (uint16_t)lane;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const uint32_t lead = on ? tab[slot] : (uint32_t)lane;
  const uint32_t lead_bin = (uint32_t)__shfl((int)sp.bin, (int)lead, 64);
  const bool grouped = on && lead_bin == sp.bin;
  if (grouped)
    (void)__hip_atomic_fetch_add(sum + lead, sp.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (on && (!grouped || lead == (uint32_t)lane)) {
    pend.v = grouped ? sum[lane] : sp.v;
    pend.o = add_packed_issue((P3_GLOBAL uint64_t*)A.acc + sp.bin, pend.v);
    pend.bin = sp.bin;
  }
}

// Sub-dispatch owning block b of a batched grid.
P3_INLINE uint32_t multi_find(const MultiArgs& M, uint32_t b) {
  uint32_t k = 0;
  while (k + 1 < M.n && b >= M.start[k + 1]) ++k;
  return k;
}

// Sub-dispatch owning block b of a batch launch: the last k with
// start[k] <= b (binary search over the device array; start[0] = 0 and b is
// below start[n]).
P3_INLINE uint32_t batch_find(const BatchArgs& M, uint32_t b) {
  const P3_CONST uint32_t* s = (const P3_CONST uint32_t*)M.start;
  uint32_t lo = 0, hi = M.n;   // s[lo] <= b < s[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// Pick of the walker's next turn: row = its previous accepted xform (or the
// fallback row), column = low 14 bits of the next draw. Tables of up to 16
// xforms with boundaries are read from the block's LDS copy of those
// (pick = number of boundaries at or below the column), others from HBM.
template <class Spec, int kB>
P3_INLINE uint32_t draw_pick(const ChaosArgs& A, const BlockLds<kB>& L, Rng& rng, int prev,
                             uint32_t& raw) {
  const uint32_t row = prev >= 0 ? (uint32_t)prev : A.fallback_row;
  raw = rng.irand();
  const uint32_t col = raw & (kGrain - 1);
  constexpr int kCmp = Spec::kRegular - 1;   // boundaries between the n xforms
  if (Spec::kRegular < kPickRows && A.pick_bounds != nullptr) {
    const uint4* b = (const uint4*)(L.pick_b + row * kPickCols);
    const uint4 q0 = b[0];
    const uint4 q1 = kCmp > 8 ? b[1] : make_uint4(0u, 0u, 0u, 0u);
    const uint32_t w[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    uint32_t f = 0u;
#pragma unroll
    for (int k = 0; k < kCmp; ++k)
      f += col >= ((w[k >> 1] >> (16 * (k & 1))) & 0xffffu) ? 1u : 0u;
    return f;
  }
  return ((const P3_CONST uint32_t*)A.distrib)[(size_t)row * kGrain + col];
}

// Persistent block of kB walkers (walkers block*kB .. +kB-1 of A). Lane t
// seeds walker t and draws its first pick; then every turn:
//   1  barrier (the block ends when no walker is left)
//   2  each wave turns the per-xform pick counts into offsets (redundantly,
//      no barrier) and lists the walkers its lanes hold in xform order
//   3  barrier
//   4  lane t runs the turn of the t-th walker in that order and draws that
//      walker's next pick, counting it into the other turn's buckets
// so the lanes of a wave mostly share one xform and its straight-line code.
// Each walker keeps its own ISAAC column and loop state; only which lane
// computes it changes, so the per-walker draw sequence (pick, turn, pick,
// ...), arithmetic and splats are those of the one-lane-per-walker loop.
template <int kB, class Spec>
P3_INLINE void chaos_binned(const ChaosArgs& A, BlockLds<kB>& L, uint32_t block) {
  const int t = threadIdx.x;
  const uint32_t w0 = block * kB;
  const uint32_t total = A.iters + A.fuse;
  const uint32_t idle = A.fallback_row + 1u;   // bucket of finished walkers

  for (int k = t; k < 2 * (kMaxXforms + 2); k += kB) (&L.bucket[0][0])[k] = 0u;
  if (t < 2) L.live[t] = L.probe[t] = 0u;
  for (int k = t; k < 256; k += kB) L.palette[k] = A.palette[k];
  if (Spec::kRegular < kPickRows && A.pick_bounds != nullptr)
    for (uint32_t k = t; k < (A.fallback_row + 1u) * kPickCols; k += kB)
      L.pick_b[k] = ((const P3_CONST uint16_t*)A.pick_bounds)[k];
  __syncthreads();
  PendingAdd pend{0ull, 0ull, kNoSplat};
  int wl = t;          // walker this lane holds
  uint32_t f = idle;   // its pick
  uint32_t pos;        // its rank among the walkers with that pick
  {  // seed walker t and draw its first pick
    const uint32_t wid = w0 + t;
    Rng rng;
    rng.m.base = (P3_LDS uint32_t*)(L.isaac + t);
    rng.m.stride = kB;
    rng.s.cnt = rng.s.a = rng.s.b = rng.s.c = 0u;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    uint32_t iter = total;
    if (wid < A.walkers) {
      if (A.isaac_in) {
        const P3_CONST uint32_t* st = (const P3_CONST uint32_t*)A.isaac_in + (size_t)wid * 36u;
        rng.s.cnt = st[0]; rng.s.a = st[1]; rng.s.b = st[2]; rng.s.c = st[3];
        for (int k = 0; k < 16; ++k) { rng.m.mem(k) = st[4 + k]; rng.m.rsl(0u, k) = st[20 + k]; }
      } else {
        isaac_seed(rng.s, rng.m, A.seed, wid);
      }
      x = rng.rand11();
      y = rng.rand11();
      z = rng.rand01();
      iter = 0u;
    }
    if (iter < total) {
      uint32_t raw;
      f = draw_pick<Spec>(A, L, rng, -1, raw);
      if (P3_GLOBAL float* tr = trace_entry(A, wid, 0u)) tr[11] = __uint_as_float(raw);
    }
    L.ra[t] = rng.s.a; L.rb[t] = rng.s.b; L.rc[t] = rng.s.c;
    L.state[t] = pack_state(rng.s.cnt, -1, 0u);
    L.x[t] = x; L.y[t] = y; L.z[t] = z;
    L.iter[t] = iter;
    L.fn[t] = (uint16_t)f;
    pos = atomicAdd(&L.bucket[0][f], 1u);
  }

  // The ISAAC top-up rhythm (isaac_top_up_turn): every walker's draws over
  // the first kProbe turns give the block's mean draws per turn, and the
  // rhythm is the number of such turns that fit in 14 of the 16 values a
  // top-up leaves ahead, at most 3 (1 until then, and for walkers with fewer
  // turns).
  constexpr uint32_t kProbe = 12u;
  uint32_t every = 1u;
  for (uint32_t p = 0, turn = 0;; p ^= 1u, ++turn) {
    {  // walkers still turning: one LDS add per wave, one barrier; count p ^ 1
       // was last read before the previous turn's second barrier
      const uint64_t busy = __ballot(f != idle);
      if ((t & 63) == 0 && busy) atomicAdd(&L.live[p], (uint32_t)__builtin_popcountll(busy));
      __syncthreads();
      if (L.live[p] == 0u) break;
      if (t == 0) L.live[p ^ 1u] = 0u;
      if (turn == kProbe + 1u) {   // the probe sums were added before the barrier
        const uint32_t n = L.probe[1], d = L.probe[0];
        const uint32_t fit = d ? (uint32_t)((14ull * n * kProbe) / d) : 3u;
        every = fit >= 3u ? 3u : (fit >= 2u ? 2u : 1u);
      }
    }
    // every wave scans the bucket counts itself (64 buckets per step,
    // carried across steps), takes its walkers' bucket offsets from the lane
    // holding each bucket, and lists its walkers in xform order; the other
    // turn's buckets (last read before the previous barrier) are cleared
    {
      const int lane = t & 63;
      uint32_t carry = 0u, mine = 0u;
      for (uint32_t k0 = 0; k0 <= idle; k0 += 64u) {
        const uint32_t k = k0 + (uint32_t)lane;
        const uint32_t v = k <= idle ? L.bucket[p][k] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const uint32_t o = __shfl_up(incl, d, 64);
          if (lane >= d) incl += o;
        }
        const uint32_t b = __shfl(carry + incl - v, (int)(f & 63u), 64);
        if ((f >> 6) == (k0 >> 6)) mine = b;
        carry += __shfl(incl, 63, 64);
      }
      L.order[mine + pos] = (uint16_t)wl;
      for (int k = t; k < kMaxXforms + 2; k += kB) L.bucket[p ^ 1u][k] = 0u;
    }
    __syncthreads();
    // run the turn of the t-th walker in that order, then its next pick
    wl = L.order[t];
    f = idle;
    Splat sp{kNoSplat, 0ull};
    if (L.iter[wl] < total) {
      Rng rng;
      rng.m.base = (P3_LDS uint32_t*)(L.isaac + wl);
      rng.m.stride = kB;
      const uint32_t st = L.state[wl];
      rng.s.cnt = st & 127u; rng.s.a = L.ra[wl]; rng.s.b = L.rb[wl]; rng.s.c = L.rc[wl];
      isaac_top_up_turn(rng.s, rng.m, turn, every);   // rounds generated together
      Walker wk{L.x[wl], L.y[wl], L.z[wl], (int)((st >> 8) & 255u) - 1, (st >> 16) & 7u,
                L.iter[wl]};
      const uint32_t wid = w0 + wl;
      run_turn<Spec>(A, (pal_t)L.palette, wk, rng, L.fn[wl], wid, sp);
      if (wk.iter < total) {
        uint32_t raw;
        f = draw_pick<Spec>(A, L, rng, wk.prev, raw);
        if (P3_GLOBAL float* tr = trace_entry(A, wid, wk.iter)) tr[11] = __uint_as_float(raw);
      }
      L.ra[wl] = rng.s.a; L.rb[wl] = rng.s.b; L.rc[wl] = rng.s.c;
      L.state[wl] = pack_state(rng.s.cnt, wk.prev, wk.consec);
      L.x[wl] = wk.x; L.y[wl] = wk.y; L.z[wl] = wk.z;
      L.iter[wl] = wk.iter;
      L.fn[wl] = (uint16_t)f;
      if (turn == kProbe) {
        atomicAdd(&L.probe[0], drawn_since_start(A, rng, wid));
        atomicAdd(&L.probe[1], 1u);
      }
    }
    splat_wave<kB>(A, L, sp, t, pend);
    pos = atomicAdd(&L.bucket[p ^ 1u][f], 1u);
  }
  settle_add(A, pend);
}

}  // namespace p3
