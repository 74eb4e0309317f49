// Device-side building blocks of the gfx950 render path: ISAAC streams held
// in LDS, the masked 28-bit rand transforms, range-guarded trig, and the
// buffer-layout constants shared by the kernels.
//
// ISAAC state per walker is [cnt, a, b, c | mem[16] | rsl[2][16]]. cnt/a/b/c
// live in VGPRs; mem and rsl are indexed by data-dependent values
// (mem[(x >> 2) & 15], rsl[--cnt]), which has no cheap per-lane VGPR form on
// CDNA, so they live in LDS as [word][walker] columns: walker w's word k is
// at k * stride + w. Two rsl rounds are held so that the next round can be
// generated ahead, at a point where the whole wave runs it together, instead
// of inside whichever branch happens to draw the 17th value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "p3_args.h"

// Address-space-qualified pointers: LDS (3), global (1) and constant (4).
// Generic pointers would compile to flat instructions; loads through a
// constant-space pointer at a wave-uniform address become scalar loads.
#define P3_LDS __attribute__((address_space(3)))
#define P3_GLOBAL __attribute__((address_space(1)))
#define P3_CONST __attribute__((address_space(4)))
#define P3_INLINE __device__ inline __attribute__((always_inline))

namespace p3 {

typedef const P3_CONST float* xrec_t;  // a 116-float xform record

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;
constexpr float kEps = 1e-10f;

// ---------------------------------------------------------------- ISAAC --

// cnt: bits 0-4 the unread values of the current rsl round (0..16), bit 5
// which of the two rounds is current, bit 6 set when the other round already
// holds the next 16 values.
struct IsaacRegs {
  uint32_t cnt, a, b, c;
};
constexpr uint32_t kIsaacUnread = 31u, kIsaacHalf = 32u, kIsaacReady = 64u;

// LDS view of one walker's mem/rsl words.
struct IsaacLds {
  P3_LDS uint32_t* base;   // &lds[walker]
  int stride;              // walkers per block
  P3_INLINE P3_LDS uint32_t& mem(int k) const { return base[k * stride]; }
  P3_INLINE P3_LDS uint32_t& rsl(uint32_t h, int k) const {
    return base[(16 + 16 * (int)h + k) * stride];
  }
};

// chaos.wgsl:243-294 / isaac.ts:35-105: one round, 16 outputs into rsl[h].
P3_INLINE void isaac_round(IsaacRegs& s, const IsaacLds& m, uint32_t h) {
  s.c += 1u;
  uint32_t a = s.a;
  uint32_t b = s.b + s.c;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t x = m.mem(k);
    if ((k & 3) == 0) a ^= a << 13;
    else if ((k & 3) == 1) a ^= a >> 6;
    else if ((k & 3) == 2) a ^= a << 2;
    else a ^= a >> 16;
    a += m.mem((k + 8) & 15);
    const uint32_t y = m.mem((x >> 2) & 15) + a + b;
    m.mem(k) = y;
    b = m.mem((y >> 6) & 15) + x;
    m.rsl(h, k) = b;
  }
  s.a = a;
  s.b = b;
}

// chaos.wgsl:297-308: pop rsl[--cnt], refilling with a round when empty
// (taking the round generated ahead when there is one).
P3_INLINE uint32_t isaac_irand(IsaacRegs& s, const IsaacLds& m) {
  uint32_t n = s.cnt & kIsaacUnread;
  uint32_t h = (s.cnt & kIsaacHalf) ? 1u : 0u;
  uint32_t ready = s.cnt & kIsaacReady;
  if (n == 0u) {
    if (ready) {
      h ^= 1u;
      ready = 0u;
    } else {
      isaac_round(s, m, h);
    }
    n = 16u;
  }
  n -= 1u;
  s.cnt = n | (h ? kIsaacHalf : 0u) | ready;
  return m.rsl(h, (int)n);
}

// Generates the next round into the other rsl half when it is free, so that
// the next 16 + unread draws need no round.
P3_INLINE void isaac_top_up(IsaacRegs& s, const IsaacLds& m) {
  if (!(s.cnt & kIsaacReady)) {
    isaac_round(s, m, (s.cnt & kIsaacHalf) ? 0u : 1u);
    s.cnt |= kIsaacReady;
  }
}

// Top-up policy of a chaos block. With every = N > 1 (a genome that draws
// few values per turn, measured by the block) the whole block generates on
// every N-th turn, each lane with a free half, so rounds run on one turn in N
// and the other turns' barriers wait for no round; with N = 1 a wave
// generates when one of its lanes has fewer than kIsaacLow values left (none
// generated ahead), again every lane with a free half. Either way a turn that
// draws more than is left still generates inline, so the policy changes only
// when rounds run.
constexpr uint32_t kIsaacLow = 8u;
P3_INLINE void isaac_top_up_turn(IsaacRegs& s, const IsaacLds& m, uint32_t turn, uint32_t every) {
  if (every > 1u) {
    if (turn % every == 0u) isaac_top_up(s, m);
    return;
  }
  const bool low = !(s.cnt & kIsaacReady) && (s.cnt & kIsaacUnread) < kIsaacLow;
  if (__any(low)) isaac_top_up(s, m);
}

P3_INLINE void isaac_mix8(uint32_t v[8]) {
  uint32_t a = v[0], b = v[1], c = v[2], d = v[3], e = v[4], f = v[5], g = v[6], h = v[7];
  a ^= b << 11; d += a; b += c;
  b ^= c >> 2;  e += b; c += d;
  c ^= d << 8;  f += c; d += e;
  d ^= e >> 16; g += d; e += f;
  e ^= f << 10; h += e; f += g;
  f ^= g >> 4;  a += f; g += h;
  g ^= h << 8;  b += g; h += a;
  h ^= a >> 9;  c += h; a += b;
  v[0] = a; v[1] = b; v[2] = c; v[3] = d; v[4] = e; v[5] = f; v[6] = g; v[7] = h;
}

// isaac.ts:194-229 (packIsaacStates, one walker) on the device: PCG32
// XSH-RS words into rsl, the two-pass golden-ratio key schedule, one round.
P3_INLINE void isaac_seed(IsaacRegs& s, const IsaacLds& m, uint32_t seed,
                                  uint32_t walker) {
  const uint32_t MUL = 747796405u, INC = 2891336453u;
  uint32_t pcg = seed ^ (walker * 2654435761u + 1u);
#pragma unroll
  for (int k = 0; k < 4; ++k) pcg = pcg * MUL + INC;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t st = pcg;
    pcg = pcg * MUL + INC;
    const uint32_t shift = (st >> 28) + 4u;
    const uint32_t word = ((st >> shift) ^ st) * 277803737u;
    m.rsl(0u, i) = (word >> 22) ^ word;
  }
  uint32_t v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = 0x9e3779b9u;
#pragma unroll
  for (int i = 0; i < 4; ++i) isaac_mix8(v);
#pragma unroll
  for (int i = 0; i < 16; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] += m.rsl(0u, i + j);
    isaac_mix8(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) m.mem(i + j) = v[j];
  }
#pragma unroll
  for (int i = 0; i < 16; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] += m.mem(i + j);
    isaac_mix8(v);
#pragma unroll
    for (int j = 0; j < 8; ++j) m.mem(i + j) = v[j];
  }
  s.a = s.b = s.c = 0u;
  isaac_round(s, m, 0u);
  s.cnt = 16u;
}

// Values a walker has drawn since its state (c0, unread0): every round adds
// 16 values and every draw takes one, so drawn = 16 (c - c0) - (unread -
// unread0), unread counting the round generated ahead.
P3_INLINE uint32_t isaac_drawn(const IsaacRegs& s, uint32_t c0, uint32_t unread0) {
  const uint32_t unread = (s.cnt & kIsaacUnread) + ((s.cnt & kIsaacReady) ? 16u : 0u);
  return 16u * (s.c - c0) - unread + unread0;
}

// A walker's RNG as the variations see it.
struct Rng {
  IsaacRegs s;
  IsaacLds m;
  P3_INLINE uint32_t irand() { return isaac_irand(s, m); }
  // chaos.wgsl:319-331: low 28 bits scaled by 1/(2^28-1) and 1/(2^27-1).
  P3_INLINE float rand01() {
    return (float)(irand() & 0x0fffffffu) * (1.0f / 268435455.0f);
  }
  P3_INLINE float rand11() {
    return (float)((int32_t)(irand() & 0x0fffffffu) - 0x07ffffff) * (1.0f / 134217727.0f);
  }
};

// ----------------------------------------------------------------- math --

// chaos.wgsl:356-362 integer hash for out-of-range trig arguments.
P3_INLINE float hash01(uint32_t h) {
  h ^= h >> 17; h *= 0xed5ad4bbu;
  h ^= h >> 11; h *= 0xac4c1b51u;
  h ^= h >> 15;
  return (float)h / 4294967296.0f;
}

// chaos.wgsl:363-384
P3_INLINE float safe_sin(float a) {
  if (fabsf(a) <= 1.0e6f) return sinf(a);
  return sinf(hash01(__float_as_uint(a)) * kTau);
}
P3_INLINE float safe_cos(float a) {
  if (fabsf(a) <= 1.0e6f) return cosf(a);
  return cosf(hash01(__float_as_uint(a)) * kTau);
}
P3_INLINE float safe_tan(float a) { return safe_sin(a) / safe_cos(a); }
P3_INLINE float safe_tanh(float a) {  // clamp to +/-20; NaN passes through
  const float c = a < -20.0f ? -20.0f : (a > 20.0f ? 20.0f : a);
  return tanhf(c);
}

// WGSL mix(a, b, t) = a * (1 - t) + b * t
P3_INLINE float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// WGSL u32(f32): saturating; NaN and negatives -> 0.
P3_INLINE uint32_t f2u_sat(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)f;
}

}  // namespace p3
