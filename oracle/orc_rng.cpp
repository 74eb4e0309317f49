// ISAAC (RANDSIZL = 4) and per-walker seeding — oracle restatement.
// Test infrastructure only: see oracle.hpp.
//
// isaac_round  follows src/isaac.ts:35-105 and src/shaders/chaos.wgsl:243-294
// isaac_irand  follows src/isaac.ts:171-179 and src/shaders/chaos.wgsl:297-308
// isaac_init   follows src/isaac.ts:108-154 (irandinit, golden-ratio mix)
// seed_walker  follows src/isaac.ts:194-229 (packIsaacStates per walker)
#include "oracle.hpp"

namespace orc {

// One ISAAC round: 16 steps; step k uses the shift pattern <<13, >>6, <<2, >>16
// (k mod 4) and pairs mem[k] with mem[(k + 8) mod 16].
void isaac_round(Isaac& s) {
  s.c += 1u;
  uint32_t a = s.a;
  uint32_t b = s.b + s.c;
  for (int k = 0; k < RANDSIZ; ++k) {
    uint32_t x = s.mem[k];
    switch (k & 3) {
      case 0: a ^= a << 13; break;
      case 1: a ^= a >> 6; break;
      case 2: a ^= a << 2; break;
      default: a ^= a >> 16; break;
    }
    a += s.mem[(k + 8) & 15];
    uint32_t y = s.mem[(x >> 2) & 15] + a + b;
    s.mem[k] = y;
    b = s.mem[(y >> 6) & 15] + x;
    s.rsl[k] = b;
  }
  s.a = a;
  s.b = b;
}

// Pops rsl[--cnt]; when the buffer is empty a fresh round is generated and
// rsl[15] is returned with cnt left at 15.
uint32_t isaac_irand(Isaac& s) {
  if (s.cnt == 0) {
    isaac_round(s);
    s.cnt = RANDSIZ - 1;
  } else {
    s.cnt -= 1;
  }
  return s.rsl[s.cnt];
}

static void mix8(uint32_t v[8]) {
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

void isaac_init(Isaac& s, bool from_rsl) {
  s.a = s.b = s.c = 0;
  uint32_t v[8];
  for (auto& x : v) x = 0x9e3779b9u;
  for (int i = 0; i < 4; ++i) mix8(v);
  if (from_rsl) {
    for (int pass = 0; pass < 2; ++pass) {
      const uint32_t* src = pass == 0 ? s.rsl : s.mem;
      for (int i = 0; i < RANDSIZ; i += 8) {
        for (int j = 0; j < 8; ++j) v[j] += src[i + j];
        mix8(v);
        for (int j = 0; j < 8; ++j) s.mem[i + j] = v[j];
      }
    }
  } else {
    for (int i = 0; i < RANDSIZ; i += 8) {
      mix8(v);
      for (int j = 0; j < 8; ++j) s.mem[i + j] = v[j];
    }
  }
  isaac_round(s);
  s.cnt = RANDSIZ;
}

// Walker w of a dispatch seeded with `seed`: a PCG32 (XSH-RS output) stream
// started at seed ^ (w * 2654435761 + 1) and advanced four times fills rsl,
// then the ISAAC key schedule runs from rsl.
void seed_walker(Isaac& s, uint32_t seed, uint32_t walker) {
  const uint32_t MUL = 747796405u, INC = 2891336453u;
  uint32_t pcg = seed ^ (walker * 2654435761u + 1u);
  for (int k = 0; k < 4; ++k) pcg = pcg * MUL + INC;
  for (int i = 0; i < RANDSIZ; ++i) {
    uint32_t st = pcg;
    pcg = pcg * MUL + INC;
    uint32_t shift = (st >> 28) + 4u;
    uint32_t word = ((st >> shift) ^ st) * 277803737u;
    s.rsl[i] = (word >> 22) ^ word;
  }
  s.cnt = 0;
  isaac_init(s, true);
}

}  // namespace orc

// ---- C ABI for the Python test harness ------------------------------------
extern "C" {

static void load(orc::Isaac& s, const uint32_t* st) {
  s.cnt = st[0]; s.a = st[1]; s.b = st[2]; s.c = st[3];
  for (int i = 0; i < 16; ++i) { s.mem[i] = st[4 + i]; s.rsl[i] = st[20 + i]; }
}
static void store(const orc::Isaac& s, uint32_t* st) {
  st[0] = s.cnt; st[1] = s.a; st[2] = s.b; st[3] = s.c;
  for (int i = 0; i < 16; ++i) { st[4 + i] = s.mem[i]; st[20 + i] = s.rsl[i]; }
}

void orc_isaac_init(uint32_t* state36, int from_rsl) {
  orc::Isaac s; load(s, state36); orc::isaac_init(s, from_rsl != 0); store(s, state36);
}
void orc_isaac_round(uint32_t* state36) {
  orc::Isaac s; load(s, state36); orc::isaac_round(s); store(s, state36);
}
void orc_isaac_draw(uint32_t* state36, uint32_t* out, int n) {
  orc::Isaac s; load(s, state36);
  for (int i = 0; i < n; ++i) out[i] = orc::isaac_irand(s);
  store(s, state36);
}
void orc_seed_walkers(uint32_t seed, uint32_t walker_base, int n, uint32_t* out36) {
  for (int w = 0; w < n; ++w) {
    orc::Isaac s;
    orc::seed_walker(s, seed, walker_base + (uint32_t)w);
    store(s, out36 + (size_t)w * 36);
  }
}
void orc_rand_transforms(const uint32_t* raw, int n, float* out01, float* out11) {
  for (int i = 0; i < n; ++i) {
    out01[i] = orc::rand01_of(raw[i]);
    out11[i] = orc::rand11_of(raw[i]);
  }
}

}  // extern "C"
