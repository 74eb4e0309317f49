// pyr3 render-path CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// This directory is a CPU restatement of pyr3's render hot path
// (chaos game -> u32 histogram scatter -> density estimation -> tonemap),
// written from the behaviour of the reference sources it cites. Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it, and only as the checker; the product (pyr3_amd, libpyr3hip.so) never
// links, imports or calls anything here.
//
// Arithmetic follows the reference: FP32 for everything the WGSL kernels do
// in f32 (compiled with -ffp-contract=off so a*b+c stays two roundings, like
// the WGSL expression order), f64 where the reference host code uses JS
// numbers (xform-pick table, kernel-norm LUT, spatial kernel).
//
// Parity pins (what each function is checked against, besides the GPU):
//   isaac_*            flam3-C golden state (tests/golden/isaac/golden.json)
//   seed_walker        closed form + randcnt/determinism/distinct-stream invariants
//   build_distrib      known answers: 50/50 flip at 8192, 3:1 split, xaos zero row
//   var_* (f64)        4,607 instrumented-flam3 rows (tests/golden/variations)
//   chaos              single-linear-xform closed forms, opacity-0 / count invariants,
//                      saturation, retry/draw-count contract on scripted streams
//   density            sum-of-weights = 1 per bucket, isolated-bucket conservation,
//                      brute-force f64 scatter formulation on tiny grids
//   tonemap            empty->background, vibrancy 0/1 closed forms, f64 restatement
#pragma once
#include <cstdint>
#include <cstddef>

namespace orc {

constexpr int RANDSIZ = 16;
constexpr int ISAAC_U32 = 36;          // [cnt, a, b, c, mem[16], rsl[16]]
constexpr int GRAIN = 16384;           // xform-pick table row length
constexpr int MAX_XFORMS = 128;        // fallback row index
constexpr int XFORM_FLOATS = 116;      // 464-byte packed xform
constexpr int MAX_VARS = 8;

struct Isaac {
  uint32_t cnt, a, b, c;
  uint32_t mem[RANDSIZ];
  uint32_t rsl[RANDSIZ];
};

void isaac_round(Isaac& s);
uint32_t isaac_irand(Isaac& s);
void isaac_init(Isaac& s, bool from_rsl);
void seed_walker(Isaac& s, uint32_t seed, uint32_t walker);

inline float rand01_of(uint32_t raw) {
  return (float)(raw & 0x0fffffffu) * (1.0f / 268435455.0f);
}
inline float rand11_of(uint32_t raw) {
  return (float)((int32_t)(raw & 0x0fffffffu) - 0x07ffffff) * (1.0f / 134217727.0f);
}

}  // namespace orc
