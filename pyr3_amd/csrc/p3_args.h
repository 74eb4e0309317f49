// Argument blocks and LDS layout shared by the host launchers and the
// device code (including the chaos kernels compiled at run time).
#pragma once

namespace p3 {

constexpr int kGrain = 16384;      // xform-pick table row length
constexpr int kMaxXforms = 128;    // regular xforms; the fallback row index
constexpr int kXformFloats = 116;  // 464-byte xform record
constexpr int kMaxVars = 8;        // variation slots per xform
constexpr int kTraceFloats = 16;   // floats per trace entry
constexpr int kChaosBlock = 640;   // walkers per chaos block (one block per CU)
constexpr int kCatalogSize = 326;  // catalog indices 0..325 (97 = pre_blur)
constexpr int kPickRows = 17;      // pick tables of up to 16 xforms (+ fallback row)
constexpr int kPickCols = 16;      //   also travel as step boundaries, 16 per row

// Catalog entries with a direct-colour side (src/variations.ts:681-692).
constexpr bool is_dc_index(unsigned idx) {
  return (idx >= 99u && idx <= 102u) || idx == 220u || idx == 265u || (idx >= 310u && idx <= 313u);
}

struct ChaosArgs {
  const float* xforms;
  const uint32_t* distrib;     // rows 0..n-1 then the fallback row
  // Null, or the same table as step boundaries (rows with n <= 16 xforms
  // whose entries never decrease along the row): entry [row][k] = first
  // column whose pick exceeds k (16384 if none), so pick = #{k : col >= it}.
  const uint16_t* pick_bounds;
  const float4* palette;       // 256 RGBA
  uint32_t* hist;              // W x H x 4
  const uint32_t* isaac_in;    // walkers x 36 or null (seed on device)
  float* trace;
  uint32_t fallback_row;
  uint32_t width, height, iters, fuse, seed, walkers;
  int32_t final_idx;
  float scale, cx, cy, cos_r, sin_r, jitter, xform_blend;
  uint32_t palette_mode;
  uint32_t trace_walkers, trace_iters;
  // splat colour modes (0 palette, 1 flow, 2 trap distance, 3 phase) and
  // their parameters; capture of the colour index into idx_sum (or null)
  uint32_t color_mode, trap_kind, trap_mode;
  float flow_strength, flow_scale;
  float trap_cx, trap_cy, trap_radius, trap_nx, trap_ny, trap_falloff, trap_freq, trap_strength;
  float phase_strength, phase_freq;
  uint32_t* idx_sum;
  // packed adds: one u64 per bin holding (R, G, B, count) as four 16-bit
  // fields, and four u64 counts per bin of the carries out of each field
  // (the last: out of bit 63)
  uint64_t* acc;
  uint64_t* carry;
  // Kernels built for several genome shapes at once (a batch's union
  // kernel): this dispatch's shape among them and the case number of its
  // xform 0 in the union's xform switch (0 and 0 in a one-shape kernel).
  uint32_t shape, case_base;
};

// Batched launch: up to kMaxSub sub-dispatches of one genome shape in one
// grid; sub-dispatch k owns blocks [start[k], start[k+1]).
constexpr int kMaxSub = 64;
struct MultiArgs {
  const ChaosArgs* args;  // n entries in device memory
  uint32_t n;
  uint32_t start[kMaxSub + 1];
};

// Batch launch of a union kernel: any number of sub-dispatches (of any of
// the union's shapes, each with its own histogram); sub-dispatch k owns
// blocks [start[k], start[k+1]) and start[] (n + 1 entries) is in device
// memory.
struct BatchArgs {
  const ChaosArgs* args;  // n entries in device memory
  const uint32_t* start;  // n + 1 entries in device memory
  uint32_t n;
};

// LDS image of a chaos block's walkers (structure of arrays, kB walkers):
// ISAAC words [48][kB] (mem 0..15, two rsl rounds 16..47), the ISAAC
// registers a, b, c, the loop state word (ISAAC unread count and flags in
// bits 0-6, previous xform + 1 in bits 8-15, retry count in bits 16-18), the
// position and colour, the turn index, the pick and the xform order; the
// splat aggregation's group sums (its hash table of lanes reuses `order`,
// dead by then): 236 bytes per walker, plus the palette (4 KiB), the pick
// boundaries (544 B), two bucket rows and two live counts. 640 walkers fit
// one CU's 160 KiB.
template <int kB>
struct BlockLds {
  float4 palette[256];
  alignas(16) uint16_t pick_b[kPickRows * kPickCols];
  uint32_t live[2];                     // walkers still turning, two turns alternating
  uint32_t probe[2];                    // values drawn and walkers at the probe turn
  uint32_t isaac[48 * kB];
  uint32_t ra[kB], rb[kB], rc[kB];
  uint32_t state[kB];
  float x[kB], y[kB], z[kB];
  uint32_t iter[kB];
  uint16_t fn[kB], order[kB];
  uint32_t bucket[2][kMaxXforms + 2];   // per-xform counts, two turns alternating
  // packed sum of each splat group (indexed by the group's leader lane)
  alignas(8) uint64_t agg_sum[kB];
};

}  // namespace p3
