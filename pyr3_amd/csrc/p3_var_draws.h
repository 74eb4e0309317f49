// Which catalog entries draw from the walker's ISAAC stream.
//
// An entry draws when its case in apply_variation / apply_variation_ext
// passes the walker's Rng (src/shaders/chaos.wgsl:7731-8107: the cases that
// take the walker index); the other entries never draw. Index 97 (pre_blur)
// draws five values per slot in the chain. The kernel generator uses it to
// skip, during the fuse, a final xform that draws nothing (its output feeds
// neither the walker nor a splat there).
#pragma once
#include <stdint.h>

namespace p3 {

constexpr uint32_t kVarDrawsMask[11] = {0x00006000u, 0x001ffa00u, 0xfc000000u, 0x01077581u, 0x080fc225u, 0x0e5065c4u, 0x0c342603u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u};

constexpr bool var_draws(unsigned idx) {
  return idx == 97u || (idx < 326u && ((kVarDrawsMask[idx >> 5] >> (idx & 31u)) & 1u) != 0u);
}

}  // namespace p3
