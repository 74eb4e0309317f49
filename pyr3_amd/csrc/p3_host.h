// Host-side helpers shared by the renderer handle (p3_api.cpp) and the
// batch object (p3_batch.cpp): decoding of the reference's uniform blocks
// into kernel arguments and the checks on uploaded blobs.
#pragma once
#include <stdint.h>

#include <cstring>
#include <string>
#include <vector>

#include "p3_launch.h"

namespace p3h {

inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Step boundaries of a pick table (n_rows x 16 entries, see ChaosArgs), or
// false when the table has more than 16 xforms or a row whose entries
// decrease somewhere (then the kernel reads the table itself).
bool pick_bounds(const uint32_t* table, int n_rows, std::vector<uint16_t>& out);

// Per-radius DE kernel sums (257 floats), computed once per process.
const float* kernel_norm_lut();

// Empty when every record's variation count and active slots are valid,
// else the message.
std::string records_error(const float* xforms, int n_slots);

// Kernel arguments of one chaos dispatch (the 32-word uniform block `u`)
// whose records, table, palette and histogram are at dxf / dds / pal / hist.
p3::ChaosArgs dispatch_args_at(const uint32_t* u, int n_regular, int32_t final_idx,
                               const float* dxf, const uint32_t* dds, const float* pal,
                               uint32_t* hist);

// DE parameters of the 8-word DEUniforms block (max/min radius, curve,
// k1, k2 and the search radius R); empty string, or the reason the block is
// refused.
std::string de_params(const uint32_t* de, p3::DeArgs& a);

// Tonemap parameters of the 24-word VizUniforms block (k1, k2, gamma,
// vibrancy, highlight power, linear range, background, transparency,
// curves mask, HSL adjust).
void tm_params(const uint32_t* viz, p3::TmArgs& t);

}  // namespace p3h
