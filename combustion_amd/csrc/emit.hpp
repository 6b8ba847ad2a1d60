// Emitters: per-state programs and mechanisms to HIP source for gfx950.
#pragma once
#include <string>

#include "ast.hpp"

namespace combustion {

struct StraightOptions {
  unsigned block_size = 256;
  bool reciprocal = false;   // a / b as a * (1/b), 1/b by v_rcp_f64 and two Newton steps (within 1 ulp of 1/b)
  unsigned waves_per_eu = 0; // occupancy hint (0: compiler default)
  bool f32 = false;          // evaluate in FP32 (inputs and outputs stay FP64 arrays); exp/log/rcp by the hardware ops
};
// Straight-line kernel: signature (const double* const* in, double* const* out, unsigned n, unsigned offset).
std::string emit_straight(const ast::Linear& l, const std::string& name, const StraightOptions& o);

// Mixed-precision rates kernel template (kernels/rates_mixed.hip), embedded at build time.
extern const char* const rates_mixed_source;

}  // namespace combustion
