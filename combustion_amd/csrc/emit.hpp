// Emitters: per-state programs and mechanisms to HIP source for gfx950.
#pragma once
#include <string>
#include <vector>

#include "ast.hpp"

namespace combustion {

struct StraightOptions {
  unsigned block_size = 256;
  unsigned min_waves = 0;    // > 0: waves per SIMD the register allocation must leave room for (__launch_bounds__)
  // a / b as a * (1/b), 1/b by v_rcp_f64 (relative error below 2^-24) and Newton steps: two give the correctly
  // rounded 1/b on every input tried, one a relative error below 1e-14 (measured 2.2e-15 over 4M random inputs).
  bool reciprocal = false;
  unsigned newton_steps = 2;
  bool f32 = false;          // evaluate in FP32 (inputs and outputs stay FP64 arrays); exp/log/rcp by the hardware ops
  // FP64 arithmetic with FP32 cores for exp and ln: e^x = 2^n · 2^f with the FP64 reduction x log2 e = n + f (|f| ≤ ½)
  // and v_exp_f32 for 2^f; ln x = (e + log2 m) ln 2 with the FP64 exponent/mantissa split and v_log_f32 for log2 m
  // (relative error of a few FP32 ulps over the whole FP64 range).
  bool fast_transcendentals = false;
  // FP64 e^x inlined as a polynomial (exp64 in hip_emit.cpp, constants from exp_coefficients.py): the argument clamped
  // to [-746, 710] (outside it e^x rounds to 0 or overflows to +inf), Cody-Waite reduction to |r| <= ln(2)/2 with
  // n taken from the low word of x log2(e) + 1.5·2^52, a degree-10 polynomial (relative error below 5e-16) and
  // v_ldexp_f64: no range branches.
  bool exp_poly = false;
  // Registers kept in LDS instead of VGPRs (one FP64 slot per register and thread, [slot][thread] so that a wave's
  // accesses are consecutive): stored where defined and read where used. Every `fence_every` instructions a
  // compiler-only memory fence (no instruction) ends the segment, so that the compiler reloads such a value in each
  // segment that uses it instead of keeping the first load live across the program. For one-lane-per-state
  // kernels whose live values exceed the register file: LDS holds what would otherwise spill to scratch.
  std::vector<int32_t> lds_registers;
  unsigned fence_every = 128;
  // Group values: with group_table >= 0, groups of `group_size` consecutive states (group g = id / group_size; the
  // group size divides the block size) share `group_stride` values, group g's at in[group_table][g · stride ..],
  // which a block stages in LDS at once. Input r with group_offset[r] >= 0 is the group's value at that offset, read
  // where first used rather than at the top (so that they are not all live at once), plus, for the state number
  // increment_state[r] within its group (id mod group_size), the group's value at increment_offset[r] (>= 0); the
  // other inputs are in[r][id]. Empty vectors: no grouped input.
  int32_t group_table = -1;
  unsigned group_size = 1, group_stride = 0;
  std::vector<int32_t> group_offset, increment_offset, increment_state;
  // With skip_table >= 0 (groups of one block each): in[skip_table] holds one int32 per group, and a block whose
  // group's value is nonzero returns before any work, writing nothing.
  int32_t skip_table = -1;
  // Output o of state id is stored at out[o][id · output_stride] (with out[o] = base + o and the stride the number of
  // outputs, state by state).
  unsigned output_stride = 1;
};
// Straight-line kernel: signature (const double* const* in, double* const* out, unsigned n, unsigned offset).
std::string emit_straight(const ast::Linear& l, const std::string& name, const StraightOptions& o);

// A per-state device function in the form of the reference's nekRK export (cuda.rs:296-318): `__NEKRK_DEVICE__`
// prefix, `dfloat` arithmetic, the first n_scalar inputs as named scalar parameters and the rest as the array
// parameter `array[]`. Outputs: DeviceReturn returns the single output; DeviceOut writes out[k];
// DeviceStrided writes out[k*n_states+id] (parameters `unsigned int id` first and `unsigned int n_states` last).
enum class DeviceOutput { Return, Out, Strided };
std::string emit_device_function(const ast::Linear& l, const std::vector<std::string>& scalar_names, const std::string& array,
                                 const std::string& name, DeviceOutput mode);

// Mixed-precision rates kernel template (kernels/rates_mixed.hip), embedded at build time.
extern const char* const rates_mixed_source;
// FP64 rates kernel for large mechanisms (kernels/rates_tiled.hip).
extern const char* const rates_tiled_source;
// Mixture-averaged transport kernel template (kernels/transport_tiled.hip).
extern const char* const transport_tiled_source;
// Surrogate network kernels (kernels/mlp.hip).
extern const char* const mlp_source;
// Surrogate inference kernel (kernels/surrogate.hip).
extern const char* const surrogate_source;

}  // namespace combustion
