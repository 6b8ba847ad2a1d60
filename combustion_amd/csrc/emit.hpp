// Emitters: per-state programs and mechanisms to HIP source for gfx950.
#pragma once
#include <string>

#include "ast.hpp"

namespace combustion {

// Straight-line kernel: signature (const double* const* in, double* const* out, unsigned n, unsigned offset).
std::string emit_straight(const ast::Linear& l, const std::string& name, unsigned block_size);

}  // namespace combustion
