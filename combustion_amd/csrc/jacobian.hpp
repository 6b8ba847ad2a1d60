// Analytic Jacobians of straight-line programs (forward-mode differentiation of the linear form, jacobian.cpp).
#pragma once
#include <vector>

#include "ast.hpp"

namespace combustion::ast {

// A program with the same inputs whose outputs are l's outputs followed by d output[rows[i]] / d input[wrt[j]] for
// each row i and column j (row-major). Derivatives follow the operations as written: min and max take the derivative
// of the operand they select (the first on ties), a select that of its chosen operand.
Linear jacobian(const Linear& l, const std::vector<int32_t>& wrt, const std::vector<int32_t>& rows);

}  // namespace combustion::ast
