// Allocation-free calls of a module: the pointer table and the forms' work memory in a static device array
// (noalloc.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "forms.hpp"

namespace combustion {

// Writes the arrays' pointer table into this GPU's static scratch and launches the form's kernels on `stream` alone.
// Device pointers only; allocates, creates and synchronizes nothing. Throws Unsupported or ArgumentError for a
// module or a state count beyond the scratch's limits.
void call_static(Form& form, const double* const* inputs, size_t n, double* const* outputs, hipStream_t stream);

}  // namespace combustion
