// Device-function export (cuda.rs:296-337): the text of the four nekRK files, and a module running them on SoA states.
#pragma once
#include <memory>
#include <string>

#include "forms.hpp"
#include "mechanism.hpp"

namespace combustion {

// which: 0 rates.c (<prefix>species_rates), 1 conductivity.c (conductivityNIVT), 2 viscosity.c (viscosityIVT),
// 3 diffusivity.c (density_diffusivity); transport fits of length D.
std::string export_text(const Mechanism& m, int which, int D, const std::string& prefix);
// The four files (prefix nekrk_) compiled by hiprtc behind one kernel: inputs [T, P, Y_0..Y_{K-1}], outputs
// [ω_0..ω_{Ka-1}, λ, μ, D_0..D_{K-1}] in SI units, like a chemistry program with rates and transport.
void compile_export(const Mechanism& m, int D);
std::unique_ptr<Form> assemble_export(const Mechanism& m, int D);

}  // namespace combustion
