// Program builders: the per-state rates program (reaction.rs) and the mixture-averaged transport program
// (transport.rs), plus the per-state SoA program that the GPU boundary evaluates.
#pragma once
#include <vector>

#include "ast.hpp"
#include "mechanism.hpp"

namespace combustion {

// reaction.rs:73
struct Temperatures {
  ast::E lnT, T, T2, T3, T4, rcpT, rcpT2;
};

// reaction.rs:312-327: net production rate of every active species (lean: ChemistryOptions::lean_rates).
std::vector<ast::E> species_rates(ast::Block& f, const Mechanism& m, const Temperatures& Ts, ast::E concentration,
                                  const std::vector<ast::E>& concentrations, bool lean = false);
// reaction.rs:356-391: inputs [P/R, R/P, T, V, n_0..n_{K-2}], outputs [T dT/dt / T * T, dV/dt, ω_0..ω_{Ka-1}].
ast::Function rates(const Mechanism& m);
// reaction.rs:329-354: inputs [P/R, T, V, n_0..n_{K-2}], outputs the net rate of progress of every reaction.
ast::Function reactions_rates(const Mechanism& m);

// transport.rs:119-137: polynomial fits in ln(T/T0) of degree D-1.
struct Polynomials {
  int D = 4;
  size_t K = 0;
  std::vector<double> conductivityIVT;   // [K][D]  λ_k/√T
  std::vector<double> VviscosityIVVT;    // [K][D]  √(η_k/√T)
  std::vector<double> diffusivityITVT;   // [K][K][D]  (P·D)_kj / T^{3/2}; entries j<k are fitted, j>k mirror them
};
Polynomials fit_polynomials(const Species& species, double T0, int D);

// Single-species physics (transport.rs:57-117), exposed for tests.
namespace transport {
double viscosity(const Species& s, size_t k, double T);
double diffusivity(const Species& s, size_t a, size_t b, double T);
double conductivity(const Species& s, size_t k, double T);
double omega22(const Species& s, size_t a, size_t b, double T);
double omega11(const Species& s, size_t a, size_t b, double T);
}  // namespace transport

// Device-function export (cuda.rs:140-350): the per-state functions of the reference's nekRK export, with its
// input names. Rates: inputs [pressure_R, lnT, T, T2, T3, T4, rcpT, concentrations_0..K-1], outputs ω_0..ω_{Ka-1}.
ast::Function rates_device_function(const Mechanism& m);
// Transport (which = 1 conductivityNIVT, 2 viscosityIVT, 3 density_diffusivity): inputs [sum_mole_proportions]
// (1) or [mean_molar_mass_VTN, VT] (3), then lnT, lnT2.. (powers 1..D-1 of ln(T/T0)) and mole_proportions_0..K-1;
// one output (1, 2) or K (3).
ast::Function transport_device_function(const Species& species, const Polynomials& p, double T0, double viscosity, double conductivity,
                                        int which);

// transport.rs:179-214: inputs [Σx, T/T0, x_0..x_{K-2}], outputs [λ/λ0, μ/η0, ρD_k/η0].
ast::Function properties(const Species& species, const Polynomials& p, double T0, double viscosity, double conductivity);

// Reference scales for transport (the role of cantera.rs:60-76): T0 = 1000 K, η0 = ρ0·(1 m²/s) and
// λ0 = cp0·η0/W̄0 for the mechanism's initial-state mixture at T0 and 1 atm.
struct TransportScales {
  double T0, viscosity, conductivity;
};
TransportScales transport_scales(const Mechanism& m, double T0 = 1000.);

// Per-state program over SoA inputs [T, P, Y_0..Y_{K-1}] (K, Pa, mass fractions). Outputs, in order:
// ω_0..ω_{Ka-1} [mol/m³/s] when rates is set; λ [W/m/K], μ [Pa s], D_0..D_{K-1} [m²/s] when transport is set;
// and, when rhs is set, dT/dt [K/s] last.
// Concentrations follow the reference's amounts/volume convention on a mass basis: n_k = Y_k/W_k and
// 1/V = (P/R)/T / Σ_j Y_j/W_j, so C_k = max(0, n_k)/V for k < K-1 and the bulk species closes C_tot = P/(R T).
struct ChemistryOptions {
  bool rates = true, transport = true, rhs = false;
  // Transport program shaped for the one-lane-per-state kernels; the same values to rounding (a few ulps per term).
  // Register-lean viscosity: v_k is evaluated again (in Horner form) where Wilke's k-th row needs it instead of
  // living beside 1/v_j through the whole sum. Issue-lean pair terms: each K×K term holds one per-pair constant,
  // so that no instruction needs two (gfx950 instructions read at most one scalar register; a second constant costs
  // two VGPR moves): Wilke's (a_kj + b_kj v_k/v_j)² as a_kj² (1 + ṽ_k/ṽ_j)² with ṽ_k = v_k W_k^(-1/4), and the
  // binary-diffusion fit c0 + c1 ℓ + … as c0 (1 + (c1/c0) ℓ + …).
  bool lean_transport = false;
  // Rates program shaped for the one-lane-per-state kernels: each reaction's products start from its own rate
  // constant, q = k ΠC^ν' − k (1/K) ΠC^ν'', so that no product of concentrations or of exp(±G/RT) factors is shared
  // between reactions (a shared one stays live from its first use to its last); the same values to rounding.
  bool lean_rates = false;
  // Transport program for FP32 evaluation: the diffusion sums S_k kept at or above FLT_MIN (see density_diffusivity).
  bool f32_range = false;
};
ast::Function chemistry(const Mechanism& m, const Polynomials* p, const TransportScales& scales, const ChemistryOptions& o);
// Constant-pressure reactor right-hand side: inputs [T, P, Y_0..Y_{K-1}], outputs [dT/dt, dY_0/dt..dY_{K-1}/dt]
// (dY_k/dt = W_k ω_k / ρ, zero for species that do not react); lean as ChemistryOptions::lean_rates.
ast::Function reactor(const Mechanism& m, bool lean = false);
// The registers of a linearized rates program that hold the clamped exp(-G_k/RT) (negative) or exp(G_k/RT) values
// (reaction.rs:236-268), in order of definition.
std::vector<int32_t> gibbs_exponentials(const ast::Linear& l, bool negative);

}  // namespace combustion
