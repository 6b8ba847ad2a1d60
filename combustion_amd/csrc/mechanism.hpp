// Mechanism model: the Cantera-YAML front end and the species/reaction tables the builders consume.
// Mirrors model.rs:5-67 (data model), yaml.rs:40-171 (parser and unit conversions) and lib.rs:8-140
// (NASA-7 pieces, Species, Reaction, new()).
#pragma once
#include <array>
#include <cmath>
#include <string>
#include <vector>

namespace combustion {

// model.rs:6-8
constexpr double kB = 1.380649e-23;  // J/K
constexpr double NA = 6.02214076e23;
constexpr double R = kB * NA;

// model.rs:5 Element, lib.rs:22-24 standard atomic weights (kg/mol)
enum class Element { H, He, C, N, O, F, Cl, Ar };
constexpr int n_elements = 8;
extern const char* const element_names[n_elements];
extern const double standard_atomic_weights[n_elements];

// lib.rs:8-19
struct NASA7 {
  double temperature_split;  // NaN for a single-range polynomial (both pieces equal)
  std::array<std::array<double, 7>, 2> pieces;
  static constexpr double reference_pressure = 101325. / R;  // lib.rs:14 (P0/R)
  const std::array<double, 7>& piece(double T) const { return pieces[T <= temperature_split ? 0 : 1]; }
  double molar_heat_capacity_at_constant_pressure_R(double T) const;
  double enthalpy_RT(double T) const;
  double gibbs_RT(double T) const;
};

// model.rs:31-38
struct RateConstant {
  double preexponential_factor;   // SI (m, mol, s)
  double temperature_exponent;
  double activation_temperature;  // K
};
struct Troe { double A, T3, T1, T2; };
struct SRI { double A, B, C, D, E; };
// One entry of a pressure-dependent Arrhenius (PLOG) table: the rate constant at pressure P (Pa).
struct PlogEntry {
  double P;
  RateConstant k;
};

// lib.rs:84-91
enum class ReactionModel { Elementary, Irreversible, ThreeBody, PressureModification, Troe, SRI };

// lib.rs:93-125
struct Reaction {
  std::string equation;
  std::vector<uint8_t> reactants, products;  // over all K species
  std::vector<int8_t> net;                   // products - reactants over the active species
  int sum_reactants = 0, sum_products = 0, sum_net = 0;
  RateConstant rate_constant{};
  ReactionModel model = ReactionModel::Elementary;
  std::vector<double> efficiencies;  // over all K species, default 1 (third-body and falloff models)
  RateConstant k0{};                 // low-pressure limit (falloff models)
  Troe troe{};
  SRI sri{};
  // Pressure-dependent Arrhenius with more than one entry (Elementary/Irreversible models), sorted by P; entries
  // of equal P are summed. Empty for the single-entry form, which is rate_constant itself (yaml.rs:120,149).
  std::vector<PlogEntry> plog;
};

enum class Geometry { Atom, Linear, Nonlinear };

// lib.rs:26-61 (SI units)
struct Species {
  std::vector<std::string> names;
  std::vector<std::array<int, n_elements>> composition;
  std::vector<double> molar_mass;  // kg/mol
  std::vector<NASA7> thermodynamics;
  std::vector<Geometry> geometry;
  std::vector<double> diameter;                 // m
  std::vector<double> well_depth_J;             // J
  std::vector<double> polarizability;           // m^3
  std::vector<double> permanent_dipole_moment;  // C m
  std::vector<double> rotational_relaxation;
  std::vector<double> internal_degrees_of_freedom;
  std::vector<double> heat_capacity_ratio;
  size_t len() const { return molar_mass.size(); }
};

// lib.rs:63-68, lib.rs:70-80
struct State {
  double pressure_R, temperature, volume;
  std::vector<double> amounts;
};

struct Mechanism {
  Species species;
  size_t active = 0;  // reacting species form a prefix (lib.rs:130-136)
  std::vector<Reaction> reactions;
  State initial_state;
  double time_step = 1e-8;  // yaml.rs:169
};

// yaml.rs:54 parse + lib.rs:127 new; throws std::runtime_error on malformed or unsupported input.
Mechanism parse_yaml(const std::string& text);

}  // namespace combustion
