#include "mechanism.hpp"

#include <algorithm>
#include <map>
#include <sstream>
#include <stdexcept>

#include "yaml.hpp"

namespace combustion {

const char* const element_names[n_elements] = {"H", "He", "C", "N", "O", "F", "Cl", "Ar"};
// lib.rs:22-24, g/mol divided by 1e3
const double standard_atomic_weights[n_elements] = {1.008 / 1e3,  4.002602 / 1e3,      12.011 / 1e3, 14.007 / 1e3,
                                                    15.999 / 1e3, 18.998403163 / 1e3, 35.45 / 1e3,  39.95 / 1e3};

// lib.rs:16-18, same association order as the reference.
double NASA7::molar_heat_capacity_at_constant_pressure_R(double T) const {
  const auto& a = piece(T);
  return a[0] + a[1] * T + a[2] * T * T + a[3] * T * T * T + a[4] * T * T * T * T;
}
double NASA7::enthalpy_RT(double T) const {
  const auto& a = piece(T);
  return a[0] + a[1] / 2. * T + a[2] / 3. * T * T + a[3] / 4. * T * T * T + a[4] / 5. * T * T * T * T + a[5] / T;
}
double NASA7::gibbs_RT(double T) const {
  const auto& a = piece(T);
  return a[0] - a[6] - a[0] * std::log(T) - a[1] / 2. * T + (1. / 3. - 1. / 2.) * a[2] * T * T +
         (1. / 4. - 1. / 3.) * a[3] * T * T * T + (1. / 5. - 1. / 4.) * a[4] * T * T * T * T + a[5] / T;
}

namespace {

using yaml::Node;
[[noreturn]] void fail(const std::string& what) { throw std::runtime_error(what); }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}
bool ends_with(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}
// str::trim_end_matches: strips every trailing repetition of the pattern.
std::string trim_end_matches(std::string s, const std::string& pattern) {
  while (ends_with(s, pattern)) s.resize(s.size() - pattern.size());
  return s;
}
std::vector<std::string> split(const std::string& s, const std::string& sep) {
  std::vector<std::string> out;
  size_t start = 0;
  while (true) {
    size_t i = s.find(sep, start);
    if (i == std::string::npos) {
      out.push_back(s.substr(start));
      return out;
    }
    out.push_back(s.substr(start, i - start));
    start = i + sep.size();
  }
}

using Side = std::vector<std::pair<std::string, int>>;  // species -> coefficient, first-appearance order

// yaml.rs:40-49
std::array<Side, 2> parse_equation(const std::string& equation) {
  std::vector<std::string> sides = equation.find("<=>") != std::string::npos ? split(equation, "<=>") : split(equation, "=>");
  if (sides.size() != 2) fail("equation '" + equation + "': expected exactly one '<=>' or '=>'");
  std::array<Side, 2> out;
  for (int s = 0; s < 2; s++) {
    std::string side = trim(sides[s]);
    side = trim_end_matches(side, "(+ M)");
    side = trim_end_matches(side, "+ M");
    side = trim_end_matches(side, "(+M)");
    for (const std::string& raw : split(side, " + ")) {
      std::string term = trim(raw);
      std::string specie = term;
      int count = 1;
      if (term.find(' ') != std::string::npos) {
        std::vector<std::string> words;
        size_t i = 0;
        while (i < term.size()) {
          while (i < term.size() && std::isspace((unsigned char)term[i])) i++;
          size_t j = i;
          while (j < term.size() && !std::isspace((unsigned char)term[j])) j++;
          if (j > i) words.push_back(term.substr(i, j - i));
          i = j;
        }
        if (words.size() != 2) fail("equation '" + equation + "': cannot parse term '" + term + "'");
        char* end = nullptr;
        long c = std::strtol(words[0].c_str(), &end, 10);
        if (*end != 0 || c < 0 || c > 255) fail("equation '" + equation + "': bad stoichiometric coefficient '" + words[0] + "'");
        count = (int)c;
        specie = words[1];
      }
      bool found = false;
      for (auto& [name, n] : out[s])
        if (name == specie) {
          n += count;
          found = true;
        }
      if (!found) out[s].emplace_back(specie, count);
    }
  }
  return out;
}

int coefficient(const Side& side, const std::string& specie) {
  for (auto& [name, n] : side)
    if (name == specie) return n;
  return 0;
}

struct RawSpecie {
  std::string name;
  std::vector<std::pair<Element, int>> composition;  // file order
  std::vector<double> temperature_ranges;
  std::vector<std::array<double, 7>> pieces;
  double well_depth_K = 0, diameter_A = 0;
  Geometry geometry = Geometry::Atom;
  double polarizability_A3 = 0, rotational_relaxation = 0, dipole_Debye = 0;
};

struct RawReaction {
  std::array<Side, 2> equation;
  std::string text;
  RateConstant rate_constant{};
  ReactionModel model = ReactionModel::Elementary;
  std::vector<std::pair<std::string, double>> efficiencies;
  RateConstant k0{};
  Troe troe{};
  SRI sri{};
  std::vector<PlogEntry> plog;
};

// A pressure: a number in Pa, or a number followed by Pa, kPa, MPa, bar, atm or torr.
double pressure_Pa(const Node& n, const std::string& where) {
  if (n.kind != Node::Scalar) fail(where + ": missing pressure");
  std::istringstream in(n.scalar);
  double v;
  std::string unit;
  if (!(in >> v)) fail(where + ": bad pressure '" + n.scalar + "'");
  in >> unit;
  if (unit.empty() || unit == "Pa") return v;
  if (unit == "kPa") return v * 1e3;
  if (unit == "MPa") return v * 1e6;
  if (unit == "bar") return v * 1e5;
  if (unit == "atm") return v * 101325.;
  if (unit == "torr") return v * (101325. / 760.);
  fail(where + ": unknown pressure unit '" + unit + "'");
}

Element element_from(const std::string& s) {
  for (int e = 0; e < n_elements; e++)
    if (s == element_names[e]) return Element(e);
  fail("unknown element '" + s + "'");
}

}  // namespace

Mechanism parse_yaml(const std::string& text) {
  Node root = yaml::parse(text);
  const Node& units = root["units"];
  // yaml.rs:58
  if (units["length"].is_null() || units["length"].str() != "cm" || units["time"].is_null() || units["time"].str() != "s" ||
      units["quantity"].is_null() || units["quantity"].str() != "mol")
    fail("units must be {length: cm, time: s, quantity: mol}");
  if (units["activation-energy"].is_null()) fail("units: missing activation-energy");
  const std::string ea_unit = units["activation-energy"].str();
  if (ea_unit != "K" && ea_unit != "cal/mol") fail("unsupported activation-energy unit '" + ea_unit + "'");

  // yaml.rs:59-96
  std::vector<RawSpecie> species;
  const Node& species_node = root["species"];
  if (species_node.kind != Node::Sequence) fail("missing species list");
  for (const Node& s : species_node.items) {
    RawSpecie r;
    r.name = s["name"].str();
    for (auto& [k, v] : s["composition"].pairs) r.composition.emplace_back(element_from(k), (int)v.i64());
    const Node& thermo = s["thermo"];
    for (const Node& t : thermo["temperature-ranges"].items) r.temperature_ranges.push_back(t.f64());
    for (const Node& p : thermo["data"].items) {
      if (p.size() != 7) fail(r.name + ": NASA-7 piece must have 7 coefficients");
      std::array<double, 7> a;
      for (int i = 0; i < 7; i++) a[i] = p[i].f64();
      r.pieces.push_back(a);
    }
    // yaml.rs:64-77: identical pieces (a6 within 3e-8) are merged into one range.
    bool duplicates = false;
    for (size_t i = 1; i < r.pieces.size(); i++)
      for (size_t j = i; j < r.pieces.size(); j++) {
        const auto &a = r.pieces[i - 1], &b = r.pieces[j];
        if (std::equal(a.begin(), a.begin() + 6, b.begin()) && std::fabs(a[6] - b[6]) < 3e-8) duplicates = true;
      }
    if (duplicates) {
      if (r.pieces.size() != 2 || r.temperature_ranges.size() != 3) fail(r.name + ": cannot merge duplicate NASA-7 pieces");
      r.temperature_ranges = {r.temperature_ranges[0], r.temperature_ranges[2]};
      r.pieces.resize(1);
    }
    const Node& tr = s["transport"];
    r.well_depth_K = tr["well-depth"].f64();
    r.diameter_A = tr["diameter"].f64();
    const std::string geometry = tr["geometry"].str();
    auto opt = [&](const char* key) { return tr[key].is_null() ? 0. : tr[key].f64(); };
    if (geometry == "atom") {
      r.geometry = Geometry::Atom;
    } else if (geometry == "linear") {
      r.geometry = Geometry::Linear;
      r.polarizability_A3 = opt("polarizability");
      r.rotational_relaxation = opt("rotational-relaxation");
    } else if (geometry == "nonlinear") {
      r.geometry = Geometry::Nonlinear;
      r.polarizability_A3 = opt("polarizability");
      r.rotational_relaxation = opt("rotational-relaxation");
      r.dipole_Debye = opt("dipole");
    } else {
      fail(r.name + ": unsupported geometry '" + geometry + "'");
    }
    species.push_back(std::move(r));
  }

  // yaml.rs:97-158
  std::vector<RawReaction> reactions;
  const Node& reactions_node = root["reactions"];
  if (reactions_node.kind != Node::Sequence) fail("missing reactions list");
  for (const Node& n : reactions_node.items) {
    RawReaction r;
    r.text = n["equation"].str();
    r.equation = parse_equation(r.text);
    const Node& type = n["type"];
    const std::string type_name = type.is_null() ? "" : type.str();
    int reactants = 0;
    for (auto& [_, c] : r.equation[0]) reactants += c;
    if (type_name == "three-body") reactants += 1;
    auto rate_constant = [&](const Node& k, int exponent) {
      if (k.kind != Node::Mapping) fail("reaction '" + r.text + "': missing rate constant");
      RateConstant c;
      c.preexponential_factor = k["A"].f64() * std::pow(1e-6, (double)exponent);
      c.temperature_exponent = k["b"].f64();
      double Ea = k["Ea"].f64();
      c.activation_temperature = ea_unit == "K" ? Ea : Ea * 4.184 / R;
      return c;
    };
    const Node* k = &n["rate-constant"];
    if (k->is_null()) k = &n["high-P-rate-constant"];
    if (k->is_null()) {
      if (type_name != "pressure-dependent-Arrhenius" || n["rate-constants"].size() < 1)
        fail("reaction '" + r.text + "': no rate constant");
      k = &n["rate-constants"][0];
    }
    r.rate_constant = rate_constant(*k, reactants - 1);
    if (type_name == "pressure-dependent-Arrhenius" && n["rate-constants"].size() > 1) {
      // Not in the reference (yaml.rs:120,149 assert a single entry): Cantera's PLOG table.
      for (size_t i = 0; i < n["rate-constants"].size(); i++) {
        const Node& e = n["rate-constants"][i];
        r.plog.push_back({pressure_Pa(e["P"], "reaction '" + r.text + "'"), rate_constant(e, reactants - 1)});
      }
      std::stable_sort(r.plog.begin(), r.plog.end(), [](const PlogEntry& a, const PlogEntry& b) { return a.P < b.P; });
      if (r.plog.front().P <= 0) fail("reaction '" + r.text + "': non-positive PLOG pressure");
    }
    for (auto& [name, v] : n["efficiencies"].pairs) r.efficiencies.emplace_back(name, v.f64());
    const bool reversible = r.text.find(" <=> ") != std::string::npos;
    const bool irreversible = !reversible && r.text.find(" => ") != std::string::npos;
    if (type_name.empty() || type_name == "elementary" || type_name == "pressure-dependent-Arrhenius") {
      if (reversible)
        r.model = ReactionModel::Elementary;
      else if (irreversible)
        r.model = ReactionModel::Irreversible;
      else
        fail("reaction '" + r.text + "': neither ' <=> ' nor ' => '");
    } else if (type_name == "three-body") {
      r.model = ReactionModel::ThreeBody;
    } else if (type_name == "falloff") {
      r.k0 = rate_constant(n["low-P-rate-constant"], reactants);
      if (!n["Troe"].is_null()) {
        const Node& t = n["Troe"];
        r.model = ReactionModel::Troe;
        r.troe = {t["A"].f64(), t["T3"].f64(), t["T1"].f64(), t["T2"].is_null() ? INFINITY : t["T2"].f64()};
      } else if (!n["SRI"].is_null()) {
        const Node& t = n["SRI"];
        r.model = ReactionModel::SRI;
        r.sri = {t["A"].f64(), t["B"].f64(), t["C"].f64(), t["D"].is_null() ? 1. : t["D"].f64(), t["E"].is_null() ? 0. : t["E"].f64()};
      } else {
        r.model = ReactionModel::PressureModification;
      }
    } else {
      fail("reaction '" + r.text + "': unsupported type '" + type_name + "'");
    }
    reactions.push_back(std::move(r));
  }

  // yaml.rs:159-162: stable partition, reacting species first.
  auto reacts = [&](const std::string& name) {
    for (auto& r : reactions)
      if (coefficient(r.equation[0], name) != coefficient(r.equation[1], name)) return true;
    return false;
  };
  std::stable_partition(species.begin(), species.end(), [&](const RawSpecie& s) { return reacts(s.name); });

  // lib.rs:127-140
  Mechanism m;
  Species& S = m.species;
  const size_t K = species.size();
  for (const RawSpecie& s : species) {
    S.names.push_back(s.name);
    std::array<int, n_elements> composition{};
    double W = 0;
    for (auto& [e, count] : s.composition) {
      composition[(int)e] += count;
      W += (double)count * standard_atomic_weights[(int)e];
    }
    S.composition.push_back(composition);
    S.molar_mass.push_back(W);
    NASA7 t;
    if (s.temperature_ranges.size() == 3 && s.pieces.size() == 2) {
      t.temperature_split = s.temperature_ranges[1];
      t.pieces = {s.pieces[0], s.pieces[1]};
    } else if (s.temperature_ranges.size() == 2 && s.pieces.size() == 1) {
      t.temperature_split = NAN;
      t.pieces = {s.pieces[0], s.pieces[0]};
    } else {
      fail(s.name + ": NASA-7 needs 1 or 2 temperature ranges with matching pieces");
    }
    S.thermodynamics.push_back(t);
    S.geometry.push_back(s.geometry);
    S.diameter.push_back(s.diameter_A * 1e-10);
    S.well_depth_J.push_back(s.well_depth_K * kB);
    S.polarizability.push_back(s.geometry == Geometry::Atom ? 0. : s.polarizability_A3 * 1e-30);
    const double Cm_per_Debye = 1e-21 / 299792458.;  // lib.rs:5-6
    S.permanent_dipole_moment.push_back(s.geometry == Geometry::Nonlinear ? s.dipole_Debye * Cm_per_Debye : 0.);
    S.rotational_relaxation.push_back(s.geometry == Geometry::Atom ? 0. : s.rotational_relaxation);
    S.internal_degrees_of_freedom.push_back(s.geometry == Geometry::Atom ? 0. : s.geometry == Geometry::Linear ? 1. : 3. / 2.);
    S.heat_capacity_ratio.push_back(1. + 2. / (s.geometry == Geometry::Atom ? 3. : s.geometry == Geometry::Linear ? 5. : 6.));
  }
  size_t active = 0;
  while (active < K && reacts(S.names[active])) active++;
  for (size_t k = active; k < K; k++)
    if (reacts(S.names[k])) fail("reacting species must precede inert ones");
  m.active = active;
  if (reactions.empty()) fail("mechanism has no reactions");

  for (const RawReaction& r : reactions) {
    for (int s = 0; s < 2; s++)
      for (auto& [name, _] : r.equation[s])
        if (std::find(S.names.begin(), S.names.end(), name) == S.names.end())
          fail("reaction '" + r.text + "': unknown species '" + name + "'");
    Reaction x;
    x.equation = r.text;
    for (size_t k = 0; k < K; k++) {
      x.reactants.push_back((uint8_t)coefficient(r.equation[0], S.names[k]));
      x.products.push_back((uint8_t)coefficient(r.equation[1], S.names[k]));
    }
    for (size_t k = 0; k < active; k++) x.net.push_back((int8_t)(x.products[k] - x.reactants[k]));
    for (size_t k = 0; k < K; k++) {
      x.sum_reactants += x.reactants[k];
      x.sum_products += x.products[k];
    }
    x.sum_net = x.sum_products - x.sum_reactants;
    x.rate_constant = r.rate_constant;
    x.model = r.model;
    if (r.model != ReactionModel::Elementary && r.model != ReactionModel::Irreversible) {
      x.efficiencies.assign(K, 1.);
      for (size_t k = 0; k < K; k++)
        for (auto& [name, e] : r.efficiencies)
          if (name == S.names[k]) x.efficiencies[k] = e;
    }
    x.k0 = r.k0;
    x.troe = r.troe;
    x.sri = r.sri;
    x.plog = r.plog;
    m.reactions.push_back(std::move(x));
  }

  // yaml.rs:163-169 and lib.rs:70-80: initial state (T = 1000 K, P = 1 atm, V = 1, listed species equimolar).
  static const char* initial[] = {"HE", "AR", "N2", "CO2", "H2", "O2", "CO", "CH2O", "CH4", "C2H6", "C2H4", "C2H2", "C3H6", "pC3H4", "aC3H4",
                                  "1-C4H8", "2-C4H8", "i-C4H8", "1,3-C4H6", "C4H4", "C4H2", "C6H6", "C7H8", "CH3OH", "CH3CHO", "C2H5OH",
                                  "CH3OCH3", "C3H8", "n-C4H10", "i-C4H10"};
  std::vector<double> proportions(K, 0.);
  double total = 0;
  for (size_t k = 0; k < K; k++) {
    for (const char* name : initial)
      if (S.names[k] == name) proportions[k] = 1.;
    total += proportions[k];
  }
  m.initial_state.temperature = 1000.;
  m.initial_state.volume = 1.;
  m.initial_state.pressure_R = 101325. / R;
  const double amount = m.initial_state.pressure_R * m.initial_state.volume / m.initial_state.temperature;
  for (size_t k = 0; k < K; k++) m.initial_state.amounts.push_back(amount * proportions[k] / total);
  return m;
}

}  // namespace combustion
