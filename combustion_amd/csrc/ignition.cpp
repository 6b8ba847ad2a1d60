// Initial states of `<mechanism>.ignition` files, read as cantera.rs:13-29 reads them: a YAML sequence whose first
// entry holds the pressure [Pa], the temperature [K] and X, a map from every species name to its (unnormalised,
// non-negative) mole proportion, and no volume. The state is the reference's State: V = 1 m³ and the amounts
// n_k = (P/R) V / T · x_k / Σ_j x_j, in the mechanism's species order (active species first).
#include <cmath>
#include <string>

#include "ignition.hpp"
#include "yaml.hpp"

namespace combustion {

namespace {
[[noreturn]] void fail(int line, const std::string& what) {
  throw yaml::Error("ignition state (line " + std::to_string(line) + "): " + what);
}
constexpr double kGasConstant = 1.380649e-23 * 6.02214076e23;  // model.rs:6-8
}  // namespace

IgnitionState ignition_state(const Mechanism& m, const std::string& text) {
  const yaml::Node doc = yaml::parse(text);
  if (doc.kind != yaml::Node::Sequence || doc.size() == 0) fail(doc.line, "expected a sequence of states");
  const yaml::Node& s = doc[0];
  if (s.kind != yaml::Node::Mapping) fail(s.line, "a state is a mapping");
  const yaml::Node &P = s["pressure"], &T = s["temperature"], &X = s["X"];
  if (P.is_null() || T.is_null() || X.is_null()) fail(s.line, "a state needs pressure, temperature and X");
  if (!s["volume"].is_null()) fail(s["volume"].line, "a state gives no volume (the reference asserts its absence)");
  if (X.kind != yaml::Node::Mapping) fail(X.line, "X is a map from species name to mole proportion");
  IgnitionState st;
  st.pressure = P.f64();
  st.temperature = T.f64();
  st.volume = 1.;
  if (!(st.pressure > 0) || !(st.temperature > 0)) fail(s.line, "pressure and temperature must be positive");
  std::vector<double> x;
  double sum = 0;
  for (const std::string& name : m.species.names) {
    const yaml::Node& v = X[name];
    if (v.is_null()) fail(X.line, "X has no entry for species " + name);
    const double xk = v.f64();
    if (!(xk >= 0) || !std::isfinite(xk)) fail(v.line, "mole proportion of " + name + " must be finite and non-negative");
    x.push_back(xk);
    sum += xk;
  }
  if (!(sum > 0)) fail(X.line, "the mole proportions sum to zero");
  const double total = st.pressure / kGasConstant * st.volume / st.temperature;
  for (double xk : x) st.amounts.push_back(total * xk / sum);
  return st;
}

}  // namespace combustion
