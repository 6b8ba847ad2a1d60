// `.ignition` state files (cantera.rs:13-29).
#pragma once
#include <string>
#include <vector>

#include "mechanism.hpp"

namespace combustion {

struct IgnitionState {
  double temperature, pressure, volume;  // K, Pa, m³ (always 1)
  std::vector<double> amounts;           // mol, one per species of the mechanism in its order
};
// Throws yaml::Error for a malformed file or a state the reference would reject.
IgnitionState ignition_state(const Mechanism& m, const std::string& text);

}  // namespace combustion
