// The per-state program compiled for the host CPU (the counterpart of ir/lib.rs:225-263 and device.rs:28-48).
#pragma once
#include <string>
#include <vector>

#include "ast.hpp"

namespace combustion {

// Straight-line C of a linear program: void <name>(const double* in, double* out), one state in AoS order.
std::string emit_host(const ast::Linear& l, const std::string& name);
// The host compiler: $COMBUSTION_AMD_HOST_CC, else ROCm's clang, else cc.
std::string host_compiler();

class HostProgram {
 public:
  explicit HostProgram(const ast::Linear& l);
  ~HostProgram();
  HostProgram(const HostProgram&) = delete;
  HostProgram& operator=(const HostProgram&) = delete;
  int n_inputs() const { return n_in_; }
  int n_outputs() const { return n_out_; }
  // SoA host arrays of n states; threads >= 1 take contiguous ranges of states.
  void call(const double* const* inputs, size_t n, double* const* outputs, int threads) const;

 private:
  using Fn = void (*)(const double*, double*);
  int n_in_, n_out_;
  void* handle_ = nullptr;
  Fn fn_ = nullptr;
};

}  // namespace combustion
