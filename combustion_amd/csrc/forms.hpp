// Kernel forms behind cb_hip_assemble: each turns a program (and, for the tiled forms, the mechanism it was
// built from) into gfx950 code and launches it over SoA state arrays.
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "ast.hpp"
#include "builders.hpp"
#include "mechanism.hpp"

namespace combustion {

struct ArgumentError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct CompileError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class ProgramKind { Other, Chemistry };

struct ChemistryDescription {
  ChemistryOptions options;
  TransportScales scales{};
  std::shared_ptr<Polynomials> polynomials;
};

struct ProgramView {
  std::shared_ptr<const Mechanism> mechanism;
  const ast::Function* function;
  const ast::Linear* linear;
  ProgramKind kind;
  const ChemistryDescription* chemistry;
};

class Form {
 public:
  virtual ~Form() = default;
  virtual int n_inputs() const = 0;
  virtual int n_outputs() const = 0;
  // inputs/outputs: tables of n_inputs()/n_outputs() arrays of n doubles, host or device memory.
  virtual void call(const double* const* inputs, size_t n, double* const* outputs, bool device, hipStream_t stream, float* kernel_ms) = 0;
  // Device memory the form's kernels need beyond the arrays and their pointer table for n states (the surrogate's
  // fallback counter and state list); 0 for the others.
  virtual size_t work_bytes(size_t n) const {
    (void)n;
    return 0;
  }
  // The kernels alone, all on `stream`, with nothing allocated, created or synchronized: in/out are device tables of
  // the arrays' device pointers and work is work_bytes(n) bytes of 256-byte aligned device memory (or null when 0).
  virtual void launch_only(const double* const* in, double* const* out, size_t n, hipStream_t stream, void* work) = 0;
};

// How SoAForm::launch runs its kernels: on the caller's stream alone (single_stream) or with a forked second
// stream of the module's; with the form's work memory given by the caller (work) or the module's own.
struct Launch {
  hipStream_t stream;
  bool single_stream;
  void* work;
};

// cb_hip_assemble's form argument: the form in the low bits, modifier flags above (cb_form, cb_form_flags).
constexpr int kFormBase = 0xF;
constexpr int kFormTiledTransport = 0x10;
constexpr int kFormTiledRates = 0x20;

// Common hiprtc options for gfx950 (packed FP32 instructions disabled).
std::vector<std::string> hip_options(bool contract);

std::string form_source(const ProgramView& p, int form, unsigned block_size);
void compile_form(const ProgramView& p, int form, unsigned block_size);
std::unique_ptr<Form> assemble_form(const ProgramView& p, int form, unsigned block_size);

// A straight-line program with explicit emission options (LDS-resident registers, say), FP contraction on.
struct StraightOptions;
StraightOptions straight_options(int form, unsigned block_size);
void compile_straight(const ProgramView& p, const StraightOptions& o);
std::unique_ptr<Form> assemble_straight(const ProgramView& p, const StraightOptions& o);

// Tiled forms (tiled.cpp).
bool tiled_applies(const ProgramView& p);
// form: 1 FP64, 2 MFMA mixed precision, 3 FP32-core straight-line rates (see cb_form), with kFormTiledTransport.
std::string tiled_source(const ProgramView& p, int form, unsigned block_size);
void compile_tiled(const ProgramView& p, int form, unsigned block_size);
std::unique_ptr<Form> assemble_tiled(const ProgramView& p, int form, unsigned block_size);

// Transport surrogate (kernels/surrogate.hip): network parameters in the Mlp layout for n_in = n_out = K+2 and
// the per-feature and per-output normalisation, training box included.
struct SurrogateWeights {
  int width = 256, depth = 3;
  std::vector<float> params, f_mean, f_rstd, lo, hi, o_mean, o_std;
  std::vector<int32_t> major;  // species whose pair terms the baselines keep
  int rates_form = 3;          // 3: FP32-core straight-line rates; 2: MFMA mixed-precision rates
};
std::unique_ptr<Form> assemble_surrogate(const ProgramView& p, const SurrogateWeights& w);
unsigned surrogate_fallback_count(Form& f, hipStream_t stream);

}  // namespace combustion
