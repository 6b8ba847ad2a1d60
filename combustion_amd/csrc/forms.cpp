#include "forms.hpp"

#include <vector>

#include "device.hpp"
#include "emit.hpp"
#include "soa.hpp"

namespace combustion {

std::vector<std::string> hip_options(bool contract) {
  return {"--offload-arch=gfx950", "-O3", "-std=c++17", contract ? "-ffp-contract=fast" : "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

namespace {

// Straight-line program, one lane per state (the cuda.rs::compile counterpart). FP contraction is off so that
// every operation rounds as in the FP64 interpreter (ast/interpret.rs).
class Straight : public SoAForm {
 public:
  Straight(const ProgramView& p, const StraightOptions& o, bool contract)
      : SoAForm(p.function->n_input, (int)p.function->output.size()), block_size_(o.block_size) {
    kernel_.load(device::compile(emit_straight(*p.linear, "program", o), "program", hip_options(contract)), "program");
  }

 protected:
  void launch(const double* const* in, double* const* out, size_t n, const Launch& how) override {
    hipStream_t stream = how.stream;
    unsigned n32 = (unsigned)n, offset = 0;
    void* args[] = {&in, &out, &n32, &offset};
    unsigned grid = (unsigned)((n + block_size_ - 1) / block_size_);
    device::check(hipModuleLaunchKernel(kernel_.function, grid, 1, 1, block_size_, 1, 1, 0, stream, args, nullptr), "launch");
  }

 private:
  unsigned block_size_;
  device::Kernel kernel_;
};

}  // namespace

// Form 0 keeps every operation as the FP64 interpreter rounds it; form 1 on a straight-line program allows FMA
// contraction, reciprocal-multiply division with one Newton step (relative error below 1e-14) and the polynomial
// exponential (a few ulps), as the FP64 chemistry kernels do; form 3 also evaluates exp and ln with FP32 cores.
StraightOptions straight_options(int form, unsigned block_size) {
  StraightOptions o;
  o.block_size = block_size ? block_size : 256;
  o.reciprocal = form == 1 || form == 3;
  o.newton_steps = form == 1 ? 1 : 2;
  o.exp_poly = form == 1;
  o.fast_transcendentals = form == 3;
  return o;
}

// Chemistry programs (without the reactor right-hand side) take the per-section kernels in forms 1-3; form 2 needs
// them, and so does the tiled-transport modifier.
bool use_tiled(const ProgramView& p, int form_flags) {
  const int form = form_flags & kFormBase;
  if (form_flags & (kFormTiledTransport | kFormTiledRates)) {
    if (form < 1 || form > 3) throw ArgumentError("the tiled-transport and tiled-rates modifiers apply to forms 1-3");
    if ((form_flags & kFormTiledRates) && form == 2) throw ArgumentError("the mixed form has its own rates kernel");
    if (!tiled_applies(p)) throw Unsupported("the tiled modifiers need a chemistry program without the reactor right-hand side");
    return true;
  }
  return form == 2 || ((form == 1 || form == 3) && tiled_applies(p));
}

void check_form(int form_flags) {
  if ((form_flags & ~(kFormBase | kFormTiledTransport | kFormTiledRates)) || (form_flags & kFormBase) > 3) throw ArgumentError("unknown form");
}

std::string form_source(const ProgramView& p, int form, unsigned block_size) {
  check_form(form);
  if (use_tiled(p, form)) return tiled_source(p, form, block_size);
  return emit_straight(*p.linear, "program", straight_options(form, block_size));
}

void compile_form(const ProgramView& p, int form, unsigned block_size) {
  check_form(form);
  if (use_tiled(p, form)) return compile_tiled(p, form, block_size);
  device::compile(emit_straight(*p.linear, "program", straight_options(form, block_size)), "program", hip_options(form != 0));
}

void compile_straight(const ProgramView& p, const StraightOptions& o) {
  device::compile(emit_straight(*p.linear, "program", o), "program", hip_options(true));
}

std::unique_ptr<Form> assemble_straight(const ProgramView& p, const StraightOptions& o) {
  return std::make_unique<Straight>(p, o, true);
}

std::unique_ptr<Form> assemble_form(const ProgramView& p, int form, unsigned block_size) {
  check_form(form);
  if (use_tiled(p, form)) return assemble_tiled(p, form, block_size);
  return std::make_unique<Straight>(p, straight_options(form, block_size), form != 0);
}

}  // namespace combustion
