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
  Straight(const ProgramView& p, unsigned block_size)
      : SoAForm(p.function->n_input, (int)p.function->output.size()), block_size_(block_size) {
    kernel_.load(device::compile(emit_straight(*p.linear, "program", block_size), "program", hip_options(false)), "program");
  }

 protected:
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) override {
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

std::string form_source(const ProgramView& p, int form, unsigned block_size) {
  switch (form) {
    case 0: return emit_straight(*p.linear, "program", block_size ? block_size : 256);
    case 1: return tiled_source(p, false, block_size);
    case 2: return tiled_source(p, true, block_size);
  }
  throw ArgumentError("unknown form");
}

void compile_form(const ProgramView& p, int form, unsigned block_size) {
  if (form == 0) {
    device::compile(emit_straight(*p.linear, "program", block_size ? block_size : 256), "program", hip_options(false));
    return;
  }
  if (form == 1 || form == 2) {
    compile_tiled(p, form == 2, block_size);
    return;
  }
  throw ArgumentError("unknown form");
}

std::unique_ptr<Form> assemble_form(const ProgramView& p, int form, unsigned block_size) {
  switch (form) {
    case 0: return std::make_unique<Straight>(p, block_size ? block_size : 256);
    case 1: return assemble_tiled(p, false, block_size);
    case 2: return assemble_tiled(p, true, block_size);
  }
  throw ArgumentError("unknown form");
}

}  // namespace combustion
