// Batched stiff integration of constant-pressure reactors on the GPU (the role of cvode/lib.rs:17-47 and the
// integration loop of plot.rs:80-118): the reactor right-hand side and its Jacobian (the analytic program, or
// difference quotients) as FP64 straight-line kernels, and the Rosenbrock stages in kernels/rosenbrock.hip.
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <vector>

#include "ast.hpp"
#include "device.hpp"
#include "forms.hpp"
#include "mechanism.hpp"

namespace combustion {

// Options (cb_integrator_create's flags): the Jacobian, the Rosenbrock kernel's packing, and the method.
constexpr int kJacobianAutomatic = 0;    // the analytic program up to kAnalyticMaxState rows, difference quotients above
constexpr int kJacobianAnalytic = 1;     // the forward-mode program of cb_build_jacobian
constexpr int kJacobianDifferences = 2;  // difference quotients from n (K+1) copies of the states (split right-hand side)
constexpr int kJacobianMask = 0xF;
constexpr int kOneSystemPerWave = 0x10;  // one reactor per Rosenbrock wave (readlane broadcasts) whatever its size
constexpr int kMethodOde23s = 0x20;      // ode23s's (2,3) pair (three phases) instead of RODAS4 (six)
constexpr int kAnalyticMaxState = 8;  // beyond, the analytic program spills (3.2 KB per lane at 11 rows) and loses to differences

// Fills the code-object cache with the integrator's kernels for a mechanism (no GPU needed).
void compile_integrator(std::shared_ptr<const Mechanism> m, int flags);

class Integrator {
 public:
  // Tolerances as CVODE's scalar pair (cvode/lib.rs:30): rtol relative; atol_T for the temperature and atol_Y for
  // the mass fractions.
  Integrator(std::shared_ptr<const Mechanism> m, double rtol, double atol_T, double atol_Y, int flags);
  ~Integrator();
  int state_size() const { return m_; }
  // Advances n reactors to t_end in place, all arrays device memory: y [K+1][n] (T, then Y_0..Y_{K-1}, SoA), P [n],
  // t [n], h [n] (each reactor's time and next step size), status [n] (0 running, 1 at t_end, 2 step size
  // collapsed; reactors not at 0 are left alone), steps [n] (incremented per attempted step). At most max_steps
  // steps; returns the steps taken by the batch.
  long advance(double* y, const double* P, double* t, double* h, int* status, unsigned* steps, size_t n, double t_end, int max_steps,
               hipStream_t stream);

 private:
  std::shared_ptr<const Mechanism> mech_;
  int m_;
  double rtol_, atol_T_, atol_Y_;
  bool differences_;
  int phases_;  // Rosenbrock phases per step: 3 (ode23s) or 6 (RODAS4)
  int parts_count_;  // the batch's parts, each on its own stream
  unsigned systems_per_wave_;
  ast::Function rhs_function_, jacobian_function_, shared_function_, varying_function_;
  ast::Linear rhs_linear_, jacobian_linear_, shared_linear_, varying_linear_;
  std::unique_ptr<Form> rhs_, jacobian_, shared_, varying_;
  device::Kernel step_;
  size_t n_ = 0;
  std::vector<void*> buffers_;
  double *jac_ = nullptr, *F_ = nullptr, *ys_ = nullptr, *kst_ = nullptr, *lu_ = nullptr, *rd_ = nullptr;  // kst_: the stage vectors
  int* perm_ = nullptr;
  unsigned* running_ = nullptr;
  double *groups_ = nullptr, *raised_T_ = nullptr;  // each reactor's values for its copies; the raised temperatures
  // The batch runs in parts_count_ parts, the first on the caller's stream and the others on sides_, so that one
  // part's latency-bound right-hand sides overlap another's Rosenbrock phases. Each part's pointer tables start at its
  // first reactor: the programs' inputs from y or ys and outputs to F or jac; for difference quotients, the shared
  // part's inputs [T, P] and outputs, the copies' inputs (the reactors' values and status) and outputs, and the
  // temperature-raised states' inputs and outputs.
  struct Part {
    size_t off = 0, n = 0;
    void **tab_y = nullptr, **tab_ys = nullptr, **tab_jac = nullptr, **tab_F = nullptr;
    void **tab_shared_in = nullptr, **tab_shared_out = nullptr, **tab_rep_in = nullptr, **tab_rep_out = nullptr;
    void **tab_t_in = nullptr, **tab_t_out = nullptr;
  };
  std::vector<Part> parts_;
  std::vector<hipStream_t> sides_;  // the streams of the parts after the first
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  hipEvent_t polled_[2] = {nullptr, nullptr};  // the last two polls' counts copied to host_count_ (pinned)
  unsigned* host_count_ = nullptr;
  const double* last_y_ = nullptr;
  const double* last_P_ = nullptr;
  const int* last_status_ = nullptr;
  void release();
  void prepare(double* y, const double* P, const int* status, size_t n);
  void join_sides(hipStream_t stream);
};

}  // namespace combustion
