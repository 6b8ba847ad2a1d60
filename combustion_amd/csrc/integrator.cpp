// Batched constant-pressure reactors on the GPU: per step, f and J (the analytic Jacobian program in one kernel, or
// difference quotients: a kernel writing each reactor's values and increments, the right-hand side split at the mass
// fractions — its temperature and pressure part once per reactor, the rest over the reactor's K+1 copies, each copy
// raising one component — and the whole right-hand side at the raised temperature, whose outputs the first
// Rosenbrock phase turns into J), the Rosenbrock phases of kernels/rosenbrock.hip (three for ode23s, six for
// RODAS4) with a right-hand-side evaluation between two phases; the batch in parts, on the caller's stream and on
// streams of the integrator's own, polled for reactors still running every few steps.
#include "integrator.hpp"

#include <algorithm>
#include <sstream>

#include "builders.hpp"
#include "emit.hpp"
#include "jacobian.hpp"

namespace combustion {

extern const char* const rosenbrock_source;

namespace {
constexpr int kPollEvery = 8;

// One wave per block for the right-hand-side and Jacobian programs: a launch over n reactors then spreads over
// every CU (2^14 reactors in blocks of 256 states would occupy 64 of the 256 CUs).
constexpr unsigned kProgramBlock = 64;

// The right-hand side as the FP64 rates kernel is built: the lean program (each rate-of-progress product started
// from its reaction's rate constant, so that fewer values stay live), and, when its peak live set exceeds two waves'
// registers, its clamped exp(-G_k/RT) values in LDS (at K=53: no scratch, against 1.2 KB per lane for the plain
// program in registers).
StraightOptions rhs_options(const ast::Linear& l) {
  StraightOptions o = straight_options(1, kProgramBlock);
  if (ast::peak_live(l) > 128) o.lds_registers = gibbs_exponentials(l, true);
  if (o.lds_registers.size() * o.block_size * 8 > 160 * 1024) o.lds_registers.clear();
  return o;
}

// Difference quotients: the copies of a reactor, K+1 rounded up to a power of two (the copies past K+1 repeat the
// state), so that a block of 64 holds whole reactors.
unsigned copies_per_reactor(int m) {
  unsigned c = 1;
  while (c < (unsigned)m) c <<= 1;
  return c;
}

// A reactor's values for its copies, reactor by reactor (stride B + 2M, M = K+1): the B values of the right-hand side's
// shared part, then T, P, Y_0..Y_{K-1}, then the increments inc_1..inc_{M-1} of the components Y_0..Y_{K-1}.
size_t group_stride(size_t B, int m) { return B + 2 * (size_t)m; }

// The part of the right-hand side shared by a reactor's copies (from T and P), writing the first B of the reactor's
// values.
StraightOptions shared_options(const ast::Linear& shared, int m) {
  StraightOptions o = straight_options(1, kProgramBlock);
  o.output_stride = (unsigned)group_stride(shared.output.size(), m);
  return o;
}

// The part that varies with the mass fractions, over the copies (a block of 64 holds whole reactors, whose values it
// stages in LDS): every input from the reactor's values, copy j = 1..M-1 adding inc_j to component j (Y_{j-1}). When
// a block holds one reactor's copies, it skips a reactor that is not running (in[1]: the reactors' status).
StraightOptions varying_options(const ast::Linear& varying, int m) {
  StraightOptions o = rhs_options(varying);
  if (copies_per_reactor(m) == o.block_size) o.skip_table = 1;
  const int32_t B = varying.n_input - (m + 1);
  o.group_table = 0;
  o.group_size = copies_per_reactor(m);
  o.group_stride = (unsigned)group_stride((size_t)B, m);
  o.group_offset.resize(varying.n_input);
  o.increment_offset.assign(varying.n_input, -1);
  o.increment_state.assign(varying.n_input, -1);
  for (int32_t r = 0; r <= m; r++) o.group_offset[r] = B + r;  // T, P, Y_0..Y_{K-1}
  for (int32_t r = 2; r <= m; r++) {                           // Y_{r-2}, component r - 1
    o.increment_offset[r] = B + m + r - 1;
    o.increment_state[r] = r - 1;
  }
  for (int32_t b = 0; b < B; b++) o.group_offset[m + 1 + b] = b;
  return o;
}

ast::Function placeholder(const ast::Linear& l) {  // the Function of a program that lives in its linear form only
  ast::Function f;
  f.n_input = l.n_input;
  f.output.assign(l.output.size(), -1);
  return f;
}
}  // namespace

// The right-hand side program, its Jacobian program (d [dT/dt, dY/dt] / d [T, Y]; P is a parameter) or its split
// for difference quotients, and the Rosenbrock kernels' source, for a mechanism.
struct IntegratorPrograms {
  int m;
  bool differences;
  ast::Function rhs_function, jacobian_function, shared_function, varying_function;
  ast::Linear rhs_linear, jacobian_linear, shared_linear, varying_linear;
  std::string step_source;
  IntegratorPrograms(const Mechanism& mech, int flags) : m((int)mech.species.len() + 1) {
    if (m > 64) throw Unsupported("the batched integrator holds a reactor's state in one wave: at most 63 species");
    const int jac = flags & kJacobianMask;
    if ((flags & ~(kJacobianMask | kOneSystemPerWave | kMethodOde23s)) || jac > kJacobianDifferences)
      throw ArgumentError("unknown integrator flags");
    differences = jac == kJacobianDifferences || (jac == kJacobianAutomatic && m > kAnalyticMaxState);
    rhs_function = reactor(mech, true);
    rhs_linear = ast::linearize(rhs_function);
    if (!differences) {
      std::vector<int32_t> wrt = {0}, rows;
      for (int k = 0; k + 1 < m; k++) wrt.push_back(2 + k);
      for (int o = 0; o < m; o++) rows.push_back(o);
      jacobian_linear = ast::jacobian(rhs_linear, wrt, rows);
      jacobian_function = placeholder(jacobian_linear);
    } else {
      ast::Split parts = ast::split_at(rhs_linear, 2);  // T and P shared, the mass fractions varying
      shared_linear = std::move(parts.shared);
      varying_linear = std::move(parts.varying);
      shared_function = placeholder(shared_linear);
      varying_function = placeholder(varying_linear);
    }
    std::ostringstream src;
    src << "#define M " << m << "\n";
    if (flags & kOneSystemPerWave) src << "#define P 1\n";
    if (!(flags & kMethodOde23s)) src << "#define RODAS 1\n";
    if (differences) src << "#define JAC_DIFFERENCES 1\n#define COPIES " << copies_per_reactor(m) << "\n";
    src << rosenbrock_source;
    step_source = src.str();
  }
};

void compile_integrator(std::shared_ptr<const Mechanism> m, int flags) {
  IntegratorPrograms p(*m, flags);
  compile_straight(ProgramView{m, &p.rhs_function, &p.rhs_linear, ProgramKind::Other, nullptr}, rhs_options(p.rhs_linear));
  if (!p.differences) compile_form(ProgramView{m, &p.jacobian_function, &p.jacobian_linear, ProgramKind::Other, nullptr}, 1, kProgramBlock);
  if (p.differences) {
    compile_straight(ProgramView{m, &p.shared_function, &p.shared_linear, ProgramKind::Other, nullptr}, shared_options(p.shared_linear, p.m));
    compile_straight(ProgramView{m, &p.varying_function, &p.varying_linear, ProgramKind::Other, nullptr},
                     varying_options(p.varying_linear, p.m));
  }
  device::compile(p.step_source, "rosenbrock", hip_options(true));
}

Integrator::Integrator(std::shared_ptr<const Mechanism> m, double rtol, double atol_T, double atol_Y, int flags)
    : mech_(std::move(m)), m_((int)mech_->species.len() + 1), rtol_(rtol), atol_T_(atol_T), atol_Y_(atol_Y) {
  if (!(rtol > 0) || !(atol_T > 0) || !(atol_Y > 0)) throw ArgumentError("tolerances must be positive");
  IntegratorPrograms p(*mech_, flags);
  differences_ = p.differences;
  phases_ = (flags & kMethodOde23s) ? 3 : 6;
  // one reactor per wave (state size above 32): each launch has work enough for the host to keep four streams fed;
  // the packed waves' launches are short, and two parts keep the launches per step within the host's rate
  parts_count_ = m_ > 32 ? 4 : 2;
  systems_per_wave_ = (flags & kOneSystemPerWave) ? 1u : 64u / (unsigned)m_;
  rhs_function_ = std::move(p.rhs_function);
  rhs_linear_ = std::move(p.rhs_linear);
  rhs_ = assemble_straight(ProgramView{mech_, &rhs_function_, &rhs_linear_, ProgramKind::Other, nullptr}, rhs_options(rhs_linear_));
  if (!differences_) {
    jacobian_function_ = std::move(p.jacobian_function);
    jacobian_linear_ = std::move(p.jacobian_linear);
    jacobian_ = assemble_form(ProgramView{mech_, &jacobian_function_, &jacobian_linear_, ProgramKind::Other, nullptr}, 1, kProgramBlock);
  } else {
    shared_function_ = std::move(p.shared_function);
    shared_linear_ = std::move(p.shared_linear);
    varying_function_ = std::move(p.varying_function);
    varying_linear_ = std::move(p.varying_linear);
    shared_ = assemble_straight(ProgramView{mech_, &shared_function_, &shared_linear_, ProgramKind::Other, nullptr},
                                shared_options(shared_linear_, m_));
    varying_ = assemble_straight(ProgramView{mech_, &varying_function_, &varying_linear_, ProgramKind::Other, nullptr},
                                 varying_options(varying_linear_, m_));
  }
  step_.load(device::compile(p.step_source, "rosenbrock", hip_options(true)), "rosenbrock");
}

Integrator::~Integrator() {
  release();
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
  for (hipEvent_t e : polled_)
    if (e) (void)hipEventDestroy(e);
  if (host_count_) (void)hipHostFree(host_count_);
  for (hipStream_t s : sides_) (void)hipStreamDestroy(s);
}

void Integrator::release() {
  for (void* p : buffers_) (void)hipFree(p);
  buffers_.clear();
  parts_.clear();
  n_ = 0;
  last_y_ = nullptr;
  last_P_ = nullptr;
  last_status_ = nullptr;
}

void Integrator::prepare(double* y, const double* P, const int* status, size_t n) {
  if (n == n_ && y == last_y_ && P == last_P_ && status == last_status_) return;
  device::check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  release();
  if (sides_.empty()) {
    sides_.resize(parts_count_ - 1);
    for (hipStream_t& s : sides_) device::check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    device::check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreateWithFlags");
    device::check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreateWithFlags");
    for (hipEvent_t& e : polled_) device::check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
    device::check(hipHostMalloc((void**)&host_count_, 2 * sizeof(unsigned), hipHostMallocDefault), "hipHostMalloc");
  }
  const size_t m = (size_t)m_;
  auto alloc = [&](size_t bytes) {
    void* p = nullptr;
    device::check(hipMalloc(&p, bytes), "hipMalloc integrator");
    buffers_.push_back(p);
    return p;
  };
  jac_ = (double*)alloc((differences_ ? m : m + m * m) * n * sizeof(double));  // f(y), then J unless by differences
  F_ = (double*)alloc(m * n * sizeof(double));
  ys_ = (double*)alloc(m * n * sizeof(double));
  kst_ = (double*)alloc((phases_ == 6 ? 5 : 3) * m * n * sizeof(double));  // RODAS4: k1..k5; ode23s: k1, k2, F1
  lu_ = (double*)alloc(m * m * n * sizeof(double));
  rd_ = (double*)alloc(m * n * sizeof(double));
  perm_ = (int*)alloc(m * n * sizeof(int));
  running_ = (unsigned*)alloc(256);  // the running counts of the last two polls
  const size_t B = differences_ ? shared_linear_.output.size() : 0, C = copies_per_reactor(m_), nr = n * C;
  double *rep_out = nullptr, *t_out = nullptr;
  if (differences_) {
    // the reactors' values for their copies (group_stride per reactor); the right-hand sides of the C copies of
    // every state (C the power of two >= K + 1; M rows of C n doubles); the temperature-raised states' outputs
    groups_ = (double*)alloc(group_stride(B, m_) * n * sizeof(double));
    rep_out = (double*)alloc(m * nr * sizeof(double));
    t_out = (double*)alloc(m * n * sizeof(double));
    raised_T_ = (double*)alloc(n * sizeof(double));
  }
  // the batch in parts_count_ parts, each on its own stream, each starting at a multiple of 8 reactors (64-byte rows)
  const int parts = parts_count_;
  std::vector<size_t> bounds;
  for (int k = 0; k <= parts; k++) bounds.push_back(k == parts ? n : std::min(n, (n * k / parts + 7) & ~(size_t)7));
  auto table = [&](const std::vector<const void*>& ptrs) {
    void** t = (void**)alloc(ptrs.size() * sizeof(void*));
    device::check(hipMemcpy(t, ptrs.data(), ptrs.size() * sizeof(void*), hipMemcpyHostToDevice), "hipMemcpy table");
    return t;
  };
  for (int k = 0; k < parts; k++) {
    const size_t off = bounds[k];
    Part q;
    q.off = off;
    q.n = bounds[k + 1] - off;
    if (q.n == 0) continue;
    // pointer tables of the programs: inputs [T, P, Y_0..Y_{K-1}] from y or ys; outputs F rows or jac rows; all
    // starting at the part's first reactor
    auto inputs = [&](const double* base) {
      std::vector<const void*> p = {base + off, P + off};
      for (size_t i = 1; i < m; i++) p.push_back(base + i * n + off);
      return p;
    };
    auto rows = [&](const double* base, size_t count, size_t ld, size_t first_value) {
      std::vector<const void*> p;
      for (size_t i = 0; i < count; i++) p.push_back(base + i * ld + first_value);
      return p;
    };
    q.tab_y = table(inputs(y));
    q.tab_ys = table(inputs(ys_));
    if (!differences_) q.tab_jac = table(rows(jac_, m + m * m, n, off));
    q.tab_F = table(rows(F_, m, n, off));
    if (differences_) {
      std::vector<const void*> shared_out;
      for (size_t b = 0; b < B; b++) shared_out.push_back(groups_ + off * group_stride(B, m_) + b);
      q.tab_shared_in = table({y + off, P + off});
      q.tab_shared_out = table(shared_out);
      q.tab_rep_in = table({groups_ + off * group_stride(B, m_), status + off});  // values, and the skip flags
      q.tab_rep_out = table(rows(rep_out, m, nr, off * C));
      std::vector<const void*> t_in = inputs(y);
      t_in[0] = raised_T_ + off;
      q.tab_t_in = table(t_in);
      q.tab_t_out = table(rows(t_out, m, n, off));
    }
    parts_.push_back(q);
  }
  n_ = n;
  last_y_ = y;
  last_P_ = P;
  last_status_ = status;
}

// The caller's stream waits for the work enqueued so far on the other parts' streams (one event, recorded again for
// each: a wait takes the event's last record at the time it is enqueued).
void Integrator::join_sides(hipStream_t stream) {
  for (hipStream_t s : sides_) {
    device::check(hipEventRecord(join_, s), "hipEventRecord");
    device::check(hipStreamWaitEvent(stream, join_, 0), "hipStreamWaitEvent");
  }
}

long Integrator::advance(double* y, const double* P, double* t, double* h, int* status, unsigned* steps, size_t n, double t_end,
                         int max_steps, hipStream_t stream) {
  if (n == 0) return 0;
  if (n >= (size_t(1) << 31)) throw ArgumentError("at most 2^31 - 1 reactors");
  prepare(y, P, status, n);
  hipFunction_t running = nullptr, clear = nullptr, copies_kernel = nullptr, stage = nullptr;
  device::check(hipModuleGetFunction(&stage, step_.module, "rosenbrock_stage"), "hipModuleGetFunction");
  device::check(hipModuleGetFunction(&running, step_.module, "rosenbrock_running"), "hipModuleGetFunction");
  device::check(hipModuleGetFunction(&clear, step_.module, "rosenbrock_clear"), "hipModuleGetFunction");
  if (differences_) device::check(hipModuleGetFunction(&copies_kernel, step_.module, "rosenbrock_copies"), "hipModuleGetFunction");
  unsigned n32 = (unsigned)n;
  size_t ld = n;
  const unsigned systems_per_wave = systems_per_wave_;
  double te = t_end, rtol = rtol_, atol0 = atol_T_, atol = atol_Y_;
  const size_t m = (size_t)m_, C = copies_per_reactor(m_);
  unsigned stride = (unsigned)group_stride(shared_linear_.output.size(), m_), shared_values = (unsigned)shared_linear_.output.size();
  // the other parts' streams start after the work already on the caller's stream
  device::check(hipEventRecord(fork_, stream), "hipEventRecord");
  for (hipStream_t s : sides_) device::check(hipStreamWaitEvent(s, fork_, 0), "hipStreamWaitEvent");
  long taken = 0, polled_steps[2] = {0, 0};
  int polls = 0;
  for (int step = 0; step < max_steps; step++) {
    for (size_t p = 0; p < parts_.size(); p++) {
      const Part& q = parts_[p];
      hipStream_t st = p == 0 ? stream : sides_[p - 1];
      const size_t off = q.off, count = q.n;
      unsigned c32 = (unsigned)count;
      if (differences_) {
        double *yq = y + off, *groups = groups_ + off * stride, *raised = raised_T_ + off;
        const double* Pq = P + off;
        void* rargs[] = {&c32, &ld, &yq, &Pq, &groups, &stride, &shared_values, &raised, &rtol, &atol0, &atol};
        const size_t fields = count * 2 * m;
        device::check(hipModuleLaunchKernel(copies_kernel, (unsigned)((fields + 255) / 256), 1, 1, 256, 1, 1, 0, st, rargs, nullptr),
                      "launch rosenbrock_copies");
        shared_->launch_only((const double* const*)q.tab_shared_in, (double* const*)q.tab_shared_out, count, st, nullptr);
        varying_->launch_only((const double* const*)q.tab_rep_in, (double* const*)q.tab_rep_out, count * C, st, nullptr);
        rhs_->launch_only((const double* const*)q.tab_t_in, (double* const*)q.tab_t_out, count, st, nullptr);
      } else {
        jacobian_->launch_only((const double* const*)q.tab_y, (double* const*)q.tab_jac, count, st, nullptr);
      }
      double *jac = jac_ + off, *F = F_ + off, *yq = y + off, *ys = ys_ + off, *kst = kst_ + off, *lu = lu_ + off * m * m,
             *rd = rd_ + off * m, *tq = t + off, *hq = h + off;
      int *perm = perm_ + off * m, *statq = status + off;
      unsigned* stepq = steps + off;
      void **rep_f = q.tab_rep_out, **rep_t = q.tab_t_out;
      for (int phase = 1; phase <= phases_; phase++) {
        void* args[] = {&phase, &c32, &ld, &jac, &rep_f, &rep_t, &F, &yq, &ys, &kst,
                        &lu, &rd, &perm, &tq, &hq, &statq, &stepq, &te, &rtol, &atol0, &atol};
        device::check(hipModuleLaunchKernel(phase == 1 ? step_.function : stage, (c32 + systems_per_wave - 1) / systems_per_wave, 1, 1, 64,
                                            1, 1, 0, st, args, nullptr),
                      "launch rosenbrock");
        if (phase < phases_) rhs_->launch_only((const double* const*)q.tab_ys, (double* const*)q.tab_F, count, st, nullptr);
      }
    }
    taken++;
    if ((step + 1) % kPollEvery == 0 || step + 1 == max_steps) {
      // the count waits for every part's steps so far; it goes to pinned memory behind an event, and the host reads
      // the previous poll's count instead, which the GPU has usually passed: the GPU never idles for the round trip.
      // Reactors that are not running are left alone, so the steps launched after a poll that found none running
      // change nothing, and the call reports the steps up to that poll.
      const int slot = polls & 1;
      join_sides(stream);
      unsigned* count = running_ + slot;
      void* cargs[] = {&count};
      device::check(hipModuleLaunchKernel(clear, 1, 1, 1, 64, 1, 1, 0, stream, cargs, nullptr), "launch rosenbrock_clear");
      void* rargs[] = {&n32, &status, &count};
      device::check(hipModuleLaunchKernel(running, (n32 + 255) / 256, 1, 1, 256, 1, 1, 0, stream, rargs, nullptr), "launch rosenbrock_running");
      // the other parts go on after the count only, so that it sees every part at this step, not later ones
      device::check(hipEventRecord(fork_, stream), "hipEventRecord");
      for (hipStream_t s : sides_) device::check(hipStreamWaitEvent(s, fork_, 0), "hipStreamWaitEvent");
      device::check(hipMemcpyAsync(host_count_ + slot, count, sizeof(unsigned), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
      device::check(hipEventRecord(polled_[slot], stream), "hipEventRecord");
      polled_steps[slot] = taken;
      if (++polls >= 2) {
        device::check(hipEventSynchronize(polled_[slot ^ 1]), "hipEventSynchronize");
        if (host_count_[slot ^ 1] == 0) {
          taken = polled_steps[slot ^ 1];
          break;
        }
      }
    }
  }
  // every step launched is complete on return
  join_sides(stream);
  device::check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return taken;
}

}  // namespace combustion
