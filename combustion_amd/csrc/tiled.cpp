// Mechanism-tiled forms. The mixed-precision form launches two kernels on the caller's stream:
//   rates_mixed (kernels/rates_mixed.hip): reaction tiles on BF16 MFMA, mechanism tables in device memory;
//   the transport section of the program emitted as straight-line FP32 code.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "device.hpp"
#include "emit.hpp"
#include "forms.hpp"
#include "soa.hpp"

namespace combustion {

namespace {

uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Tables of kernels/rates_mixed.hip (layouts described there).
struct MixedTables {
  int K, KA, KB, AB, RT, KE;
  std::vector<uint16_t> A_fwd, A_rev, A_nu;
  std::vector<double> rate, k0;
  std::vector<float> troe, eff;
  std::vector<int32_t> slots, tile_info;
  std::vector<double> rcp_W, gibbs, split;
  double ln_clamp_pos, ln_clamp_neg;
};

int model_rank(ReactionModel m) {
  switch (m) {
    case ReactionModel::Elementary:
    case ReactionModel::Irreversible: return 0;
    case ReactionModel::ThreeBody: return 1;
    case ReactionModel::PressureModification: return 2;
    case ReactionModel::Troe: return 3;
    case ReactionModel::SRI: throw Unsupported("SRI falloff is not implemented (reaction.rs:194-203)");
  }
  return 0;
}

MixedTables mixed_tables(const Mechanism& m) {
  MixedTables t;
  const Species& sp = m.species;
  t.K = (int)sp.len();
  t.KA = (int)m.active;
  const int R = (int)m.reactions.size();
  t.KB = (t.K + 15) / 16;
  t.AB = (t.KA + 15) / 16;
  t.RT = (R + 15) / 16;
  // Reactions sorted by model so that most 16-reaction tiles hold one model.
  std::vector<int> order(R);
  for (int r = 0; r < R; r++) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return model_rank(m.reactions[a].model) < model_rank(m.reactions[b].model); });
  const int RP = t.RT * 16;
  t.A_fwd.assign((size_t)t.RT * t.KB * 64 * 4, 0);
  t.A_rev.assign((size_t)t.RT * t.KB * 64 * 4, 0);
  t.A_nu.assign((size_t)t.RT * t.AB * 64 * 4, 0);
  // A operand of v_mfma_f32_16x16x16_bf16: lane L holds A[row L%16][col 4 (L/16) + i].
  for (int rt = 0; rt < t.RT; rt++)
    for (int L = 0; L < 64; L++)
      for (int i = 0; i < 4; i++) {
        const int row = L % 16, col = 4 * (L / 16) + i;
        for (int kb = 0; kb < t.KB; kb++) {
          const int r = rt * 16 + row, k = kb * 16 + col;
          if (r < R && k < t.K) {
            const Reaction& x = m.reactions[order[r]];
            const size_t at = (((size_t)rt * t.KB + kb) * 64 + L) * 4 + i;
            t.A_fwd[at] = bf16_bits((float)x.reactants[k]);
            t.A_rev[at] = bf16_bits((float)x.products[k]);
          }
        }
        for (int a = 0; a < t.AB; a++) {
          const int k = a * 16 + row, r = rt * 16 + col;
          if (r < R && k < t.KA)
            t.A_nu[(((size_t)rt * t.AB + a) * 64 + L) * 4 + i] = bf16_bits((float)m.reactions[order[r]].net[k]);
        }
      }
  t.KE = (t.K + 3) / 4;
  t.rate.assign(4 * RP, 0.);
  t.k0.assign(4 * RP, 0.);
  t.troe.assign(4 * RP, 0.f);
  const int zero_row = t.KA * 16 * 8;  // byte offset of the zero row of the Gibbs table (rows of 16 doubles)
  t.slots.assign(8 * RP, zero_row);
  for (int r = 0; r < RP; r++) {
    t.slots[8 * r] = 5 | (8 << 8);
    t.slots[8 * r + 7] = 0;
  }
  t.tile_info.assign(t.RT, 0);
  t.eff.assign((size_t)t.RT * t.KE * 64, 0.f);
  for (int r = 0; r < R; r++) {
    const Reaction& x = m.reactions[order[r]];
    const RateConstant& k = x.rate_constant;
    int type = 0;
    switch (x.model) {
      case ReactionModel::Elementary: type = 0; break;
      case ReactionModel::Irreversible: type = 1; break;
      case ReactionModel::ThreeBody: type = 2; break;
      case ReactionModel::PressureModification: type = 3; break;
      case ReactionModel::Troe: type = 4; break;
      default: throw Unsupported("reaction model not supported by the mixed form");
    }
    t.rate[4 * r] = std::log(k.preexponential_factor);
    t.rate[4 * r + 1] = k.temperature_exponent;
    t.rate[4 * r + 2] = k.activation_temperature;
    int kinf_const = 0;
    if (type >= 3) {
      // Troe factors A∞ out of k0 (reaction.rs:163-170)
      t.k0[4 * r] = std::log(type == 4 ? x.k0.preexponential_factor / k.preexponential_factor : x.k0.preexponential_factor);
      t.k0[4 * r + 1] = x.k0.temperature_exponent;
      t.k0[4 * r + 2] = x.k0.activation_temperature;
    }
    if (type == 4) {
      kinf_const = k.temperature_exponent == 0. && k.activation_temperature == 0.;
      t.troe[4 * r] = (float)x.troe.A;
      t.troe[4 * r + 1] = (float)(1. / x.troe.T3);
      t.troe[4 * r + 2] = x.troe.T1 < 1e30 ? (float)(1. / x.troe.T1) : 0.f;
      t.troe[4 * r + 3] = std::isfinite(x.troe.T2) ? (float)x.troe.T2 : 0.f;
    }
    if (-x.sum_net < -8 || -x.sum_net > 7) throw Unsupported("change in moles out of range: " + x.equation);
    t.slots[8 * r] = type | (kinf_const << 4) | ((8 - x.sum_net) << 8);
    int plus = 0, minus = 0;
    if (type != 1)
      for (size_t s = 0; s < x.net.size(); s++) {
        const int nu = x.net[s];
        if (!nu) continue;
        if (std::abs(nu) > 2) throw Unsupported("|ν| > 2 in reaction " + x.equation);
        const int slot = (int)s * 16 * 8 | (std::abs(nu) == 2);
        if (nu > 0) {
          if (plus == 3) throw Unsupported("more than 3 products change in reaction " + x.equation);
          t.slots[8 * r + 1 + plus++] = slot;
        } else {
          if (minus == 3) throw Unsupported("more than 3 reactants change in reaction " + x.equation);
          t.slots[8 * r + 4 + minus++] = slot;
        }
      }
    const int rt = r / 16;
    if (type >= 2) t.tile_info[rt] |= 1;
    if (type >= 3) t.tile_info[rt] |= 2;
    if (type == 4) t.tile_info[rt] |= 4;
    // A operand of v_mfma_f32_16x16x4_f32: lane L holds A[row L%16][col L/16] = e - 1 of species 4 kk + L/16.
    if (type >= 2)
      for (int kk = 0; kk < t.KE; kk++)
        for (int L = 0; L < 64; L++) {
          const int row = L % 16, s = kk * 4 + L / 16;
          if (rt * 16 + row == r && s < t.K) t.eff[((size_t)rt * t.KE + kk) * 64 + L] = (float)(x.efficiencies[s] - 1.);
        }
  }
  for (int k = 0; k < t.K; k++) t.rcp_W.push_back(1. / sp.molar_mass[k]);
  for (int k = 0; k < t.KA; k++) {
    const NASA7& n = sp.thermodynamics[k];
    for (int p = 0; p < 2; p++) {
      const auto& a = n.pieces[p];
      const double g[7] = {a[0] - a[6], -a[0], -a[1] / 2., (1. / 3. - 1. / 2.) * a[2], (1. / 4. - 1. / 3.) * a[3], (1. / 5. - 1. / 4.) * a[4], a[5]};
      t.gibbs.insert(t.gibbs.end(), g, g + 7);
    }
    t.split.push_back(std::isnan(n.temperature_split) ? INFINITY : n.temperature_split);
  }
  t.ln_clamp_pos = std::log(6981463048192.);  // reaction.rs:244-256
  t.ln_clamp_neg = std::log((double)cbrtf(3.40282347e38f));
  return t;
}

template <class T>
T* upload(const std::vector<T>& v) {
  T* d = nullptr;
  device::check(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)), "hipMalloc");
  if (!v.empty()) device::check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

std::string rates_mixed_text(const MixedTables& t) {
  std::ostringstream s;
  s << "#define K " << t.K << "\n#define KA " << t.KA << "\n#define KB " << t.KB << "\n#define AB " << t.AB << "\n#define RT " << t.RT
    << "\n#define KE " << t.KE << "\n"
    << rates_mixed_source;
  return s.str();
}

ast::Linear transport_program(const ProgramView& p) {
  ChemistryOptions o;
  o.rates = false;
  o.transport = true;
  return ast::linearize(chemistry(*p.mechanism, p.chemistry->polynomials.get(), p.chemistry->scales, o));
}

StraightOptions f32_options(unsigned block_size) {
  StraightOptions o;
  o.block_size = block_size ? block_size : 256;
  o.f32 = true;
  return o;
}

void check_program(const ProgramView& p) {
  if (p.kind != ProgramKind::Chemistry) throw Unsupported("tiled forms need a program built by cb_build_chemistry");
  if (p.chemistry->options.rhs) throw Unsupported("tiled forms do not evaluate the reactor right-hand side");
}

class Mixed : public SoAForm {
 public:
  Mixed(const ProgramView& p, unsigned block_size) : SoAForm(p.function->n_input, (int)p.function->output.size()) {
    check_program(p);
    const auto& o = p.chemistry->options;
    if (o.rates) {
      t_ = mixed_tables(*p.mechanism);
      rates_.load(device::compile(rates_mixed_text(t_), "rates_mixed", hip_options(true)), "rates_mixed");
      d_fwd_ = upload(t_.A_fwd);
      d_rev_ = upload(t_.A_rev);
      d_nu_ = upload(t_.A_nu);
      d_rate_ = upload(t_.rate);
      d_k0_ = upload(t_.k0);
      d_troe_ = upload(t_.troe);
      d_info_ = upload(t_.slots);
      d_tile_ = upload(t_.tile_info);
      d_eff_ = upload(t_.eff);
      d_rcp_W_ = upload(t_.rcp_W);
      d_gibbs_ = upload(t_.gibbs);
      d_split_ = upload(t_.split);
      n_rates_ = t_.KA;
    }
    if (o.transport) {
      const StraightOptions so = f32_options(block_size);
      transport_block_ = so.block_size;
      transport_.load(device::compile(emit_straight(transport_program(p), "transport_f32", so), "transport_f32", hip_options(true)), "transport_f32");
      has_transport_ = true;
    }
  }
  ~Mixed() override {
    for (void* ptr : {(void*)d_fwd_, (void*)d_rev_, (void*)d_nu_, (void*)d_rate_, (void*)d_k0_, (void*)d_troe_, (void*)d_info_,
                      (void*)d_tile_, (void*)d_eff_, (void*)d_rcp_W_, (void*)d_gibbs_, (void*)d_split_})
      if (ptr) (void)hipFree(ptr);
  }

 protected:
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) override {
    unsigned n32 = (unsigned)n;
    if (n_rates_) {
      const unsigned states_per_block = 16 * 4;
      unsigned grid = (unsigned)((n + states_per_block - 1) / states_per_block);
      double pos = t_.ln_clamp_pos, neg = t_.ln_clamp_neg;
      void* args[] = {&in, &out, &n32, &d_fwd_, &d_rev_, &d_nu_, &d_rate_, &d_k0_, &d_troe_, &d_info_, &d_tile_, &d_eff_,
                      &d_rcp_W_, &d_gibbs_, &d_split_, &pos, &neg};
      device::check(hipModuleLaunchKernel(rates_.function, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr), "launch rates_mixed");
    }
    if (has_transport_) {
      double* const* tout = out + n_rates_;
      unsigned offset = 0;
      void* args[] = {&in, &tout, &n32, &offset};
      unsigned grid = (unsigned)((n + transport_block_ - 1) / transport_block_);
      device::check(hipModuleLaunchKernel(transport_.function, grid, 1, 1, transport_block_, 1, 1, 0, stream, args, nullptr), "launch transport_f32");
    }
  }

 private:
  MixedTables t_;
  device::Kernel rates_, transport_;
  int n_rates_ = 0;
  bool has_transport_ = false;
  unsigned transport_block_ = 256;
  uint16_t *d_fwd_ = nullptr, *d_rev_ = nullptr, *d_nu_ = nullptr;
  double *d_rate_ = nullptr, *d_k0_ = nullptr;
  float *d_troe_ = nullptr, *d_eff_ = nullptr;
  int32_t *d_info_ = nullptr, *d_tile_ = nullptr;
  double *d_rcp_W_ = nullptr, *d_gibbs_ = nullptr, *d_split_ = nullptr;
};

}  // namespace

std::string tiled_source(const ProgramView& p, bool mixed, unsigned block_size) {
  check_program(p);
  if (!mixed) throw Unsupported("the tiled FP64 form is not built yet");
  std::string s;
  if (p.chemistry->options.rates) s += rates_mixed_text(mixed_tables(*p.mechanism));
  if (p.chemistry->options.transport) s += "\n// ---- transport (FP32 straight-line) ----\n" + emit_straight(transport_program(p), "transport_f32", f32_options(block_size));
  return s;
}

void compile_tiled(const ProgramView& p, bool mixed, unsigned block_size) {
  check_program(p);
  if (!mixed) throw Unsupported("the tiled FP64 form is not built yet");
  if (p.chemistry->options.rates) device::compile(rates_mixed_text(mixed_tables(*p.mechanism)), "rates_mixed", hip_options(true));
  if (p.chemistry->options.transport)
    device::compile(emit_straight(transport_program(p), "transport_f32", f32_options(block_size)), "transport_f32", hip_options(true));
}

std::unique_ptr<Form> assemble_tiled(const ProgramView& p, bool mixed, unsigned block_size) {
  if (!mixed) throw Unsupported("the tiled FP64 form is not built yet");
  return std::make_unique<Mixed>(p, block_size);
}

}  // namespace combustion
