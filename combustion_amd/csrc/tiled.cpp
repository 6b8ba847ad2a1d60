// Mechanism-tiled forms, each launching its kernels on the caller's stream:
//   fp64  (chemistry programs): the rates section as straight-line FP64 code, then transport_tiled in FP64;
//   mixed: rates_mixed (reaction tiles on BF16 MFMA, mechanism tables in device memory), then transport_tiled
//          in FP32.
// transport_tiled (kernels/transport_tiled.hip) evaluates the Wilke sums as three K×K MFMA contractions and the
// binary-diffusion polynomials as (pairs × D)·(D × states) MFMA tiles.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <sstream>

#include "device.hpp"
#include "emit.hpp"
#include "forms.hpp"
#include "mlp.hpp"
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
  int K, KA, KB, AB, RT, KE, WAVES;
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
    case ReactionModel::SRI: throw Unsupported("the MFMA rates kernel has no SRI falloff model (the other forms do)");
  }
  return 0;
}

constexpr double kLog2e = 1.4426950408889634;

MixedTables mixed_tables(const Mechanism& m) {
  MixedTables t;
  const Species& sp = m.species;
  t.K = (int)sp.len();
  t.KA = (int)m.active;
  const int R = (int)m.reactions.size();
  t.KB = (t.K + 31) / 32;
  t.AB = (t.KA + 15) / 16;
  t.KE = (t.K + 3) / 4;
  // Rows: the reactions sorted by model, each model group padded to whole 16-row halves so that every half holds one
  // model (the kernel then branches on the half, not on the row); then padded to whole 32-row tiles.
  std::vector<int> order(R);
  for (int r = 0; r < R; r++) order[r] = r;
  for (const Reaction& r : m.reactions)
    if (!r.plog.empty()) throw Unsupported("the MFMA rates kernel has no pressure-dependent Arrhenius tables (the other forms do)");
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return model_rank(m.reactions[a].model) < model_rank(m.reactions[b].model); });
  std::vector<int> rows;  // padded row -> reaction (or -1)
  for (size_t i = 0; i < order.size();) {
    const int rank = model_rank(m.reactions[order[i]].model);
    while (i < order.size() && model_rank(m.reactions[order[i]].model) == rank) rows.push_back(order[i++]);
    while (rows.size() % 16) rows.push_back(-1);
  }
  while (rows.size() % 32 || rows.empty()) rows.push_back(-1);
  t.RT = (int)rows.size() / 32;
  const int HT = 2 * t.RT, RP = 32 * t.RT;
  // A operands of v_mfma_f32_16x16x32_bf16: lane L, element j is A[row L%16][k = 8 (L/16) + j].
  t.A_fwd.assign((size_t)HT * t.KB * 64 * 8, 0);
  t.A_rev.assign((size_t)HT * t.KB * 64 * 8, 0);
  for (int ht = 0; ht < HT; ht++)
    for (int kb = 0; kb < t.KB; kb++)
      for (int L = 0; L < 64; L++)
        for (int j = 0; j < 8; j++) {
          const int r = 16 * ht + L % 16, k = 32 * kb + 8 * (L / 16) + j;
          if (rows[r] < 0 || k >= t.K) continue;
          const Reaction& x = m.reactions[rows[r]];
          const size_t at = (((size_t)ht * t.KB + kb) * 64 + L) * 8 + j;
          t.A_fwd[at] = bf16_bits((float)x.reactants[k]);
          t.A_rev[at] = bf16_bits((float)x.products[k]);
        }
  // ν·q: A[species row][reaction k] with k of (L/16, j) = j < 4 ? 4 (L/16) + j : 16 + 4 (L/16) + j - 4, the rows the
  // lane's two 16-reaction halves produce.
  t.A_nu.assign((size_t)t.RT * t.AB * 64 * 8, 0);
  for (int rt = 0; rt < t.RT; rt++)
    for (int a = 0; a < t.AB; a++)
      for (int L = 0; L < 64; L++)
        for (int j = 0; j < 8; j++) {
          const int g = L / 16, k = 16 * a + L % 16;
          const int r = 32 * rt + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
          if (rows[r] >= 0 && k < t.KA)
            t.A_nu[(((size_t)rt * t.AB + a) * 64 + L) * 8 + j] = bf16_bits((float)m.reactions[rows[r]].net[k]);
        }
  t.rate.assign(4 * RP, 0.);
  t.k0.assign(4 * RP, 0.);
  t.troe.assign(4 * RP, 0.f);
  // G tables: row k of the plus table at byte k·128, of the minus table at (KA+1+k)·128; row KA (zero) pads.
  const int row_bytes = 16 * 8, zero = t.KA * row_bytes;
  const int minus0 = (t.KA + 1) * row_bytes;
  if (2 * (t.KA + 1) * row_bytes > 65535) throw Unsupported("too many species for the mixed kernel's 16-bit slots");
  t.slots.assign(4 * RP, 0);
  for (int r = 0; r < RP; r++) {
    t.slots[4 * r] = 5;
    for (int w = 1; w < 4; w++) t.slots[4 * r + w] = zero | (zero << 16);
  }
  t.tile_info.assign(HT, 5);  // 5: a half of padding rows only
  t.eff.assign((size_t)HT * t.KE * 64, 0.f);
  for (int r = 0; r < RP; r++) {
    if (rows[r] < 0) continue;
    const Reaction& x = m.reactions[rows[r]];
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
    t.rate[4 * r] = std::log2(k.preexponential_factor);
    t.rate[4 * r + 1] = k.temperature_exponent;
    t.rate[4 * r + 2] = k.activation_temperature * kLog2e;
    t.rate[4 * r + 3] = -x.sum_net;
    int kinf_const = 0;
    if (type >= 3) {
      // Troe factors A∞ out of k0 (reaction.rs:163-170)
      t.k0[4 * r] = std::log2(type == 4 ? x.k0.preexponential_factor / k.preexponential_factor : x.k0.preexponential_factor);
      t.k0[4 * r + 1] = x.k0.temperature_exponent;
      t.k0[4 * r + 2] = x.k0.activation_temperature * kLog2e;
    }
    if (type == 4) {
      kinf_const = k.temperature_exponent == 0. && k.activation_temperature == 0.;
      t.troe[4 * r] = (float)x.troe.A;
      t.troe[4 * r + 1] = (float)(1. / x.troe.T3);
      t.troe[4 * r + 2] = x.troe.T1 < 1e30 ? (float)(1. / x.troe.T1) : 0.f;
      t.troe[4 * r + 3] = std::isfinite(x.troe.T2) ? (float)x.troe.T2 : NAN;  // NaN: no T2 term (a finite T2, 0 included, adds exp(-T2/T))
    }
    t.slots[4 * r] = type | (kinf_const << 4);
    std::vector<int> offsets;
    if (type != 1)
      for (size_t s = 0; s < x.net.size(); s++) {
        const int nu = x.net[s];
        for (int c = 0; c < std::abs(nu); c++) offsets.push_back(nu > 0 ? (int)s * row_bytes : minus0 + (int)s * row_bytes);
      }
    if (offsets.size() > 6) throw Unsupported("more than 6 stoichiometric changes in reaction " + x.equation);
    while (offsets.size() < 6) offsets.push_back(zero);
    for (int w = 0; w < 3; w++) t.slots[4 * r + 1 + w] = offsets[2 * w] | (offsets[2 * w + 1] << 16);
    // the half's model: elementary and irreversible rows share halves (type 0), the others have halves of their own
    const int ht = r / 16;
    t.tile_info[ht] = type <= 1 ? 0 : type;
    // A operand of v_mfma_f32_16x16x4_f32: lane L holds A[row L%16][col L/16] = e - 1 of species 4 kk + L/16.
    if (type >= 2)
      for (int kk = 0; kk < t.KE; kk++)
        for (int L = 0; L < 64; L++) {
          const int row = L % 16, s = kk * 4 + L / 16;
          if (ht * 16 + row == r && s < t.K) t.eff[((size_t)ht * t.KE + kk) * 64 + L] = (float)(x.efficiencies[s] - 1.);
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
  // LDS per wave (struct Lds of rates_mixed.hip): as many waves per block (up to 4) as fit in 160 KiB.
  const size_t staging = (size_t)t.KB * 3 * 64 * 8 * 2, gtables = (size_t)2 * (t.KA + 1) * 16 * 8;
  const size_t per_wave = std::max(staging, gtables) + (size_t)t.KE * 4 * 16 * 4 + 16 * (3 * 8 + 2 * 4);
  t.WAVES = (int)std::min<size_t>(4, 163840 / per_wave);
  if (t.WAVES < 1) throw Unsupported("mechanism too large for the mixed rates kernel's LDS");
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
    << "\n#define KE " << t.KE << "\n#define WAVES " << t.WAVES << "\n"
    << rates_mixed_source;
  return s.str();
}

// Tables of kernels/rates_tiled.hip (layouts described there): rows sorted by model, each model's rows padded to whole
// rounds of G = 64/NS so that a round is of one model.
struct TiledRatesTables {
  int K, KA, ROUNDS, NS;
  std::vector<double> rate, k0, troe, eff_value, rcp_W, nasa, split;
  std::vector<int32_t> info, ops, eff_species;
};

TiledRatesTables rates_tiled_tables(const Mechanism& m) {
  TiledRatesTables t;
  const Species& sp = m.species;
  t.K = (int)sp.len();
  t.KA = (int)m.active;
  // states per wave: the largest power of two up to 8 whose LDS rows (C, E+, E-, ω and six more) fit 11 KiB, for
  // about 15 waves per CU
  const size_t rows_lds = (size_t)t.K + 3 * (size_t)t.KA + 6;
  const size_t per_state = rows_lds * 8;
  t.NS = 8;
  while (t.NS > 1 && t.NS * per_state > 11264) t.NS /= 2;
  if (t.NS * per_state > 64 * 1024) throw Unsupported("mechanism too large for the tiled FP64 rates kernel's LDS");
  const int R = (int)m.reactions.size(), G = 64 / t.NS;
  auto model_code = [](const Reaction& x) {
    if (!x.plog.empty()) throw Unsupported("the tiled FP64 rates kernel has no pressure-dependent Arrhenius tables (the straight-line forms do)");
    switch (x.model) {
      case ReactionModel::Elementary: return 0;
      case ReactionModel::Irreversible: return 1;
      case ReactionModel::ThreeBody: return 2;
      case ReactionModel::PressureModification: return 3;
      case ReactionModel::Troe: return 4;
      default: throw Unsupported("the tiled FP64 rates kernel has no SRI falloff model (the straight-line forms do)");
    }
  };
  // entries of a reaction's factor and target lists (see ops below): forward, reverse, targets
  auto entries = [&](const Reaction& x) {
    int nf = 0, nr = 1, na = 0;
    for (int sp = 0; sp < t.K; sp++) nf += x.reactants[sp];
    if (model_code(x) != 1) {
      for (size_t sp = 0; sp < x.net.size(); sp++) nr += std::abs((int)x.net[sp]);
      for (int sp = 0; sp < t.K; sp++) nr += x.products[sp];
    }
    for (size_t sp = 0; sp < x.net.size(); sp++) na += x.net[sp] != 0;
    return std::array<int, 3>{nf, nr, na};
  };
  // rows sorted by model, and within a model by their list lengths, so that a round's rows have similar counts
  std::vector<int> order(R);
  for (int r = 0; r < R; r++) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int ma = model_code(m.reactions[a]), mb = model_code(m.reactions[b]);
    if (ma != mb) return ma < mb;
    return entries(m.reactions[a]) < entries(m.reactions[b]);
  });
  std::vector<int> rows;
  for (size_t i = 0; i < order.size();) {
    const int code = model_code(m.reactions[order[i]]);
    while (i < order.size() && model_code(m.reactions[order[i]]) == code) rows.push_back(order[i++]);
    while (rows.size() % G) rows.push_back(-1);
  }
  if (rows.empty()) rows.assign(G, -1);
  t.ROUNDS = (int)rows.size() / G;
  const size_t RP = rows.size();
  t.rate.assign(4 * RP, 0.);
  t.k0.assign(4 * RP, 0.);
  t.troe.assign(4 * RP, 0.);
  t.info.assign(4 * RP, 0);
  // LDS rows of the kernel, as byte offsets of the row (a lane adds 8 × its state column)
  const int row_C = 0, row_Ep = t.K, row_En = t.K + t.KA, row_W = t.K + 2 * t.KA, row_spare = t.K + 3 * t.KA;
  const int row_one = row_spare + 1, row_zero = row_spare + 2, row_C0 = row_spare + 3;
  auto offset = [&](int lds_row) { return (uint32_t)(lds_row * t.NS * 8); };
  t.ops.assign(12 * RP, 0);
  auto put = [&](size_t r, int half, uint32_t value) {
    uint32_t& word = reinterpret_cast<uint32_t&>(t.ops[12 * r + half / 2]);
    const uint32_t shift = 16 * (half % 2);
    word = (word & ~(0xFFFFu << shift)) | (value << shift);
  };
  auto fill_padding = [&](size_t r) {
    for (int h = 0; h < 16; h++) put(r, h, offset(row_one));
    for (int h = 16; h < 22; h++) put(r, h, offset(row_spare));
    t.ops[12 * r + 11] = 0x888888;  // ν + 8 = 8 for the six targets
  };
  for (size_t r = 0; r < RP; r++) {
    fill_padding(r);
    if (rows[r] < 0) {
      t.info[4 * r] = 5;
      continue;
    }
    const Reaction& x = m.reactions[rows[r]];
    const int code = model_code(x);
    const RateConstant& k = x.rate_constant;
    if (code != 1 && (-x.sum_net < -2 || -x.sum_net > 1))
      throw Unsupported("unsupported change in the number of moles: " + x.equation);
    t.rate[4 * r] = std::log(k.preexponential_factor);
    t.rate[4 * r + 1] = k.temperature_exponent;
    t.rate[4 * r + 2] = k.activation_temperature;
    int kinf_const = 0;
    if (code >= 3) {
      // Troe factors A∞ out of k0 (reaction.rs:163-170); Lindemann takes both limits whole
      t.k0[4 * r] = std::log(code == 4 ? x.k0.preexponential_factor / k.preexponential_factor : x.k0.preexponential_factor);
      t.k0[4 * r + 1] = x.k0.temperature_exponent;
      t.k0[4 * r + 2] = x.k0.activation_temperature;
      t.k0[4 * r + 3] = k.preexponential_factor;
    }
    if (code == 4) {
      kinf_const = k.temperature_exponent == 0. && k.activation_temperature == 0.;
      t.troe[4 * r] = x.troe.A;
      t.troe[4 * r + 1] = x.troe.T3;
      t.troe[4 * r + 2] = x.troe.T1;
      t.troe[4 * r + 3] = std::isfinite(x.troe.T2) ? x.troe.T2 : NAN;
    }
    t.info[4 * r] = code | (kinf_const << 4);
    if (code >= 2) {
      t.info[4 * r + 1] = (int32_t)t.eff_species.size();
      for (int s = 0; s < t.K; s++)
        if (x.efficiencies[s] != 1.) {
          t.eff_species.push_back(s);
          t.eff_value.push_back(x.efficiencies[s] - 1.);
        }
      t.info[4 * r + 2] = (int32_t)t.eff_species.size() - t.info[4 * r + 1];
    }
    // forward factors: the reactants' concentrations with multiplicity (halves 0-3)
    int h = 0;
    for (int sp = 0; sp < t.K; sp++)
      for (int c = 0; c < x.reactants[sp]; c++) {
        if (h == 4) throw Unsupported("more than 4 reactants in reaction " + x.equation);
        put(r, h++, offset(row_C + sp));
      }
    // reverse factors (halves 4-15): E± per unit of net ν, the products' concentrations, (P0/RT)^(-Σν)
    std::vector<uint32_t> rev;
    if (code == 1) {
      rev.push_back(offset(row_zero));
    } else {
      for (size_t sp = 0; sp < x.net.size(); sp++)
        for (int c = 0; c < std::abs((int)x.net[sp]); c++) rev.push_back(offset((x.net[sp] > 0 ? row_Ep : row_En) + (int)sp));
      for (int sp = 0; sp < t.K; sp++)
        for (int c = 0; c < x.products[sp]; c++) rev.push_back(offset(row_C + sp));
      const int power = -x.sum_net;  // 1, 0, -1 or -2 (checked above)
      rev.push_back(offset(power == 1 ? row_C0 : power == 0 ? row_one : power == -1 ? row_C0 + 1 : row_C0 + 2));
    }
    if (rev.size() > 12) throw Unsupported("more than 12 reverse-rate factors in reaction " + x.equation);
    for (size_t i = 0; i < rev.size(); i++) put(r, 4 + (int)i, rev[i]);
    // targets of ν_k q (halves 16-21) and their ν + 8 (4 bits each, word 11)
    int n = 0;
    uint32_t nus = 0x888888;
    for (size_t sp = 0; sp < x.net.size(); sp++) {
      const int nu = x.net[sp];
      if (!nu) continue;
      if (n == 6 || std::abs(nu) > 7) throw Unsupported("too many or too large stoichiometric changes in reaction " + x.equation);
      put(r, 16 + n, offset(row_W + (int)sp));
      nus = (nus & ~(0xFu << (4 * n))) | ((uint32_t)(nu + 8) << (4 * n));
      n++;
    }
    t.ops[12 * r + 11] = (int32_t)nus;
  }
  if (t.eff_species.empty()) {  // keep the device arrays non-empty
    t.eff_species.push_back(0);
    t.eff_value.push_back(0.);
  }
  for (int k = 0; k < t.K; k++) t.rcp_W.push_back(1. / sp.molar_mass[k]);
  for (int k = 0; k < t.KA; k++) {
    const NASA7& nasa = sp.thermodynamics[k];
    for (int p = 0; p < 2; p++) {
      const auto& a = nasa.pieces[p];
      const double g[7] = {a[0] - a[6], -a[0], -a[1] / 2., (1. / 3. - 1. / 2.) * a[2], (1. / 4. - 1. / 3.) * a[3], (1. / 5. - 1. / 4.) * a[4], a[5]};
      t.nasa.insert(t.nasa.end(), g, g + 7);
    }
    t.split.push_back(std::isnan(nasa.temperature_split) ? INFINITY : nasa.temperature_split);
  }
  // each round's largest list lengths, in every row of the round (info word 3: forward | reverse << 4 | targets << 8),
  // so that the kernel skips the padding beyond them with wave-uniform branches
  for (int round = 0; round < t.ROUNDS; round++) {
    std::array<int, 3> most{0, 0, 0};
    for (int g = 0; g < G; g++) {
      const int row = rows[(size_t)round * G + g];
      if (row < 0) continue;
      const auto e = entries(m.reactions[row]);
      for (int i = 0; i < 3; i++) most[i] = std::max(most[i], e[i]);
    }
    for (int g = 0; g < G; g++) t.info[4 * ((size_t)round * G + g) + 3] = most[0] | most[1] << 4 | most[2] << 8;
  }
  if (rows_lds * t.NS * 8 > 65535) throw Unsupported("mechanism too large for the tiled FP64 rates kernel's LDS offsets");
  return t;
}

std::string rates_tiled_text(const TiledRatesTables& t) {
  std::ostringstream s;
  s << "#define K " << t.K << "\n#define KA " << t.KA << "\n#define ROUNDS " << t.ROUNDS << "\n#define NS " << t.NS << "\n"
    << rates_tiled_source;
  return s.str();
}

// Tables of kernels/transport_tiled.hip: unscaled fits (transport.rs:125-136), W^1/4, the Wilke matrix a² and the
// diffusion-fit tiles in the lane layout of v_mfma_{f64,f32}_16x16x4 (A[row L%16][col L/16]).
struct TransportTables {
  int K, KP, DK, NTILE;
  double T0;
  std::vector<double> rcp_W, W, fit_c, fit_v, wilke, dcoef;
};

TransportTables transport_tables(const Mechanism& m, const Polynomials& p, double T0) {
  TransportTables t;
  const Species& sp = m.species;
  t.K = (int)sp.len();
  t.KP = (t.K + 15) / 16 * 16;
  t.DK = (p.D + 3) / 4 * 4;
  if (t.DK > 8) throw Unsupported("transport fits longer than 8 coefficients");
  t.T0 = T0;
  const int K = t.K, KP = t.KP, D = p.D, DK = t.DK;
  for (int k = 0; k < K; k++) {
    t.W.push_back(sp.molar_mass[k]);
    t.rcp_W.push_back(1. / sp.molar_mass[k]);
    for (int d = 0; d < DK; d++) {
      t.fit_c.push_back(d < D ? p.conductivityIVT[k * D + d] : 0.);
      t.fit_v.push_back(d < D ? p.VviscosityIVVT[k * D + d] : 0.);
    }
  }
  // Wilke (transport.rs:148-158): Φ_kj = (a_kj + b_kj v_k/v_j)², a = (8 (1 + W_k/W_j))^-1/4, b = a (W_j/W_k)^1/4,
  // stored as [row tile mt][column block kk][a², a b, b²][lane].
  t.wilke.assign((size_t)(KP / 16) * (KP / 4) * 3 * 64, 0.);
  for (int mt = 0; mt < KP / 16; mt++)
    for (int kk = 0; kk < KP / 4; kk++)
      for (int L = 0; L < 64; L++) {
        const int k = 16 * mt + L % 16, j = 4 * kk + L / 16;
        if (k >= K || j >= K) continue;
        const double W_k = sp.molar_mass[k], W_j = sp.molar_mass[j];
        const double a = std::sqrt(1. / std::sqrt(8.) * 1. / std::sqrt(1. + W_k / W_j));
        const double b = a * std::sqrt(std::sqrt(W_j / W_k));
        const size_t at = (((size_t)mt * (KP / 4) + kk) * 3) * 64 + L;
        t.wilke[at] = a * a;
        t.wilke[at + 64] = a * b;
        t.wilke[at + 128] = b * b;
      }
  // Diffusion tiles: 16 consecutive j < k for one k, rows k = 1..K-1 in order and j0 ascending within a row (the
  // kernel finds row k's tiles at Σ_{k' < k} ⌈k'/16⌉); A[m][d] = coefficient d of d_{k, j0+m} (zero past j = k-1).
  t.NTILE = 0;
  for (int k = 1; k < K; k++)
    for (int j0 = 0; j0 < k; j0 += 16) {
      t.NTILE++;
      for (int blk = 0; blk < DK / 4; blk++)
        for (int L = 0; L < 64; L++) {
          const int j = j0 + L % 16, d = 4 * blk + L / 16;
          t.dcoef.push_back(j < k && d < D ? p.diffusivityITVT[((size_t)k * K + j) * D + d] : 0.);
        }
    }
  return t;
}

std::string transport_text(const TransportTables& t, bool f64) {
  std::ostringstream s;
  s << "#define K " << t.K << "\n#define KP " << t.KP << "\n#define NTILE " << t.NTILE << "\n#define DK " << t.DK << "\n#define F64 "
    << (f64 ? 1 : 0) << "\n"
    << transport_tiled_source;
  return s.str();
}

std::vector<float> to_f32(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

const Polynomials& polynomials(const ProgramView& p) {
  if (!p.chemistry->polynomials) throw Unsupported("the program has no transport fits");
  return *p.chemistry->polynomials;
}

// transport_tiled with its tables in device memory; writes [λ, μ, D_0..D_{K-1}] to out[0..K+2).
class TransportTiled {
 public:
  TransportTiled(const ProgramView& p, bool f64) : t_(transport_tables(*p.mechanism, polynomials(p), p.chemistry->scales.T0)), f64_(f64) {
    kernel_.load(device::compile(transport_text(t_, f64), "transport_tiled", hip_options(true)), "transport_tiled");
    d_rcp_W_ = upload(t_.rcp_W);
    d_W_ = upload(t_.W);
    if (f64) {
      d_c_ = upload(t_.fit_c);
      d_v_ = upload(t_.fit_v);
      d_wilke_ = upload(t_.wilke);
      d_dcoef_ = upload(t_.dcoef);
    } else {
      d_c_ = upload(to_f32(t_.fit_c));
      d_v_ = upload(to_f32(t_.fit_v));
      d_wilke_ = upload(to_f32(t_.wilke));
      d_dcoef_ = upload(to_f32(t_.dcoef));
    }
  }
  ~TransportTiled() {
    for (void* ptr : {(void*)d_rcp_W_, (void*)d_W_, d_c_, d_v_, d_wilke_, d_dcoef_})
      if (ptr) (void)hipFree(ptr);
  }
  // Every state, or (with idx and count) the count states listed in idx.
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream, const unsigned* idx = nullptr,
              const unsigned* count = nullptr) {
    unsigned n32 = (unsigned)n;
    double T0 = t_.T0;
    void* args[] = {&in, &out, &n32, &d_rcp_W_, &d_W_, &d_c_, &d_v_, &d_wilke_, &d_dcoef_, &T0, &idx, &count};
    const unsigned states_per_block = 16;  // 4 waves share 16 states
    unsigned grid = (unsigned)((n + states_per_block - 1) / states_per_block);
    device::check(hipModuleLaunchKernel(kernel_.function, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr), "launch transport");
  }

 private:
  TransportTables t_;
  bool f64_;
  device::Kernel kernel_;
  double *d_rcp_W_ = nullptr, *d_W_ = nullptr;
  void *d_c_ = nullptr, *d_v_ = nullptr, *d_wilke_ = nullptr, *d_dcoef_ = nullptr;
};

void check_program(const ProgramView& p) {
  if (p.kind != ProgramKind::Chemistry) throw Unsupported("tiled forms need a program built by cb_build_chemistry");
  if (p.chemistry->options.rhs) throw Unsupported("tiled forms do not evaluate the reactor right-hand side");
}

constexpr size_t kMaxStraightRatesSpecies = 64;

// The rates section of a chemistry program as straight-line FP64 code (contraction and reciprocal division).
std::string rates_straight_text(const ProgramView& p, bool fast) {
  // Above 64 species the straight-line rates kernel spills kilobytes per lane to scratch (6.9 KB at K=160) and its
  // code object runs to megabytes; the one run of the K=160 FP64 form on a GPU ended in a memory aperture violation
  // during assembly whose cause was not established, so these mechanisms are refused here.
  if (p.mechanism->species.len() > kMaxStraightRatesSpecies)
    throw Unsupported("the straight-line FP64 rates kernel supports at most " + std::to_string(kMaxStraightRatesSpecies) +
                      " species (this mechanism has " + std::to_string(p.mechanism->species.len()) + ")");
  ChemistryOptions o;
  o.transport = false;
  o.lean_rates = true;
  StraightOptions so;
  so.reciprocal = true;
  so.newton_steps = 1;
  so.fast_transcendentals = fast;
  so.exp_poly = true;
  const ast::Linear l = ast::linearize(chemistry(*p.mechanism, nullptr, p.chemistry->scales, o));
  // A program that needs more than the 128 doubles of two waves per SIMD runs at one wave with its values in VGPRs,
  // AGPRs and scratch; its clamped exp(-G_k/RT) values then live in LDS (K_a slots of 2 KB per block of 256). At
  // K=53 (about 230 live doubles with the lean rates form) the kernel then needs no scratch and takes 1.25 ms per 2^20
  // states against 1.30 ms with the values in registers; keeping some exp(+G_k/RT) values there too, a fence every
  // 64 or 256 instructions, or two waves per SIMD with 40 slots (1.6 ms, 1.2 KB of scratch) were not faster, nor
  // was reading the FP64 constants from a __constant__ table by scalar loads (3.2 k s_mov per wave fall to 1.8 k
  // SALU instructions, but 1.48 ms against 1.18 ms: one wave per SIMD leaves the loads' latency exposed).
  // Beyond 80 slots the values stay in registers.
  if (ast::peak_live(l) > 128) so.lds_registers = gibbs_exponentials(l, true);
  if (so.lds_registers.size() * so.block_size * 8 > 160 * 1024) so.lds_registers.clear();
  return emit_straight(l, fast ? "rates_fast" : "rates_fp64", so);
}

// rates_tiled with its tables in device memory; writes ω to out[0..Ka).
class RatesTiled {
 public:
  explicit RatesTiled(const Mechanism& m) : t_(rates_tiled_tables(m)) {
    kernel_.load(device::compile(rates_tiled_text(t_), "rates_tiled", hip_options(true)), "rates_tiled");
    d_rate_ = upload(t_.rate);
    d_k0_ = upload(t_.k0);
    d_troe_ = upload(t_.troe);
    d_info_ = upload(t_.info);
    d_ops_ = upload(t_.ops);
    d_eff_species_ = upload(t_.eff_species);
    d_eff_value_ = upload(t_.eff_value);
    d_rcp_W_ = upload(t_.rcp_W);
    d_nasa_ = upload(t_.nasa);
    d_split_ = upload(t_.split);
  }
  ~RatesTiled() {
    for (void* ptr : {(void*)d_rate_, (void*)d_k0_, (void*)d_troe_, (void*)d_info_, (void*)d_ops_, (void*)d_eff_species_,
                      (void*)d_eff_value_, (void*)d_rcp_W_, (void*)d_nasa_, (void*)d_split_})
      if (ptr) (void)hipFree(ptr);
  }
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) {
    unsigned n32 = (unsigned)n;
    double pos = 6981463048192., neg = (double)cbrtf(3.40282347e38f);  // reaction.rs:244-256
    void* args[] = {&in, &out, &n32, &d_rate_, &d_k0_, &d_troe_, &d_info_, &d_ops_, &d_eff_species_, &d_eff_value_, &d_rcp_W_,
                    &d_nasa_, &d_split_, &pos, &neg};
    const unsigned states_per_block = (unsigned)t_.NS;
    device::check(hipModuleLaunchKernel(kernel_.function, (unsigned)((n + states_per_block - 1) / states_per_block), 1, 1, 64, 1, 1, 0,
                                        stream, args, nullptr),
                  "launch rates_tiled");
  }

 private:
  TiledRatesTables t_;
  device::Kernel kernel_;
  double *d_rate_ = nullptr, *d_k0_ = nullptr, *d_troe_ = nullptr, *d_eff_value_ = nullptr, *d_rcp_W_ = nullptr, *d_nasa_ = nullptr,
         *d_split_ = nullptr;
  int32_t *d_info_ = nullptr, *d_ops_ = nullptr, *d_eff_species_ = nullptr;
};

// rates_mixed with its tables in device memory; writes ω to out[0..Ka).
class RatesMixed {
 public:
  explicit RatesMixed(const Mechanism& m) : t_(mixed_tables(m)) {
    kernel_.load(device::compile(rates_mixed_text(t_), "rates_mixed", hip_options(true)), "rates_mixed");
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
  }
  ~RatesMixed() {
    for (void* ptr : {(void*)d_fwd_, (void*)d_rev_, (void*)d_nu_, (void*)d_rate_, (void*)d_k0_, (void*)d_troe_, (void*)d_info_,
                      (void*)d_tile_, (void*)d_eff_, (void*)d_rcp_W_, (void*)d_gibbs_, (void*)d_split_})
      if (ptr) (void)hipFree(ptr);
  }
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) {
    unsigned n32 = (unsigned)n;
    const unsigned states_per_block = 16 * t_.WAVES;
    unsigned grid = (unsigned)((n + states_per_block - 1) / states_per_block);
    double pos = t_.ln_clamp_pos, neg = t_.ln_clamp_neg;
    void* args[] = {&in, &out, &n32, &d_fwd_, &d_rev_, &d_nu_, &d_rate_, &d_k0_, &d_troe_, &d_info_, &d_tile_, &d_eff_,
                    &d_rcp_W_, &d_gibbs_, &d_split_, &pos, &neg};
    device::check(hipModuleLaunchKernel(kernel_.function, grid, 1, 1, 64 * t_.WAVES, 1, 1, 0, stream, args, nullptr), "launch rates_mixed");
  }

 private:
  MixedTables t_;
  device::Kernel kernel_;
  uint16_t *d_fwd_ = nullptr, *d_rev_ = nullptr, *d_nu_ = nullptr;
  double *d_rate_ = nullptr, *d_k0_ = nullptr;
  float *d_troe_ = nullptr, *d_eff_ = nullptr;
  int32_t *d_info_ = nullptr, *d_tile_ = nullptr;
  double *d_rcp_W_ = nullptr, *d_gibbs_ = nullptr, *d_split_ = nullptr;
};

// The rates section as straight-line FP64 code (with FP32 cores for exp and ln when fast); writes ω to out[0..Ka).
class RatesStraight {
 public:
  RatesStraight(const ProgramView& p, bool fast) {
    const char* name = fast ? "rates_fast" : "rates_fp64";
    kernel_.load(device::compile(rates_straight_text(p, fast), name, hip_options(true)), name);
  }
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) {
    unsigned n32 = (unsigned)n, offset = 0, block = 256;
    void* args[] = {&in, &out, &n32, &offset};
    unsigned grid = (unsigned)((n + block - 1) / block);
    device::check(hipModuleLaunchKernel(kernel_.function, grid, 1, 1, block, 1, 1, 0, stream, args, nullptr), "launch rates_fp64");
  }

 private:
  device::Kernel kernel_;
};

// Mechanisms up to 64 species take the transport section as straight-line code (one lane per state): at K=53 it
// issues fewer instructions per state than the tiled kernel, whose MFMA contractions do not pay for the
// reciprocals and reductions around them. Larger mechanisms (a K² program) take transport_tiled.
// Mechanisms above kMaxStraightRatesSpecies (or any, with kFormTiledRates) take rates_tiled in forms 1 and 3.
bool tiled_rates(const ProgramView& p, int form) {
  return p.mechanism->species.len() > kMaxStraightRatesSpecies || (form & kFormTiledRates);
}

// form | kFormTiledTransport sends every mechanism through transport_tiled (the MFMA contractions) instead.
bool straight_transport(const ProgramView& p, int form) { return p.mechanism->species.len() <= 64 && !(form & kFormTiledTransport); }

std::string transport_straight_text(const ProgramView& p, bool f64) {
  ChemistryOptions o;
  o.rates = false;
  o.lean_transport = true;
  o.f32_range = !f64;
  StraightOptions so;
  so.reciprocal = f64;
  so.newton_steps = 1;
  so.f32 = !f64;
  // FP64: the lean transport program fits two waves per SIMD (243 VGPRs, no AGPRs). FP32: three waves (168 VGPRs,
  // one spilled register) beat two (246 VGPRs) by 1.5 % on the fast form's step at K=53.
  so.min_waves = f64 ? 2 : 3;
  return emit_straight(ast::linearize(chemistry(*p.mechanism, p.chemistry->polynomials.get(), p.chemistry->scales, o)),
                       f64 ? "transport_fp64" : "transport_f32", so);
}

class TransportStraight {
 public:
  TransportStraight(const ProgramView& p, bool f64) {
    const char* name = f64 ? "transport_fp64" : "transport_f32";
    kernel_.load(device::compile(transport_straight_text(p, f64), name, hip_options(true)), name);
  }
  void launch(const double* const* in, double* const* out, size_t n, hipStream_t stream) {
    unsigned n32 = (unsigned)n, offset = 0, block = 256;
    void* args[] = {&in, &out, &n32, &offset};
    unsigned grid = (unsigned)((n + block - 1) / block);
    device::check(hipModuleLaunchKernel(kernel_.function, grid, 1, 1, block, 1, 1, 0, stream, args, nullptr), "launch transport");
  }

 private:
  device::Kernel kernel_;
};

class Tiled : public SoAForm {
 public:
  // form 1: FP64 rates and transport; 2: MFMA mixed rates, FP32 transport; 3: FP32-core rates, FP32 transport.
  Tiled(const ProgramView& p, int form_flags) : SoAForm(p.function->n_input, (int)p.function->output.size()) {
    check_program(p);
    const auto& o = p.chemistry->options;
    const int form = form_flags & kFormBase;
    const bool f64 = form == 1;
    if (o.rates) {
      n_rates_ = (int)p.mechanism->active;
      if (form == 2)
        mixed_ = std::make_unique<RatesMixed>(*p.mechanism);
      else if (tiled_rates(p, form_flags))
        tiled_rates_ = std::make_unique<RatesTiled>(*p.mechanism);
      else
        fp64_ = std::make_unique<RatesStraight>(p, form == 3);
    }
    if (o.transport) {
      if (straight_transport(p, form_flags))
        straight_ = std::make_unique<TransportStraight>(p, f64);
      else
        transport_ = std::make_unique<TransportTiled>(p, f64);
    }
  }

  ~Tiled() override {
    if (fork_) (void)hipEventDestroy(fork_);
    if (join_) (void)hipEventDestroy(join_);
    if (side_) (void)hipStreamDestroy(side_);
  }

 protected:
  // The rates and transport kernels read the same inputs and write disjoint outputs: the transport kernel runs on a
  // second stream of the module's (default priority), forked from and joined back into the caller's stream, so that
  // the two kernels' blocks share the CUs at the ends of each launch (2.27 against 2.32 ms per 2^20 states at K=53).
  void launch(const double* const* in, double* const* out, size_t n, const Launch& how) override {
    hipStream_t stream = how.stream;
    const bool both = (mixed_ || fp64_ || tiled_rates_) && (straight_ || transport_) && !how.single_stream;
    hipStream_t ts = stream;
    if (both) {
      if (!side_) {
        device::check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        device::check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreateWithFlags");
        device::check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreateWithFlags");
      }
      device::check(hipEventRecord(fork_, stream), "hipEventRecord");
      device::check(hipStreamWaitEvent(side_, fork_, 0), "hipStreamWaitEvent");
      ts = side_;
    }
    // The longer rates kernel goes first (the other order measured 2.29 against 2.25 ms).
    if (mixed_) mixed_->launch(in, out, n, stream);
    if (fp64_) fp64_->launch(in, out, n, stream);
    if (tiled_rates_) tiled_rates_->launch(in, out, n, stream);
    if (straight_) straight_->launch(in, out + n_rates_, n, ts);
    if (transport_) transport_->launch(in, out + n_rates_, n, ts);
    if (both) {
      device::check(hipEventRecord(join_, side_), "hipEventRecord");
      device::check(hipStreamWaitEvent(stream, join_, 0), "hipStreamWaitEvent");
    }
  }

 private:
  std::unique_ptr<RatesMixed> mixed_;
  std::unique_ptr<RatesStraight> fp64_;
  std::unique_ptr<RatesTiled> tiled_rates_;
  std::unique_ptr<TransportStraight> straight_;
  std::unique_ptr<TransportTiled> transport_;
  int n_rates_ = 0;
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
};

std::vector<uint16_t> to_bf16(const std::vector<float>& v) {
  std::vector<uint16_t> out;
  for (float x : v) out.push_back(bf16_bits(x));
  return out;
}

std::string surrogate_text(int K, const MlpLayout& lay, int DK, int NM) {
  std::ostringstream s;
  s << "#define K " << K << "\n#define IN " << lay.in_pad << "\n#define WIDTH " << lay.width << "\n#define DEPTH " << lay.depth
    << "\n#define OUT " << lay.out_pad << "\n#define DK " << DK << "\n#define NM " << NM << "\n"
    << surrogate_source;
  return s.str();
}

// FP32-core straight-line rates (or the MFMA mixed rates), then the transport surrogate with the plain FP64 transport
// kernel on the states outside its box.
class Surrogate : public SoAForm {
 public:
  Surrogate(const ProgramView& p, const SurrogateWeights& w) : SoAForm(p.function->n_input, (int)p.function->output.size()) {
    check_program(p);
    const auto& o = p.chemistry->options;
    if (!o.transport) throw Unsupported("the surrogate form needs the transport section");
    const Mechanism& m = *p.mechanism;
    const int K = (int)m.species.len(), NF = K + 2;
    const MlpLayout lay(NF, w.width, w.depth, NF);
    if (lay.width % 64) throw Unsupported("the surrogate kernel needs a hidden width that is a multiple of 64");
    if ((long)w.params.size() != lay.n_params) throw ArgumentError("surrogate parameters do not match the network size");
    for (const auto* v : {&w.f_mean, &w.f_rstd, &w.lo, &w.hi, &w.o_mean, &w.o_std})
      if ((int)v->size() != NF) throw ArgumentError("surrogate normalisation arrays need K+2 entries");
    if (o.rates) {
      n_rates_ = (int)m.active;
      if (w.rates_form == 2)
        mixed_ = std::make_unique<RatesMixed>(m);
      else
        rates_ = std::make_unique<RatesStraight>(p, true);
    }
    // The fallback for states outside the network's box: the plain FP64 transport kernel (the tiled one, which
    // evaluates a list of states in place).
    transport_ = std::make_unique<TransportTiled>(p, true);
    const Polynomials& poly = polynomials(p);
    const int D = poly.D, DK = D;
    if (D > 8) throw Unsupported("transport fits longer than 8 coefficients");
    // Baselines (kernels/surrogate.hip): every species' conductivity and viscosity fits, 1/W, and per species k the
    // Wilke coefficients and diffusion fits against the major species j = major[i] (slot NM: the bulk species b
    // for Wilke, r(k) for diffusion).
    const int b = K - 1, NM = (int)w.major.size();
    if (NM < 1 || NM > K) throw ArgumentError("the number of major species is 1..K");
    {
      // LDS of kernels/surrogate.hip: activation rows, outputs, partial sums
      const long row = (std::max(lay.in_pad + 2 * (2 * NM + 1), (long)lay.width) + 15) / 16 * 16 + 8;
      const long lds = 64 * row * 2 + (long)NF * 64 * 4 + 4 * 3 * 64 * 4 + 4 * 64 * 4;
      if (lds > 160 * 1024) throw Unsupported("the surrogate kernel's LDS does not fit 160 KB for this mechanism and network");
    }
    std::vector<int32_t> slot(K, -1);
    int other = -1;
    for (int i = 0; i < NM; i++) {
      const int j = w.major[i];
      if (j < 0 || j >= K || slot[j] >= 0) throw ArgumentError("major species must be distinct species indices");
      slot[j] = i;
      if (j != b && other < 0) other = j;
    }
    if (other < 0) throw ArgumentError("the major species need one besides the bulk species");
    const auto& W = m.species.molar_mass;
    std::vector<float> fc, fv, rw, wa, wb, df;
    for (int k = 0; k < K; k++) {
      for (int d = 0; d < D; d++) {
        fc.push_back((float)poly.conductivityIVT[k * D + d]);
        fv.push_back((float)poly.VviscosityIVVT[k * D + d]);
      }
      rw.push_back((float)(1. / W[k]));
      for (int i = 0; i <= NM; i++) {
        const int j = i < NM ? w.major[i] : b;
        const double a = 1. / std::pow(8. * (1. + W[k] / W[j]), 0.25);
        wa.push_back((float)a);
        wb.push_back((float)(a * std::pow(W[j] / W[k], 0.25)));
        const int jd = i < NM ? w.major[i] : (k == b ? other : b);
        for (int d = 0; d < D; d++) df.push_back((float)poly.diffusivityITVT[((size_t)k * K + jd) * D + d]);
      }
    }
    std::vector<long> offsets;
    for (const auto& L : lay.layers) {
      offsets.push_back(L.w);
      offsets.push_back(L.b);
    }
    kernel_.load(device::compile(surrogate_text(K, lay, DK, NM), "surrogate_transport", hip_options(true)), "surrogate_transport");
    device::check(hipModuleGetFunction(&clear_, kernel_.module, "surrogate_clear"), "hipModuleGetFunction surrogate_clear");
    d_w16_ = upload(to_bf16(w.params));
    d_params_ = upload(w.params);
    d_offsets_ = upload(offsets);
    d_f_mean_ = upload(w.f_mean);
    d_f_rstd_ = upload(w.f_rstd);
    d_lo_ = upload(w.lo);
    d_hi_ = upload(w.hi);
    d_o_mean_ = upload(w.o_mean);
    d_o_std_ = upload(w.o_std);
    d_bc_ = upload(fc);
    d_bv_ = upload(fv);
    d_major_ = upload(w.major);
    d_slot_ = upload(slot);
    d_wa_ = upload(wa);
    d_wb_ = upload(wb);
    d_bd_ = upload(df);
    d_rw_ = upload(rw);
    T0_ = (float)p.chemistry->scales.T0;
    device::check(hipMalloc(&d_count_, sizeof(unsigned)), "hipMalloc");
  }
  ~Surrogate() override {
    for (void* ptr : {(void*)d_w16_, (void*)d_params_, (void*)d_offsets_, (void*)d_f_mean_, (void*)d_f_rstd_, (void*)d_lo_,
                      (void*)d_hi_, (void*)d_o_mean_, (void*)d_o_std_, (void*)d_bc_, (void*)d_bv_, (void*)d_bd_, (void*)d_rw_, (void*)d_count_,
                      (void*)d_fallback_, (void*)d_major_, (void*)d_slot_, (void*)d_wa_, (void*)d_wb_})
      if (ptr) (void)hipFree(ptr);
    if (fork_) (void)hipEventDestroy(fork_);
    if (join_) (void)hipEventDestroy(join_);
    if (side_) (void)hipStreamDestroy(side_);
  }
  // States the last call sent to the plain transport kernel (waits for the stream).
  unsigned fallback_count(hipStream_t stream) {
    unsigned c = 0;
    device::check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    device::check(hipMemcpy(&c, d_count_, sizeof c, hipMemcpyDeviceToHost), "hipMemcpy");
    return c;
  }

 protected:
  // Work memory: the fallback counter (one unsigned in the first 256 bytes), then the list of fallback states.
  size_t work_bytes(size_t n) const override { return 256 + n * sizeof(unsigned); }

  void launch(const double* const* in, double* const* out, size_t n, const Launch& how) override {
    hipStream_t stream = how.stream;
    unsigned *count = d_count_, *list = d_fallback_;
    if (how.work) {
      count = (unsigned*)how.work;
      list = (unsigned*)((char*)how.work + 256);
    } else if (n > capacity_) {
      // an earlier call on another stream may still fill the list
      device::check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (d_fallback_) device::check(hipFree(d_fallback_), "hipFree");
      d_fallback_ = nullptr;
      device::check(hipMalloc(&d_fallback_, n * sizeof(unsigned)), "hipMalloc");
      capacity_ = n;
      list = d_fallback_;
    }
    // The rates kernel on the caller's stream; the surrogate and its fallback on a side stream forked from and joined
    // back into it (the two read the same inputs and write disjoint outputs), as in the tiled forms.
    hipStream_t ts = stream;
    const bool fork = (rates_ || mixed_) && !how.single_stream;
    if (fork) {
      if (!side_) {
        device::check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        device::check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreateWithFlags");
        device::check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreateWithFlags");
      }
      device::check(hipEventRecord(fork_, stream), "hipEventRecord");
      device::check(hipStreamWaitEvent(side_, fork_, 0), "hipStreamWaitEvent");
      ts = side_;
    }
    if (rates_) rates_->launch(in, out, n, stream);
    if (mixed_) mixed_->launch(in, out, n, stream);
    double* const* tout = out + n_rates_;
    void* clear_args[] = {&count};
    device::check(hipModuleLaunchKernel(clear_, 1, 1, 1, 64, 1, 1, 0, ts, clear_args, nullptr), "launch surrogate_clear");
    unsigned n32 = (unsigned)n;
    void* args[] = {&in, &tout, &n32, &d_w16_, &d_params_, &d_offsets_, &d_f_mean_, &d_f_rstd_, &d_lo_, &d_hi_,
                    &d_o_mean_, &d_o_std_, &d_bc_, &d_bv_, &d_major_, &d_slot_, &d_wa_, &d_wb_, &d_bd_, &d_rw_, &T0_,
                    &list, &count};
    const unsigned per_block = 64;
    device::check(hipModuleLaunchKernel(kernel_.function, (unsigned)((n + per_block - 1) / per_block), 1, 1, 256, 1, 1, 0, ts, args,
                                        nullptr),
                  "launch surrogate_transport");
    transport_->launch(in, tout, n, ts, list, count);
    if (fork) {
      device::check(hipEventRecord(join_, side_), "hipEventRecord");
      device::check(hipStreamWaitEvent(stream, join_, 0), "hipStreamWaitEvent");
    }
  }

 private:
  std::unique_ptr<RatesStraight> rates_;
  std::unique_ptr<RatesMixed> mixed_;
  std::unique_ptr<TransportTiled> transport_;
  int n_rates_ = 0;
  device::Kernel kernel_;
  uint16_t* d_w16_ = nullptr;
  float *d_params_ = nullptr, *d_f_mean_ = nullptr, *d_f_rstd_ = nullptr, *d_lo_ = nullptr, *d_hi_ = nullptr, *d_o_mean_ = nullptr,
        *d_o_std_ = nullptr, *d_bc_ = nullptr, *d_bv_ = nullptr, *d_bd_ = nullptr, *d_rw_ = nullptr, *d_wa_ = nullptr, *d_wb_ = nullptr;
  int32_t *d_major_ = nullptr, *d_slot_ = nullptr;
  long* d_offsets_ = nullptr;
  float T0_ = 1000.f;
  unsigned *d_count_ = nullptr, *d_fallback_ = nullptr;
  size_t capacity_ = 0;
  hipFunction_t clear_ = nullptr;
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
};

}  // namespace

bool tiled_applies(const ProgramView& p) { return p.kind == ProgramKind::Chemistry && !p.chemistry->options.rhs; }

std::string tiled_source(const ProgramView& p, int form_flags, unsigned) {
  check_program(p);
  const int form = form_flags & kFormBase;
  const bool mixed = form != 1;
  std::string s;
  if (p.chemistry->options.rates)
    s += form == 2                      ? rates_mixed_text(mixed_tables(*p.mechanism))
         : tiled_rates(p, form_flags) ? rates_tiled_text(rates_tiled_tables(*p.mechanism))
                                      : rates_straight_text(p, form == 3);
  if (p.chemistry->options.transport)
    s += "\n// ---- transport ----\n" + (straight_transport(p, form_flags) ? transport_straight_text(p, !mixed)
                                                         : transport_text(transport_tables(*p.mechanism, polynomials(p), p.chemistry->scales.T0), !mixed));
  return s;
}

void compile_tiled(const ProgramView& p, int form_flags, unsigned) {
  check_program(p);
  const int form = form_flags & kFormBase;
  const bool mixed = form != 1;
  if (p.chemistry->options.rates) {
    if (form == 2)
      device::compile(rates_mixed_text(mixed_tables(*p.mechanism)), "rates_mixed", hip_options(true));
    else if (tiled_rates(p, form_flags))
      device::compile(rates_tiled_text(rates_tiled_tables(*p.mechanism)), "rates_tiled", hip_options(true));
    else
      device::compile(rates_straight_text(p, form == 3), form == 3 ? "rates_fast" : "rates_fp64", hip_options(true));
  }
  if (p.chemistry->options.transport) {
    if (straight_transport(p, form_flags))
      device::compile(transport_straight_text(p, !mixed), mixed ? "transport_f32" : "transport_fp64", hip_options(true));
    else
      device::compile(transport_text(transport_tables(*p.mechanism, polynomials(p), p.chemistry->scales.T0), !mixed), "transport_tiled",
                      hip_options(true));
  }
}

std::unique_ptr<Form> assemble_tiled(const ProgramView& p, int form, unsigned) { return std::make_unique<Tiled>(p, form); }

std::unique_ptr<Form> assemble_surrogate(const ProgramView& p, const SurrogateWeights& w) { return std::make_unique<Surrogate>(p, w); }

unsigned surrogate_fallback_count(Form& f, hipStream_t stream) {
  auto* s = dynamic_cast<Surrogate*>(&f);
  if (!s) throw ArgumentError("not a surrogate module");
  return s->fallback_count(stream);
}

}  // namespace combustion
