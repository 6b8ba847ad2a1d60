// Rates builder: restates reaction.rs:53-391 on the C++ IR.
#include <cfloat>
#include <cmath>
#include <optional>

#include "builders.hpp"

namespace combustion {

using ast::Block;
using ast::E;

namespace {

// reaction.rs:244-251: the largest f32 m with m*m*m <= f32::MAX (evaluated in f32), and cbrt(f32::MAX).
float max_cube_root_f32() {
  volatile float max = cbrtf(FLT_MAX);
  while ((float)((float)(max * max) * max) > FLT_MAX) max = nextafterf(max, -INFINITY);
  return max;
}
const double exp_clamp = (double)max_cube_root_f32();   // 6981463048192
const double exp_neg_clamp = (double)cbrtf(FLT_MAX);     // 6981463572480 (glibc cbrtf, correctly rounded)
const double flt_max = (double)FLT_MAX;
const double flt_min = (double)FLT_MIN;  // f32::MIN_POSITIVE

// reaction.rs:53-63 (non-negative exponents only: reactant and product coefficients are u8)
E product_of_exponentiations(Block& f, const std::vector<E>& C, const std::vector<uint8_t>& n, const char* name) {
  std::optional<E> p;
  for (size_t k = 0; k < C.size(); k++)
    for (int i = 0; i < n[k]; i++) p = p ? f.defv(f.mul(*p, C[k]), name) : C[k];
  if (!p) throw ast::Error("reaction without species on one side");
  return *p;
}

// reaction.rs:75-81
E dot(Block& f, const std::vector<E>& a, const std::vector<E>& x) {
  std::vector<E> terms;
  for (size_t i = 0; i < a.size() && i < x.size(); i++) terms.push_back(f.mul(a[i], x[i]));
  return f.sum(terms);
}
std::vector<E> consts(Block& f, const double* a, size_t n) {
  std::vector<E> out;
  for (size_t i = 0; i < n; i++) out.push_back(f.c(a[i]));
  return out;
}
E molar_heat_capacity_R(Block& f, const std::array<double, 7>& a, const Temperatures& T) {
  return dot(f, consts(f, a.data(), 5), {f.c(1.), T.T, T.T2, T.T3, T.T4});
}
E enthalpy_RT(Block& f, const std::array<double, 7>& a, const Temperatures& T) {
  E e = f.add(f.c(a[0]), f.mul(f.c(a[1] / 2.), T.T));
  e = f.add(e, f.mul(f.c(a[2] / 3.), T.T2));
  e = f.add(e, f.mul(f.c(a[3] / 4.), T.T3));
  e = f.add(e, f.mul(f.c(a[4] / 5.), T.T4));
  return f.add(e, f.mul(f.c(a[5]), T.rcpT));
}
std::array<double, 7> gibbs0_RT(const std::array<double, 7>& a) {
  return {a[0] - a[6], -a[0], -a[1] / 2., (1. / 3. - 1. / 2.) * a[2], (1. / 4. - 1. / 3.) * a[3], (1. / 5. - 1. / 4.) * a[4], a[5]};
}
std::vector<E> gibbs_basis(Block& f, const Temperatures& T) { return {f.c(1.), T.lnT, T.T, T.T2, T.T3, T.T4, T.rcpT}; }

// reaction.rs:83-110: one Select per temperature-split bucket (buckets in first-appearance order).
std::vector<E> thermodynamics(Block& f, const std::vector<NASA7>& species, E (*expression)(Block&, const std::array<double, 7>&, const Temperatures&),
                              const Temperatures& T, const char* debug) {
  std::vector<E> out(species.size(), -1);
  std::vector<double> splits;
  std::vector<std::vector<size_t>> buckets;
  for (size_t k = 0; k < species.size(); k++) {
    double s = species[k].temperature_split;
    size_t b = 0;
    for (; b < splits.size(); b++)
      if ((std::isnan(s) && std::isnan(splits[b])) || s == splits[b]) break;
    if (b == splits.size()) {
      splits.push_back(s);
      buckets.emplace_back();
    }
    buckets[b].push_back(k);
  }
  for (size_t b = 0; b < splits.size(); b++) {
    if (std::isnan(splits[b])) {
      for (size_t k : buckets[b]) {
        E e = expression(f, species[k].pieces[0], T);
        out[k] = f.is_leaf(e) ? e : f.defv(e, "");
      }
    } else {
      std::vector<E> t, fe;
      std::vector<int32_t> results;
      for (size_t k : buckets[b]) {
        int32_t id = f.value(std::string(debug) + "[" + std::to_string(k) + "]");
        results.push_back(id);
        out[k] = f.v(id);
        t.push_back(expression(f, species[k].pieces[0], T));
        fe.push_back(expression(f, species[k].pieces[1], T));
      }
      f.select(f.le(T.T, f.c(splits[b])), t, fe, results);
    }
  }
  return out;
}

// reaction.rs:113-139
E arrhenius(Block& f, const RateConstant& k, const Temperatures& T) {
  const double beta = k.temperature_exponent;
  std::optional<E> factor;
  double remainder;
  if (beta <= -1.5) {
    factor = T.rcpT2;
    remainder = beta + 2.;
  } else if (beta <= -0.5) {
    factor = T.rcpT;
    remainder = beta + 1.;
  } else if (beta >= 3.5) {
    factor = T.T4;
    remainder = beta - 4.;
  } else if (beta >= 2.5) {
    factor = T.T3;
    remainder = beta - 3.;
  } else if (beta >= 1.5) {
    factor = T.T2;
    remainder = beta - 2.;
  } else if (beta >= 0.5) {
    factor = T.T;
    remainder = beta - 1.;
  } else {
    remainder = beta;
  }
  std::vector<E> x;
  if ((float)remainder != 0.f) x.push_back((float)remainder == 1.f ? T.lnT : f.defv(f.mul(f.c(remainder), T.lnT), "β'lnT"));
  if (k.activation_temperature != 0.) x.push_back(f.defv(f.mul(f.c(-k.activation_temperature), T.rcpT), "-Ea/kT"));
  auto sum = f.sigma(x);
  E e = sum ? f.defv(f.exp(f.add(*sum, f.c(std::log(k.preexponential_factor)))), "exp(β'lnT-Ea/kT)") : f.c(k.preexponential_factor);
  E k_ = factor ? f.mul(*factor, e) : e;
  return f.min(f.c(flt_max), k_);
}

// Pressure-dependent Arrhenius with several entries (not in the reference, whose yaml.rs:120,149 assert one):
// Cantera's PLOG rate, ln k interpolated linearly in ln P between the two table pressures around P, and the first
// or last entry's k beyond the table; entries of equal pressure are summed. ln P = ln(P/(R T)) + ln T + ln R from
// the state's total concentration. The result is clamped at FLT_MAX like every Arrhenius rate (reaction.rs:138).
E plog_rate(Block& f, const Reaction& r, const Temperatures& T, E concentration) {
  std::vector<std::pair<double, std::vector<RateConstant>>> levels;
  for (const PlogEntry& e : r.plog) {
    if (levels.empty() || levels.back().first != e.P) levels.push_back({e.P, {}});
    levels.back().second.push_back(e.k);
  }
  std::vector<E> lnk;
  for (auto& [P, ks] : levels) {
    if (ks.size() == 1 && ks[0].preexponential_factor > 0) {
      const RateConstant& k = ks[0];
      std::vector<E> t = {f.c(std::log(k.preexponential_factor))};
      if (k.temperature_exponent != 0.) t.push_back(f.mul(f.c(k.temperature_exponent), T.lnT));
      if (k.activation_temperature != 0.) t.push_back(f.mul(f.c(-k.activation_temperature), T.rcpT));
      lnk.push_back(f.defv(f.sum(t), "ln k(P)"));
    } else {
      std::vector<E> t;
      for (const RateConstant& k : ks)
        t.push_back(f.mul(f.c(k.preexponential_factor),
                          f.exp(f.sub(f.mul(f.c(k.temperature_exponent), T.lnT), f.mul(f.c(k.activation_temperature), T.rcpT)))));
      lnk.push_back(f.defv(f.ln(f.sum(t)), "ln k(P)"));
    }
  }
  E v = lnk.back();
  if (levels.size() > 1) {
    E lnP = f.defv(f.add(f.add(f.ln(concentration), T.lnT), f.c(std::log(R))), "lnP");
    // from the top: v = ln k beyond the table, then each interval (ln P_i, ln P_{i+1}] and below the table
    for (int i = (int)levels.size() - 2; i >= -1; i--) {
      E t;
      if (i >= 0) {
        const double lo = std::log(levels[i].first), hi = std::log(levels[i + 1].first);
        t = f.add(lnk[i], f.mul(f.sub(lnk[i + 1], lnk[i]), f.mul(f.sub(lnP, f.c(lo)), f.c(1. / (hi - lo)))));
      } else {
        t = lnk[0];
      }
      // lnP <= ln P_{i+1}: below the interval's top (i = -1: below the table)
      const double edge = std::log(levels[i + 1].first);
      int32_t id = f.value("ln k(P)");
      f.select(f.le(lnP, f.c(edge)), {t}, {v}, {id});
      v = f.v(id);
    }
  }
  return f.min(f.c(flt_max), f.exp(v));
}

// reaction.rs:141-145
E efficiency(Block& f, const std::vector<double>& efficiencies, E concentration, const std::vector<E>& C) {
  std::vector<E> non_one;
  for (size_t k = 0; k < efficiencies.size() && k < C.size(); k++) {
    double e = efficiencies[k];
    if (e == 1.) continue;
    non_one.push_back(e - 1. == 1. ? C[k] : f.defv(f.mul(f.c(e - 1.), C[k]), "fC"));
  }
  return non_one.empty() ? concentration : f.add(concentration, f.sum(non_one));
}

// reaction.rs:40-48: numerator and optional denominator of the forward rate constant.
struct Ratio {
  E n;
  std::optional<E> d;
};

// reaction.rs:149-205
Ratio forward_rate_constant(Block& f, const Reaction& r, const Temperatures& T, E concentration, const std::vector<E>& C) {
  switch (r.model) {
    case ReactionModel::Elementary:
    case ReactionModel::Irreversible:
      if (!r.plog.empty()) return {f.edef(plog_rate(f, r, T, concentration), "k"), std::nullopt};
      return {f.edef(arrhenius(f, r.rate_constant, T), "k"), std::nullopt};
    case ReactionModel::ThreeBody:
      return {f.mul(arrhenius(f, r.rate_constant, T), efficiency(f, r.efficiencies, concentration, C)), std::nullopt};
    case ReactionModel::PressureModification: {
      E M = efficiency(f, r.efficiencies, concentration, C);
      E k0 = arrhenius(f, r.k0, T);
      E C_k0 = f.defv(f.mul(M, k0), "C_k0");
      E k_inf = f.edef(arrhenius(f, r.rate_constant, T), "k_inf");
      return {f.mul(C_k0, k_inf), f.add(f.add(C_k0, k_inf), f.c(flt_min))};
    }
    case ReactionModel::Troe: {
      E M = efficiency(f, r.efficiencies, concentration, C);
      const double factor = r.rate_constant.preexponential_factor;
      RateConstant k0 = r.k0, k_inf = r.rate_constant;
      k0.preexponential_factor /= factor;
      k_inf.preexponential_factor = 1.;
      E k0v = f.defv(arrhenius(f, k0, T), "k_0");
      E k_inf_e, Pr;
      if (k_inf.temperature_exponent == 0. && k_inf.activation_temperature == 0.) {
        Pr = f.defv(f.mul(M, k0v), "Pr");
        k_inf_e = f.c(factor);
      } else {
        E kv = f.edef(arrhenius(f, k_inf, T), "k_inf");
        Pr = f.defv(f.div(f.mul(M, k0v), f.add(kv, f.c(flt_min))), "Pr");
        k_inf_e = f.mul(f.c(factor), kv);
      }
      const Troe& t = r.troe;
      std::vector<E> terms;
      terms.push_back(f.mul(f.c(1. - t.A), f.defv(f.exp(f.div(T.T, f.c(-t.T3))), "exp(-T/T3)")));
      terms.push_back(t.T1 < 1e30 ? f.mul(f.c(t.A), f.defv(f.exp(f.div(T.T, f.c(-t.T1))), "exp(-T/T1)")) : f.c(t.A));
      if (std::isfinite(t.T2)) terms.push_back(f.defv(f.exp(f.mul(f.c(-t.T2), T.rcpT)), "exp(-T2/T)"));
      E Fcent = f.sum(terms);
      E lnFcent = f.f64(Fcent) ? f.c(*f.f64(Fcent)) : f.defv(f.ln(Fcent), "lnFcent");
      E Cc = f.sub(f.mul(f.c(-0.67), lnFcent), f.c(0.4 * std::log(10.)));
      E N = f.add(f.mul(f.c(-1.27), lnFcent), f.c(0.75 * std::log(10.)));
      E lnPrC = f.defv(f.add(f.ln(f.add(Pr, f.c(flt_min))), Cc), "lnPr+C");
      E f1 = f.defv(f.div(lnPrC, f.add(f.mul(f.c(-0.14), lnPrC), N)), "f1");
      E F = f.exp(f.div(lnFcent, f.add(f.mul(f1, f1), f.c(1.))));
      return {f.mul(f.mul(F, k_inf_e), Pr), f.add(Pr, f.c(1.))};
    }
    case ReactionModel::SRI: {
      // Not in the reference (reaction.rs:194-203 is unimplemented!): the SRI form of Stewart, Larson & Golden
      // (1989) as Cantera documents it, F = d [a exp(-b/T) + exp(-T/c)]^X T^e with X = 1/(1 + log10(Pr)²)
      // (exp(-T/c) dropped when c = 0), k = k∞ Pr/(1 + Pr) F, Pr = M k0/k∞ guarded like the Troe branch.
      E M = efficiency(f, r.efficiencies, concentration, C);
      E k0 = f.defv(arrhenius(f, r.k0, T), "k_0");
      E k_inf = f.edef(arrhenius(f, r.rate_constant, T), "k_inf");
      E Pr = f.defv(f.div(f.mul(M, k0), f.add(k_inf, f.c(flt_min))), "Pr");
      E l = f.mul(f.ln(f.add(Pr, f.c(flt_min))), f.c(1. / std::log(10.)));
      E X = f.defv(f.div(f.c(1.), f.add(f.c(1.), f.sq(l))), "X");
      const SRI& s = r.sri;
      std::vector<E> terms;
      E first = f.exp(f.mul(f.c(-s.B), T.rcpT));
      terms.push_back(s.A == 1. ? first : f.mul(f.c(s.A), first));
      if (s.C != 0.) terms.push_back(f.exp(f.div(T.T, f.c(-s.C))));
      E base = f.defv(f.sum(terms), "SRI base");
      E F = f.exp(f.mul(X, f.ln(base)));
      if (s.E != 0.) F = f.mul(F, f.exp(f.mul(f.c(s.E), T.lnT)));
      if (s.D != 1.) F = f.mul(f.c(s.D), F);
      return {f.mul(f.mul(f.defv(F, "F"), k_inf), Pr), f.add(Pr, f.c(1.))};
    }
  }
  throw ast::Error("unknown reaction model");
}

// reaction.rs:207-310: builds the net rate of progress of each reaction in turn, caching the clamped
// exponentials of the standard Gibbs energies across reactions.
class Reactions {
 public:
  Reactions(Block& f, const Mechanism& m, const Temperatures& T, E concentration, const std::vector<E>& C, bool lean = false)
      : f_(f), m_(m), T_(T), concentration_(concentration), C_(C), lean_(lean), g_(m.active, -1), exp_(m.active, -1), exp_neg_(m.active, -1) {
    rcp_C0_ = f.defv(f.mul(f.c(1. / NASA7::reference_pressure), T.T), "RT/P0");
    C0_ = f.defv(f.mul(f.c(NASA7::reference_pressure), T.rcpT), "P0/RT");
  }

  E operator()(const Reaction& r) {
    Block& f = f_;
    Ratio k = forward_rate_constant(f, r, T_, concentration_, C_);
    if (lean_) return lean(r, k);
    E Rf = product_of_exponentiations(f, C_, r.reactants, "Rf");
    E Rnet = Rf;
    if (r.model != ReactionModel::Irreversible) {
      std::optional<E> p;
      for (size_t s = 0; s < r.net.size(); s++) {
        int n = r.net[s];
        if (n == 0) continue;
        E t = n > 0 ? exp(s) : exp_neg(s);
        for (int i = 0; i < std::abs(n); i++) p = p ? f.mul(*p, t) : t;
      }
      if (!p) throw ast::Error("reversible reaction without net change: " + r.equation);
      E rcp_K;
      switch (-r.sum_net) {
        case 0: rcp_K = *p; break;
        case 1: rcp_K = f.defv(f.mul(C0_, *p), "K"); break;
        case -1: rcp_K = f.defv(f.mul(rcp_C0_, *p), "K"); break;
        case -2: rcp_K = f.defv(f.mul(f.mul(rcp_C0_, rcp_C0_), *p), "K"); break;
        default: throw ast::Error("unsupported change in the number of moles: " + r.equation);
      }
      E Rr = f.defv(f.mul(rcp_K, product_of_exponentiations(f, C_, r.products, "Rr")), "Rr/K");
      Rnet = f.defv(f.sub(Rf, Rr), "R");
    }
    E q = f.mul(Rnet, k.n);
    if (k.d && !(f.f64(*k.d) && *f.f64(*k.d) == 1.)) q = f.div(q, *k.d);
    return f.defv(q, "cR");
  }

 private:
  // The same rate of progress with every product started from the reaction's own rate constant:
  // q = k ΠC^ν' − k (1/K) ΠC^ν''. A product of concentrations or of exp(±G/RT) factors that two reactions have in
  // common (C_H C_O2, say) is then not one shared value that stays live between its first and last use (about 70
  // such values at K=53), at no extra multiplication; the rounding order differs from the reference's by a few ulps.
  E lean(const Reaction& r, const Ratio& k) {
    Block& f = f_;
    E A = k.n;
    for (size_t s = 0; s < C_.size(); s++)
      for (int i = 0; i < r.reactants[s]; i++) A = f.mul(A, C_[s]);
    E q = A;
    if (r.model != ReactionModel::Irreversible) {
      E B = k.n;
      switch (-r.sum_net) {
        case 0: break;
        case 1: B = f.mul(B, C0_); break;
        case -1: B = f.mul(B, rcp_C0_); break;
        case -2: B = f.mul(f.mul(B, rcp_C0_), rcp_C0_); break;
        default: throw ast::Error("unsupported change in the number of moles: " + r.equation);
      }
      bool net = false;
      for (size_t s = 0; s < r.net.size(); s++) {
        const int n = r.net[s];
        for (int i = 0; i < std::abs(n); i++) B = f.mul(B, n > 0 ? exp(s) : exp_neg(s));
        net = net || n != 0;
      }
      if (!net) throw ast::Error("reversible reaction without net change: " + r.equation);
      for (size_t s = 0; s < C_.size(); s++)
        for (int i = 0; i < r.products[s]; i++) B = f.mul(B, C_[s]);
      q = f.sub(A, B);
    }
    if (k.d && !(f.f64(*k.d) && *f.f64(*k.d) == 1.)) q = f.div(q, *k.d);
    return f.defv(q, "cR");
  }

  Block& f_;
  const Mechanism& m_;
  Temperatures T_;
  E concentration_;
  const std::vector<E>& C_;
  bool lean_;
  E rcp_C0_, C0_;
  std::vector<E> g_, exp_, exp_neg_;

  E gibbs(size_t k) {
    if (g_[k] >= 0) return g_[k];
    Block& f = f_;
    const NASA7& n = m_.species.thermodynamics[k];
    E e;
    if (std::isnan(n.temperature_split)) {
      auto a = gibbs0_RT(n.pieces[0]);
      e = dot(f, consts(f, a.data(), 7), gibbs_basis(f, T_));
    } else if (lean_) {
      // Lean form: both pieces evaluated, then the value selected (the same value as selecting the coefficients
      // first, reaction.rs:85-107): each piece's constants enter its FMAs as scalar operands, where selecting seven
      // coefficient pairs per species costs a VGPR move and a select per 32-bit half.
      auto a = gibbs0_RT(n.pieces[0]), b = gibbs0_RT(n.pieces[1]);
      const std::vector<E> basis = gibbs_basis(f, T_);
      E low = dot(f, consts(f, a.data(), 7), basis), high = dot(f, consts(f, b.data(), 7), basis);
      const int32_t id = f.value("Gibbs0_RT" + std::to_string(k));
      f.select(f.le(T_.T, f.c(n.temperature_split)), {low}, {high}, {id});
      return g_[k] = f.v(id);
    } else {
      auto a = gibbs0_RT(n.pieces[0]), b = gibbs0_RT(n.pieces[1]);
      std::vector<E> t, fe, piece;
      std::vector<int32_t> results;
      for (int i = 0; i < 7; i++) {
        int32_t id = f.value("g" + std::to_string(k) + "_" + std::to_string(i));
        results.push_back(id);
        piece.push_back(f.v(id));
        t.push_back(f.c(a[i]));
        fe.push_back(f.c(b[i]));
      }
      f.select(f.le(T_.T, f.c(n.temperature_split)), t, fe, results);
      e = dot(f, piece, gibbs_basis(f, T_));
    }
    return g_[k] = f.defv(e, "Gibbs0_RT" + std::to_string(k));
  }
  E exp(size_t k) {
    if (exp_[k] < 0) {
      if (lean_) lean_exponentials(k);
      else exp_[k] = f_.defv(f_.min(f_.c(exp_clamp), f_.exp(gibbs(k))), "expGibbs0_RT" + std::to_string(k));
    }
    return exp_[k];
  }
  E exp_neg(size_t k) {
    if (exp_neg_[k] < 0) {
      if (lean_) lean_exponentials(k);
      else exp_neg_[k] = f_.defv(f_.min(f_.c(exp_neg_clamp), f_.exp(f_.neg(gibbs(k)))), "exp_negGibbs0_RT" + std::to_string(k));
    }
    return exp_neg_[k];
  }
  // Lean form: one exponential per species, G = exp(G_k/RT), and exp(-G_k/RT) as 1/G (a reciprocal instead of a
  // second exponential; the same value to a few ulps, and the same clamped results where either side overflows).
  // Both clamped values are defined together, so that G is dead right after them.
  void lean_exponentials(size_t k) {
    Block& f = f_;
    E G = f.defv(f.exp(gibbs(k)), "expGibbs0_RT_unclamped" + std::to_string(k));
    exp_[k] = f.defv(f.min(f.c(exp_clamp), G), "expGibbs0_RT" + std::to_string(k));
    exp_neg_[k] = f.defv(f.min(f.c(exp_neg_clamp), f.div(f.c(1.), G)), "exp_negGibbs0_RT" + std::to_string(k));
  }
};

// reaction.rs:362-373 after the inputs are in scope: temperature powers and concentrations.
struct Prologue {
  Temperatures T;
  E concentration;
  std::vector<E> C;  // nonbulk then bulk
};
Prologue prologue(Block& f, E T, E pressure_R, E molar_density, const std::vector<E>& nonbulk_amounts) {
  Prologue p;
  E rcpT = f.defv(f.div(f.c(1.), T), "rcpT");
  E lnT = f.defv(f.ln(T), "lnT");
  E T2 = f.defv(f.mul(T, T), "T2");
  std::vector<E> nonbulk;
  for (E n : nonbulk_amounts) nonbulk.push_back(f.defv(f.mul(molar_density, f.max(f.c(0.), n)), "nonbulk_concentrations"));
  E T3 = f.defv(f.mul(T, T2), "T3");
  p.concentration = f.defv(f.mul(pressure_R, rcpT), "concentration");
  E rcpT2 = f.defv(f.mul(rcpT, rcpT), "rcpT2");
  E T4 = f.defv(f.mul(T2, T2), "T4");
  p.T = {lnT, T, T2, T3, T4, rcpT, rcpT2};
  E bulk = f.defv(f.sub(p.concentration, f.sum(nonbulk)), "bulk_concentration");
  p.C = nonbulk;
  p.C.push_back(bulk);
  return p;
}

std::vector<std::string> names(std::initializer_list<const char*> head, size_t n, const char* array) {
  std::vector<std::string> v(head.begin(), head.end());
  for (size_t i = 0; i < n; i++) v.push_back(std::string(array) + "[" + std::to_string(i) + "]");
  return v;
}

}  // namespace

std::vector<int32_t> gibbs_exponentials(const ast::Linear& l, bool negative) {
  const double bound = negative ? exp_neg_clamp : exp_clamp;
  std::vector<char> clamp(l.n_registers, 0);
  std::vector<int32_t> out;
  for (const ast::Instruction& i : l.code) {
    if (i.op == ast::LOp::Const && i.k == bound) clamp[i.dst] = 1;
    if (i.op == ast::LOp::Min && (clamp[i.a] || clamp[i.b])) out.push_back(i.dst);
  }
  return out;
}

std::vector<E> species_rates(Block& f, const Mechanism& m, const Temperatures& T, E concentration, const std::vector<E>& C,
                             bool lean) {
  std::vector<std::optional<E>> sums(m.active);
  Reactions cR(f, m, T, concentration, C, lean);
  for (const Reaction& r : m.reactions) {
    E q = cR(r);
    for (size_t k = 0; k < m.active; k++) {
      int nu = r.net[k];
      if (nu == 0) continue;
      const std::string name = "rates_" + m.species.names[k];
      auto& s = sums[k];
      if (nu == 1)
        s = s ? f.defv(f.add(*s, q), name) : q;
      else if (nu == -1)
        s = s ? f.defv(f.sub(*s, q), name) : f.defv(f.neg(q), name);
      else {
        E nq = f.defv(f.mul(f.c((double)nu), q), "νcR");
        s = s ? f.defv(f.add(*s, nq), name) : nq;
      }
    }
  }
  std::vector<E> out;
  for (size_t k = 0; k < m.active; k++) {
    if (!sums[k]) throw ast::Error("active species without net production: " + m.species.names[k]);
    out.push_back(*sums[k]);
  }
  return out;
}

// Reactor right-hand side of reaction.rs:375-386 from the rates and the temperature powers.
static std::pair<E, E> energy_terms(Block& f, const Mechanism& m, const Temperatures& T, const std::vector<E>& C, const std::vector<E>& omega) {
  std::vector<NASA7> active(m.species.thermodynamics.begin(), m.species.thermodynamics.begin() + m.active);
  std::vector<E> h = thermodynamics(f, active, enthalpy_RT, T, "enthalpy_RT");
  E energy_rate_RT = dot(f, omega, h);
  std::vector<E> cp = thermodynamics(f, m.species.thermodynamics, molar_heat_capacity_R, T, "molar_heat_capacity_at_CP_R");
  E Cp = dot(f, cp, C);
  E dtT_T = f.defv(f.div(f.neg(energy_rate_RT), Cp), "dT_T");
  const double W_bulk = m.species.molar_mass.back();
  std::vector<E> reduced;
  for (size_t k = 0; k < m.species.len(); k++) reduced.push_back(f.c(1. - m.species.molar_mass[k] / W_bulk));
  E dtE = dot(f, reduced, omega);
  return {dtT_T, dtE};
}

ast::Function rates(const Mechanism& m) {
  const size_t K = m.species.len();
  Block f(names({"pressure_", "rcp_pressure_", "T", "volume"}, K - 1, "active_amounts"));
  E pressure_R = f.v(0), rcp_pressure_R = f.v(1), T = f.v(2), volume = f.v(3);
  std::vector<E> amounts;
  for (size_t k = 0; k + 1 < K; k++) amounts.push_back(f.v(4 + (int32_t)k));
  E molar_density = f.defv(f.div(f.c(1.), volume), "molar_density");
  Prologue p = prologue(f, T, pressure_R, molar_density, amounts);
  std::vector<E> omega = species_rates(f, m, p.T, p.concentration, p.C);
  auto [dtT_T, dtE] = energy_terms(f, m, p.T, p.C, omega);
  E dtV = f.mul(volume, f.add(dtT_T, f.mul(f.mul(rcp_pressure_R, T), dtE)));
  std::vector<E> out = {f.mul(T, dtT_T), dtV};
  out.insert(out.end(), omega.begin(), omega.end());
  return f.finish(out);
}

// The rates function of the device-function export (cuda.rs:210-226): ω from the temperature powers and all K
// concentrations.
ast::Function rates_device_function(const Mechanism& m) {
  const size_t K = m.species.len();
  Block f(names({"pressure_R", "lnT", "T", "T2", "T3", "T4", "rcpT"}, K, "concentrations"));
  Temperatures T{f.v(1), f.v(2), f.v(3), f.v(4), f.v(5), f.v(6), -1};
  T.rcpT2 = f.defv(f.mul(T.rcpT, T.rcpT), "rcpT2");
  E concentration = f.defv(f.mul(f.v(0), T.rcpT), "concentration");
  std::vector<E> C;
  for (size_t k = 0; k < K; k++) C.push_back(f.v(7 + (int32_t)k));
  return f.finish(species_rates(f, m, T, concentration, C));
}

ast::Function reactions_rates(const Mechanism& m) {
  const size_t K = m.species.len();
  Block f(names({"pressure_", "T", "volume"}, K - 1, "active_amounts"));
  E pressure_R = f.v(0), T = f.v(1), volume = f.v(2);
  std::vector<E> amounts;
  for (size_t k = 0; k + 1 < K; k++) amounts.push_back(f.v(3 + (int32_t)k));
  E molar_density = f.defv(f.div(f.c(1.), volume), "molar_density");
  Prologue p = prologue(f, T, pressure_R, molar_density, amounts);
  Reactions cR(f, m, p.T, p.concentration, p.C);
  std::vector<E> out;
  for (const Reaction& r : m.reactions) out.push_back(cR(r));
  return f.finish(out);
}

// Transport section of the per-state program (transport.cpp).
std::vector<E> transport_outputs(Block& f, const Species& s, const Polynomials& p, const TransportScales& scales, E N, E theta, bool lean,
                                 bool f32_range, const std::vector<E>& x, E& mean_molar_massN);

ast::Function chemistry(const Mechanism& m, const Polynomials* poly, const TransportScales& scales, const ChemistryOptions& o) {
  const size_t K = m.species.len();
  if (K < 2) throw ast::Error("need at least two species");
  if (o.transport && !poly) throw ast::Error("transport requested without fits");
  Block f(names({"T", "P"}, K, "Y"));
  E T = f.v(0), P = f.v(1);
  // Mass-basis amounts n_k = Y_k/W_k and molar density 1/V = (P/R)/T / Σ n_k.
  std::vector<E> n;
  for (size_t k = 0; k < K; k++) n.push_back(f.defv(f.mul(f.v(2 + (int32_t)k), f.c(1. / m.species.molar_mass[k])), "amounts"));
  E pressure_R = f.defv(f.mul(P, f.c(1. / R)), "pressure_R");
  E sum_n = f.defv(f.sum(n), "Σamounts");
  E rcpT0 = f.defv(f.div(f.c(1.), T), "rcpT");  // the same definition the prologue reuses
  E molar_density = f.defv(f.div(f.mul(pressure_R, rcpT0), sum_n), "molar_density");
  std::vector<E> nonbulk(n.begin(), n.end() - 1);
  Prologue p = prologue(f, T, pressure_R, molar_density, nonbulk);
  std::vector<E> out;
  std::vector<E> omega;
  if (o.rates || o.rhs) omega = species_rates(f, m, p.T, p.concentration, p.C, o.lean_rates);
  if (o.rates) out.insert(out.end(), omega.begin(), omega.end());
  if (o.transport) {
    E theta = f.defv(f.mul(T, f.c(1. / scales.T0)), "T/T0");
    E W_N;
    std::vector<E> tr = transport_outputs(f, m.species, *poly, scales, p.concentration, theta, o.lean_transport, o.f32_range,
                                          std::vector<E>(p.C.begin(), p.C.end() - 1), W_N);
    out.insert(out.end(), tr.begin(), tr.end());
  }
  if (o.rhs) {
    auto [dtT_T, dtE] = energy_terms(f, m, p.T, p.C, omega);
    (void)dtE;
    out.push_back(f.mul(T, dtT_T));
  }
  return f.finish(out);
}

// The constant-pressure reactor's right-hand side in the variables a batch integrator steps (the role of the
// reference's cvode/lib.rs:17-47 right-hand side and plot.rs:80-118): inputs [T, P, Y_0..Y_{K-1}], outputs
// [dT/dt, dY_0/dt..dY_{K-1}/dt] with dT/dt the energy balance of reaction.rs:375-380 and dY_k/dt = W_k ω_k / ρ,
// ρ = P / (R T Σ_j Y_j/W_j) (zero for the species that do not react).
ast::Function reactor(const Mechanism& m, bool lean) {
  const size_t K = m.species.len();
  if (K < 2) throw ast::Error("need at least two species");
  Block f(names({"T", "P"}, K, "Y"));
  E T = f.v(0), P = f.v(1);
  std::vector<E> n;
  for (size_t k = 0; k < K; k++) n.push_back(f.defv(f.mul(f.v(2 + (int32_t)k), f.c(1. / m.species.molar_mass[k])), "amounts"));
  E pressure_R = f.defv(f.mul(P, f.c(1. / R)), "pressure_R");
  E sum_n = f.defv(f.sum(n), "Σamounts");
  E rcpT0 = f.defv(f.div(f.c(1.), T), "rcpT");
  E molar_density = f.defv(f.div(f.mul(pressure_R, rcpT0), sum_n), "molar_density");
  std::vector<E> nonbulk(n.begin(), n.end() - 1);
  Prologue p = prologue(f, T, pressure_R, molar_density, nonbulk);
  std::vector<E> omega = species_rates(f, m, p.T, p.concentration, p.C, lean);
  auto [dtT_T, dtE] = energy_terms(f, m, p.T, p.C, omega);
  (void)dtE;
  E rcp_rho = f.defv(f.div(f.mul(T, sum_n), pressure_R), "1/ρ");
  std::vector<E> out = {f.mul(T, dtT_T)};
  for (size_t k = 0; k < K; k++)
    out.push_back(k < m.active ? f.mul(f.mul(f.c(m.species.molar_mass[k]), omega[k]), rcp_rho) : f.c(0.));
  return f.finish(out);
}

}  // namespace combustion
