// Transport: kinetic-theory species properties, their polynomial fits, and the mixture-averaged program.
// Restates transport.rs:14-214 (and its collision-integral tables, transport/collision_integrals.rs:1-82).
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include "builders.hpp"
#include "collision_integrals.hpp"

namespace combustion {

using ast::Block;
using ast::E;

namespace {

constexpr double pi = 3.14159265358979323846;
constexpr double light_speed = 299792458.;
constexpr double fine_structure = 7.2973525693e-3;
constexpr double Planck = 6.62607015e-34;
constexpr double electron_charge = 1.602176634e-19;
const double mu0 = 2. * fine_structure * Planck / (electron_charge * electron_charge * light_speed);  // transport.rs:43
const double epsilon0 = 1. / (light_speed * light_speed * mu0);                                       // transport.rs:44

double sq(double x) { return x * x; }
double cb(double x) { return x * x * x; }
double powi(double x, int k) {
  double y = 1.;
  for (int i = 0; i < k; i++) y *= x;
  return y;
}

// Least squares min ||A c - b|| by Householder QR (A: n x d, row major). The reference solves the same
// problem by SVD (transport.rs:24-31); both return the unique minimiser of a full-rank problem.
std::vector<double> least_squares(std::vector<double> A, std::vector<double> b, int n, int d) {
  for (int j = 0; j < d; j++) {
    double norm = 0;
    for (int i = j; i < n; i++) norm += A[i * d + j] * A[i * d + j];
    norm = std::sqrt(norm);
    if (norm == 0) throw std::runtime_error("rank-deficient fit");
    double alpha = A[j * d + j] > 0 ? -norm : norm;
    std::vector<double> v(n, 0.);
    for (int i = j; i < n; i++) v[i] = A[i * d + j];
    v[j] -= alpha;
    double vv = 0;
    for (int i = j; i < n; i++) vv += v[i] * v[i];
    if (vv == 0) continue;
    for (int c = j; c < d; c++) {
      double s = 0;
      for (int i = j; i < n; i++) s += v[i] * A[i * d + c];
      s = 2 * s / vv;
      for (int i = j; i < n; i++) A[i * d + c] -= s * v[i];
    }
    double s = 0;
    for (int i = j; i < n; i++) s += v[i] * b[i];
    s = 2 * s / vv;
    for (int i = j; i < n; i++) b[i] -= s * v[i];
  }
  std::vector<double> x(d);
  for (int j = d - 1; j >= 0; j--) {
    double s = b[j];
    for (int c = j + 1; c < d; c++) s -= A[j * d + c] * x[c];
    x[j] = s / A[j * d + j];
  }
  return x;
}

// transport.rs:24-31 weighted_regression: rows scaled by w.
std::vector<double> weighted_regression(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& w, int D) {
  const int n = (int)x.size();
  std::vector<double> A(n * D), b(n);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < D; k++) A[i * D + k] = w[i] * powi(x[i], k);
    b[i] = w[i] * y[i];
  }
  return least_squares(A, b, n, D);
}

// transport.rs:48-51: degree-6 fit in δ* of each T* row, evaluated to degree 5.
template <int N>
std::array<std::array<double, 6>, N> delta_fits(const double (&table)[N][8]) {
  std::array<std::array<double, 6>, N> out;
  std::vector<double> x(collision::header_delta_star, collision::header_delta_star + 8), w(8, 1.);
  for (int r = 0; r < N; r++) {
    std::vector<double> y(table[r], table[r] + 8);
    auto p = weighted_regression(x, y, w, 7);
    for (int i = 0; i < 6; i++) out[r][i] = p[i];
  }
  return out;
}
const auto omega22_fits = delta_fits<37>(collision::omega22);
const auto a_star_fits = delta_fits<39>(collision::a_star);

// transport.rs:14-17
double quadratic_interpolation(const double x[3], const double y[3], double x0) {
  return ((x[1] - x[0]) * (y[2] - y[1]) - (y[1] - y[0]) * (x[2] - x[1])) / ((x[1] - x[0]) * (x[2] - x[0]) * (x[2] - x[1])) * (x0 - x[0]) * (x0 - x[1]) +
         ((y[1] - y[0]) / (x[1] - x[0])) * (x0 - x[1]) + y[1];
}

struct Physics {
  const Species& s;
  // transport.rs:58-61
  double reduced_mass(size_t a, size_t b) const {
    const auto& W = s.molar_mass;
    return W[a] / NA * W[b] / NA / (W[a] / NA + W[b] / NA);
  }
  // transport.rs:62-68
  double chi(size_t a, size_t b) const {
    const auto& mu = s.permanent_dipole_moment;
    if ((mu[a] > 0.) == (mu[b] > 0.)) return 1.;
    size_t polar = mu[a] != 0. ? a : b, non_polar = mu[a] != 0. ? b : a;
    return 1. + 1. / 4. * s.polarizability[non_polar] / cb(s.diameter[non_polar]) *
                    sq(mu[polar] / std::sqrt(4. * pi * epsilon0 * s.well_depth_J[polar] * cb(s.diameter[polar]))) *
                    std::sqrt(s.well_depth_J[polar] / s.well_depth_J[non_polar]);
  }
  double interaction_well_depth(size_t a, size_t b) const { return std::sqrt(s.well_depth_J[a] * s.well_depth_J[b]) * sq(chi(a, b)); }
  double T_star(size_t a, size_t b, double T) const { return T * kB / interaction_well_depth(a, b); }
  double reduced_dipole_moment(size_t a, size_t b) const {
    return s.permanent_dipole_moment[a] * s.permanent_dipole_moment[b] /
           (8. * pi * epsilon0 * std::sqrt(s.well_depth_J[a] * s.well_depth_J[b]) * cb((s.diameter[a] + s.diameter[b]) / 2.));
  }
  double reduced_diameter(size_t a, size_t b) const { return (s.diameter[a] + s.diameter[b]) / 2. * std::pow(chi(a, b), -1. / 6.); }
  // transport.rs:82-96
  template <int I0, int N>
  double collision_integral(const double (&table)[N][8], const std::array<std::array<double, 6>, N>& fit, size_t a, size_t b, double T) const {
    const double ln_T_star = std::log(T_star(a, b, T));
    double header_ln[39];
    for (int i = 0; i < 39; i++) header_ln[i] = std::log(collision::header_T_star[i]);
    int p = -1;
    for (int i = 1; i < 39; i++)
      if (ln_T_star < header_ln[i]) {
        p = i - 1;
        break;
      }
    if (p < 0) throw std::runtime_error("reduced temperature above the collision-integral table");
    int start = std::min(p, I0 + N - 3);
    if (start - I0 < 0) throw std::runtime_error("reduced temperature below the collision-integral table");
    const double* h = header_ln + start;
    if (!(h[0] <= ln_T_star && ln_T_star <= h[2])) throw std::runtime_error("reduced temperature outside the interpolation stencil");
    const double delta = reduced_dipole_moment(a, b);
    if (!(collision::header_delta_star[0] <= delta && delta <= collision::header_delta_star[7]))
      throw std::runtime_error("reduced dipole moment outside the collision-integral table");
    double y[3];
    for (int r = 0; r < 3; r++) {
      if (delta == 0.) {
        y[r] = table[start - I0 + r][0];
      } else {
        const auto& P = fit[start - I0 + r];
        double v = 0;
        for (int k = 0; k < 6; k++) v += P[k] * powi(delta, k);
        y[r] = v;
      }
    }
    double image = quadratic_interpolation(h, y, ln_T_star);
    if (!(image > 0.)) throw std::runtime_error("non-positive collision integral");
    return image;
  }
  double omega22(size_t a, size_t b, double T) const { return collision_integral<1, 37>(collision::omega22, omega22_fits, a, b, T); }
  double viscosity(size_t k, double T) const {
    return 5. / 16. * std::sqrt(pi * s.molar_mass[k] / NA * kB * T) / (omega22(k, k, T) * pi * sq(s.diameter[k]));
  }
  double omega11(size_t a, size_t b, double T) const {
    return omega22(a, b, T) / collision_integral<0, 39>(collision::a_star, a_star_fits, a, b, T);
  }
  double diffusivity(size_t a, size_t b, double T) const {
    return 3. / 16. * std::sqrt(2. * pi / reduced_mass(a, b)) * std::pow(kB * T, 3. / 2.) / (pi * sq(reduced_diameter(a, b)) * omega11(a, b, T));
  }
  // transport.rs:106-116
  double conductivity(size_t k, double T) const {
    const double f_internal = s.molar_mass[k] / NA / (kB * T) * diffusivity(k, k, T) / viscosity(k, T);
    auto fz = [](double Ts) { return 1. + std::pow(pi, 3. / 2.) / std::sqrt(Ts) * (1. / 2. + 1. / Ts) + (1. / 4. * sq(pi) + 2.) / Ts; };
    const double dof = s.internal_degrees_of_freedom[k];
    const double c1 = 2. / pi * (5. / 2. - f_internal) /
                      (s.rotational_relaxation[k] * fz(T_star(k, k, 298.)) / fz(T_star(k, k, T)) + 2. / pi * (5. / 3. * dof + f_internal));
    const double f_translation = 5. / 2. * (1. - c1 * dof / (3. / 2.));
    const double f_rotation = f_internal * (1. + c1);
    const double Cv_internal = s.thermodynamics[k].molar_heat_capacity_at_constant_pressure_R(T) - 5. / 2. - dof;
    return (viscosity(k, T) / (s.molar_mass[k] / NA)) * kB * (f_translation * 3. / 2. + f_rotation * dof + f_internal * Cv_internal);
  }
};

}  // namespace

namespace transport {
double viscosity(const Species& s, size_t k, double T) { return Physics{s}.viscosity(k, T); }
double diffusivity(const Species& s, size_t a, size_t b, double T) { return Physics{s}.diffusivity(a, b, T); }
double conductivity(const Species& s, size_t k, double T) { return Physics{s}.conductivity(k, T); }
double omega22(const Species& s, size_t a, size_t b, double T) { return Physics{s}.omega22(a, b, T); }
double omega11(const Species& s, size_t a, size_t b, double T) { return Physics{s}.omega11(a, b, T); }
}  // namespace transport

// transport.rs:125-136
Polynomials fit_polynomials(const Species& species, double T0, int D) {
  if (D < 2 || D > 8) throw std::runtime_error("polynomial length must be in [2, 8]");
  Physics ph{species};
  const size_t K = species.len();
  const int N = 50;
  std::vector<double> T(N), x(N);
  for (int n = 0; n < N; n++) {
    T[n] = 300. + (double)n / (double)(N - 1) * (3000. - 300.);
    x[n] = std::log(T[n] / T0);
  }
  auto fit = [&](auto&& y_of_T) {
    std::vector<double> y(N), w(N);
    for (int n = 0; n < N; n++) {
      y[n] = y_of_T(T[n]);
      w[n] = 1. / sq(y[n]);
    }
    return weighted_regression(x, y, w, D);
  };
  Polynomials p;
  p.D = D;
  p.K = K;
  p.conductivityIVT.resize(K * D);
  p.VviscosityIVVT.resize(K * D);
  p.diffusivityITVT.assign(K * K * D, 0.);
  for (size_t k = 0; k < K; k++) {
    auto c = fit([&](double t) { return ph.conductivity(k, t) / std::sqrt(t); });
    auto v = fit([&](double t) { return std::sqrt(ph.viscosity(k, t) / std::sqrt(t)); });
    for (int i = 0; i < D; i++) {
      p.conductivityIVT[k * D + i] = c[i];
      p.VviscosityIVVT[k * D + i] = v[i];
    }
  }
  for (size_t k = 0; k < K; k++)
    for (size_t j = 0; j < k; j++) {
      auto d = fit([&](double t) { return ph.diffusivity(k, j, t) / (t * std::sqrt(t)); });
      for (int i = 0; i < D; i++) {
        p.diffusivityITVT[(k * K + j) * D + i] = d[i];
        p.diffusivityITVT[(j * K + k) * D + i] = d[i];
      }
    }
  return p;
}

TransportScales transport_scales(const Mechanism& m, double T0) {
  const Species& s = m.species;
  const size_t K = s.len();
  std::vector<double> X(K, 0.);
  double total = 0;
  for (size_t k = 0; k < K; k++) total += m.initial_state.amounts[k];
  if (!(total > 0)) {
    X[K - 1] = 1.;  // no listed initial species: the bulk species alone
  } else {
    for (size_t k = 0; k < K; k++) X[k] = m.initial_state.amounts[k] / total;
  }
  double W = 0, cp = 0;
  for (size_t k = 0; k < K; k++) {
    W += s.molar_mass[k] * X[k];
    cp += s.thermodynamics[k].molar_heat_capacity_at_constant_pressure_R(T0) * X[k];
  }
  const double concentration = 101325. / R / T0;
  const double viscosity = concentration * W * 1.;
  const double conductivity = cp * R / W * viscosity;
  return {T0, viscosity, conductivity};
}

namespace {

E dot_consts(Block& f, const double* P, const std::vector<E>& basis) {
  std::vector<E> terms;
  for (size_t m = 0; m < basis.size(); m++) terms.push_back(f.mul(f.c(P[m]), basis[m]));
  return f.sum(terms);
}

// transport.rs:141-146
E conductivityNIVT(Block& f, const std::vector<double>& P, int D, E N, const std::vector<E>& lnT, const std::vector<E>& x) {
  std::optional<E> A, B;
  for (size_t k = 0; k < x.size(); k++) {
    E c = f.defv(dot_consts(f, &P[k * D], lnT), "c" + std::to_string(k));
    E a = f.mul(x[k], c), b = f.div(x[k], c);
    if (!A) {
      A = a;
      B = b;
    } else {
      A = f.defv(f.add(*A, a), "a" + std::to_string(k));
      B = f.defv(f.add(*B, b), "b" + std::to_string(k));
    }
  }
  return f.add(f.div(*A, N), f.div(N, *B));
}

// Σ_m P[m] lnT^m in Horner form (lnT[1] = ln T): a node distinct from dot_consts' power-basis sum.
E horner(Block& f, const double* P, int D, const std::vector<E>& lnT) {
  E e = f.c(P[D - 1]);
  for (int m = D - 2; m >= 0; m--) e = f.add(f.mul(e, lnT[1]), f.c(P[m]));
  return e;
}

// transport.rs:148-158
E viscosityIVT(Block& f, const std::vector<double>& W, const std::vector<double>& P, int D, const std::vector<E>& lnT, const std::vector<E>& x,
               bool lean) {
  const size_t K = x.size();
  std::vector<E> v, r;
  if (lean && D >= 2) {
    // ṽ_k = v_k W_k^(-1/4): the fit's coefficients scaled at build time, so that Wilke's term for the pair (k, j)
    // is a_kj² (1 + ṽ_k/ṽ_j)² with a_kj² = 1/√8 /√(1 + W_k/W_j), and v_k² = √W_k ṽ_k² in the numerator.
    std::vector<double> Q(P.begin(), P.begin() + K * D);
    for (size_t k = 0; k < K; k++)
      for (int m = 0; m < D; m++) Q[k * D + m] *= 1. / std::sqrt(std::sqrt(W[k]));
    for (size_t k = 0; k < K; k++) r.push_back(f.defv(f.div(f.c(1.), dot_consts(f, &Q[k * D], lnT)), "rcp_VviscosityIVVT" + std::to_string(k)));
    for (size_t k = 0; k < K; k++) v.push_back(horner(f, &Q[k * D], D, lnT));
    std::vector<E> outer;
    for (size_t k = 0; k < K; k++) {
      std::vector<E> inner;
      for (size_t j = 0; j < K; j++) {
        const double a2 = 1. / std::sqrt(8.) * 1. / std::sqrt(1. + W[k] / W[j]);
        // x_j (1 + ṽ_k/ṽ_j)² first: x_j a_kj² would be one node shared by every row k of the same molar mass (CH2 and
        // CH2(S), say), which would keep the K products of the first such row live until the last one.
        inner.push_back(f.mul(f.mul(x[j], f.sq(f.add(f.c(1.), f.mul(v[k], r[j])))), f.c(a2)));
      }
      outer.push_back(f.div(f.mul(f.mul(x[k], f.c(std::sqrt(W[k]))), f.sq(v[k])), f.sum(inner)));
    }
    return f.sum(outer);
  } else {
    for (size_t k = 0; k < K; k++) v.push_back(f.defv(dot_consts(f, &P[k * D], lnT), "VviscosityIVVT" + std::to_string(k)));
    for (size_t k = 0; k < K; k++) r.push_back(f.defv(f.div(f.c(1.), v[k]), "rcp_VviscosityIVVT" + std::to_string(k)));
  }
  std::vector<E> outer;
  for (size_t k = 0; k < K; k++) {
    std::vector<E> inner;
    for (size_t j = 0; j < K; j++) {
      const double Va = std::sqrt(1. / std::sqrt(8.) * 1. / std::sqrt(1. + W[k] / W[j]));
      const double Vb = Va * std::sqrt(std::sqrt(W[j] / W[k]));
      inner.push_back(f.mul(x[j], f.sq(f.add(f.c(Va), f.mul(f.mul(f.c(Vb), v[k]), r[j])))));
    }
    outer.push_back(f.div(f.mul(x[k], f.sq(v[k])), f.sum(inner)));
  }
  return f.sum(outer);
}

// transport.rs:162-177
std::vector<E> density_diffusivity(Block& f, const std::vector<double>& W, const std::vector<double>& P, int D, E mean_molar_mass_VTN, E VT,
                                   const std::vector<E>& lnT, const std::vector<E>& x, bool lean = false, bool f32_range = false) {
  const size_t K = x.size();
  std::vector<std::optional<E>> S(K);
  auto accumulate = [&](size_t k, size_t j, E Rkj) {
    E t = f.mul(x[j], Rkj);
    S[k] = f.defv(S[k] ? f.add(*S[k], t) : t, "S" + std::to_string(k) + "_" + std::to_string(j));
    t = f.mul(x[k], Rkj);
    S[j] = f.defv(S[j] ? f.add(*S[j], t) : t, "S" + std::to_string(j) + "_" + std::to_string(k));
  };
  // Lean form: 1/(c0 + c1 ℓ + …) as (1/c0)/y with y = 1 + (c1/c0) ℓ + … (about 1 over the fits' range), and the
  // reciprocals of kBatch consecutive pairs' y from one reciprocal of their product (prefix products, then
  // back substitution: 3 multiplications per pair instead of a v_rcp_f64 and its Newton step).
  constexpr size_t kBatch = 4;  // 8 measured 1 % faster, within noise, at 255 VGPRs instead of 243
  struct Pending {
    size_t k, j;
    E y;
    double scale;
  };
  std::vector<Pending> batch;
  auto flush = [&]() {
    if (batch.empty()) return;
    const size_t n = batch.size();
    std::vector<E> prefix = {batch[0].y};
    for (size_t i = 1; i < n; i++) prefix.push_back(f.mul(prefix[i - 1], batch[i].y));
    E inv = f.div(f.c(1.), prefix[n - 1]);
    std::vector<E> R(n);
    for (size_t i = n - 1; i > 0; i--) {
      R[i] = f.mul(inv, prefix[i - 1]);
      inv = f.mul(inv, batch[i].y);
    }
    R[0] = inv;
    for (size_t i = 0; i < n; i++)
      accumulate(batch[i].k, batch[i].j,
                 f.defv(f.mul(f.c(batch[i].scale), R[i]), "R" + std::to_string(batch[i].k) + "_" + std::to_string(batch[i].j)));
    batch.clear();
  };
  for (size_t k = 0; k < K; k++)
    for (size_t j = 0; j < k; j++) {
      const double* c = &P[(k * K + j) * D];
      if (lean && D >= 2 && c[0] != 0.) {
        std::vector<E> terms = {f.c(1.)};
        for (int m = 1; m < D; m++) terms.push_back(f.mul(f.c(c[m] / c[0]), lnT[m]));
        batch.push_back({k, j, f.sum(terms), 1. / c[0]});
        if (batch.size() == kBatch) flush();
      } else {
        flush();
        accumulate(k, j, f.defv(f.div(f.c(1.), dot_consts(f, c, lnT)), "R" + std::to_string(k) + "_" + std::to_string(j)));
      }
    }
  flush();
  std::vector<E> out;
  for (size_t k = 0; k < K; k++) {
    // FP32 evaluation: S_k of a species whose partners are all traces below ~1e-38 underflows to 0 (x/0 = ±inf);
    // FLT_MIN instead keeps D_k finite, within the error floor 1/S_k of such a state. A negative S_k (fits
    // extrapolated beyond their temperature range) is kept as it is.
    E s = *S[k];
    if (f32_range) {
      const int32_t id = f.value("S_k in FP32 range");
      f.select(f.le(f.c(0.), s), {f.max(s, f.c((double)FLT_MIN))}, {s}, {id});
      s = f.v(id);
    }
    out.push_back(f.div(f.sub(mean_molar_mass_VTN, f.mul(f.mul(VT, f.c(W[k])), x[k])), s));
  }
  return out;
}

struct Scaled {
  std::vector<double> conductivity, viscosity, diffusivity;
};
// transport.rs:180-182
Scaled scale(const Polynomials& p, double T0, double viscosity, double conductivity) {
  Scaled s{p.conductivityIVT, p.VviscosityIVVT, p.diffusivityITVT};
  for (double& c : s.viscosity) c = (std::sqrt(std::sqrt(T0)) / std::sqrt(viscosity)) * c;
  for (double& c : s.conductivity) c = (std::sqrt(T0) / (2. * conductivity)) * c;
  for (double& c : s.diffusivity) c = (std::sqrt(T0) / (R * viscosity)) * c;
  return s;
}

// transport.rs:185-209 from the mixture inputs already in scope; returns [λ', μ', ρD'_k...] before the caller's scaling.
std::vector<E> properties_body(Block& f, const Species& species, const Polynomials& p, const Scaled& sc, E N, E theta,
                               const std::vector<E>& nonbulk, E& mean_molar_massN, bool lean = false, bool f32_range = false) {
  const size_t K = species.len();
  const int D = p.D;
  E lnT = f.defv(f.ln(theta), "lnT");
  std::vector<E> powers;
  E x = f.c(1.);
  for (int m = 0; m < D; m++) {
    powers.push_back(x);
    if (m + 1 < D) x = f.edef(f.mul(x, lnT), "x.lnT");
  }
  std::vector<E> X = nonbulk;
  X.push_back(f.defv(f.sub(N, f.sum(nonbulk)), "mole_proportions" + std::to_string(K - 1)));
  E VT = f.defv(f.sqrt(theta), "sqrtT");
  std::vector<E> terms;
  for (size_t k = 0; k < K; k++) terms.push_back(f.mul(f.c(species.molar_mass[k]), X[k]));
  mean_molar_massN = f.defv(f.sum(terms), "mean_molar_massN");
  E mean_molar_mass_VTN = f.defv(f.mul(mean_molar_massN, VT), "mean_molar_mass_VTN");
  std::vector<E> out;
  out.push_back(f.mul(conductivityNIVT(f, sc.conductivity, D, N, powers, X), VT));
  out.push_back(f.mul(VT, viscosityIVT(f, species.molar_mass, sc.viscosity, D, powers, X, lean)));
  auto d = density_diffusivity(f, species.molar_mass, sc.diffusivity, D, mean_molar_mass_VTN, VT, powers, X, lean, f32_range);
  out.insert(out.end(), d.begin(), d.end());
  return out;
}

}  // namespace

ast::Function properties(const Species& species, const Polynomials& p, double T0, double viscosity, double conductivity) {
  const size_t K = species.len();
  std::vector<std::string> names = {"sum_mole_proportions", "temperature"};
  for (size_t k = 0; k + 1 < K; k++) names.push_back("nonbulk_mole_proportions[" + std::to_string(k) + "]");
  Block f(names);
  std::vector<E> nonbulk;
  for (size_t k = 0; k + 1 < K; k++) nonbulk.push_back(f.v(2 + (int32_t)k));
  E W_N;
  auto out = properties_body(f, species, p, scale(p, T0, viscosity, conductivity), f.v(0), f.v(1), nonbulk, W_N);
  return f.finish(out);
}

// The transport functions of the device-function export (cuda.rs:150-208): inputs as the reference names them,
// lnT^m for m = 1..D-1 (ln of T/T0) and all K mole proportions, coefficients scaled as in transport.rs:180-182.
ast::Function transport_device_function(const Species& species, const Polynomials& p, double T0, double viscosity, double conductivity,
                                        int which) {
  const size_t K = species.len();
  const int D = p.D;
  std::vector<std::string> names;
  if (which == 1) names.push_back("sum_mole_proportions");
  if (which == 3) {
    names.push_back("mean_molar_mass_VTN");
    names.push_back("VT");
  }
  const size_t first_power = names.size();
  for (int m = 1; m < D; m++) names.push_back(m == 1 ? "lnT" : "lnT" + std::to_string(m));
  const size_t first_x = names.size();
  for (size_t k = 0; k < K; k++) names.push_back("mole_proportions" + std::to_string(k));
  Block f(names);
  std::vector<E> powers = {f.c(1.)};
  for (int m = 1; m < D; m++) powers.push_back(f.v((int32_t)(first_power + m - 1)));
  std::vector<E> x;
  for (size_t k = 0; k < K; k++) x.push_back(f.v((int32_t)(first_x + k)));
  const Scaled sc = scale(p, T0, viscosity, conductivity);
  std::vector<E> out;
  if (which == 1)
    out.push_back(conductivityNIVT(f, sc.conductivity, D, f.v(0), powers, x));
  else if (which == 2)
    out.push_back(viscosityIVT(f, species.molar_mass, sc.viscosity, D, powers, x, false));
  else if (which == 3)
    out = density_diffusivity(f, species.molar_mass, sc.diffusivity, D, f.v(0), f.v(1), powers, x);
  else
    throw ast::Error("transport device function: which is 1, 2 or 3");
  return f.finish(out);
}

// Transport section of the per-state program: physical λ, μ and D_k = ρD_k/ρ with ρ = Σ_k W_k C_k.
std::vector<E> transport_outputs(Block& f, const Species& s, const Polynomials& p, const TransportScales& scales, E N, E theta, bool lean,
                                 bool f32_range, const std::vector<E>& x, E& mean_molar_massN) {
  auto out = properties_body(f, s, p, scale(p, scales.T0, scales.viscosity, scales.conductivity), N, theta, x, mean_molar_massN, lean,
                             f32_range);
  out[0] = f.mul(f.c(scales.conductivity), out[0]);
  out[1] = f.mul(f.c(scales.viscosity), out[1]);
  E scale_D = f.defv(f.div(f.c(scales.viscosity), mean_molar_massN), "η0/ρ");
  for (size_t k = 2; k < out.size(); k++) out[k] = f.mul(out[k], scale_D);
  return out;
}

}  // namespace combustion
