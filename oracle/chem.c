/* ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only to check results or to time a CPU baseline, never as the product path.
 *
 * Closed-form FP64 restatement of the per-state chemistry, written from textbook formulas rather than from
 * the reference's expression rearrangements, so that it pins them:
 *   rates      reaction.rs:113-327 (Arrhenius with T^m factoring -> A T^b exp(-Ea/T); Troe in ln form -> the
 *              log10 form of Gilbert, Luther & Troe 1983; 1/Kc from clamped exp(±G/RT) products)
 *   NASA-7     lib.rs:16-18 h/RT and s/R integrals, g = h/RT - s/R (reaction.rs:78 coefficient transform)
 *   transport  transport.rs:141-214 (scaled polynomial algebra -> mixture-averaged formulas: Mathur-Saxena
 *              conductivity, Wilke viscosity, D_km = (1 - Y_k) / Σ_{j≠k} X_j / D_kj)
 * The saturation constants of the reference are reproduced because they are part of its numerics:
 * min(FLT_MAX, k) (reaction.rs:138), min(6981463048192, e^g), min(cbrt(FLT_MAX) = 6981463572480, e^-g)
 * (reaction.rs:244-262; f32::cbrt is the platform cbrtf, which on glibc is correctly rounded: 6981463572480)
 * and the FLT_MIN guards (reaction.rs:158,177,189).
 * The mechanism and the transport fits come in as plain arrays (see oracle/__init__.py).
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define KB 1.380649e-23
#define NAV 6.02214076e23
#define RGAS (KB * NAV)
#define P0_R (101325. / RGAS)

typedef struct {
  int K, Ka, R, D;
  const double* W;          /* [K] kg/mol */
  const double* split;      /* [K] NaN: single range */
  const double* nasa;       /* [K][2][7] */
  const int* reactants;     /* [R][K] */
  const int* products;      /* [R][K] */
  const int* model;         /* [R] 0 elementary 1 irreversible 2 three-body 3 Lindemann 4 Troe 5 SRI */
  const double* rate;       /* [R][3] A (SI), b, Ea/R (K) */
  const double* k0;         /* [R][3] */
  const double* troe;       /* [R][4] A T3 T1 T2 */
  const double* eff;        /* [R][K] */
  const double* fit_c;      /* [K][D] λ_k/√T in ln(T/T0) */
  const double* fit_v;      /* [K][D] √(η_k/√T) */
  const double* fit_d;      /* [K][K][D] P D_kj / T^1.5 (entries j<k used) */
  double T0;
  const double* sri;        /* [R][5] a b c d e */
  const int* plog_offset;   /* [R+1] rows of plog per reaction (more than one: a PLOG table) */
  const double* plog;       /* [rows][4] P (Pa), A (SI), b, Ea/R (K); sorted by P within a reaction */
} oracle_mechanism;

/* Cantera's PLOG rate (not in the reference): k at each table pressure (entries of equal P summed), ln k linear in
   ln P between the two table pressures around P, the end values beyond the table; clamped at FLT_MAX. */
static double plog_rate(const double* e, int n, double T, double P) {
  double Pl[64], kl[64];
  int L = 0;
  for (int i = 0; i < n; i++) {
    double k = e[4 * i + 1] * pow(T, e[4 * i + 2]) * exp(-e[4 * i + 3] / T);
    if (L && Pl[L - 1] == e[4 * i]) kl[L - 1] += k;
    else { Pl[L] = e[4 * i]; kl[L] = k; L++; }
  }
  double k;
  if (P <= Pl[0]) k = kl[0];
  else if (P > Pl[L - 1]) k = kl[L - 1];
  else {
    int i = 0;
    while (!(P <= Pl[i + 1])) i++;
    double lk = log(kl[i]) + (log(kl[i + 1]) - log(kl[i])) * (log(P) - log(Pl[i])) / (log(Pl[i + 1]) - log(Pl[i]));
    k = exp(lk);
  }
  return fmin((double)FLT_MAX, k);
}

/* Gibbs energy g = h/RT - s/R of species k (textbook NASA-7 integrals). */
static double gibbs_RT(const oracle_mechanism* m, int k, double T) {
  const double* a = m->nasa + (size_t)k * 14 + (T <= m->split[k] ? 0 : 7);
  double h = a[0] + a[1] * T / 2 + a[2] * T * T / 3 + a[3] * T * T * T / 4 + a[4] * T * T * T * T / 5 + a[5] / T;
  double s = a[0] * log(T) + a[1] * T + a[2] * T * T / 2 + a[3] * T * T * T / 3 + a[4] * T * T * T * T / 4 + a[6];
  return h - s;
}
static double enthalpy_RT(const oracle_mechanism* m, int k, double T) {
  const double* a = m->nasa + (size_t)k * 14 + (T <= m->split[k] ? 0 : 7);
  return a[0] + a[1] * T / 2 + a[2] * T * T / 3 + a[3] * T * T * T / 4 + a[4] * T * T * T * T / 5 + a[5] / T;
}
static double cp_R(const oracle_mechanism* m, int k, double T) {
  const double* a = m->nasa + (size_t)k * 14 + (T <= m->split[k] ? 0 : 7);
  return a[0] + a[1] * T + a[2] * T * T + a[3] * T * T * T + a[4] * T * T * T * T;
}

static double arrhenius(const double* k, double T) {
  double v = k[0] * pow(T, k[1]) * exp(-k[2] / T);
  return fmin((double)FLT_MAX, v);
}

static double ipow(double x, int n) {
  double y = 1;
  for (int i = 0; i < n; i++) y *= x;
  return y;
}

/* Concentrations of the reference's amounts/volume convention on a mass basis: C_k = max(0, Y_k/W_k) ρ
 * for k < K-1, bulk = C_tot - Σ; returns C_tot. */
static double concentrations(const oracle_mechanism* m, double T, double P, const double* Y, double* C) {
  const int K = m->K;
  double sum_n = 0;
  for (int k = 0; k < K; k++) sum_n += Y[k] / m->W[k];
  const double C_tot = P / RGAS / T;
  const double density = C_tot / sum_n;
  double s = 0;
  for (int k = 0; k < K - 1; k++) {
    C[k] = density * fmax(0., Y[k] / m->W[k]);
    s += C[k];
  }
  C[K - 1] = C_tot - s;
  return C_tot;
}

/* ω_k (k < Ka), their term scales Σ_r |ν_kr| k_f (R_f + R_r), optionally q_r and its scale |k_f| (R_f + R_r), and dT/dt
   (reaction.rs:375-380) with its scale T Σ_k s_k |h_k/RT| / Σ_k cp_k/R C_k (s_k the term scale of ω_k). */
void oracle_rates(const oracle_mechanism* m, double T, double P, const double* Y, double* omega, double* scale, double* q_out,
                  double* q_scale, double* dTdt, double* dTdt_scale) {
  const int K = m->K, Ka = m->Ka, R = m->R;
  double C[1024], Ep[1024], En[1024];
  const double C_tot = concentrations(m, T, P, Y, C);
  const double exp_clamp = 6981463048192., exp_neg_clamp = 6981463572480.;
  for (int k = 0; k < Ka; k++) {
    double g = gibbs_RT(m, k, T);
    Ep[k] = fmin(exp_clamp, exp(g));
    En[k] = fmin(exp_neg_clamp, exp(-g));
    omega[k] = 0;
    scale[k] = 0;
  }
  const double C0 = P0_R / T; /* P0/RT */
  for (int r = 0; r < R; r++) {
    const int* nu_f = m->reactants + (size_t)r * K;
    const int* nu_r = m->products + (size_t)r * K;
    const int model = m->model[r];
    double kf;
    double M = C_tot;
    if (model >= 2)
      for (int k = 0; k < K; k++) M += (m->eff[(size_t)r * K + k] - 1.) * C[k];
    const int n_plog = m->plog_offset ? m->plog_offset[r + 1] - m->plog_offset[r] : 0;
    if (model <= 1 && n_plog > 1) {
      kf = plog_rate(m->plog + 4 * m->plog_offset[r], n_plog, T, C_tot * RGAS * T);
    } else if (model <= 1) {
      kf = arrhenius(m->rate + 3 * r, T);
    } else if (model == 2) {
      kf = arrhenius(m->rate + 3 * r, T) * M;
    } else if (model == 3) {
      double k0 = arrhenius(m->k0 + 3 * r, T), kinf = arrhenius(m->rate + 3 * r, T);
      kf = M * k0 * kinf / (M * k0 + kinf + (double)FLT_MIN);
    } else if (model == 5) {
      /* SRI falloff (Stewart, Larson & Golden 1989, the form Cantera documents for its SRI rate):
         F = d [a exp(-b/T) + exp(-T/c)]^X T^e, X = 1/(1 + log10(Pr)^2); exp(-T/c) is dropped when c = 0 */
      double k0 = arrhenius(m->k0 + 3 * r, T), kinf = arrhenius(m->rate + 3 * r, T);
      double Pr = M * k0 / (kinf + (double)FLT_MIN);
      const double* s = m->sri + 5 * r;
      double l = log10(Pr + (double)FLT_MIN);
      double X = 1. / (1. + l * l);
      double base = s[0] * exp(-s[1] / T) + (s[2] != 0. ? exp(-T / s[2]) : 0.);
      double F = s[3] * pow(base, X) * pow(T, s[4]);
      kf = kinf * Pr / (1. + Pr) * F;
    } else {
      const double Ainf = m->rate[3 * r];
      double k0n[3] = {m->k0[3 * r] / Ainf, m->k0[3 * r + 1], m->k0[3 * r + 2]};
      double kin[3] = {1., m->rate[3 * r + 1], m->rate[3 * r + 2]};
      double k0 = arrhenius(k0n, T);
      double Pr = kin[1] == 0. && kin[2] == 0. ? M * k0 : M * k0 / (arrhenius(kin, T) + (double)FLT_MIN);
      double kinf = kin[1] == 0. && kin[2] == 0. ? Ainf : Ainf * arrhenius(kin, T);
      const double* t = m->troe + 4 * r;
      double Fcent = (1. - t[0]) * exp(-T / t[1]) + (t[2] < 1e30 ? t[0] * exp(-T / t[2]) : t[0]) + (isfinite(t[3]) ? exp(-t[3] / T) : 0.);
      double lf = log10(Fcent);
      double c = -0.4 - 0.67 * lf, n = 0.75 - 1.27 * lf;
      double lp = log10(Pr + (double)FLT_MIN) + c;
      double f1 = lp / (n - 0.14 * lp);
      double F = pow(10., lf / (1. + f1 * f1));
      kf = kinf * Pr / (1. + Pr) * F;
    }
    double Rf = 1, Rr = 0;
    for (int k = 0; k < K; k++) Rf *= ipow(C[k], nu_f[k]);
    if (model != 1) {
      double rcpK = 1;
      int sum_nu = 0;
      for (int k = 0; k < K; k++) sum_nu += nu_r[k] - nu_f[k];
      for (int k = 0; k < Ka; k++) {
        int nu = nu_r[k] - nu_f[k];
        rcpK *= nu > 0 ? ipow(Ep[k], nu) : ipow(En[k], -nu);
      }
      rcpK *= pow(C0, -sum_nu);
      Rr = rcpK;
      for (int k = 0; k < K; k++) Rr *= ipow(C[k], nu_r[k]);
    }
    double q = kf * (Rf - Rr);
    if (q_out) q_out[r] = q;
    if (q_scale) q_scale[r] = fabs(kf) * (Rf + Rr);
    for (int k = 0; k < Ka; k++) {
      int nu = nu_r[k] - nu_f[k];
      if (!nu) continue;
      omega[k] += nu * q;
      scale[k] += abs(nu) * fabs(kf) * (Rf + Rr);
    }
  }
  if (dTdt) {
    double num = 0, num_scale = 0, den = 0;
    for (int k = 0; k < Ka; k++) {
      const double h = enthalpy_RT(m, k, T);
      num += omega[k] * h;
      num_scale += scale[k] * fabs(h);
    }
    for (int k = 0; k < K; k++) den += cp_R(m, k, T) * C[k];
    *dTdt = -T * num / den;
    if (dTdt_scale) *dTdt_scale = T * num_scale / fabs(den);
  }
}

static double poly(const double* c, int D, double x) {
  double y = 0, p = 1;
  for (int i = 0; i < D; i++) {
    y += c[i] * p;
    p *= x;
  }
  return y;
}

/* λ, μ, D_k (and the floor for D_k: 1 / Σ_{j≠k} X_j / D_kj). */
void oracle_transport(const oracle_mechanism* m, double T, double P, const double* Y, double* lambda, double* mu, double* Dk, double* Dscale) {
  const int K = m->K, D = m->D;
  double C[1024], X[1024], lam[1024], eta[1024];
  const double C_tot = concentrations(m, T, P, Y, C);
  const double l = log(T / m->T0), sT = sqrt(T);
  double Wmean = 0;
  for (int k = 0; k < K; k++) {
    X[k] = C[k] / C_tot;
    Wmean += m->W[k] * X[k];
    lam[k] = poly(m->fit_c + (size_t)k * D, D, l) * sT;
    double v = poly(m->fit_v + (size_t)k * D, D, l);
    eta[k] = v * v * sT;
  }
  double a = 0, b = 0;
  for (int k = 0; k < K; k++) {
    a += X[k] * lam[k];
    b += X[k] / lam[k];
  }
  *lambda = 0.5 * (a + 1. / b);
  double num = 0;
  for (int k = 0; k < K; k++) {
    double den = 0;
    for (int j = 0; j < K; j++) {
      double phi = 1. + sqrt(eta[k] / eta[j]) * pow(m->W[j] / m->W[k], 0.25);
      den += X[j] * phi * phi / sqrt(8. * (1. + m->W[k] / m->W[j]));
    }
    num += X[k] * eta[k] / den;
  }
  *mu = num;
  const double T15 = T * sT;
  for (int k = 0; k < K; k++) {
    double s = 0;
    for (int j = 0; j < K; j++) {
      if (j == k) continue;
      const double* c = m->fit_d + ((size_t)(j < k ? k : j) * K + (j < k ? j : k)) * D;
      double Dkj = poly(c, D, l) * T15 / P;
      s += X[j] / Dkj;
    }
    double Yk = m->W[k] * X[k] / Wmean;
    Dk[k] = (1. - Yk) / s;
    if (Dscale) Dscale[k] = 1. / s;
  }
}

/* Batch over SoA states on `threads` host threads: outputs follow the per-state program's order
 * [ω (Ka), λ, μ, D (K)] for the sections requested (rates, transport); scales likewise. */
typedef struct {
  const oracle_mechanism* m;
  int rates, transport;
  const double *T, *P, *Y;
  size_t n, begin, end;
  double** out;
  double** scale;
} job;

static void* run(void* arg) {
  job* j = (job*)arg;
  const oracle_mechanism* m = j->m;
  const int K = m->K, Ka = m->Ka;
  double Y[1024], omega[1024], sc[1024], Dk[1024], Ds[1024];
  for (size_t i = j->begin; i < j->end; i++) {
    for (int k = 0; k < K; k++) Y[k] = j->Y[(size_t)k * j->n + i];
    int o = 0;
    if (j->rates) {
      oracle_rates(m, j->T[i], j->P[i], Y, omega, sc, NULL, NULL, NULL, NULL);
      for (int k = 0; k < Ka; k++, o++) {
        j->out[o][i] = omega[k];
        if (j->scale) j->scale[o][i] = sc[k];
      }
    }
    if (j->transport) {
      double lambda, mu;
      oracle_transport(m, j->T[i], j->P[i], Y, &lambda, &mu, Dk, Ds);
      j->out[o][i] = lambda;
      if (j->scale) j->scale[o][i] = lambda;
      o++;
      j->out[o][i] = mu;
      if (j->scale) j->scale[o][i] = mu;
      o++;
      for (int k = 0; k < K; k++, o++) {
        j->out[o][i] = Dk[k];
        if (j->scale) j->scale[o][i] = Ds[k];
      }
    }
  }
  return NULL;
}

int oracle_batch(const oracle_mechanism* m, int rates, int transport, const double* T, const double* P, const double* Y, size_t n,
                 double** out, double** scale, int threads) {
  if (m->K > 1024) return -1;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  job jobs[256];
  for (int t = 0; t < threads; t++) {
    jobs[t] = (job){m, rates, transport, T, P, Y, n, n * t / threads, n * (t + 1) / threads, out, scale};
    if (threads > 1)
      pthread_create(&tid[t], NULL, run, &jobs[t]);
    else
      run(&jobs[t]);
  }
  if (threads > 1)
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
  return 0;
}

/* Per-reaction rates of progress (reaction.rs:329 reactions_rates) for one state. */
int oracle_reaction_rates(const oracle_mechanism* m, double T, double P, const double* Y, double* q, double* q_scale) {
  double omega[1024], scale[1024];
  if (m->K > 1024) return -1;
  oracle_rates(m, T, P, Y, omega, scale, q, q_scale, NULL, NULL);
  return 0;
}

/* ω (and term scales), dT/dt (and scale), and the relative volume rate (dV/dt)/V of the constant-pressure reactor
   (reaction.rs:381-385): dT/dt / T + (R/P) T Σ_k (1 - W_k/W_bulk) ω_k, with scale |dT/dt|_scale / T + (R/P) T Σ_k
   |1 - W_k/W_bulk| s_k. */
int oracle_rhs(const oracle_mechanism* m, double T, double P, const double* Y, double* omega, double* scale, double* dTdt,
               double* dTdt_scale, double* dlnV, double* dlnV_scale) {
  double own_scale[1024];
  if (m->K > 1024) return -1;
  if (!scale) scale = own_scale;
  double dT_scale = 0;
  oracle_rates(m, T, P, Y, omega, scale, NULL, NULL, dTdt, &dT_scale);
  if (dTdt_scale) *dTdt_scale = dT_scale;
  const double Wb = m->W[m->K - 1];
  double s = 0, s_scale = 0;
  for (int k = 0; k < m->Ka; k++) {
    s += (1. - m->W[k] / Wb) * omega[k];
    s_scale += fabs(1. - m->W[k] / Wb) * scale[k];
  }
  const double T_pR = T * RGAS / P;
  if (dlnV) *dlnV = *dTdt / T + T_pR * s;
  if (dlnV_scale) *dlnV_scale = dT_scale / T + T_pR * s_scale;
  return 0;
}
