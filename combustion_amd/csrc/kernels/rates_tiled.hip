// Net production rates in FP64 for mechanisms too large for a straight-line kernel (hand-written; the host prepends the
// mechanism sizes). A wave owns NS states and splits the reactions over G = 64/NS lane groups (lane = NS group +
// state); NS is the largest power of two up to 8 whose per-state arrays fit 11 KiB of LDS per wave, so that about
// 15 waves share a CU (NS = 8 at 10 species, 4 at 53, 2 at 160):
//   1. per state, spread over the groups: concentrations C_k (reaction.rs:362-373 on a mass basis), the clamped
//      exponentials E+_k = min(c+, e^g_k) and E-_k = min(c-, 1/e^g_k) of the Gibbs functions g_k = G°_k/RT of the
//      state's NASA-7 piece (reaction.rs:236-268), and zeroed production rates, all in LDS ([species][state]);
//   2. rounds of G reactions, one per group, every reaction of a round of one model (the host sorts and pads them),
//      the next round's table rows loaded while the current one computes:
//      k_f from its Arrhenius exponent exp(ln A + β ln T - θ/T) clamped at FLT_MAX (reaction.rs:113-139), the third-
//      body sum M = C_tot + Σ (e_k - 1) C_k (reaction.rs:141-145), Lindemann or Troe falloff (reaction.rs:149-193);
//      q = k_f Π C^ν' - k_f (1/K_c) Π C^ν'' with 1/K_c = Π E±^|ν| (P0/RT)^(-Σν) (reaction.rs:271-308);
//      ω_k += ν_k q by LDS atomic adds (reaction.rs:312-327; the order of the additions differs from the reference's);
//   3. the rates of the wave's states written out.
// Tables (row = sorted, padded reaction; see rates_tiled_tables in tiled.cpp):
//   rate[row] = (ln A, β, θ = Ea/R, unused); k0[row] = (ln A0 (Troe: ln(A0/A∞)), β0, θ0, A∞); troe[row] = (α, T3, T1,
//   T2) (T1 >= 1e30: no α e^(-T/T1) term, α instead; T2 NaN: no e^(-T2/T) term);
//   info[row] = (model | kinf_const << 4, first efficiency, efficiency count, the round's largest list lengths:
//   forward | reverse << 4 | targets << 8), model 0 elementary, 1 irreversible
//   (the same path: its first reverse factor is the row of zeros), 2 three-body, 3 Lindemann, 4 Troe, 5 padding;
//   eff_species/eff_value = the (e_k - 1) lists;
//   ops[row] = 24 16-bit LDS byte offsets of rows of the per-state arrays (a lane adds 8 × its state): 4 forward
//   factors (the reactants' C, with multiplicity), 12 reverse factors (E+ for each unit of a product-side ν, E- for
//   each unit of a reactant-side ν, the products' C, and the row of (P0/RT)^(-Σν): 1, C0, 1/C0 or 1/C0²; an
//   irreversible reaction's first reverse factor is the row of zeros), padded with the row of ones; 6 targets of
//   ν_k q (the ω_k rows, padded with a row nobody reads), their ν + 8 in 4 bits each in the last word;
//   nasa[k] = the 2 x 7 coefficients of g_k (low piece first), split[k] the switch temperature (+inf: one piece).
// LDS rows ([row][state], NS doubles each): C_0..C_{K-1}, E+_0.., E-_0.., ω_0.., the spare ω row, ones, zeros,
// C0, 1/C0, 1/C0² (K + 3 KA + 6 rows).
// Required defines: K (species), KA (active species), ROUNDS (rounds of G reactions), NS (states per wave: 1, 2, 4, 8).

#define G (64 / NS)
#define RGAS (1.380649e-23 * 6.02214076e23)
#define P0_R (101325.0 / RGAS)
#define FLT_MAX_D 3.4028234663852886e38
#define FLT_MIN_D 1.1754943508222875e-38

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// e^x: Cody-Waite reduction by ln 2 in two parts, a degree-10 polynomial on |r| <= ln(2)/2 and an exponent add
// (combustion_amd/csrc/exp_coefficients.py; relative error below 5e-16), arguments clamped to [-746, 710].
__device__ __forceinline__ double exp64(double x) {
  x = __builtin_fmin(__builtin_fmax(x, -746.0), 710.0);
  const double t = __builtin_fma(x, 0x1.71547652b82fep+0, 0x1.8p+52);
  const double n = t - 0x1.8p+52;
  double r = __builtin_fma(-n, 0x1.62e42fee00000p-1, x);
  r = __builtin_fma(-n, 0x1.a39ef35793c76p-33, r);
  double p = 0x1.28a2c0a7209fbp-22;
  p = __builtin_fma(p, r, 0x1.72faf024b693bp-19);
  p = __builtin_fma(p, r, 0x1.a019a66a75dd4p-16);
  p = __builtin_fma(p, r, 0x1.a01978c6baf81p-13);
  p = __builtin_fma(p, r, 0x1.6c16c17f43a58p-10);
  p = __builtin_fma(p, r, 0x1.1111112dd67c5p-7);
  p = __builtin_fma(p, r, 0x1.55555555520afp-5);
  p = __builtin_fma(p, r, 0x1.555555554b757p-3);
  p = __builtin_fma(p, r, 0x1.0000000000005p-1);
  p = __builtin_fma(p, r, 0x1.000000000001ep+0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(p, (int)(unsigned)__double_as_longlong(t));
}

// Sum over the G groups of a state (lanes col, col + NS, ..., col + 64 - NS).
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int m = NS; m < 64; m <<= 1) v += __shfl_xor(v, m);
  return v;
}

// The table row of one reaction that every round reads.
struct Row {
  i32x4 info, o0, o1, o2;
  f64x4 rate;
};
__device__ __forceinline__ Row load_row(int row, const i32x4* __restrict__ info, const f64x4* __restrict__ rate,
                                        const i32x4* __restrict__ ops) {
  Row r;
  r.info = info[row];
  r.rate = rate[row];
  r.o0 = ops[3 * row];
  r.o1 = ops[3 * row + 1];
  r.o2 = ops[3 * row + 2];
  return r;
}
// The double at LDS byte offset half h of word w (low half first), in this lane's state column.
__device__ __forceinline__ double lds_at(const char* column, int w, int h) {
  return *(const double*)(column + ((w >> (16 * h)) & 0xFFFF));
}

#define ROWS (K + 3 * KA + 6)
#define ROW_SPARE (K + 3 * KA)
#define ROW_ONE (ROW_SPARE + 1)
#define ROW_ZERO (ROW_SPARE + 2)
#define ROW_C0 (ROW_SPARE + 3)
struct Lds {
  double v[ROWS][NS];
};

// at most 128 registers: four waves per SIMD, the 16 per CU that the LDS sizing allows
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
rates_tiled(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
            const f64x4* __restrict__ rate, const f64x4* __restrict__ k0, const f64x4* __restrict__ troe,
            const i32x4* __restrict__ info, const i32x4* __restrict__ ops, const int* __restrict__ eff_species,
            const double* __restrict__ eff_value,
            const double* __restrict__ rcp_W, const double* __restrict__ nasa, const double* __restrict__ split,
            double clamp_pos, double clamp_neg) {
  __shared__ Lds L;
  const int lane = threadIdx.x, col = lane & (NS - 1), grp = lane / NS;
  // XCD-aware state ranges: consecutive blocks go to the 8 XCDs in turn, so block b takes state group
  // xcd·per + min(xcd, rem) + b/8 (a bijection of [0, gridDim.x)), and the blocks of one XCD cover one contiguous
  // range: the NS·8 B that a block reads of each 128 B line come from its own XCD's L2, not once per XCD from HBM
  const unsigned nb = gridDim.x, per = nb / 8, rem = nb % 8, xcd = blockIdx.x % 8;
  const unsigned group = xcd * per + min(xcd, rem) + blockIdx.x / 8;
  const unsigned s = group * NS + col;
  const bool valid = s < n;
  // the first round's rows, in flight during the per-state stage
  Row cur = load_row(grp, info, rate, ops);
  // states past n compute on a harmless stand-in and write nothing
  const double T = valid ? in[0][s] : 1000.0, P = valid ? in[1][s] : 101325.0;
  const double rcpT = 1.0 / T, lnT = log(T), T2 = T * T;
  const double C_tot = P * (1.0 / RGAS) * rcpT;

  // 1. concentrations: n_k = Y_k / W_k, C_k = max(0, n_k) C_tot / Σn (bulk species: C_tot - Σ the others)
  double sum_n = 0.0;
  for (int k = grp; k < K; k += G) sum_n += (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k];
  sum_n = group_sum(sum_n);
  const double density = C_tot / sum_n;
  double nonbulk = 0.0;
  for (int k = grp; k < K - 1; k += G) {
    const double c = density * fmax(0.0, (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k]);
    L.v[k][col] = c;
    nonbulk += c;
  }
  nonbulk = group_sum(nonbulk);
  if (grp == 0) L.v[K - 1][col] = C_tot - nonbulk;
  // Gibbs functions of the active species: g = a0 - a6 - a0 ln T - a1/2 T + ... + a5/T, as reaction.rs:78-81 arranges
  // the coefficients (the host stores them so arranged), one piece by T <= split
  for (int k = grp; k < KA; k += G) {
    const double* a = nasa + 14 * k + (T <= split[k] ? 0 : 7);
    const double g = a[0] + a[1] * lnT + a[2] * T + a[3] * T2 + a[4] * (T2 * T) + a[5] * (T2 * T2) + a[6] * rcpT;
    const double e = exp64(g);
    L.v[K + k][col] = fmin(clamp_pos, e);
    L.v[K + KA + k][col] = fmin(clamp_neg, 1.0 / e);
    L.v[K + 2 * KA + k][col] = 0.0;
  }
  if (grp == 0) {
    const double rcpC0 = T * (1.0 / P0_R);
    L.v[ROW_SPARE][col] = 0.0;
    L.v[ROW_ONE][col] = 1.0;
    L.v[ROW_ZERO][col] = 0.0;
    L.v[ROW_C0][col] = P0_R * rcpT;
    L.v[ROW_C0 + 1][col] = rcpC0;
    L.v[ROW_C0 + 2][col] = rcpC0 * rcpC0;
  }
  __syncthreads();
  const char* column = (const char*)&L.v[0][col];
  // 2. rounds of G reactions
  for (int j = 0; j < ROUNDS; j++) {
    const int row = G * j + grp;
    const Row r = cur;
    if (j + 1 < ROUNDS) cur = load_row(row + G, info, rate, ops);
    // the round's list lengths (the same in each of its rows) and model (5 on the padding rows that end a model's
    // last round)
    const int lengths = __builtin_amdgcn_readfirstlane(r.info[3]);
    const int nf = lengths & 15, nr = (lengths >> 4) & 15, na = (lengths >> 8) & 15;
    const int model = r.info[0] & 15;
    if (model == 5) continue;
    const f64x4 c = r.rate;
    double kf = fmin(FLT_MAX_D, exp64(c[0] + c[1] * lnT - c[2] * rcpT));
    if (model >= 2) {
      double M = C_tot;
      for (int e = r.info[1]; e < r.info[1] + r.info[2]; e++) M = __builtin_fma(eff_value[e], L.v[eff_species[e]][col], M);
      if (model == 2) {
        kf *= M;
      } else {
        const f64x4 c0 = k0[row];
        const double k0v = fmin(FLT_MAX_D, exp64(c0[0] + c0[1] * lnT - c0[2] * rcpT));
        if (model == 3) {
          kf = M * k0v * kf / (M * k0v + kf + FLT_MIN_D);  // Lindemann (reaction.rs:153-158)
        } else {
          // Troe (reaction.rs:160-193): k0 and k∞ with A∞ factored out
          const bool kinf_const = (r.info[0] >> 4) & 1;
          const double A_inf = c0[3];
          const double kinf = kinf_const ? 1.0 : fmin(FLT_MAX_D, exp64(c[1] * lnT - c[2] * rcpT));
          const double Pr = kinf_const ? M * k0v : M * k0v / (kinf + FLT_MIN_D);
          const f64x4 t = troe[row];
          double Fcent = (1.0 - t[0]) * exp64(-T / t[1]) + (t[2] < 1e30 ? t[0] * exp64(-T / t[2]) : t[0]);
          if (t[3] == t[3]) Fcent += exp64(-t[3] * rcpT);
          const double lnF = log(Fcent);
          const double C = -0.67 * lnF - 0.4 * 2.302585092994046, N = -1.27 * lnF + 0.75 * 2.302585092994046;
          const double lp = log(Pr + FLT_MIN_D) + C;
          const double f1 = lp / (-0.14 * lp + N);
          const double F = exp64(lnF / (f1 * f1 + 1.0));
          kf = F * A_inf * kinf * Pr / (Pr + 1.0);
        }
      }
    }
    // q = k_f Π C^ν' - k_f (1/K_c) Π C^ν'' from the row's lists of LDS rows, up to the round's lengths (a row's
    // entries past its own lengths are the row of ones)
    double f = kf, b = kf;
#pragma unroll
    for (int h = 0; h < 4; h++)
      if (h < nf) f *= lds_at(column, r.o0[h >> 1], h & 1);
#pragma unroll
    for (int h = 0; h < 12; h++)
      if (h < nr) b *= lds_at(column, h < 4 ? r.o0[2 + (h >> 1)] : r.o1[(h - 4) >> 1], h & 1);
    const double q = f - b;
    // ω_k += ν_k q for the row's targets (a row's targets past its own count are the spare row, with ν = 0)
#pragma unroll
    for (int h = 0; h < 6; h++) {
      if (h < na) {
        const int w = h < 4 ? r.o2[h >> 1] : r.o2[2];
        const int off = (w >> (16 * (h & 1))) & 0xFFFF;
        const double nu = (double)(((r.o2[3] >> (4 * h)) & 15) - 8);
        atomicAdd((double*)(column + off), nu * q);
      }
    }
  }
  __syncthreads();
  // 3. outputs
  if (valid)
    for (int k = grp; k < KA; k += G) out[k][s] = L.v[K + 2 * KA + k][col];
}
