// Net production rates in mixed precision for gfx950 (hand-written; the host prepends the mechanism sizes).
//
// Each wave owns NS = 16 states and walks the reactions in tiles of 16 (sorted by model so tiles are uniform):
//   1. per state and species (4 lanes per state): concentrations C_k (FP64), ln C_k as three BF16 terms in the
//      MFMA B layout, Gibbs g_k = G°/RT (FP64) and per-state scalars, all in LDS;
//   2. per reaction tile: ln R_f = ν'·ln C and ln Π C^ν'' = ν''·ln C as split-BF16 MFMA products (exact small-integer
//      orders, FP32 accumulation, the leading BF16 term accumulated apart from the two trailing ones);
//   3. elementwise, in the MFMA output layout (lane = 4 reactions × 1 state): ln k (FP64 Arrhenius exponent,
//      FP32 falloff/Troe), ln(1/Kc) = Σν·min(±g, ln clamp) - Σν ln C0 in FP64, and
//      q = exp(ln k_f R_f) - exp(ln k_f R_r) with FP64 argument reduction around the FP32 v_exp_f32 and an FP64
//      subtraction;
//   4. ω += ν·q as split-BF16 MFMA products (FP32 accumulation) under a per-state block-floating-point exponent,
//      since |q| exceeds the FP32 range on some states.
// Required defines: K, KA, KB (species k-blocks of 16, over K), AB (active-species tiles of 16), RT (reaction tiles),
// KE (species blocks of 4 for the third-body contraction), WAVES (waves per block, sized to the LDS).

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

#define NS 16
#define LOG2E 1.4426950408889634
#define LN2F 0.6931471805599453f
#define LN10 2.302585092994046
#define LN_FLT_MAX 88.72283905206835
#define LN_FLT_MIN (-87.33654475055310)
#define C_FLOOR 1e-30f

// Reaction tables (sorted reaction r): rate[4r..] = ln A, β, Ea/R of k (Troe: ln A∞ and the exponent of k∞/A∞);
// k0[4r..] = ln A0 (Troe: ln(A0/A∞)), β0, Ea0/R; troe[4r..] = α, 1/T3, 1/T1 (0: constant α term), T2 (0: absent);
// slots[8r..] = info, then 3 "plus" and 3 "minus" ΔG slots, then 0. info = model | kinf_const << 4 | (8 - Σν) << 8
// (model 0 elementary, 1 irreversible, 2 three-body, 3 Lindemann, 4 Troe, 5 padding). A slot is the byte offset of
// the species' row in the Gibbs table (the row after the last species is zero) | (|ν| == 2): plus slots hold the
// species with ν > 0, minus slots those with ν < 0 (a reaction has at most 3 species per side).
// tile_info[t] = 1: third bodies in the tile, 2: falloff, 4: Troe; eff: FP32 A operands of (e - 1) per tile.
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// x = hi + mid + lo with each term a BF16 (24 significant bits in all).
__device__ __forceinline__ void split3(float x, unsigned short& h, unsigned short& m, unsigned short& l) {
  h = bf16_bits(x);
  float r = x - bf16_value(h);
  m = bf16_bits(r);
  r -= bf16_value(m);
  l = bf16_bits(r);
}

// e^L for FP64 L: FP64 reduction to 2^n · 2^f with |f| <= 1/2, the 2^f core by v_exp_f32.
__device__ __forceinline__ double exp_wide(double L) {
  double x = L * LOG2E;
  double n = __builtin_rint(x);
  float f = (float)(x - n);
  n = __builtin_fmin(__builtin_fmax(n, -1100.0), 1100.0);
  return __builtin_ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
}
__device__ __forceinline__ int binary_exponent(double x) {
  return (int)((__double_as_longlong(x) >> 52) & 0x7FF) - 1023;
}
__device__ __forceinline__ float ln_f32(float x) { return __builtin_amdgcn_logf(x) * LN2F; }
__device__ __forceinline__ float exp_f32(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float softplus(float x) { return x > 0.f ? x + ln_f32(1.f + exp_f32(-x)) : ln_f32(1.f + exp_f32(x)); }

struct Lds {
  unsigned short lnC[KB][3][64][4];  // B operands: [k-block][split][lane][i] = ln C[16 kb + 4 (lane/16) + i][lane % 16]
  double g[KA + 1][NS];  // Gibbs energies; the last row is zero (padding slots)
  float C[KE * 4][NS];
  double lnT[NS], rcpT[NS], lnC0[NS];
  float T[NS], Ctot[NS];
};

extern "C" __global__ void __launch_bounds__(64 * WAVES)
rates_mixed(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
            const bf16x4* __restrict__ A_fwd, const bf16x4* __restrict__ A_rev, const bf16x4* __restrict__ A_nu,
            const f64x4* __restrict__ rate, const f64x4* __restrict__ k0, const f32x4* __restrict__ troe,
            const i32x4* __restrict__ slots, const int* __restrict__ tile_info,
            const float* __restrict__ eff, const double* __restrict__ rcp_W, const double* __restrict__ gibbs,
            const double* __restrict__ split, double ln_clamp_pos, double ln_clamp_neg) {
  __shared__ Lds lds_all[WAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Lds& L = lds_all[wave];
  const int col = lane & 15, grp = lane >> 4;
  const unsigned s = (blockIdx.x * WAVES + wave) * NS + col;
  const bool valid = s < n;

  // 1. Per state (col) and species group (grp): concentrations, ln C, Gibbs energies.
  const double T = valid ? in[0][s] : 1000.0;
  const double P = valid ? in[1][s] : 101325.0;
  const double rcpT = 1.0 / T, lnT = log(T);
  double sum_n = 0;
  for (int k = grp; k < K; k += 4) sum_n += (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k];
  sum_n += __shfl_xor(sum_n, 16);
  sum_n += __shfl_xor(sum_n, 32);
  const double C_tot = P * (1.0 / (1.380649e-23 * 6.02214076e23)) * rcpT;
  const double density = C_tot / sum_n;
  double partial = 0;
  for (int k = grp; k < K - 1; k += 4) {
    const double Ck = density * fmax(0.0, (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k]);
    partial += Ck;
    L.C[k][col] = (float)Ck;
  }
  partial += __shfl_xor(partial, 16);
  partial += __shfl_xor(partial, 32);
  for (int k = K + grp; k < KE * 4; k += 4) L.C[k][col] = 0.f;
  if (grp == 0) {
    L.C[K - 1][col] = (float)(C_tot - partial);
    L.lnT[col] = lnT;
    L.rcpT[col] = rcpT;
    L.lnC0[col] = log(101325.0 / (1.380649e-23 * 6.02214076e23)) - lnT;
    L.T[col] = (float)T;
    L.Ctot[col] = (float)C_tot;
  }
  for (int k = grp; k < KA; k += 4) {
    const double* a = gibbs + (size_t)k * 14 + (T <= split[k] ? 0 : 7);
    double g = a[0] + a[1] * lnT + a[2] * T + a[3] * (T * T) + a[4] * (T * T * T) + a[5] * (T * T * T * T) + a[6] * rcpT;
    L.g[k][col] = g;
  }
  if (grp == 0) L.g[KA][col] = 0.0;
  __syncthreads();
  for (int k = grp; k < KB * 16; k += 4) {
    const float c = k < K ? L.C[k][col] : 0.f;
    const float x = k < K ? ln_f32(fmaxf(c, C_FLOOR)) : 0.f;
    unsigned short h, m, l;
    split3(x, h, m, l);
    const int kb = k >> 4, r = k & 15, lane_b = ((r >> 2) << 4) | col, i = r & 3;
    L.lnC[kb][0][lane_b][i] = h;
    L.lnC[kb][1][lane_b][i] = m;
    L.lnC[kb][2][lane_b][i] = l;
  }
  __syncthreads();

  bf16x4 Bh[KB], Bm[KB], Bl[KB];
  for (int kb = 0; kb < KB; kb++) {
    Bh[kb] = *(const bf16x4*)L.lnC[kb][0][lane];
    Bm[kb] = *(const bf16x4*)L.lnC[kb][1][lane];
    Bl[kb] = *(const bf16x4*)L.lnC[kb][2][lane];
  }
  const double my_lnT = L.lnT[col], my_rcpT = L.rcpT[col], my_lnC0 = L.lnC0[col];
  const float my_T = L.T[col], my_Ctot = L.Ctot[col];

  f32x4 om[AB];
  for (int a = 0; a < AB; a++) om[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
  int E = -100000;  // block exponent of this lane's state column

  const char* gcol = (const char*)&L.g[0][col];
  for (int rt = 0; rt < RT; rt++) {
    const int tinfo = tile_info[rt];
    // This lane's reactions r = rt*16 + 4*grp + i: tables first, so that the loads overlap the MFMAs.
    f64x4 kr[4];
    i32x4 sp[4], sm[4];
    for (int i = 0; i < 4; i++) {
      const int r = rt * 16 + 4 * grp + i;
      kr[i] = rate[r];
      sp[i] = slots[2 * r];
      sm[i] = slots[2 * r + 1];
    }
    // 2. ln R_f and ln Π C^ν'' for the 16 reactions of the tile.
    f32x4 fh = {0.f, 0.f, 0.f, 0.f}, fl = fh, rh = fh, rl = fh;
    for (int kb = 0; kb < KB; kb++) {
      const bf16x4 af = A_fwd[(rt * KB + kb) * 64 + lane];
      const bf16x4 ar = A_rev[(rt * KB + kb) * 64 + lane];
      fh = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, Bh[kb], fh, 0, 0, 0);
      rh = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ar, Bh[kb], rh, 0, 0, 0);
      fl = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, Bm[kb], fl, 0, 0, 0);
      rl = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ar, Bm[kb], rl, 0, 0, 0);
      fl = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, Bl[kb], fl, 0, 0, 0);
      rl = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ar, Bl[kb], rl, 0, 0, 0);
    }
    // Third-body concentrations M - C_tot = (e - 1)·C as FP32 MFMA products.
    f32x4 Mv = {0.f, 0.f, 0.f, 0.f};
    if (tinfo & 1) {
      const float* A = eff + (size_t)rt * KE * 64;
      for (int kk = 0; kk < KE; kk++) Mv = __builtin_amdgcn_mfma_f32_16x16x4f32(A[kk * 64 + lane], L.C[kk * 4 + grp][col], Mv, 0, 0, 0);
    }
    // 3. Elementwise: log2 of k_f R_f and k_f R_r (FP64), then one block exponent per state.
    double xf[4], xr[4];
    int type[4];
    for (int i = 0; i < 4; i++) {
      const int r = rt * 16 + 4 * grp + i;
      const int info = sp[i][0];
      type[i] = info & 15;
      double lnk = fmin(kr[i][0] + kr[i][1] * my_lnT - kr[i][2] * my_rcpT, LN_FLT_MAX);
      if (tinfo & 1) {
        const double lnM = (double)ln_f32(my_Ctot + Mv[i]);
        if (type[i] == 2) lnk += lnM;
        if ((tinfo & 2) && type[i] >= 3) {
          // falloff: ln k_f = ln k∞ + ln Pr - ln(1 + Pr) (+ ln F for Troe), Pr = M k0 / k∞
          const f64x4 c0 = k0[r];
          const double lnk0 = fmin(c0[0] + c0[1] * my_lnT - c0[2] * my_rcpT, LN_FLT_MAX);
          const bool troe_row = type[i] == 4;
          const double lnkinf = troe_row ? ((info >> 4) & 1 ? 0.0 : fmin(kr[i][1] * my_lnT - kr[i][2] * my_rcpT, LN_FLT_MAX)) : lnk;
          const double lnPr = lnM + lnk0 - lnkinf;
          const double lnPr_1p = lnPr > 0 ? -(double)ln_f32(1.f + exp_f32((float)-lnPr)) : lnPr - (double)ln_f32(1.f + exp_f32((float)lnPr));
          double lnF = 0.0;
          if ((tinfo & 4) && troe_row) {
            const f32x4 t = troe[r];
            float Fcent = (1.f - t[0]) * exp_f32(-my_T * t[1]) + (t[2] != 0.f ? t[0] * exp_f32(-my_T * t[2]) : t[0]);
            if (t[3] != 0.f) Fcent += exp_f32(-t[3] * __builtin_amdgcn_rcpf(my_T));
            const float lnFcent = ln_f32(Fcent);
            const float c = -0.4f * (float)LN10 - 0.67f * lnFcent;
            const float N = 0.75f * (float)LN10 - 1.27f * lnFcent;
            const float u = fmaxf((float)lnPr, (float)LN_FLT_MIN) + c;
            const float f1 = u * __builtin_amdgcn_rcpf(N - 0.14f * u);
            lnF = (double)(lnFcent * __builtin_amdgcn_rcpf(f1 * f1 + 1.f));
          }
          lnk = (troe_row ? kr[i][0] : 0.0) + lnkinf + lnPr_1p + lnF;
        }
      }
      // ln(1/Kc) = Σ_{ν>0} ν min(g, ln c+) + Σ_{ν<0} |ν| min(-g, ln c-) + (-Σν) ln C0
      double lnK = (double)(((info >> 8) & 255) - 8) * my_lnC0;
      for (int j = 1; j < 4; j++) {
        const int w = sp[i][j];
        lnK += __builtin_ldexp(fmin(*(const double*)(gcol + (w & ~1)), ln_clamp_pos), w & 1);
        const int v = sm[i][j - 1];
        lnK -= __builtin_ldexp(fmax(*(const double*)(gcol + (v & ~1)), -ln_clamp_neg), v & 1);
      }
      xf[i] = (lnk + (double)fh[i] + (double)fl[i]) * LOG2E;
      xr[i] = type[i] == 1 ? -1e300 : (lnk + (double)rh[i] + (double)rl[i] + lnK) * LOG2E;
      if (type[i] == 5) xf[i] = xr[i] = -1e300;
    }
    // 4. ω += ν·q under a per-state block exponent (|q| exceeds the FP32 range on some states).
    double xm = fmax(fmax(fmax(xf[0], xr[0]), fmax(xf[1], xr[1])), fmax(fmax(xf[2], xr[2]), fmax(xf[3], xr[3])));
    int e = xm > -1e299 ? (int)__builtin_ceil(xm) : -100000;
    e = max(e, __shfl_xor(e, 16));
    e = max(e, __shfl_xor(e, 32));
    if (e > E) {
      if (E > -100000) {
        const float f = __builtin_ldexpf(1.f, max(E - e, -126));
        for (int a = 0; a < AB; a++) om[a] *= f;
      }
      E = e;
    }
    const double Ed = (double)E;
    f32x4 qs;
    for (int i = 0; i < 4; i++) {
      const float a = __builtin_amdgcn_exp2f((float)(xf[i] - Ed));
      const float b = __builtin_amdgcn_exp2f((float)(xr[i] - Ed));
      qs[i] = (float)((double)a - (double)b);  // forward minus reverse in FP64
    }
    // q = hi + mid + lo in BF16
    bf16x2 h01 = __builtin_convertvector((f32x2){qs[0], qs[1]}, bf16x2), h23 = __builtin_convertvector((f32x2){qs[2], qs[3]}, bf16x2);
    f32x4 r1 = qs - (f32x4){(float)h01[0], (float)h01[1], (float)h23[0], (float)h23[1]};
    bf16x2 m01 = __builtin_convertvector((f32x2){r1[0], r1[1]}, bf16x2), m23 = __builtin_convertvector((f32x2){r1[2], r1[3]}, bf16x2);
    f32x4 r2 = r1 - (f32x4){(float)m01[0], (float)m01[1], (float)m23[0], (float)m23[1]};
    bf16x2 l01 = __builtin_convertvector((f32x2){r2[0], r2[1]}, bf16x2), l23 = __builtin_convertvector((f32x2){r2[2], r2[3]}, bf16x2);
    bf16x4 qh, qm, ql;
    __builtin_memcpy(&qh, &h01, 4);
    __builtin_memcpy((char*)&qh + 4, &h23, 4);
    __builtin_memcpy(&qm, &m01, 4);
    __builtin_memcpy((char*)&qm + 4, &m23, 4);
    __builtin_memcpy(&ql, &l01, 4);
    __builtin_memcpy((char*)&ql + 4, &l23, 4);
    for (int a = 0; a < AB; a++) {
      const bf16x4 an = A_nu[(rt * AB + a) * 64 + lane];
      om[a] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(an, qh, om[a], 0, 0, 0);
      om[a] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(an, qm, om[a], 0, 0, 0);
      om[a] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(an, ql, om[a], 0, 0, 0);
    }
  }
  if (!valid) return;
  for (int a = 0; a < AB; a++)
    for (int i = 0; i < 4; i++) {
      const int k = 16 * a + 4 * grp + i;
      if (k < KA) out[k][s] = E > -100000 ? __builtin_ldexp((double)om[a][i], E) : 0.0;
    }
}
