// Net production rates in mixed precision for gfx950 (hand-written; the host prepends the mechanism sizes).
//
// Each wave owns NS = 16 states and walks the reactions in tiles of 32 (sorted by model, each model's rows padded to
// whole 16-row halves),
// every quantity of a reaction being a base-2 logarithm until the rate of progress:
//   1. per state and species (4 lanes per state): concentrations C_k, log2 C_k as three BF16 terms in the MFMA B
//      layout, the clamped Gibbs tables G+_k = min(g_k, ln c+) log2 e and G-_k = min(-g_k, ln c-) log2 e (FP64,
//      reaction.rs:236-268) and per-state scalars, in LDS;
//   2. per reaction tile, two 16-row halves: log2 R_f = ν'·log2 C and log2 Π C^ν'' = ν''·log2 C as split-BF16
//      products on v_mfma_f32_16x16x32_bf16 (exact small-integer orders, FP32 accumulation, the leading BF16 term
//      accumulated apart from the two trailing ones), and the third-body sums M - C_tot = (e - 1)·C on
//      v_mfma_f32_16x16x4f32;
//   3. elementwise in the MFMA output layout (a lane holds 4 reactions × 1 state per half): x_f = log2 k_f R_f and
//      x_r = log2 k_f R_r in FP64 (Arrhenius exponent, FP32 falloff/Troe terms, log2(1/Kc) = Σ G± over the
//      reaction's slots + (-Σν) log2 C0);
//   4. q = 2^x_f - 2^x_r under a per-state block exponent E (|q| exceeds the FP32 range on some states): the
//      difference d = x_r - x_f is taken in FP64, and near equilibrium (|d| < 1/4) q = -2^(x_f-E) (2^d - 1) with
//      2^d - 1 by its series, so the cancellation of R_f - R_r costs no accuracy; elsewhere q = 2^(x_f-E) - 2^(x_r-E);
//   5. ω += ν·q as split-BF16 products on v_mfma_f32_16x16x32_bf16 (FP32 accumulation, 32 reactions per product).
// Required defines: K, KA, KB (species blocks of 32), AB (active-species tiles of 16), RT (reaction tiles of 32),
// KE (species blocks of 4 for the third-body contraction), WAVES (waves per block, sized to the LDS).

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

#define NS 16
#define LOG2E 1.4426950408889634
#define LN2 0.6931471805599453
#define LN10F 2.302585092994046f
#define LOG2_FLT_MAX 127.99999991400867  // log2(FLT_MAX): the reference's min(f32::MAX, k), reaction.rs:138
#define LOG2_FLT_MIN (-126.0f)          // log2(FLT_MIN): the + f32::MIN_POSITIVE guards, reaction.rs:177,189
#define LOG2_ABSENT (-16384.f)          // log2 C of an absent species (C <= 0): exact in BF16; 2^x of any product with it is 0
#define NO_TERM (-1e300)

// Reaction tables (sorted reaction r):
//   rate[r] = (log2 A, β, Ea/R log2 e, -Σν) of k (Troe: log2 A∞ and the exponent of k∞/A∞);
//   k0[r] = (log2 A0 (Troe: log2(A0/A∞)), β0, Ea0/R log2 e, 0);  troe[r] = (α, 1/T3, 1/T1 (0: constant α), T2 (NaN: absent));
//   slots[r] = (info, 6 ΔG slots as 16-bit byte offsets into the G tables, two per word): info = model |
//   kinf_const << 4 (model 0 elementary, 1 irreversible, 2 three-body, 3 Lindemann, 4 Troe, 5 padding); a species
//   with ν = ±2 takes two slots, unused slots point at a zero row;
//   tile_info[16-row half] = the half's model (0 elementary and irreversible, 2, 3, 4; 5 padding only): the rows of
//   a half share one model, so branches on it are uniform;  eff: FP32 A operands of (e - 1) per half.
__device__ __forceinline__ u16 bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf16_value(u16 b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float log2f_hw(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float exp2f_hw(float x) { return __builtin_amdgcn_exp2f(x); }
// log2 C in FP32 over the whole FP64 range: the hardware logarithm on FP32-normal values; below FLT_MIN, the binary
// exponent plus the logarithm of the mantissa; C <= 0 (a species absent from the state) as LOG2_ABSENT, so that the
// reactions it enters get zero rates, as the reference's products of concentrations do
__device__ __forceinline__ float log2_concentration(double c) {
  if (!(c > 0.0)) return LOG2_ABSENT;
  if (c >= 1.1754943508222875e-38) return log2f_hw((float)c);
  int e;
  const double m = __builtin_frexp(c, &e);
  return (float)e + log2f_hw((float)m);
}
// log2(1 + 2^x) in FP32
__device__ __forceinline__ float softplus2(float x) {
  return x > 0.f ? x + log2f_hw(1.f + exp2f_hw(-x)) : log2f_hw(1.f + exp2f_hw(x));
}
// 2^d - 1 for |d| < 1/4: expm1(u), u = d ln 2, by its Taylor series to u^7 (relative error below 1e-9)
__device__ __forceinline__ float exp2m1_small(float d) {
  const float u = d * 0.6931471805599453f;
  float p = 1.f / 5040.f;
  p = p * u + 1.f / 720.f;
  p = p * u + 1.f / 120.f;
  p = p * u + 1.f / 24.f;
  p = p * u + 1.f / 6.f;
  p = p * u + 0.5f;
  p = p * u + 1.f;
  return p * u;
}

struct Lds {
  union {
    u16 lnC[KB][3][64][8];     // B operands: [block][split][lane][j] = log2 C[32 kb + 8 (lane/16) + j][lane % 16]
    double G[2][KA + 1][NS];   // clamped Gibbs tables (plus, minus); row KA of each is zero
  };
  float C[KE * 4][NS];
  double log2T[NS], rcpT[NS], log2C0[NS];
  float T[NS], Ctot[NS];
};

// x_f, x_r of this lane's 4 reactions (rows 4 grp + i) of one 16-row half.
__device__ __forceinline__ void half_tile(int h, int rt, int lane, int grp, int col, const Lds& L, const char* gcol,
                                          const bf16x8 (&Bh)[KB], const bf16x8 (&Bm)[KB], const bf16x8 (&Bl)[KB],
                                          const bf16x8* __restrict__ A_fwd, const bf16x8* __restrict__ A_rev,
                                          const f64x4* __restrict__ rate, const f64x4* __restrict__ k0,
                                          const f32x4* __restrict__ troe, const i32x4* __restrict__ slots,
                                          const int* __restrict__ tile_info, const float* __restrict__ eff,
                                          double my_log2T, double my_rcpT, double my_log2C0, float my_T, float my_Ctot,
                                          double (&xf)[4], double (&xr)[4]) {
  const int ht = 2 * rt + h;
  const int htype = tile_info[ht];  // the half's model, uniform over the wave
  if (htype == 5) {                 // padding rows only
    #pragma unroll
    for (int i = 0; i < 4; i++) xf[i] = xr[i] = NO_TERM;
    return;
  }
  f64x4 kr[4];
  i32x4 sl[4];
  #pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = 16 * ht + 4 * grp + i;
    kr[i] = rate[r];
    sl[i] = slots[r];
  }
  f32x4 fh = {0.f, 0.f, 0.f, 0.f}, fl = fh, rh = fh, rl = fh;
  for (int kb = 0; kb < KB; kb++) {
    const bf16x8 af = A_fwd[(ht * KB + kb) * 64 + lane];
    const bf16x8 ar = A_rev[(ht * KB + kb) * 64 + lane];
    fh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, Bh[kb], fh, 0, 0, 0);
    rh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, Bh[kb], rh, 0, 0, 0);
    fl = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, Bm[kb], fl, 0, 0, 0);
    rl = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, Bm[kb], rl, 0, 0, 0);
    fl = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, Bl[kb], fl, 0, 0, 0);
    rl = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar, Bl[kb], rl, 0, 0, 0);
  }
  f32x4 Mv = {0.f, 0.f, 0.f, 0.f};
  if (htype >= 2) {
    const float* A = eff + (size_t)ht * KE * 64;
    for (int kk = 0; kk < KE; kk++) Mv = __builtin_amdgcn_mfma_f32_16x16x4f32(A[kk * 64 + lane], L.C[kk * 4 + grp][col], Mv, 0, 0, 0);
  }
  #pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = 16 * ht + 4 * grp + i;
    const int info = sl[i][0];
    const int rtype = info & 15;  // the row's model: the half's, 1 (irreversible, in elementary halves) or 5 (padding)
    double lk = fmin(kr[i][0] + kr[i][1] * my_log2T - kr[i][2] * my_rcpT, LOG2_FLT_MAX);
    if (htype == 2) {
      lk += (double)log2f_hw(my_Ctot + Mv[i]);
    } else if (htype >= 3) {
      // falloff: log2 k_f = log2 k∞ + log2(Pr / (1 + Pr)) (+ log2 F for Troe), Pr = M k0 / k∞ (reaction.rs:149-193)
      const double lM = (double)log2f_hw(my_Ctot + Mv[i]);
      const f64x4 c0 = k0[r];
      const double lk0 = fmin(c0[0] + c0[1] * my_log2T - c0[2] * my_rcpT, LOG2_FLT_MAX);
      const double lkinf = htype == 4 ? ((info >> 4) & 1 ? 0.0 : fmin(kr[i][1] * my_log2T - kr[i][2] * my_rcpT, LOG2_FLT_MAX)) : lk;
      const double lPr = lM + lk0 - lkinf;
      const double lfrac = -(double)softplus2((float)-lPr);
      double lF = 0.0;
      if (htype == 4) {
        const f32x4 t = troe[r];
        const float e3 = exp2f_hw(-my_T * t[1] * (float)LOG2E);
        float Fcent = (1.f - t[0]) * e3 + (t[2] != 0.f ? t[0] * exp2f_hw(-my_T * t[2] * (float)LOG2E) : t[0]);
        if (t[3] == t[3]) Fcent += exp2f_hw(-t[3] * __builtin_amdgcn_rcpf(my_T) * (float)LOG2E);  // T2 given (not NaN)
        const float lnFcent = log2f_hw(Fcent) * (float)LN2;
        const float c = -0.4f * LN10F - 0.67f * lnFcent;
        const float N = 0.75f * LN10F - 1.27f * lnFcent;
        const float u = fmaxf((float)lPr, LOG2_FLT_MIN) * (float)LN2 + c;
        const float f1 = u * __builtin_amdgcn_rcpf(N - 0.14f * u);
        lF = (double)(lnFcent * __builtin_amdgcn_rcpf(f1 * f1 + 1.f) * (float)LOG2E);
      }
      lk = (htype == 4 ? kr[i][0] : 0.0) + lkinf + lfrac + lF;
    }
    // log2(1/Kc) = Σ over the slots of G± + (-Σν) log2 C0
    double lK = kr[i][3] * my_log2C0;
    #pragma unroll
    for (int w = 1; w < 4; w++) {
      const unsigned pair = (unsigned)sl[i][w];
      lK += *(const double*)(gcol + (pair & 0xFFFFu));
      lK += *(const double*)(gcol + (pair >> 16));
    }
    xf[i] = rtype == 5 ? NO_TERM : lk + (double)fh[i] + (double)fl[i];
    xr[i] = rtype == 5 || rtype == 1 ? NO_TERM : lk + (double)rh[i] + (double)rl[i] + lK;
  }
}

extern "C" __global__ void __launch_bounds__(64 * WAVES)
rates_mixed(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
            const bf16x8* __restrict__ A_fwd, const bf16x8* __restrict__ A_rev, const bf16x8* __restrict__ A_nu,
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

  // 1. Per state (col) and species group (grp): concentrations, log2 C, then the Gibbs tables.
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
    L.log2T[col] = lnT * LOG2E;
    L.rcpT[col] = rcpT;
    L.log2C0[col] = (log(101325.0 / (1.380649e-23 * 6.02214076e23)) - lnT) * LOG2E;
    L.T[col] = (float)T;
    L.Ctot[col] = (float)C_tot;
  }
  __syncthreads();
  for (int k = grp; k < KB * 32; k += 4) {
    // the FP64 concentration again (L.C holds it in FP32, which flushes values below 1.4e-45)
    const double c = k < K - 1 ? density * fmax(0.0, (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k]) : C_tot - partial;
    const float x = k < K ? log2_concentration(c) : 0.f;
    const u16 h = bf16_bits(x);
    const float r1 = x - bf16_value(h);
    const u16 m = bf16_bits(r1);
    const u16 l = bf16_bits(r1 - bf16_value(m));
    const int kb = k >> 5, r = k & 31, lane_b = ((r >> 3) << 4) | col, j = r & 7;
    L.lnC[kb][0][lane_b][j] = h;
    L.lnC[kb][1][lane_b][j] = m;
    L.lnC[kb][2][lane_b][j] = l;
  }
  __syncthreads();
  bf16x8 Bh[KB], Bm[KB], Bl[KB];
  for (int kb = 0; kb < KB; kb++) {
    Bh[kb] = *(const bf16x8*)L.lnC[kb][0][lane];
    Bm[kb] = *(const bf16x8*)L.lnC[kb][1][lane];
    Bl[kb] = *(const bf16x8*)L.lnC[kb][2][lane];
  }
  __syncthreads();  // the G tables reuse the B staging area
  for (int k = grp; k < KA; k += 4) {
    const double* a = gibbs + (size_t)k * 14 + (T <= split[k] ? 0 : 7);
    const double g = a[0] + a[1] * lnT + a[2] * T + a[3] * (T * T) + a[4] * (T * T * T) + a[5] * (T * T * T * T) + a[6] * rcpT;
    L.G[0][k][col] = fmin(g, ln_clamp_pos) * LOG2E;
    L.G[1][k][col] = fmin(-g, ln_clamp_neg) * LOG2E;
  }
  if (grp == 0) {
    L.G[0][KA][col] = 0.0;
    L.G[1][KA][col] = 0.0;
  }
  __syncthreads();
  const double my_log2T = L.log2T[col], my_rcpT = L.rcpT[col], my_log2C0 = L.log2C0[col];
  const float my_T = L.T[col], my_Ctot = L.Ctot[col];
  const char* gcol = (const char*)&L.G[0][0][col];

  f32x4 om[AB];
  #pragma unroll
  for (int a = 0; a < AB; a++) om[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
  int E = -100000;  // block exponent of this lane's state column

  for (int rt = 0; rt < RT; rt++) {
    double xf0[4], xr0[4], xf1[4], xr1[4];  // the two halves' exponents (register arrays: constant indices only)
    half_tile(0, rt, lane, grp, col, L, gcol, Bh, Bm, Bl, A_fwd, A_rev, rate, k0, troe, slots, tile_info, eff, my_log2T, my_rcpT,
              my_log2C0, my_T, my_Ctot, xf0, xr0);
    half_tile(1, rt, lane, grp, col, L, gcol, Bh, Bm, Bl, A_fwd, A_rev, rate, k0, troe, slots, tile_info, eff, my_log2T, my_rcpT,
              my_log2C0, my_T, my_Ctot, xf1, xr1);
    double xf[8], xr[8];
    #pragma unroll
    for (int i = 0; i < 4; i++) {
      xf[i] = xf0[i], xr[i] = xr0[i];
      xf[4 + i] = xf1[i], xr[4 + i] = xr1[i];
    }
    // block exponent: the largest 2^x of the tile over the state's 4 lanes
    double xm = fmax(xf[0], xr[0]);
    #pragma unroll
    for (int i = 1; i < 8; i++) xm = fmax(xm, fmax(xf[i], xr[i]));
    int e = xm > -1e299 ? (int)__builtin_ceil(xm) : -100000;
    e = max(e, __shfl_xor(e, 16));
    e = max(e, __shfl_xor(e, 32));
    if (e > E) {
      if (E > -100000) {
        const float f = __builtin_ldexpf(1.f, max(E - e, -126));
        #pragma unroll
        for (int a = 0; a < AB; a++) om[a] *= f;
      }
      E = e;
    }
    const double Ed = (double)E;
    float q[8];
    #pragma unroll
    for (int i = 0; i < 8; i++) {
      const float a = exp2f_hw((float)(xf[i] - Ed));
      const float d = (float)(xr[i] - xf[i]);  // FP64 difference of the two exponents
      q[i] = fabsf(d) < 0.25f ? -a * exp2m1_small(d) : a - exp2f_hw((float)(xr[i] - Ed));
    }
    // q = hi + mid + lo in BF16, the B operand of ν·q (element j of lane group g: reaction j < 4 ? 4g + j : 16 + 4g + j - 4)
    bf16x8 qh, qm, ql;
    #pragma unroll
    for (int i = 0; i < 8; i += 2) {
      const f32x2 v = {q[i], q[i + 1]};
      const bf16x2 h = __builtin_convertvector(v, bf16x2);
      const f32x2 r1 = v - (f32x2){(float)h[0], (float)h[1]};
      const bf16x2 m = __builtin_convertvector(r1, bf16x2);
      const f32x2 r2 = r1 - (f32x2){(float)m[0], (float)m[1]};
      const bf16x2 l = __builtin_convertvector(r2, bf16x2);
      qh[i] = h[0], qh[i + 1] = h[1];
      qm[i] = m[0], qm[i + 1] = m[1];
      ql[i] = l[0], ql[i + 1] = l[1];
    }
    #pragma unroll
    for (int a = 0; a < AB; a++) {
      const bf16x8 an = A_nu[(rt * AB + a) * 64 + lane];
      om[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(an, qh, om[a], 0, 0, 0);
      om[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(an, qm, om[a], 0, 0, 0);
      om[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(an, ql, om[a], 0, 0, 0);
    }
  }
  if (!valid) return;
  #pragma unroll
  for (int a = 0; a < AB; a++)
    #pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = 16 * a + 4 * grp + i;
      if (k < KA) out[k][s] = E > -100000 ? __builtin_ldexp((double)om[a][i], E) : 0.0;
    }
}
