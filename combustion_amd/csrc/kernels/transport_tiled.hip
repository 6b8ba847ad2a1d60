// Mixture-averaged transport for gfx950 (hand-written; the host prepends the sizes and the precision), for
// mechanisms too large for the straight-line transport program (K > 64) and for lists of states.
//
// A block of WAVES waves owns NS = 16 states; species, Wilke row tiles and diffusion tiles are split over the
// waves. With x_k = C_k the concentrations and the fits in ℓ = ln(T/T0) (transport.rs:125-136):
//   λ = ½ √T (Σ x_k c_k / N + N / Σ x_k / c_k)                                  (transport.rs:141-146)
//   μ = √T Σ_k x_k v_k² / den_k,  den_k = Σ_j x_j (a_kj + b_kj v_k / v_j)²       (transport.rs:148-158)
//       = (A2·x)_k + 2 v_k (AB·(x/v))_k + v_k² (B2·(x/v²))_k: three K×K contractions on MFMA (16x16x4, FP64 or
//       FP32), species × states, one 16-row tile of k per wave at a time
//   ρD_k = √T (Σ_j W_j x_j - W_k x_k) / (R S_k),  S_k = Σ_{j≠k} x_j / d_kj(ℓ)     (transport.rs:162-177)
//       d_kj for j < k on MFMA: (16 pairs × 4 coefficients)·(ℓ⁰..ℓ³ × 16 states) per tile, a tile being 16
//       consecutive j of one k; 1/d on VALU (in FP64 four at a time from the reciprocal of their product). Row by row (k = 1..K-1), the waves split row k's tiles by j block
//       (block b always to wave b mod WAVES): each adds x_k/d_kj into its register sums for S_j (no other wave
//       touches S_j) and sums x_j/d_kj across its lanes; after a barrier the owner of k's block adds the waves' row
//       sums, in wave order, into S_k in LDS, and after the last row each wave adds its register sums there. One S
//       array serves the block (it reuses the 1/v_k array of the viscosity stage), and every state adds the terms in
//       the same order.
// Outputs λ [W/m/K], μ [Pa s], D_k = ρD_k/ρ [m²/s] with ρ = Σ W_k x_k.
// Required defines: K, KP (K rounded up to 16), NTILE (diffusion tiles, K-1 rows of ⌈k/16⌉), DK (fit length
// rounded up to 4: 4 or 8),
// F64 (1: FP64 arithmetic and v_mfma_f64_16x16x4f64; 0: FP32 arithmetic and v_mfma_f32_16x16x4f32).
// With a state list (idx, count) the kernel evaluates only the listed states, in place.

// Output row of (lane group g, element i) in the 16x16 MFMA result: FP32 holds rows 4g..4g+3 in a lane,
// the FP64 instruction interleaves them (rows g, g+4, g+8, g+12).
#if F64
#define ROW(g, i) ((g) + 4 * (i))
typedef double real;
typedef double r4 __attribute__((ext_vector_type(4)));
#define MFMA(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0)
__device__ __forceinline__ real rcp(real b) {
  real r = __builtin_amdgcn_rcp(b);
  real e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ real sqrt_real(double x) { return __builtin_sqrt(x); }
// Four reciprocals from one: 1/d0 = d1 (d2 d3) / (d0 d1 d2 d3) and so on (a few ulps; the d are binary-diffusion
// fits, far from the ends of the FP64 range)
__device__ __forceinline__ void rcp4(const real (&d)[4], real (&r)[4]) {
  const real p01 = d[0] * d[1], p23 = d[2] * d[3];
  const real q = rcp(p01 * p23);
  const real q01 = q * p23, q23 = q * p01;
  r[0] = q01 * d[1];
  r[1] = q01 * d[0];
  r[2] = q23 * d[3];
  r[3] = q23 * d[2];
}
#else
#define ROW(g, i) (4 * (g) + (i))
typedef float real;
typedef float r4 __attribute__((ext_vector_type(4)));
#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0)
__device__ __forceinline__ real rcp(real b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ real sqrt_real(double x) { return __builtin_sqrtf((float)x); }
__device__ __forceinline__ void rcp4(const real (&d)[4], real (&r)[4]) {
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = rcp(d[i]);
}
#endif

#define NS 16
#define WAVES 4
#define RGAS (1.380649e-23 * 6.02214076e23)
#define KG ((K + 15) / 16)  // species per lane group (16 groups: 4 per wave)
#define QB ((KP / 16 + WAVES - 1) / WAVES)  // diffusion column blocks per wave

struct Lds {
  real x[KP][NS];  // concentrations (0 past K)
  union {
    real r[KP][NS];  // 1 / v_k (1 past K), in the viscosity stage
    real S[KP][NS];  // Σ_{j≠k} x_j / d_kj, in the diffusion stage
  };
  real pw[DK][NS];                 // ℓ^d
  real part[2][WAVES][NS];         // the waves' sums over one row of the diffusion stage (by row parity)
  double red[WAVES][4][NS];        // per-wave partial reductions
};

// Sum over the 4 lane groups of a wave (lanes col, col+16, col+32, col+48), then over the waves in order.
__device__ __forceinline__ double block_sum(Lds& L, int slot, double v, int wave, int grp, int col) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  if (grp == 0) L.red[wave][slot][col] = v;
  __syncthreads();
  double s = L.red[0][slot][col];
  for (int w = 1; w < WAVES; w++) s += L.red[w][slot][col];
  return s;
}

// three waves per SIMD (at most 168 registers), which the LDS of K = 160 (44 KiB per block) also allows
extern "C" __global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(3)))
transport_tiled(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
                const double* __restrict__ rcp_W, const double* __restrict__ W, const real* __restrict__ fit_c,
                const real* __restrict__ fit_v, const real* __restrict__ wilke, const real* __restrict__ dcoef,
                double T0, const unsigned* __restrict__ idx, const unsigned* __restrict__ count) {
  __shared__ Lds L;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4, g16 = 4 * wave + grp;
  // With idx, the block's states are idx[slot] for the first *count slots (the surrogate's fallback list).
  const unsigned ns = idx ? *count : n;
  if (blockIdx.x * NS >= ns) return;  // uniform over the block
  const unsigned slot = blockIdx.x * NS + col;
  const bool valid = slot < ns;
  const unsigned s = valid ? (idx ? idx[slot] : slot) : 0;

  // 1. Concentrations (amounts n_k = Y_k/W_k, C_k = max(0, n_k) C_tot/Σn, bulk C_tot - Σ), fits and the
  //    conductivity sums; this lane's species are k = g16 + 16 m.
  const double T = valid ? in[0][s] : 1000.0;
  const double P = valid ? in[1][s] : 101325.0;
  double nk[KG];
  double part = 0;
#pragma unroll
  for (int m = 0; m < KG; m++) {
    const int k = g16 + 16 * m;
    nk[m] = k < K ? (valid ? in[2 + k][s] : 1.0 / K) * rcp_W[k] : 0.0;
    part += nk[m];
  }
  const double sum_n = block_sum(L, 0, part, wave, grp, col);
  const double C_tot = P / RGAS / T;
  const double density = C_tot / sum_n;
  part = 0;
#pragma unroll
  for (int m = 0; m < KG; m++) {
    const int k = g16 + 16 * m;
    if (k < K - 1) part += density * fmax(0.0, nk[m]);
  }
  const double nonbulk = block_sum(L, 1, part, wave, grp, col);
  const double l = log(T / T0);
  if (g16 < DK) {
    real p = 1;
    for (int d = 0; d < g16; d++) p *= (real)l;
    L.pw[g16][col] = p;
  }
  real pw[DK];
  pw[0] = 1;
#pragma unroll
  for (int d = 1; d < DK; d++) pw[d] = pw[d - 1] * (real)l;
  double A = 0, B = 0, WN = 0;
#pragma unroll
  for (int m = 0; m < KG; m++) {
    const int k = g16 + 16 * m;
    real x = 0, r = 1;
    if (k < K) {
      const double Ck = k < K - 1 ? density * fmax(0.0, nk[m]) : C_tot - nonbulk;
      x = (real)Ck;
      WN += W[k] * Ck;
      real c = 0, v = 0;
#pragma unroll
      for (int d = 0; d < DK; d++) {
        c += fit_c[k * DK + d] * pw[d];
        v += fit_v[k * DK + d] * pw[d];
      }
      A += (double)(x * c);
      B += (double)(x * rcp(c));
      r = rcp(v);
    }
    L.x[k][col] = x;
    L.r[k][col] = r;
  }
  // (block_sum's barrier also publishes x, r and pw)
  A = block_sum(L, 2, A, wave, grp, col);
  B = block_sum(L, 3, B, wave, grp, col);
  WN = block_sum(L, 0, WN, wave, grp, col);

  // 2. Viscosity: rows k = 16 mt + ROW(grp, i) of the MFMA output, state = col.
  real mu = 0;
  for (int mt = wave; mt < KP / 16; mt += WAVES) {
    r4 d1 = {0, 0, 0, 0}, d2 = d1, d3 = d1;
    const real* a = wilke + (size_t)mt * (KP / 4) * 3 * 64 + lane;
#pragma unroll 4
    for (int kk = 0; kk < KP / 4; kk++) {
      const real x = L.x[4 * kk + grp][col], r = L.r[4 * kk + grp][col];
      const real y = x * r;
      d1 = MFMA(a[(kk * 3 + 0) * 64], x, d1);
      d2 = MFMA(a[(kk * 3 + 1) * 64], y, d2);
      d3 = MFMA(a[(kk * 3 + 2) * 64], y * r, d3);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = 16 * mt + ROW(grp, i);
      if (k < K) {
        const real v = rcp(L.r[k][col]);
        const real den = d1[i] + 2 * v * d2[i] + v * v * d3[i];
        mu += L.x[k][col] * v * v * rcp(den);
      }
    }
  }

  // 3. Binary diffusion: S (over the 1/v_k array) cleared once every wave is past the viscosity stage; tile
  //    t = (k, j0) gives rows j = j0 + ROW(grp, i); the tiles of row k start at t0 = Σ_{k' < k} ⌈k'/16⌉.
  __syncthreads();
#pragma unroll
  for (int m = 0; m < KG; m++) L.S[g16 + 16 * m][col] = 0;
  __syncthreads();
  const real lp0 = L.pw[grp][col];
  const real lp1 = DK > 4 ? L.pw[(4 + grp) % DK][col] : (real)0;
  // the column sums S_j += x_k / d_kj of this wave's j blocks (jb = wave + WAVES q) stay in registers until the end
  real acc[QB][4];
#pragma unroll
  for (int q = 0; q < QB; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) acc[q][i] = 0;
  int t0 = 0;
  for (int k = 1; k < K; k++) {
    const int nb = (k + 15) / 16;
    const real xk = L.x[k][col];
    real sk = 0;
#pragma unroll
    for (int q = 0; q < QB; q++) {
      const int jb = wave + WAVES * q;
      if (jb >= nb) break;  // uniform
      const int t = t0 + jb, j0 = 16 * jb;
      r4 d = {0, 0, 0, 0};
      d = MFMA(dcoef[(size_t)t * (DK / 4) * 64 + lane], lp0, d);
      if (DK > 4) d = MFMA(dcoef[((size_t)t * (DK / 4) + 1) * 64 + lane], lp1, d);
      real dv[4], R[4];
#pragma unroll
      for (int i = 0; i < 4; i++) dv[i] = j0 + ROW(grp, i) < k ? d[i] : (real)1;  // pairs past the diagonal: 1
      rcp4(dv, R);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int j = j0 + ROW(grp, i);
        if (j < k) {
          sk += L.x[j][col] * R[i];
          acc[q][i] += xk * R[i];
        }
      }
    }
    sk += __shfl_xor(sk, 16);
    sk += __shfl_xor(sk, 32);
    if (grp == 0) L.part[k & 1][wave][col] = sk;
    __syncthreads();
    // the wave that owns k's block (and so every other update of S_k) adds the row's sums; the next row's sums go
    // to the other buffer, and this buffer is rewritten only after the next barrier, which this wave reaches later
    if (wave == (k / 16) % WAVES && grp == 0) {
      real sum = L.part[k & 1][0][col];
#pragma unroll
      for (int w = 1; w < WAVES; w++) sum += L.part[k & 1][w][col];
      L.S[k][col] += sum;
    }
    t0 += nb;
  }
  // each S_j has one owner wave: its column sums join the row sums already there
#pragma unroll
  for (int q = 0; q < QB; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int j = 16 * (wave + WAVES * q) + ROW(grp, i);
      if (j < KP) L.S[j][col] += acc[q][i];
    }
  const double mu_all = block_sum(L, 1, (double)mu, wave, grp, col);

  // 4. Outputs: λ and μ by one lane per state, D_k for this lane's species.
  if (!valid) return;
  const double sqrtT = (double)sqrt_real(T);
  if (g16 == 0) {
    out[0][s] = 0.5 * sqrtT * (A / C_tot + C_tot / B);
    out[1][s] = sqrtT * mu_all;
  }
#pragma unroll
  for (int m = 0; m < KG; m++) {
    const int k = g16 + 16 * m;
    if (k < K) out[2 + k][s] = sqrtT * (WN - W[k] * (double)L.x[k][col]) / (RGAS * (double)L.S[k][col]) / WN;
  }
}
