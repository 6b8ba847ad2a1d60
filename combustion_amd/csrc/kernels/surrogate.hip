// Transport surrogate inference for gfx950 (hand-written; the host prepends the sizes).
//
// A block of 4 waves evaluates NB = 128 states. Per state, from the inputs (T, P, Y_0..Y_{K-1}):
//   mole fractions X_k = (Y_k/W_k) / Σ_j Y_j/W_j and the O(K) baselines, from the transport fits in ℓ = ln(T/T0)
//   (transport.rs:125-136; λ_k = c_k(ℓ)√T, η_k = v_k(ℓ)²√T, D_kb = d_kb(ℓ) T^1.5/P against the bulk species b):
//     λ_b = ½ (Σ X_k λ_k + 1/Σ X_k/λ_k)   the mixture rule of transport.rs:141-146 itself, which is O(K)
//     μ_b = Σ X_k η_k                       linear mixing instead of Wilke's O(K²) sum
//     D_b,k = D_kb (1 - Y_k)/(1 - X_k)      the mixture-averaged rule with every D_kj replaced by D_kb (b' = 0 for k = b)
//   features f = ((T, ln P, X_0..X_{K-1}) - mean)/std, zero padded to IN;
//   hidden layers h = SiLU(W h + b) of WIDTH; output r = W h + b of OUT (first K+2 valid);
//   [λ, μ, D_0..D_{K-1}] = baseline · exp(r σ + m), with (m, σ) the training set's standardisation.
// Layers run on v_mfma_f32_16x16x16bf16_1k: A = W rows (weights, BF16, read from memory), B = activations
// (features × states, BF16 in LDS), FP32 accumulators; wave w owns a quarter of the output features of every layer
// for all NB states, and the block rewrites the activation buffer in place after a barrier.
// States whose features leave the training box [lo, hi] are not written: their index goes to fallback[] (count in
// fallback_count, zeroed by the caller) for the plain transport kernel.
// Required defines: K, IN (features padded to 64), WIDTH (hidden width, multiple of 64), DEPTH (hidden layers),
// OUT (outputs padded to 64), DK (fit length, at most 8).

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

#define NB 128
#define HSTRIDE (WIDTH + 8)  // activation row stride (BF16 elements), padded against bank conflicts
#define NF (K + 2)

__device__ __forceinline__ u16 to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z * 1.4426950408889634f));
}
__device__ __forceinline__ float poly(const float* __restrict__ c, float l) {
  float p = c[DK - 1];
  for (int j = DK - 2; j >= 0; j--) p = p * l + c[j];
  return p;
}

struct Lds {
  u16 h[NB][HSTRIDE];
  float lnT[NB], lnP[NB], l[NB], rN[NB], lam[NB], mu[NB];
  int inside[NB];
};

// One layer: z[o][s] = Σ_i W[o][i] h[s][i] for this wave's output rows (RT row tiles of 16 from row0).
template <int RT, int NIN>
__device__ __forceinline__ void layer(const u16* __restrict__ W, int row0, Lds& L, f32x4 (&acc)[RT][NB / 16], int lane) {
  for (int r = 0; r < RT; r++)
    for (int c = 0; c < NB / 16; c++) acc[r][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int m = lane & 15, q = 4 * (lane >> 4);
#pragma unroll 2
  for (int k = 0; k < NIN; k += 16) {
    bf16x4 a[RT];
    for (int r = 0; r < RT; r++) a[r] = *(const bf16x4*)(W + (size_t)(row0 + 16 * r + m) * NIN + k + q);
    for (int c = 0; c < NB / 16; c++) {
      const bf16x4 b = *(const bf16x4*)&L.h[16 * c + m][k + q];
      for (int r = 0; r < RT; r++) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[r], b, acc[r][c], 0, 0, 0);
    }
  }
}

extern "C" __global__ void __launch_bounds__(256)
surrogate_transport(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
                    const u16* __restrict__ w16, const float* __restrict__ params, const long* __restrict__ offsets,
                    const float* __restrict__ f_mean, const float* __restrict__ f_rstd, const float* __restrict__ lo,
                    const float* __restrict__ hi, const float* __restrict__ o_mean, const float* __restrict__ o_std,
                    const float* __restrict__ fit_c, const float* __restrict__ fit_v, const float* __restrict__ fit_db,
                    const float* __restrict__ rcp_W, float T0, unsigned* __restrict__ fallback,
                    unsigned* __restrict__ fallback_count) {
  __shared__ Lds L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned s0 = blockIdx.x * NB;

  // 1. Per state (one thread each): mole fractions, baselines, features and the box test.
  if (tid < NB) {
    const int st = tid;
    const unsigned s = s0 + st;
    const bool valid = s < n;
    int inside = valid;
    const double Td = valid ? in[0][s] : 1000.0, Pd = valid ? in[1][s] : 101325.0;
    const float T = (float)Td, lnP = (float)log(Pd), l = (float)log(Td / (double)T0);
    float N = 0.f;
    for (int k = 0; k < K; k++) N += (valid ? (float)in[2 + k][s] : 1.f / K) * rcp_W[k];
    const float rN = 1.f / N, sqrtT = __builtin_sqrtf(T);
    float a = 0.f, b = 0.f, m = 0.f;
    for (int k = 0; k < K; k++) {
      const float X = (valid ? (float)in[2 + k][s] : 1.f / K) * rcp_W[k] * rN;
      const float lam = poly(fit_c + k * DK, l) * sqrtT, v = poly(fit_v + k * DK, l);
      a += X * lam;
      b += X * __builtin_amdgcn_rcpf(lam);
      m += X * v * v;
      const int f = 2 + k;
      if (!(X >= lo[f] && X <= hi[f])) inside = 0;
      L.h[st][f] = to_bf16((X - f_mean[f]) * f_rstd[f]);
    }
    if (!(T >= lo[0] && T <= hi[0]) || !(lnP >= lo[1] && lnP <= hi[1])) inside = 0;
    L.h[st][0] = to_bf16((T - f_mean[0]) * f_rstd[0]);
    L.h[st][1] = to_bf16((lnP - f_mean[1]) * f_rstd[1]);
    for (int f = NF; f < IN; f++) L.h[st][f] = 0;
    L.lnT[st] = __builtin_logf(T);
    L.lnP[st] = lnP;
    L.l[st] = l;
    L.rN[st] = rN;
    L.lam[st] = 0.5f * (a + __builtin_amdgcn_rcpf(b));
    L.mu[st] = m * sqrtT;
    L.inside[st] = inside;
  }
  __syncthreads();

  // 2. Hidden layers, each wave WIDTH/4 output features for all NB states.
  constexpr int RH = WIDTH / 64;
  for (int d = 0; d < DEPTH; d++) {
    const long w_off = offsets[2 * d], b_off = offsets[2 * d + 1];
    f32x4 acc[RH][NB / 16];
    if (d == 0)
      layer<RH, IN>(w16 + w_off, wave * (WIDTH / 4), L, acc, lane);
    else
      layer<RH, WIDTH>(w16 + w_off, wave * (WIDTH / 4), L, acc, lane);
    __syncthreads();  // every wave has read the previous activations
    for (int r = 0; r < RH; r++) {
      const int o = wave * (WIDTH / 4) + 16 * r + 4 * (lane >> 4);
      const float b0 = params[b_off + o], b1 = params[b_off + o + 1], b2 = params[b_off + o + 2], b3 = params[b_off + o + 3];
      for (int c = 0; c < NB / 16; c++) {
        const int st = 16 * c + (lane & 15);
        const f32x4 z = acc[r][c];
        u16 v[4] = {to_bf16(silu(z[0] + b0)), to_bf16(silu(z[1] + b1)), to_bf16(silu(z[2] + b2)), to_bf16(silu(z[3] + b3))};
        __builtin_memcpy(&L.h[st][o], v, 8);
      }
    }
    __syncthreads();
  }

  // 3. Output layer (OUT/4 features per wave), baseline and exponential; fallback list for states outside the box.
  constexpr int RO = OUT / 64;
  f32x4 acc[RO][NB / 16];
  layer<RO, WIDTH>(w16 + offsets[2 * DEPTH], wave * (OUT / 4), L, acc, lane);
  const long b_off = offsets[2 * DEPTH + 1];
  for (int r = 0; r < RO; r++)
    for (int c = 0; c < NB / 16; c++) {
      const int st = 16 * c + (lane & 15);
      const unsigned s = s0 + st;
      if (s >= n || !L.inside[st]) continue;
      for (int i = 0; i < 4; i++) {
        const int o = wave * (OUT / 4) + 16 * r + 4 * (lane >> 4) + i;
        if (o >= NF) continue;
        float lnbase;
        if (o == 0) {
          lnbase = __builtin_logf(L.lam[st]);
        } else if (o == 1) {
          lnbase = __builtin_logf(L.mu[st]);
        } else {
          const int k = o - 2;
          const float Yk = (float)in[2 + k][s], Xk = Yk * rcp_W[k] * L.rN[st];
          lnbase = __builtin_logf(poly(fit_db + k * DK, L.l[st])) + 1.5f * L.lnT[st] - L.lnP[st] +
                   __builtin_logf(fmaxf(1.f - Yk, 1e-30f)) - __builtin_logf(fmaxf(1.f - Xk, 1e-30f));
        }
        const float y = (acc[r][c][i] + params[b_off + o]) * o_std[o] + o_mean[o] + lnbase;
        out[o][s] = (double)__builtin_expf(y);
      }
    }
  if (tid < NB) {
    const unsigned s = s0 + tid;
    if (s < n && !L.inside[tid]) fallback[atomicAdd(fallback_count, 1u)] = s;
  }
}
