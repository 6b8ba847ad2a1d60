// Transport surrogate inference for gfx950 (hand-written; the host prepends the sizes).
//
// A block of 4 waves evaluates NB = 128 states. Per state: features f = ((T, ln P, Y_0..Y_{K-1}) - mean)/std, zero
// padded to IN; hidden layers h = SiLU(W h + b) of WIDTH; output r = W h + b of OUT (first K+2 valid). Outputs are
// [λ, μ, D_0..D_{K-1}] = base · exp(r·σ + m): the network predicts the log of each property relative to a cheap
// baseline (the bulk species' conductivity and viscosity fits and the binary diffusion coefficients against the
// bulk species, transport.rs:125-136 fits, D_kb = d_kb(ℓ) T^1.5 / P), standardised by (m, σ).
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

struct Lds {
  u16 h[NB][HSTRIDE];
  float lnT[NB], lnP[NB], l[NB];
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
                    const float* __restrict__ base_c, const float* __restrict__ base_v, const float* __restrict__ base_d,
                    float T0, unsigned* __restrict__ fallback, unsigned* __restrict__ fallback_count) {
  __shared__ Lds L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned s0 = blockIdx.x * NB;

  // 1. Features (state = tid % NB, features split over the two halves of the block) and the box test.
  if (tid < NB) {
    const unsigned s = s0 + tid;
    const bool valid = s < n;
    const double T = valid ? in[0][s] : 1000.0, P = valid ? in[1][s] : 101325.0;
    L.lnT[tid] = (float)log(T);
    L.lnP[tid] = (float)log(P);
    L.l[tid] = (float)log(T / (double)T0);
    L.inside[tid] = valid;
  }
  __syncthreads();
  for (int idx = tid; idx < NB * IN; idx += 256) {
    const int st = idx % NB, f = idx / NB;
    const unsigned s = s0 + st;
    float v = 0.f;
    if (f < NF && s < n) {
      const float x = f == 0 ? (float)in[0][s] : f == 1 ? L.lnP[st] : (float)in[f][s];
      if (!(x >= lo[f] && x <= hi[f])) L.inside[st] = 0;  // benign race: every writer stores 0
      v = (x - f_mean[f]) * f_rstd[f];
    }
    L.h[st][f] = to_bf16(v);
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
      const float l = L.l[st], lnT = L.lnT[st], lnP = L.lnP[st];
      for (int i = 0; i < 4; i++) {
        const int o = wave * (OUT / 4) + 16 * r + 4 * (lane >> 4) + i;
        if (o >= NF) continue;
        const float* cf = o == 0 ? base_c : o == 1 ? base_v : base_d + (size_t)(o - 2) * DK;
        float p = cf[DK - 1];
        for (int j = DK - 2; j >= 0; j--) p = p * l + cf[j];
        // ln base: λ = c √T, μ = v² √T, D = d T^1.5 / P
        const float lnbase = o == 0 ? __builtin_logf(p) + 0.5f * lnT
                           : o == 1 ? 2.f * __builtin_logf(p) + 0.5f * lnT
                                    : __builtin_logf(p) + 1.5f * lnT - lnP;
        const float y = (acc[r][c][i] + params[b_off + o]) * o_std[o] + o_mean[o] + lnbase;
        out[o][s] = (double)__builtin_expf(y);
      }
    }
  if (tid < NB) {
    const unsigned s = s0 + tid;
    if (s < n && !L.inside[tid]) fallback[atomicAdd(fallback_count, 1u)] = s;
  }
}
