// Transport surrogate inference for gfx950 (hand-written; the host prepends the sizes).
//
// A block of 4 waves evaluates NB = 64 states. Per state, from the inputs (T, P, Y_0..Y_{K-1}):
//   mole fractions X_k = (Y_k/W_k) / Σ_j Y_j/W_j and the O(K·NM) baselines, from the transport fits in ℓ = ln(T/T0)
//   (transport.rs:125-136; λ_k = c_k(ℓ)√T, v_k = v_k(ℓ) with η_k = v_k²√T, D_kj = d_kj(ℓ) T^1.5/P): the mixture rules
//   with the pair sums restricted to the NM major species S and the rest X_rest = 1 - Σ_S X_j given to the bulk
//   species b (X'_k = X_rest - [k ∉ S] X_k; r(k) = b, or another major species for k = b):
//     λ_b = ½ (Σ X_k λ_k + 1/Σ X_k/λ_k)                                 transport.rs:141-146 itself, O(K)
//     μ_b = √T Σ_k X_k v_k² / den_k,  den_k = Σ_S X_j φ_kj + X'_k φ_kb + [k ∉ S] X_k,  φ_kj = (a_kj + b_kj v_k/v_j)²
//                                                                        (Wilke, transport.rs:148-158)
//     D_b,k = T^1.5 (1 - Y_k) / (P S_k),  S_k = Σ_{S, j≠k} X_j/d_kj + X'_k/d_k,r(k)   (transport.rs:162-177)
//   features f = ((T, ln P, X_0..X_{K-1}) - mean)/std, zero padded to IN;
//   hidden layers h = SiLU(W h + b) of WIDTH; output r = W h + b of OUT (first K+2 valid);
//   [λ, μ, D_0..D_{K-1}] = baseline · exp(r σ + m), with (m, σ) the training set's standardisation.
// Stage 1 puts one lane per state and splits the species k over the 4 waves (coalesced loads of Y_k; the tables of
// a species are wave-uniform); the major species' X_j and 1/v_j go through LDS (the activation rows' unused columns)
// to every wave, and partial sums meet in LDS in wave order. Layers run on v_mfma_f32_16x16x32_bf16: A = 16 weight
// rows × 32 inputs (16 B per lane from memory, the next step's fragment loaded while the current one multiplies),
// B = 32 inputs × 16 states (16 B per lane from the activation rows in LDS), FP32 accumulators; wave w owns a quarter of each layer's output
// rows for all NB states, and the block rewrites the activation rows in place between two barriers. The output
// layer's values go through LDS so that every output row is stored as 64 consecutive doubles.
// States whose features leave the training box [lo, hi] are not written: their index goes to fallback[] (count in
// fallback_count, zeroed by the caller) for the plain transport kernel.
// Required defines: K, IN (inputs padded to 64), WIDTH (hidden width, multiple of 64), DEPTH (hidden layers),
// OUT (outputs padded to 64), DK (fit length, at most 8), NM (major species).
// Tables: major[NM] (species indices), slot[K] (position of k in S or -1), wa/wb [K][NM+1] (Wilke a_kj, b_kj for
// j = major[i], slot NM for b), dfit [K][NM+1][DK] (d_kj for j = major[i], slot NM for r(k)).

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

#define NB 64
#define NF (K + 2)
#define KW ((K + 3) / 4)  // species per wave in stage 1
// Activation row (BF16 elements): the inputs and the stage-1 scratch, or the hidden width, rounded up to 16 and plus 8,
// an odd number of 16-byte units, so that the 16 rows of a B fragment start in distinct bank groups.
#define ROW_NEED (IN + 2 * (2 * NM + 1) > WIDTH ? IN + 2 * (2 * NM + 1) : WIDTH)
#define HS ((ROW_NEED + 15) / 16 * 16 + 8)

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
  u16 h[NB][HS];         // activations: state-major rows
  float y[NF][NB];       // ln baselines, then the outputs
  float part[4][3][NB];  // per-wave partial sums: Σn, then (Σ Xλ, Σ X/λ, Σ Xη)
  int inside[4][NB];
};  // 52 KB at K=53, WIDTH=256: three blocks per CU

// Stage-1 scratch in the activation row of state st, past the IN feature columns: X_j of the major species at
// [0, NM), 1/v_j at [NM, 2 NM), 1/v_b at 2 NM (floats).
__device__ __forceinline__ void put(Lds& L, int st, int i, float x) { __builtin_memcpy(&L.h[st][IN + 2 * i], &x, 4); }
__device__ __forceinline__ float get(const Lds& L, int st, int i) {
  float x;
  __builtin_memcpy(&x, &L.h[st][IN + 2 * i], 4);
  return x;
}

// z[o][s] = Σ_i W[o][i] h[s][i] for the RT row tiles of 16 from row0 and the 4 column tiles of 16 states.
template <int RT, int NIN>
__device__ __forceinline__ void layer(const u16* __restrict__ W, int row0, Lds& L, f32x4 (&acc)[RT][4], int lane) {
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int m = lane & 15, q = 8 * (lane >> 4);
  const u16* wp = W + (size_t)(row0 + m) * NIN + q;
  bf16x8 a[RT];
#pragma unroll
  for (int r = 0; r < RT; r++) a[r] = *(const bf16x8*)(wp + (size_t)16 * r * NIN);
#pragma unroll
  for (int k = 0; k < NIN; k += 32) {
    bf16x8 an[RT];
    if (k + 32 < NIN) {
#pragma unroll
      for (int r = 0; r < RT; r++) an[r] = *(const bf16x8*)(wp + (size_t)16 * r * NIN + k + 32);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const bf16x8 b = *(const bf16x8*)&L.h[16 * c + m][k + q];
#pragma unroll
      for (int r = 0; r < RT; r++) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b, acc[r][c], 0, 0, 0);
    }
    if (k + 32 < NIN) {
#pragma unroll
      for (int r = 0; r < RT; r++) a[r] = an[r];
    }
  }
}

extern "C" __global__ void __launch_bounds__(256, 3)  // 3 blocks (12 waves) per CU: 168 VGPRs, 52 KB LDS
surrogate_transport(const double* const* __restrict__ in, double* const* __restrict__ out, unsigned n,
                    const u16* __restrict__ w16, const float* __restrict__ params, const long* __restrict__ offsets,
                    const float* __restrict__ f_mean, const float* __restrict__ f_rstd, const float* __restrict__ lo,
                    const float* __restrict__ hi, const float* __restrict__ o_mean, const float* __restrict__ o_std,
                    const float* __restrict__ fit_c, const float* __restrict__ fit_v, const int* __restrict__ major,
                    const int* __restrict__ slot, const float* __restrict__ wa, const float* __restrict__ wb,
                    const float* __restrict__ dfit, const float* __restrict__ rcp_W, float T0, unsigned* __restrict__ fallback,
                    unsigned* __restrict__ fallback_count) {
  __shared__ Lds L;
  // The wave index in an SGPR: every table indexed by a species k = wave + 4j is then read with scalar loads.
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const unsigned s0 = blockIdx.x * NB;

  // 1. One lane per state; wave w takes the species k = w + 4j.
  const unsigned s = s0 + lane;
  const bool valid = s < n;
  const double Td = valid ? in[0][s] : 1000.0, Pd = valid ? in[1][s] : 101325.0;
  const float T = (float)Td, lnP = (float)log(Pd), l = (float)log(Td / (double)T0), lnT = __builtin_logf(T);
  const float sqrtT = __builtin_sqrtf(T);
  float nk[KW];  // Y_k / W_k
  float part = 0.f;
#pragma unroll
  for (int j = 0; j < KW; j++) {
    const int k = wave + 4 * j;
    nk[j] = 0.f;
    if (k < K) {
      nk[j] = (valid ? (float)in[2 + k][s] : 1.f / K) * rcp_W[k];
      part += nk[j];
    }
  }
  L.part[wave][0][lane] = part;
  __syncthreads();
  const float rN = 1.f / (L.part[0][0][lane] + L.part[1][0][lane] + L.part[2][0][lane] + L.part[3][0][lane]);
  float a = 0.f, b = 0.f, mu = 0.f;
  float xk[KW], vk[KW];
  int inside = valid;
#pragma unroll
  for (int j = 0; j < KW; j++) {
    const int k = wave + 4 * j;
    xk[j] = 0.f;
    vk[j] = 1.f;
    if (k < K) {
      const float X = nk[j] * rN;
      const float lam = poly(fit_c + k * DK, l) * sqrtT, v = poly(fit_v + k * DK, l);
      xk[j] = X;
      vk[j] = v;
      a += X * lam;
      b += X * __builtin_amdgcn_rcpf(lam);
      const int f = 2 + k;
      if (!(X >= lo[f] && X <= hi[f])) inside = 0;
      L.h[lane][f] = to_bf16((X - f_mean[f]) * f_rstd[f]);
      const int i = slot[k];
      if (i >= 0) {
        put(L, lane, i, X);
        put(L, lane, NM + i, __builtin_amdgcn_rcpf(v));
      }
      if (k == K - 1) put(L, lane, 2 * NM, __builtin_amdgcn_rcpf(v));
    }
  }
  __syncthreads();  // every wave has read part[*][0] and written its major species

  // The pair sums over the major species: their X_j and 1/v_j in registers, then per own species k (wave-uniform)
  // one pass over its table rows (contiguous, so they arrive in a few wide scalar loads).
  float Xm[NM], rvm[NM];
  float x_rest = 1.f;
#pragma unroll
  for (int i = 0; i < NM; i++) {
    Xm[i] = get(L, lane, i);
    rvm[i] = get(L, lane, NM + i);
    x_rest -= Xm[i];
  }
  x_rest = fmaxf(x_rest, 0.f);
  const float rv_b = get(L, lane, 2 * NM);
#pragma unroll
  for (int j = 0; j < KW; j++) {
    const int k = wave + 4 * j;
    if (k < K) {
      const float* __restrict__ ak = wa + k * (NM + 1);
      const float* __restrict__ bk = wb + k * (NM + 1);
      const float* __restrict__ dk = dfit + k * (NM + 1) * DK;
      const int own = slot[k];
      float den = 0.f, S = 0.f;
#pragma unroll
      for (int i = 0; i < NM; i++) {
        const float t = ak[i] + bk[i] * (vk[j] * rvm[i]);
        den += Xm[i] * t * t;
        if (i != own) S += Xm[i] * __builtin_amdgcn_rcpf(poly(dk + i * DK, l));
      }
      const bool minor = own < 0;
      const float X = xk[j], Xo = fmaxf(x_rest - (minor ? X : 0.f), 0.f);
      const float t = ak[NM] + bk[NM] * (vk[j] * rv_b);
      den += Xo * t * t + (minor ? X : 0.f);
      S += Xo * __builtin_amdgcn_rcpf(poly(dk + NM * DK, l));
      mu += X * vk[j] * vk[j] * __builtin_amdgcn_rcpf(den);
      const float Yk = nk[j] / rcp_W[k];
      L.y[2 + k][lane] = __builtin_logf(fmaxf(1.f - Yk, 1e-30f)) - __builtin_logf(fmaxf(S, 1e-30f)) + 1.5f * lnT - lnP;
    }
  }
  L.part[wave][0][lane] = a;
  L.part[wave][1][lane] = b;
  L.part[wave][2][lane] = mu;
  L.inside[wave][lane] = inside;
  __syncthreads();  // every wave has read the stage-1 scratch
  for (int c = NF + wave; c < IN; c += 4) L.h[lane][c] = 0;
  if (wave == 0) {
    float A = 0.f, B = 0.f, Mu = 0.f;
    int in_box = (T >= lo[0] && T <= hi[0]) && (lnP >= lo[1] && lnP <= hi[1]);
    for (int w = 0; w < 4; w++) {
      A += L.part[w][0][lane];
      B += L.part[w][1][lane];
      Mu += L.part[w][2][lane];
      in_box &= L.inside[w][lane];
    }
    L.inside[0][lane] = in_box;
    L.y[0][lane] = __builtin_logf(0.5f * (A + __builtin_amdgcn_rcpf(B)));
    L.y[1][lane] = __builtin_logf(Mu * sqrtT);
    L.h[lane][0] = to_bf16((T - f_mean[0]) * f_rstd[0]);
    L.h[lane][1] = to_bf16((lnP - f_mean[1]) * f_rstd[1]);
  }
  __syncthreads();

  // 2. Hidden layers, each wave WIDTH/4 output rows for all NB states.
  constexpr int RH = WIDTH / 64;
  for (int d = 0; d < DEPTH; d++) {
    const long w_off = offsets[2 * d], b_off = offsets[2 * d + 1];
    f32x4 acc[RH][4];
    if (d == 0)
      layer<RH, IN>(w16 + w_off, wave * (WIDTH / 4), L, acc, lane);
    else
      layer<RH, WIDTH>(w16 + w_off, wave * (WIDTH / 4), L, acc, lane);
    __syncthreads();  // every wave has read the previous activations
#pragma unroll
    for (int r = 0; r < RH; r++) {
      const int o = wave * (WIDTH / 4) + 16 * r + 4 * (lane >> 4);
      const float b0 = params[b_off + o], b1 = params[b_off + o + 1], b2 = params[b_off + o + 2], b3 = params[b_off + o + 3];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int st = 16 * c + (lane & 15);
        const f32x4 z = acc[r][c];
        // (__bf16) rounds to nearest even like to_bf16, two values per v_cvt_pk_bf16_f32
        const bf16x4 v = {(__bf16)silu(z[0] + b0), (__bf16)silu(z[1] + b1), (__bf16)silu(z[2] + b2), (__bf16)silu(z[3] + b3)};
        __builtin_memcpy(&L.h[st][o], &v, 8);
      }
    }
    __syncthreads();
  }

  // 3. Output layer (OUT/4 rows per wave): y = exp(r σ + m + ln baseline), through LDS.
  constexpr int RO = OUT / 64;
  {
    f32x4 acc[RO][4];
    layer<RO, WIDTH>(w16 + offsets[2 * DEPTH], wave * (OUT / 4), L, acc, lane);
    const long b_off = offsets[2 * DEPTH + 1];
#pragma unroll
    for (int r = 0; r < RO; r++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int o = wave * (OUT / 4) + 16 * r + 4 * (lane >> 4) + i;
        if (o >= NF) continue;
        const float bo = params[b_off + o], so = o_std[o], mo = o_mean[o];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int st = 16 * c + (lane & 15);
          L.y[o][st] = __builtin_expf((acc[r][c][i] + bo) * so + mo + L.y[o][st]);
        }
      }
  }
  __syncthreads();
  const int in_box = L.inside[0][lane];
  if (valid && in_box)
    for (int o = wave; o < NF; o += 4) out[o][s] = (double)L.y[o][lane];
  if (wave == 0 && valid && !in_box) fallback[atomicAdd(fallback_count, 1u)] = s;
}

// Clears the fallback counter before a launch: a kernel rather than a memset, so that a captured graph replays it.
extern "C" __global__ void surrogate_clear(unsigned* fallback_count) {
  if (threadIdx.x == 0) fallback_count[threadIdx.x] = 0u;
}
