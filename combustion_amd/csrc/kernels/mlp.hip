// Dense layers of the transport surrogate (training): BF16 operands, FP32 accumulation on
// v_mfma_f32_16x16x16bf16_1k, 64×64 output tiles per 256-thread block (each wave a 32×32 quarter), k in steps of 32
// staged through LDS. One kernel serves the forward product, the backward data product and the weight gradient:
// C[m][n] = Σ_k A(m, k) B(n, k), with A(m, k) = A[m sam + k sak] and B(n, k) = B[n sbn + k sbk], so transposed
// operands are strides, and the epilogue is chosen by `mode`:
//   0  hidden layer forward: z = C + bias[n] (FP32 out) and h = SiLU(z) (BF16 out)
//   1  output layer forward: z = C + bias[n] (FP32 out)
//   2  backward data: g = C · SiLU'(zprev[m][n]) (BF16 out), the previous layer's pre-activation gradient
//   3  weight gradient: C over the k range [kz·kchunk, (kz+1)·kchunk) of blockIdx.z, FP32 out into partial[kz][m][n]
// Dimensions N and Kd are multiples of 64 and 32 (the host pads features); M is guarded.

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

#define LDSK 40  // 32 k plus 8 of padding per row

__device__ __forceinline__ u16 to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(u16 b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-z * 1.4426950408889634f)); }

// Loads rows [r0, r0+64) × k [k0, k0+32) of a strided BF16 matrix into dst[row][k], zero outside [0, R) × [klo, khi).
__device__ __forceinline__ void load_tile(u16 (*dst)[LDSK], const u16* __restrict__ src, long sr, long sk, long r0, long R,
                                          long k0, long klo, long khi) {
  const int t = threadIdx.x;
  if (sk == 1) {
    const int row = t >> 2, kc = (t & 3) * 8;
    const long r = r0 + row, k = k0 + kc;
    u16 v[8];
    if (r < R && k >= klo && k + 8 <= khi) {
      const uint4 q = *(const uint4*)(src + r * sr + k);
      __builtin_memcpy(v, &q, 16);
    } else {
      for (int i = 0; i < 8; i++) v[i] = (r < R && k + i >= klo && k + i < khi) ? src[r * sr + k + i] : (u16)0;
    }
    for (int i = 0; i < 8; i++) dst[row][kc + i] = v[i];
  } else {  // sr == 1: 8 consecutive rows at one k
    const int kc = t >> 3, row = (t & 7) * 8;
    const long k = k0 + kc, r = r0 + row;
    u16 v[8];
    if (k >= klo && k < khi && r + 8 <= R) {
      const uint4 q = *(const uint4*)(src + r + k * sk);
      __builtin_memcpy(v, &q, 16);
    } else {
      for (int i = 0; i < 8; i++) v[i] = (k >= klo && k < khi && r + i < R) ? src[r + i + k * sk] : (u16)0;
    }
    for (int i = 0; i < 8; i++) dst[row + i][kc] = v[i];
  }
}

extern "C" __global__ void __launch_bounds__(256)
mlp_gemm(int mode, long M, long N, long Kd, const u16* __restrict__ A, long sam, long sak, const u16* __restrict__ B, long sbn,
         long sbk, const float* __restrict__ bias, const float* __restrict__ zprev, float* __restrict__ zout,
         u16* __restrict__ hout, long ldc, float* __restrict__ partial, long kchunk) {
  __shared__ u16 As[64][LDSK], Bs[64][LDSK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long m0 = (long)blockIdx.x * 64, n0 = (long)blockIdx.y * 64;
  long klo = 0, khi = Kd;
  if (mode == 3) {
    klo = (long)blockIdx.z * kchunk;
    khi = klo + kchunk < Kd ? klo + kchunk : Kd;
  }
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (long k0 = klo & ~31L; k0 < khi; k0 += 32) {
    load_tile(As, A, sam, sak, m0, M, k0, klo, khi);
    load_tile(Bs, B, sbn, sbk, n0, N, k0, klo, khi);
    __syncthreads();
    for (int kk = 0; kk < 32; kk += 16) {
      bf16x4 a[2], b[2];
      for (int i = 0; i < 2; i++) a[i] = *(const bf16x4*)&As[wm * 32 + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
      for (int j = 0; j < 2; j++) b[j] = *(const bf16x4*)&Bs[wn * 32 + j * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        const long m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r, n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        const float c = acc[i][j][r];
        if (mode == 0) {
          const float z = c + bias[n];
          zout[m * ldc + n] = z;
          hout[m * ldc + n] = to_bf16(z * sigmoid(z));
        } else if (mode == 1) {
          zout[m * ldc + n] = c + bias[n];
        } else if (mode == 2) {
          const float z = zprev[m * ldc + n], s = sigmoid(z);
          hout[m * ldc + n] = to_bf16(c * s * (1.f + z * (1.f - s)));
        } else {
          partial[((long)blockIdx.z * M + m) * N + n] = c;
        }
      }
}

// out[i] = Σ_z partial[z][i] (z in order: deterministic).
extern "C" __global__ void mlp_reduce(const float* __restrict__ partial, long n, int parts, float* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < parts; z++) s += partial[(long)z * n + i];
  out[i] = s;
}

// Partial column sums of dz (BF16 gradients, FP32 sums) over the row chunk [z·chunk, (z+1)·chunk) of blockIdx.z:
// a 256-thread block takes 64 columns (the lane) and 4 interleaved row streams (the wave), combined in LDS in wave
// order, into partial[z][c]; mlp_reduce then adds the chunks in order (a fixed summation order for every run).
extern "C" __global__ void __launch_bounds__(256)
mlp_column_sum(const u16* __restrict__ dz, long M, long ld, long cols, long chunk, float* __restrict__ partial) {
  __shared__ float part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long c = (long)blockIdx.x * 64 + lane;
  const long m0 = (long)blockIdx.z * chunk, m1 = m0 + chunk < M ? m0 + chunk : M;
  float s = 0.f;
  if (c < cols)
#pragma unroll 8
    for (long m = m0 + wave; m < m1; m += 4) s += from_bf16(dz[m * ld + c]);
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && c < cols) partial[(long)blockIdx.z * cols + c] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// x (FP32) -> BF16, elementwise.
extern "C" __global__ void mlp_to_bf16(const float* __restrict__ x, long n, u16* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = to_bf16(x[i]);
}

// Mean squared error over the first `valid` of ld outputs: dz = 2 (y - t) / (M valid) as BF16 (zero on padding).
extern "C" __global__ void mlp_mse_grad(const float* __restrict__ y, const float* __restrict__ t, long M, long ld, long valid,
                                        u16* __restrict__ dz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * ld) return;
  const long c = i % ld;
  dz[i] = c < valid ? to_bf16(2.f * (y[i] - t[i]) / (float)(M * valid)) : (u16)0;
}

// Adam (Kingma & Ba 2015) on FP32 master parameters, refreshing their BF16 copies.
extern "C" __global__ void mlp_adam(float* __restrict__ p, u16* __restrict__ p16, const float* __restrict__ g, float* __restrict__ m1,
                                    float* __restrict__ m2, long n, float lr, float beta1, float beta2, float eps, float c1,
                                    float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = beta1 * m1[i] + (1.f - beta1) * gi;
  const float b = beta2 * m2[i] + (1.f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float q = p[i] - lr * (a / c1) / (__builtin_sqrtf(b / c2) + eps);
  p[i] = q;
  p16[i] = to_bf16(q);
}
