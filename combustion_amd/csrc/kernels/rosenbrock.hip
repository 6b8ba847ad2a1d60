// One step of the L-stable (2,3) Rosenbrock pair of Shampine & Reichelt ("The MATLAB ODE Suite", SIAM J. Sci.
// Comput. 18, 1997; the method of ode23s) for a batch of independent systems y' = f(y) of size M <= 64, in three
// phases between evaluations of f (and of f with its Jacobian J) by other kernels:
//   phase 1 (after [f(y), J(y)]):  W = I - h d J factored (LU, partial pivoting); k1 = W⁻¹ f(y); ys = y + h k1 / 2
//   phase 2 (after F1 = f(ys)):    k2 = W⁻¹ (F1 - k1) + k1; ys = y + h k2
//   phase 3 (after F2 = f(ys)):    k3 = W⁻¹ (F2 - e32 (k2 - F1) - 2 (k1 - f(y))); error = h (k1 - 2 k2 + k3) / 6;
//                                   accept (y = ys, t += h) when max_i |error_i| / (atol_i + rtol max(|y_i|, |ys_i|)) <= 1,
//                                   and the next h from that ratio, clipped to the end time
// with d = 1/(2+√2) and e32 = 6+√2. A wave owns one system: lane i holds row i of W (M doubles in registers); the
// pivot row of each elimination step is broadcast with readlane (the pivot lane is wave-uniform), and the LU factors
// and pivot order are kept in global memory for phases 2 and 3. Systems whose status is not 0 (running) are skipped.
// Layouts: vectors are SoA, component i of system s at [i * n + s]; jac holds f(y) in rows 0..M-1 and J row-major in
// rows M..M+M²-1 (the Jacobian program's outputs); lu[s][i][j]; perm[s][k] = the row chosen at step k.
// Required define: M.

#define D_ROS 0.29289321881345248   // 1 / (2 + √2)
#define E32 7.4142135623730950      // 6 + √2
#define H_MIN 1e-30

// The value a lane holds, for a wave-uniform lane index (two v_readlane_b32).
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}

// Solves W x = b with the factors in registers: lane i holds row i (a[k] the multiplier l_ik for k < pos, U_i[k] for
// k >= pos), pos the step at which row i was the pivot, perm[k] the pivot row of step k. Lane i returns x_i.
__device__ __forceinline__ double solve(const double (&a)[M], int pos, const int (&perm)[M], double b, int lane) {
#pragma unroll
  for (int k = 0; k < M; k++) {  // forward: L z = P b (unit lower triangular)
    const int p = perm[k];
    const double z = lane_value(b, p);
    if (pos > k) b = __builtin_fma(-a[k], z, b);
  }
  double x = 0.0;
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {  // backward: U x = z
    const int p = perm[k];
    const double xk = lane_value(b, p) / lane_value(a[k], p);
    if (pos < k) b = __builtin_fma(-a[k], xk, b);
    if (lane == k) x = xk;
  }
  return x;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

extern "C" __global__ void __launch_bounds__(64)
rosenbrock(int phase, unsigned n, const double* __restrict__ jac, const double* __restrict__ F, double* __restrict__ y,
           double* __restrict__ ys, double* __restrict__ k1, double* __restrict__ k2, double* __restrict__ F1,
           double* __restrict__ lu, int* __restrict__ perm_all, double* __restrict__ t, double* __restrict__ h,
           int* __restrict__ status, unsigned* __restrict__ steps, double t_end, double rtol, double atol0, double atol) {
  const unsigned s = blockIdx.x;
  const int lane = threadIdx.x;
  if (s >= n || status[s] != 0) return;  // uniform over the wave
  const bool row = lane < M;
  const int i = row ? lane : 0;
  double hs = h[s];
  if (phase == 1) {
    const double remaining = t_end - t[s];
    if (remaining <= 1e-12 * fabs(t_end)) {
      if (lane == 0) status[s] = 1;
      return;
    }
    hs = fmin(hs, remaining);  // the step ends at t_end at the latest
    if (lane == 0) h[s] = hs;
  }
  int* perm_s = perm_all + (size_t)s * M;
  double* lu_s = lu + (size_t)s * M * M;
  double a[M];
  int perm[M];  // wave-uniform
  int pos = -1;

  if (phase == 1) {
    // W = I - h d J, factored with partial pivoting
#pragma unroll
    for (int j = 0; j < M; j++) a[j] = (row ? -hs * D_ROS * jac[(size_t)(M + i * M + j) * n + s] : 0.0) + (row && i == j ? 1.0 : 0.0);
#pragma unroll
    for (int k = 0; k < M; k++) {
      // pivot: the largest |a[k]| over the rows not yet chosen (ties to the lowest lane)
      double key = row && pos < 0 ? fabs(a[k]) : -1.0;
      int who = lane;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ok = __shfl_xor(key, o);
        const int ow = __shfl_xor(who, o);
        if (ok > key || (ok == key && ow < who)) {
          key = ok;
          who = ow;
        }
      }
      const int p = __builtin_amdgcn_readfirstlane(who);
      perm[k] = p;
      if (lane == 0) perm_s[k] = p;
      if (lane == p) pos = k;
      // the pivot row's entries are read from lane p by every lane (outside the branch: the pivot lane's row is final)
      const double pivot = lane_value(a[k], p);
      const bool eliminate = row && pos < 0;
      const double l = eliminate ? a[k] / pivot : 0.0;
      if (eliminate) a[k] = l;
#pragma unroll
      for (int j = k + 1; j < M; j++) {
        const double pj = lane_value(a[j], p);
        if (eliminate) a[j] = __builtin_fma(-l, pj, a[j]);
      }
    }
    if (row) {
#pragma unroll
      for (int j = 0; j < M; j++) lu_s[i * M + j] = a[j];
    }
    const double f0 = row ? jac[(size_t)i * n + s] : 0.0;
    const double x = solve(a, pos, perm, f0, lane);
    if (row) {
      k1[(size_t)i * n + s] = x;
      ys[(size_t)i * n + s] = y[(size_t)i * n + s] + 0.5 * hs * x;
    }
    return;
  }

  // phases 2 and 3: the factors and each row's pivot step back from memory
#pragma unroll
  for (int j = 0; j < M; j++) a[j] = row ? lu_s[i * M + j] : 0.0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    perm[k] = __builtin_amdgcn_readfirstlane(perm_s[k]);
    if (perm[k] == lane) pos = k;
  }
  const double x1 = row ? k1[(size_t)i * n + s] : 0.0;
  if (phase == 2) {
    const double f1 = row ? F[(size_t)i * n + s] : 0.0;
    const double x = solve(a, pos, perm, f1 - x1, lane) + x1;
    if (row) {
      F1[(size_t)i * n + s] = f1;
      k2[(size_t)i * n + s] = x;
      ys[(size_t)i * n + s] = y[(size_t)i * n + s] + hs * x;
    }
    return;
  }
  const double x2 = row ? k2[(size_t)i * n + s] : 0.0;
  const double f0 = row ? jac[(size_t)i * n + s] : 0.0;
  const double f1 = row ? F1[(size_t)i * n + s] : 0.0;
  const double f2 = row ? F[(size_t)i * n + s] : 0.0;
  const double x3 = solve(a, pos, perm, f2 - E32 * (x2 - f1) - 2.0 * (x1 - f0), lane);
  const double y0 = row ? y[(size_t)i * n + s] : 0.0, y1 = row ? ys[(size_t)i * n + s] : 0.0;
  const double err = hs * (x1 - 2.0 * x2 + x3) * (1.0 / 6.0);
  const double scale = (i == 0 ? atol0 : atol) + rtol * fmax(fabs(y0), fabs(y1));
  double ratio = row ? fabs(err) / scale : 0.0;
  if (!(ratio == ratio)) ratio = 1e300;  // a NaN (a singular W, say) rejects the step
  ratio = wave_max(ratio);
  const bool accept = ratio <= 1.0;
  const double factor = fmin(5.0, fmax(0.2, 0.9 * pow(fmax(ratio, 1e-12), -1.0 / 3.0)));
  double t_new = t[s];
  if (accept) {
    if (row) y[(size_t)i * n + s] = y1;
    t_new += hs;
  }
  const double remaining = t_end - t_new;
  double h_new = hs * (accept ? factor : fmin(factor, 0.5));
  if (lane == 0) {
    t[s] = t_new;
    steps[s] += 1;
    if (remaining <= 1e-12 * fabs(t_end)) {
      status[s] = 1;  // reached the end time
    } else {
      h_new = fmin(h_new, remaining);
      if (h_new < H_MIN) status[s] = 2;  // the step size collapsed
      h[s] = h_new;
    }
  }
}

// The number of systems still running (status 0), added to *count (cleared by the caller's launch of clear_count).
extern "C" __global__ void __launch_bounds__(256) rosenbrock_running(unsigned n, const int* __restrict__ status, unsigned* __restrict__ count) {
  const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool running = s < n && status[s] == 0;
  const unsigned long long ballot = __ballot(running);
  if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(count, (unsigned)__popcll(ballot));
}

extern "C" __global__ void __launch_bounds__(64) rosenbrock_clear(unsigned* __restrict__ count) {
  if (threadIdx.x == 0) *count = 0u;
}
