// One step of a Rosenbrock method for a batch of independent systems y' = f(y) of size M <= 64, in phases between
// evaluations of f (and of f with its Jacobian J) by other kernels. Two methods, W = I - h γ J factored once per step:
// the L-stable (2,3) pair of Shampine & Reichelt ("The MATLAB ODE Suite", SIAM J. Sci. Comput. 18, 1997; the method
// of ode23s), γ = d, in three phases:
//   phase 1 (after [f(y), J(y)]):  W factored (LU, partial pivoting); k1 = W⁻¹ f(y); ys = y + h k1 / 2
//   phase 2 (after F1 = f(ys)):    k2 = W⁻¹ (F1 - k1) + k1; ys = y + h k2
//   phase 3 (after F2 = f(ys)):    k3 = W⁻¹ (F2 - e32 (k2 - F1) - 2 (k1 - f(y))); error = h (k1 - 2 k2 + k3) / 6,
//                                   y_new = ys
// with d = 1/(2+√2) and e32 = 6+√2; or, with RODAS defined, the stiffly accurate order-4 method RODAS4 of Hairer &
// Wanner with its embedded order-3 solution (γ = 1/4), in six phases:
//   phase 1 (after [f(y), J(y)]):  W factored; k1 = W⁻¹ γ h f(y); ys = Y_2 = y + a_21 k1
//   phase I = 2..6 (after f(Y_I)): k_I = W⁻¹ γ (h f(Y_I) + Σ_{j<I} c_Ij k_j); ys = Y_{I+1} = y + Σ_{j<=I} a_{I+1,j} k_j
//                                   (I < 6); error = k6, y_new = Y_6 + k6
// The last phase accepts (y = y_new, t += h) when the RMS over i of error_i / (atol_i + rtol max(|y_i|, |y_new,i|))
// (CVODE's weighted RMS norm) is at most 1, and takes the next h from that ratio (0.9 ratio^(-1/q), q the order of the error estimate's local error, within
// [0.2, 5] for ode23s and [0.2, 6] for RODAS4, at most 0.5 after a rejection), clipped to the end time.
// A wave holds P = ⌊64/M⌋ systems, M lanes each: lane M p + i holds row i of
// system p's W (M doubles in registers). With one system per wave the pivot row of each elimination step is
// wave-uniform and broadcast with readlane; with several, every lane reads its system's pivot row with a cross-lane
// permute (ds_bpermute), the pivot search and the error norm loop over the system's M lanes, and lanes whose system
// is finished or past the batch compute along without storing. The LU factors and pivot order are kept in global
// memory for the later phases, with the reciprocal of each row's pivot, so that the solves multiply instead of divide.
// Systems whose status is not 0 (running) are left alone. Blocks take their systems in XCD-aware order: the 8 XCDs
// receive blocks in turn, and block b works on system group xcd·⌊B/8⌋ + min(xcd, B mod 8) + ⌊b/8⌋ (B blocks), so
// the systems that share a cache line of the SoA vectors run under one L2.
// Layouts: a launch covers systems 0..n-1 of pointers that start at its first system; vectors are SoA with row
// stride ld >= n, component i of system s at [i * ld + s]; jac holds f(y) in rows 0..M-1 and J row-major in rows
// M..M+M²-1 (the Jacobian program's outputs); lu[s][j][i] (column-major by system), rdp[s][i] the reciprocal pivots;
// perm[s][k] = the row chosen at step k; kst holds the stage vectors, M ld doubles each (ode23s: k1, k2, F1;
// RODAS4: k1..k5).
// With JAC_DIFFERENCES defined (and COPIES, a power of two >= M), J is formed in phase 1 from difference quotients
// instead: the right-hand side is evaluated for COPIES copies of every state, system by system (copy 0 and copies past
// M-1 the state, copy j = 1..M-1 with component j raised by inc_j, from the values rosenbrock_copies writes), and at
// the state's temperature raised by inc_0; phase 1 reads the right-hand sides of the copies, rep_f[i][COPIES s + c]
// (row i of system s is COPIES consecutive doubles), and at the raised temperature, rep_t[i][s]:
// J_ij = (f_i(copy j) - f_i(y)) / inc_j, J_i0 = (f_i(raised T) - f_i(y)) / inc_0; for ode23s's phase 3, f(y) goes to jac
// rows 0..M-1.
// Required define: M; P (systems per wave) defaults to ⌊64/M⌋. Launch: ⌈n / P⌉ blocks of 64 threads.

#define D_ROS 0.29289321881345248   // 1 / (2 + √2)
#define E32 7.4142135623730950      // 6 + √2
#define H_MIN 1e-30
#ifndef P
#define P (64 / M)  // systems per wave
#endif

// Block b's group of P systems (the XCD-aware order above).
__device__ __forceinline__ unsigned system_group() {
  const unsigned b = blockIdx.x, blocks = gridDim.x, xcd = b & 7u, per = blocks >> 3, rem = blocks & 7u;
  return xcd * per + min(xcd, rem) + (b >> 3);
}

// The largest value over the wave (in every lane): DPP row shifts, then the row broadcasts of lanes 15 and 31.
__device__ __forceinline__ unsigned wave_max(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// The value a lane holds, for a wave-uniform lane index (two v_readlane_b32).
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __builtin_bit_cast(long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}

// The value that row r of this lane's system holds (r uniform over the wave when P = 1; base = the system's first
// lane). Every lane of the wave takes part.
__device__ __forceinline__ double row_value(double v, int r, int base) {
#if P == 1
  (void)base;
  return lane_value(v, r);
#else
  return __shfl(v, base + r);
#endif
}

// Solves W x = b with the factors in registers: lane i holds row i (a[k] the multiplier l_ik for k < pos, U_i[k] for
// k >= pos) and rd = 1 / U_i[pos], pos the step at which row i was the pivot, perm[k] the pivot row of step k. Lane i
// returns x_i.
__device__ __forceinline__ double solve(const double (&a)[M], int pos, const int (&perm)[M], double rd, double b, int i, int base) {
#pragma unroll
  for (int k = 0; k < M; k++) {  // forward: L z = P b (unit lower triangular)
    const double z = row_value(b, perm[k], base);
    if (pos > k) b = __builtin_fma(-a[k], z, b);
  }
  double x = 0.0;
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {  // backward: U x = z
    const double xk = row_value(b * rd, perm[k], base);  // row perm[k]'s b is final once the steps after k are done
    if (pos < k) b = __builtin_fma(-a[k], xk, b);
    if (i == k) x = xk;
  }
  return x;
}

// Difference quotients (the dense Jacobian of the reference's CVODE path, cvode/lib.rs:17-47): component j of a
// system's state is increased by inc_j = √ε max(|y_j|, atol_j / rtol).
__device__ __forceinline__ double increment(double yj, int j, double rtol, double atol0, double atol) {
  return 1.4901161193847656e-08 * fmax(fabs(yj), (j == 0 ? atol0 : atol) / rtol);
}

// The sum of a value over this lane's system, added in row order whatever the packing (every lane takes part).
__device__ __forceinline__ double system_sum(double v, int base) {
  double sum = 0.0;
#pragma unroll
  for (int r = 0; r < M; r++) sum += row_value(v, r, base);
  return sum;
}

#ifdef RODAS
// The stage coefficients of RODAS4 in its transformed form (Hairer & Wanner, "Solving Ordinary Differential
// Equations II", 2nd ed., Springer 1996, section VI.4, and their code RODAS): a_ij of the stage points, c_ij of the
// stage right-hand sides (stage 6's point is stage 5's combination plus k5: a_6j = a_5j for j < 5, a_65 = 1).
__device__ __forceinline__ constexpr double rodas_a(int i, int j) {
  return i == 2 ? 1.544
       : i == 3 ? (j == 1 ? 0.9466785280815826 : 0.2557011698983284)
       : i == 4 ? (j == 1 ? 3.314825187068521 : j == 2 ? 2.896124015972201 : 0.9986419139977817)
       : j == 1 ? 1.221224509226641 : j == 2 ? 6.019134481288629 : j == 3 ? 12.53708332932087 : j == 4 ? -0.6878860361058950 : 1.0;
}
__device__ __forceinline__ constexpr double rodas_c(int i, int j) {
  return i == 2 ? -5.6688
       : i == 3 ? (j == 1 ? -2.430093356833875 : -0.2063599157091915)
       : i == 4 ? (j == 1 ? -0.1073529058151375 : j == 2 ? -9.594562251023355 : -20.47028614809616)
       : i == 5 ? (j == 1 ? 7.496443313967647 : j == 2 ? -10.24680431464352 : j == 3 ? -33.99990352819905 : 11.70890893206160)
       : j == 1 ? 8.083246795921522 : j == 2 ? -7.981132988064893 : j == 3 ? -31.52159432874371 : j == 4 ? 16.31930543123136
                : -6.058818238834054;
}
#define GAMMA 0.25
#define ERROR_ORDER 4.0  // the embedded estimate's local error is O(h^4)
#define FACTOR_MAX 6.0

// Stage I = 2..6 from f(Y_I) in F and the stages before it: k_I = W⁻¹ γ (h f(Y_I) + Σ_{j<I} c_Ij k_j); for I < 6
// stores k_I and the next stage point Y_{I+1} = y + Σ_{j<=I} a_{I+1,j} k_j in ys; returns k_I.
template <int I>
__device__ __forceinline__ double rodas_stage(const double (&a)[M], int pos, const int (&perm)[M], double rd, int i, int base, bool row,
                                              bool active, size_t ld, unsigned s, unsigned sv, const double* __restrict__ F,
                                              const double* __restrict__ y, double* __restrict__ ys, double* __restrict__ kst, double hs) {
  const int ir = row ? i : 0;
  const size_t mn = (size_t)M * ld;
  double r = row ? hs * F[(size_t)ir * ld + sv] : 0.0;
  double z = I < 6 && row ? y[(size_t)ir * ld + sv] : 0.0;
#pragma unroll
  for (int j = 1; j < I; j++) {
    const double kj = row ? kst[(j - 1) * mn + (size_t)ir * ld + sv] : 0.0;
    r = __builtin_fma(rodas_c(I, j), kj, r);
    if (I < 6) z = __builtin_fma(rodas_a(I + 1, j), kj, z);
  }
  const double x = solve(a, pos, perm, rd, GAMMA * r, i, base);
  if (I < 6 && active) {
    kst[(I - 1) * mn + (size_t)i * ld + s] = x;
    ys[(size_t)i * ld + s] = __builtin_fma(rodas_a(I + 1, I), x, z);
  }
  return x;
}
#else
#define GAMMA D_ROS
#define ERROR_ORDER 3.0  // the error estimate's local error is O(h^3)
#define FACTOR_MAX 5.0
#endif

// Phase 1 (FIRST) or a later phase (phase >= 2): two kernels, so that the later phases, which hold no LU in
// progress, are compiled to their own (smaller) register count.
template <bool FIRST>
__device__ __forceinline__ void rosenbrock_phase(int phase, unsigned n, size_t ld, double* __restrict__ jac, const double* const* __restrict__ rep_f,
                                                 const double* const* __restrict__ rep_t, const double* __restrict__ F,
                                                 double* __restrict__ y, double* __restrict__ ys, double* __restrict__ kst,
                                                 double* __restrict__ lu, double* __restrict__ rdp, int* __restrict__ perm_all, double* __restrict__ t,
                                                 double* __restrict__ h, int* __restrict__ status, unsigned* __restrict__ steps,
                                                 double t_end, double rtol, double atol0, double atol) {
  const int lane = threadIdx.x;
  const int sys = lane / M, i = lane - sys * M, base = sys * M;
  const bool row = sys < P;  // lanes past P M hold no row (with P = 1: the lanes past M)
  const unsigned s = system_group() * P + (unsigned)(row ? sys : 0);
  const bool valid = row && s < n;
  const unsigned sv = valid ? s : 0u;  // an in-bounds system for the loads of lanes that store nothing
#if P == 1
  if (s >= n || status[s] != 0) return;  // uniform over the wave
  bool active = row;
#else
  bool active = valid && status[sv] == 0;
#endif
  double hs = h[sv];
  if (FIRST) {
    const double remaining = t_end - t[sv];
    if (active && remaining <= 1e-12 * fabs(t_end)) {
      if (i == 0) status[s] = 1;
      active = false;
    }
    hs = fmin(hs, remaining);  // the step ends at t_end at the latest
    if (active && i == 0) h[s] = hs;
  }
#if P == 1
  if (FIRST && !__builtin_amdgcn_readfirstlane((int)active)) return;  // the system is at t_end (lane 0 is a row)
#endif
  int* perm_s = perm_all + (size_t)sv * M;
  double* lu_s = lu + (size_t)sv * M * M;
  double* rd_s = rdp + (size_t)sv * M;
  double a[M];
  int perm[M];  // the same in every lane of a system (wave-uniform when P = 1)
  int pos = -1;
  const int ir = row ? i : 0;

  if (FIRST) {
    // W = I - h γ J, factored with partial pivoting
#ifdef JAC_DIFFERENCES
    const double* Fi = rep_f[ir] + (size_t)sv * COPIES;  // row i of the system's copies
    const double f0 = row ? Fi[0] : 0.0;
    const double fT = row ? rep_t[ir][sv] : 0.0;  // row i at the raised temperature
    const double rinc_i = row ? 1.0 / increment(y[(size_t)ir * ld + sv], i, rtol, atol0, atol) : 0.0;
#ifndef RODAS
    if (active) jac[(size_t)i * ld + s] = f0;  // f(y), for phase 3
#endif
#else
    const double f0 = row ? jac[(size_t)ir * ld + sv] : 0.0;
#endif
#pragma unroll
    for (int j = 0; j < M; j++) {
#ifdef JAC_DIFFERENCES
      const double Jij = ((j == 0 ? fT : Fi[j]) - f0) * row_value(rinc_i, j, base);
#else
      const double Jij = jac[(size_t)(M + ir * M + j) * ld + sv];
#endif
      a[j] = (row ? -hs * GAMMA * Jij : 0.0) + (row && i == j ? 1.0 : 0.0);
    }
    double diag = 1.0;  // this row's pivot, U_i[pos]
#pragma unroll
    for (int k = 0; k < M; k++) {
      // pivot: the largest |a[k]| over the rows not yet chosen (ties to the lowest row)
      const double key = row && pos < 0 ? fabs(a[k]) : -1.0;
#if P == 1
      // the high words of |a| (plus one; 0 for rows already chosen) order the rows as |a| does up to ties, which
      // the rare case resolves over the tied lanes in ascending order
      const unsigned hi = row && pos < 0 ? (unsigned)(__builtin_bit_cast(unsigned long long, key) >> 32) + 1u : 0u;
      unsigned long long tied = __builtin_amdgcn_ballot_w64(hi == wave_max(hi));
      int p = (int)__builtin_ctzll(tied);
      if (tied & (tied - 1)) {
        double best = -1.0;
        while (tied) {
          const int l = (int)__builtin_ctzll(tied);
          tied &= tied - 1;
          const double v = lane_value(key, l);
          if (v > best) {
            best = v;
            p = l;
          }
        }
      }
#else
      double best = -2.0;
      int p = 0;
#pragma unroll
      for (int r = 0; r < M; r++) {
        const double kr = __shfl(key, base + r);
        if (kr > best) {
          best = kr;
          p = r;
        }
      }
#endif
      perm[k] = p;
      if (active && i == 0) perm_s[k] = p;
      if (row && i == p) pos = k;
      // the pivot row's entries are read from its lane by every lane (outside the branch: the pivot row is final)
      const double pivot = row_value(a[k], p, base);
      if (row && i == p) diag = pivot;
      const bool eliminate = row && pos < 0;
      const double l = eliminate ? a[k] / pivot : 0.0;
      if (eliminate) a[k] = l;
#pragma unroll
      for (int j = k + 1; j < M; j++) {
        const double pj = row_value(a[j], p, base);
        a[j] = __builtin_fma(-l, pj, a[j]);  // l = 0 leaves the rows already chosen as they are (finite pj)
      }
    }
    const double rd = 1.0 / diag;
    if (active) {
#pragma unroll
      for (int j = 0; j < M; j++) lu_s[j * M + i] = a[j];
      rd_s[i] = rd;
    }
#ifdef RODAS
    const double x = solve(a, pos, perm, rd, GAMMA * hs * f0, i, base);  // k1
    if (active) {
      kst[(size_t)i * ld + s] = x;
      ys[(size_t)i * ld + s] = __builtin_fma(rodas_a(2, 1), x, y[(size_t)i * ld + s]);
    }
#else
    const double x = solve(a, pos, perm, rd, f0, i, base);  // k1
    if (active) {
      kst[(size_t)i * ld + s] = x;
      ys[(size_t)i * ld + s] = y[(size_t)i * ld + s] + 0.5 * hs * x;
    }
#endif
    return;
  }

  // later phases: the factors and each row's pivot step back from memory
#pragma unroll
  for (int j = 0; j < M; j++) a[j] = row ? lu_s[j * M + ir] : 0.0;
  const double rd = row ? rd_s[ir] : 1.0;
#pragma unroll
  for (int k = 0; k < M; k++) {
#if P == 1
    perm[k] = __builtin_amdgcn_readfirstlane(perm_s[k]);
#else
    perm[k] = perm_s[k];
#endif
    if (row && perm[k] == i) pos = k;
  }
#ifdef RODAS
  switch (phase) {
    case 2: rodas_stage<2>(a, pos, perm, rd, i, base, row, active, ld, s, sv, F, y, ys, kst, hs); return;
    case 3: rodas_stage<3>(a, pos, perm, rd, i, base, row, active, ld, s, sv, F, y, ys, kst, hs); return;
    case 4: rodas_stage<4>(a, pos, perm, rd, i, base, row, active, ld, s, sv, F, y, ys, kst, hs); return;
    case 5: rodas_stage<5>(a, pos, perm, rd, i, base, row, active, ld, s, sv, F, y, ys, kst, hs); return;
    default: break;
  }
  // phase 6: y_new = Y_6 + k6, and k6 is the error estimate
  const double k6 = rodas_stage<6>(a, pos, perm, rd, i, base, row, active, ld, s, sv, F, y, ys, kst, hs);
  const double y0 = row ? y[(size_t)ir * ld + sv] : 0.0;
  const double y1 = (row ? ys[(size_t)ir * ld + sv] : 0.0) + k6;
  const double err = k6;
#else
  const double x1 = row ? kst[(size_t)ir * ld + sv] : 0.0;
  double* k2 = kst + (size_t)M * ld;
  double* F1 = kst + 2 * (size_t)M * ld;
  if (phase == 2) {
    const double f1 = row ? F[(size_t)ir * ld + sv] : 0.0;
    const double x = solve(a, pos, perm, rd, f1 - x1, i, base) + x1;
    if (active) {
      F1[(size_t)i * ld + s] = f1;
      k2[(size_t)i * ld + s] = x;
      ys[(size_t)i * ld + s] = y[(size_t)i * ld + s] + hs * x;
    }
    return;
  }
  const double x2 = row ? k2[(size_t)ir * ld + sv] : 0.0;
  const double f0 = row ? jac[(size_t)ir * ld + sv] : 0.0;
  const double f1 = row ? F1[(size_t)ir * ld + sv] : 0.0;
  const double f2 = row ? F[(size_t)ir * ld + sv] : 0.0;
  const double x3 = solve(a, pos, perm, rd, f2 - E32 * (x2 - f1) - 2.0 * (x1 - f0), i, base);
  const double y0 = row ? y[(size_t)ir * ld + sv] : 0.0, y1 = row ? ys[(size_t)ir * ld + sv] : 0.0;
  const double err = hs * (x1 - 2.0 * x2 + x3) * (1.0 / 6.0);
#endif
  // step control: accept when the scaled error's root mean square over the components is at most 1 (the weighted
  // RMS norm of CVODE); the next h from it
  const double scale = (i == 0 ? atol0 : atol) + rtol * fmax(fabs(y0), fabs(y1));
  double q = row ? err / scale : 0.0;
  q *= q;
  if (!(q == q)) q = 1e300;  // a NaN (a singular W, say) rejects the step
  const double ratio = sqrt(system_sum(q, base) * (1.0 / M));
  if (!active) return;
  const bool accept = ratio <= 1.0;
  const double factor = fmin(FACTOR_MAX, fmax(0.2, 0.9 * pow(fmax(ratio, 1e-12), -1.0 / ERROR_ORDER)));
  double t_new = t[s];
  if (accept) {
    y[(size_t)i * ld + s] = y1;
    t_new += hs;
  }
  const double remaining = t_end - t_new;
  double h_new = hs * (accept ? factor : fmin(factor, 0.5));
  if (i == 0) {
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

extern "C" __global__ void __launch_bounds__(64)
rosenbrock(int phase, unsigned n, size_t ld, double* __restrict__ jac, const double* const* __restrict__ rep_f,
           const double* const* __restrict__ rep_t, const double* __restrict__ F,
           double* __restrict__ y, double* __restrict__ ys, double* __restrict__ kst,
           double* __restrict__ lu, double* __restrict__ rdp, int* __restrict__ perm_all, double* __restrict__ t, double* __restrict__ h,
           int* __restrict__ status, unsigned* __restrict__ steps, double t_end, double rtol, double atol0, double atol) {
  rosenbrock_phase<true>(phase, n, ld, jac, rep_f, rep_t, F, y, ys, kst, lu, rdp, perm_all, t, h, status, steps, t_end, rtol, atol0, atol);
}

// Phases 2 and later (phase >= 2).
extern "C" __global__ void __launch_bounds__(64)
rosenbrock_stage(int phase, unsigned n, size_t ld, double* __restrict__ jac, const double* const* __restrict__ rep_f,
                 const double* const* __restrict__ rep_t, const double* __restrict__ F,
                 double* __restrict__ y, double* __restrict__ ys, double* __restrict__ kst,
                 double* __restrict__ lu, double* __restrict__ rdp, int* __restrict__ perm_all, double* __restrict__ t, double* __restrict__ h,
                 int* __restrict__ status, unsigned* __restrict__ steps, double t_end, double rtol, double atol0, double atol) {
  rosenbrock_phase<false>(phase, n, ld, jac, rep_f, rep_t, F, y, ys, kst, lu, rdp, perm_all, t, h, status, steps, t_end, rtol, atol0, atol);
}

#ifdef JAC_DIFFERENCES
// Each system's values for its copies, after the `shared` values of the right-hand side's part shared by the copies:
// groups[s stride + shared + f] for f = 0..2M-1 is T, P, Y_0..Y_{M-2} (f < M + 1), then inc_j of component j = f - M
// (j = 1..M-1); and raised_T[s] = T + inc_0 for the whole right-hand side at the raised temperature. One thread per
// (system, f).
extern "C" __global__ void __launch_bounds__(256)
rosenbrock_copies(unsigned n, size_t ld, const double* __restrict__ y, const double* __restrict__ P_in, double* __restrict__ groups,
                  unsigned stride, unsigned shared, double* __restrict__ raised_T, double rtol, double atol0, double atol) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * 2 * M) return;
  const unsigned s = (unsigned)(idx / (2 * M)), f = (unsigned)(idx - (size_t)s * (2 * M));
  double v;
  if (f == 0) {
    v = y[s];
    raised_T[s] = v + increment(v, 0, rtol, atol0, atol);
  } else if (f == 1) {
    v = P_in[s];
  } else if (f <= M) {
    v = y[(size_t)(f - 1) * ld + s];
  } else {
    const int j = (int)f - M;
    v = increment(y[(size_t)j * ld + s], j, rtol, atol0, atol);
  }
  groups[(size_t)s * stride + shared + f] = v;
}
#endif

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
