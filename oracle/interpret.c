/* ORACLE — test infrastructure only (see chem.c).
 *
 * FP64 interpreter of a program's linear SSA form (cb_program_linear), restating ast/interpret.rs:52-148:
 * f64 min/max (Rust's NaN-ignoring f64::min/max = C fmin/fmax), <=, exp/ln from the host libm, sq = x*x,
 * Select evaluates to the chosen branch's value, and every produced value is checked for NaN
 * (ast/interpret.rs:110,122) — the first NaN stops the state and is reported.
 * Op codes: 0 const,1 neg,2 min,3 max,4 add,5 sub,6 le,7 mul,8 div,9 sqrt,10 exp,11 ln,12 sq,13 select.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* Evaluates n_states states (SoA in[n_inputs][n_states] -> out[n_outputs][n_states]).
 * Returns 0, or 1 + the index of the first state that produced a NaN (its instruction in *nan_at). */
long interpret(const uint8_t* op, const int32_t* dst, const int32_t* a, const int32_t* b, const int32_t* c, const double* k, long n_code,
               int32_t n_inputs, int32_t n_registers, const int32_t* outputs, int32_t n_outputs, const double* in, double* out, long n_states,
               long* nan_at) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n_registers);
  if (!r) return -1;
  for (long s = 0; s < n_states; s++) {
    for (int32_t i = 0; i < n_inputs; i++) r[i] = in[(size_t)i * n_states + s];
    for (long i = 0; i < n_code; i++) {
      double x;
      switch (op[i]) {
        case 0: x = k[i]; break;
        case 1: x = -r[a[i]]; break;
        case 2: x = fmin(r[a[i]], r[b[i]]); break;
        case 3: x = fmax(r[a[i]], r[b[i]]); break;
        case 4: x = r[a[i]] + r[b[i]]; break;
        case 5: x = r[a[i]] - r[b[i]]; break;
        case 6: x = r[a[i]] <= r[b[i]] ? 1. : 0.; break;
        case 7: x = r[a[i]] * r[b[i]]; break;
        case 8: x = r[a[i]] / r[b[i]]; break;
        case 9: x = sqrt(r[a[i]]); break;
        case 10: x = exp(r[a[i]]); break;
        case 11: x = log(r[a[i]]); break;
        case 12: x = r[a[i]] * r[a[i]]; break;
        case 13: x = r[a[i]] != 0. ? r[b[i]] : r[c[i]]; break;
        default: free(r); return -2;
      }
      if (isnan(x)) {
        if (nan_at) *nan_at = i;
        free(r);
        return 1 + s;
      }
      r[dst[i]] = x;
    }
    for (int32_t o = 0; o < n_outputs; o++) out[(size_t)o * n_states + s] = r[outputs[o]];
  }
  free(r);
  return 0;
}
