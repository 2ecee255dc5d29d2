/*
 * kd_exact.h — the greedy sweep's boundary test in the CPU's summation order.
 *
 * The reference sweep (kdigest/tdigest.mojo:414-435) tests candidate i of the
 * cluster that starts at s with
 *     q_left  = weight_so_far / stored_n
 *     q_right = ((weight_so_far + current.weight) + candidate.weight) / stored_n
 * where current.weight is the running sum w_s + w_(s+1) + ... (one rounding
 * per record, centroid.mojo:34-36) and weight_so_far the running sum of the
 * finished clusters' weights.  The device paths locate boundaries with a
 * closed-form weight prefix, which is exact when every weight is an integer
 * (or a power of two with no prior records) but otherwise differs from those
 * running sums in the last bits; near a boundary, and always at the last
 * candidate, where q_right is about 1 (K1's asin(2q - 1) is NaN just above
 * it), the test can then come out differently.  These helpers give the
 * running sums exactly so that every boundary can be checked, and the test
 * itself with the reference's operand order.
 *
 * Shared by the host (tests, via the oracle build) and the device.
 */
#ifndef KD_EXACT_H
#define KD_EXACT_H

#include <stdint.h>

#include "kd_math.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 2^e for an integer e in the normal range, by its bit pattern. */
KD_HD double kd_pow2(int e) { return kd_bits2d((uint64_t)(e + 1023) << 52); }

/* Binary exponent e of a positive normal x: 2^e <= x < 2^(e+1). */
KD_HD int kd_exponent(double x) { return (int)((kd_d2bits(x) >> 52) & 0x7ff) - 1023; }

/* L sequential additions W = W + x (round to nearest even), for finite
 * x > 0 and W >= 0: what a loop of L adds gives, in O(binades crossed)
 * steps.  Within a binade [2^e, 2^(e+1)) every sum is a multiple of
 * u = 2^(e-52) and W + x rounds to W + q u or W + (q+1) u, the choice fixed
 * by x's remainder below u (and, for an exact half, by the parity of W / u,
 * which is the same after the first step); so after one step the increment
 * d is constant until a sum reaches 2^(e+1). */
KD_HD double kd_repeat_add(double W, double x, int64_t L) {
  while (L > 0) {
    if (!(W >= 2.0 * x) || L < 3 || !(W >= 0x1p-960)) {
      W = W + x;
      L--;
      continue;
    }
    const int e = kd_exponent(W);
    W = W + x;
    L--;
    if (kd_exponent(W) != e) continue;
    const double d = (W + x) - W;
    if (d == 0.0) return W; /* x below half an ulp: every further add is a no-op */
    /* Units of u = 2^(e-52): scaling by powers of two is exact, and every
     * integer below 2^53 (and every product m * D below it) is an exact
     * double, so the step count is found in doubles. */
    const double u = kd_pow2(e - 52), ui = kd_pow2(52 - e);
    /* Steps k = 0, 1, ... with W + k d + x < 2^(e+1), all in units of u:
     * k * D <= A - floor(x / u) - 1. */
    const double A = (kd_pow2(e + 1) - W) * ui;
    const double D = d * ui;
    const double lim = A - __builtin_floor(x * ui) - 1.0;
    if (lim < 0.0 || !(D > 0.0)) continue;
    /* floor(lim / D) from a rounded quotient, corrected to the exact one. */
    double m = __builtin_floor(lim / D);
    while (m > 0.0 && m * D > lim) m -= 1.0;
    while ((m + 1.0) * D <= lim) m += 1.0;
    m += 1.0;
    if (m > (double)L) m = (double)L;
    W = W + (m * D) * u;
    L -= (int64_t)m;
  }
  return W;
}

/* A running sum S = S + x (x >= 0) inside the binade [2^e, 2^(e+1)) of S,
 * counted in units u = 2^(e-52) (ui = 1/u): S = A u with A an integer in
 * [2^52, 2^53).  While A + floor(x/u) < 2^53 the rounded sum is (A + r) u
 * with r the integer nearest to x/u, whatever A is, except at an exact
 * half, where the tie goes to the even sum.  So a stretch of such steps is
 * an integer prefix sum, which a block can form in parallel.  Returns r and
 * *fl = floor(x/u), or -1 when the step needs the plain addition (an exact
 * half, or x/u >= 2^52).  x * ui is exact (a power-of-two scaling) unless it
 * underflows, and then r = 0 either way. */
KD_HD int64_t kd_sum_units(double x, double ui, int64_t* fl) {
  const double q = x * ui;
  if (!(q < 0x1p52)) return -1;
  const double f = __builtin_floor(q);
  const double d = q - f; /* exact */
  if (d == 0.5) return -1;
  *fl = (int64_t)f;
  return (int64_t)f + (d > 0.5 ? 1 : 0);
}

/* The sums S_i = S_(i-1) + x_i, i = 0..m-1 (x_i >= 0), computed round by
 * round as a block does it: within a round every step is counted in units of
 * the round's binade (kd_sum_units); the first step that the units do not
 * cover (a tie, an increment of 2^52 units, a sum leaving the binade) is
 * added plainly and the next round starts after it (a sum that reached
 * 2^(e+1) exactly makes the next step such a one).  Starting values below 2^-960 (zero
 * included) pass zeros through and add the first nonzero x plainly.  With
 * stop_on, the first nonzero x whose sum exceeds thr is not consumed: its
 * index is returned and the exceeding sum goes to *stop_val.  Returns the
 * number of x consumed (m without a stop); *S is the sum after them and
 * out[i] = S_i when out is given.  Equal to the sequential loop bit for bit
 * (the host build is tested against it). */
KD_HD int64_t kd_running_sum(double* S, const double* x, int64_t m, int stop_on, double thr,
                             double* out, double* stop_val) {
  double s = *S;
  int64_t i = 0;
  while (i < m) {
    if (!(s >= 0x1p-960)) {
      while (i < m && x[i] == 0.0) {
        if (out) out[i] = s;
        i++;
      }
      if (i == m) break;
      const double ns = s + x[i];
      if (stop_on && ns > thr) {
        *stop_val = ns;
        *S = s;
        return i;
      }
      if (out) out[i] = ns;
      s = ns;
      i++;
      continue;
    }
    const int e = kd_exponent(s);
    const double u = kd_pow2(e - 52), ui = kd_pow2(52 - e);
    uint64_t A = (uint64_t)(s * ui);
    int plain = 0;
    while (i < m) {
      int64_t fl = 0;
      const int64_t r = kd_sum_units(x[i], ui, &fl);
      if (r < 0 || A + (uint64_t)fl >= (1ull << 53)) {
        plain = 1;
        break;
      }
      const uint64_t A1 = A + (uint64_t)r;
      const double s1 = (double)A1 * u;
      if (stop_on && x[i] != 0.0 && s1 > thr) {
        *stop_val = s1;
        *S = (double)A * u;
        return i;
      }
      if (out) out[i] = s1;
      A = A1;
      i++;
    }
    s = (double)A * u;
    if (plain) {
      const double ns = s + x[i];
      if (stop_on && x[i] != 0.0 && ns > thr) {
        *stop_val = ns;
        *S = s;
        return i;
      }
      if (out) out[i] = ns;
      s = ns;
      i++;
    }
  }
  *S = s;
  return m;
}

/* The sweep's test at a candidate of weight w: true when it closes the
 * current cluster, i.e. NOT (k(q_right) - k(q_left) <= 1) with the
 * reference's operands (a NaN span closes it, as on the CPU).  The tie rule
 * (equal means merge) is the caller's. */
KD_HD int kd_sweep_closes(int scale, double nrm, double stored_n, double wsf, double cur,
                          double w) {
  const double q_left = wsf / stored_n;
  const double q_right = (wsf + cur + w) / stored_n;
  const double span = kd_k(scale, q_right, nrm) - kd_k(scale, q_left, nrm);
  return !(span <= 1.0);
}

#ifdef __cplusplus
}
#endif

#endif /* KD_EXACT_H */
