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
