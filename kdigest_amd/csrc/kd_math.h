/*
 * kd_math.h — portable FP64 log / asin and the t-digest scale functions.
 *
 * One source compiled twice: by gcc for host code (the host digest layer and
 * the CPU checker) and by hipcc for gfx950 device code.  Both sides build with
 * -ffp-contract=off, use only IEEE +,-,*,/ and a correctly rounded sqrt, so a
 * k() value computed on the GPU is bit-identical to the one computed on the
 * CPU for the same (q, normalizer).  That is what lets the greedy-sweep cluster
 * boundaries (k(q_right) - k(q_left) <= 1, kdigest/tdigest.mojo:426-429) agree
 * exactly between the device path and the CPU restatement.
 *
 * Method (published): the argument-reduction + minimax-polynomial scheme of
 * Sun Microsystems' fdlibm (1993) for log (e_log.c: x = 2^k (1+f),
 * s = f/(2+f), log(1+f) = 2s + s*R(s^2)) and for asin (e_asin.c: rational
 * approximation on |x| < 0.5, and asin(x) = pi/2 - 2 asin(sqrt((1-x)/2))
 * above).  The coefficients are the published fdlibm minimax coefficients,
 * given here by their IEEE-754 bit patterns.  Accuracy is < 1 ulp; the tests
 * compare against glibc on millions of arguments.
 *
 * Scale functions restate kdigest/scale.mojo:32-192 (K0..K3 normalizer, k, q
 * and the q clamp to [1e-15, 1 - 1e-15] for K2/K3).
 */
#ifndef KD_MATH_H
#define KD_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define KD_HD __host__ __device__ static inline
#else
#define KD_HD static inline
#endif

#ifdef __cplusplus
extern "C++" {
#endif

KD_HD double kd_bits2d(uint64_t b) {
  double d;
  __builtin_memcpy(&d, &b, sizeof d);
  return d;
}
KD_HD uint64_t kd_d2bits(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, sizeof b);
  return b;
}

/* Correctly rounded square root on both sides (IEEE sqrt instruction on x86,
 * the IEEE-compliant llvm.sqrt.f64 lowering on gfx950). */
KD_HD double kd_sqrt(double x) { return __builtin_sqrt(x); }

/* ---------------------------------------------------------------- log ---- */
KD_HD double kd_log(double x) {
  const double ln2_hi = kd_bits2d(0x3fe62e42fee00000ull);
  const double ln2_lo = kd_bits2d(0x3dea39ef35793c76ull);
  const double c1 = kd_bits2d(0x3fe5555555555593ull);
  const double c2 = kd_bits2d(0x3fd999999997fa04ull);
  const double c3 = kd_bits2d(0x3fd2492494229359ull);
  const double c4 = kd_bits2d(0x3fcc71c51d8e78afull);
  const double c5 = kd_bits2d(0x3fc7466496cb03deull);
  const double c6 = kd_bits2d(0x3fc39a09d078c69full);
  const double c7 = kd_bits2d(0x3fc2f112df3e5244ull);

  uint64_t bits = kd_d2bits(x);
  int64_t e = 0;
  if (x != x) return x;                      /* NaN */
  if (x <= 0.0) return (x == 0.0) ? -1.0 / 0.0 : (x - x) / 0.0;
  if (bits >= 0x7ff0000000000000ull) return x; /* +inf */
  if (bits < 0x0010000000000000ull) {        /* subnormal: scale by 2^54 */
    x = x * kd_bits2d(0x4350000000000000ull);
    bits = kd_d2bits(x);
    e = -54;
  }
  /* x = 2^e * m with m in [sqrt(2)/2, sqrt(2)). */
  uint32_t hi = (uint32_t)(bits >> 32) & 0x000fffffu;
  e += (int64_t)((bits >> 52) & 0x7ff) - 1023;
  /* 0x6a09e is the high mantissa word of sqrt(2): mantissas at or above it
   * are halved so that m stays below sqrt(2). */
  uint32_t up = (hi + 0x95f64u) & 0x100000u;
  uint64_t mb = ((uint64_t)(hi | (up ^ 0x3ff00000u)) << 32) | (bits & 0xffffffffull);
  e += (int64_t)(up >> 20);
  double f = kd_bits2d(mb) - 1.0;
  double dk = (double)e;

  double s = f / (2.0 + f);
  double z = s * s;
  double w = z * z;
  /* R(z) = c1 z + c2 z^2 + ... + c7 z^7, split even/odd for latency. */
  double t_odd = z * (c1 + w * (c3 + w * (c5 + w * c7)));
  double t_even = w * (c2 + w * (c4 + w * c6));
  double R = t_odd + t_even;
  double hfsq = 0.5 * f * f;
  /* log(1+f) = f - (hfsq - s*(hfsq+R)); add k*ln2 in two pieces. */
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

/* --------------------------------------------------------------- asin ---- */
KD_HD double kd_asin_rat(double t) {
  /* t * P(t)/Q(t): the rational part of asin(x) = x + x*t*P/Q, t = x^2. */
  const double p0 = kd_bits2d(0x3fc5555555555555ull);
  const double p1 = kd_bits2d(0xbfd4d61203eb6f7dull);
  const double p2 = kd_bits2d(0x3fc9c1550e884455ull);
  const double p3 = kd_bits2d(0xbfa48228b5688f3bull);
  const double p4 = kd_bits2d(0x3f49efe07501b288ull);
  const double p5 = kd_bits2d(0x3f023de10dfdf709ull);
  const double q1 = kd_bits2d(0xc0033a271c8a2d4bull);
  const double q2 = kd_bits2d(0x40002ae59c598ac8ull);
  const double q3 = kd_bits2d(0xbfe6066c1b8d0159ull);
  const double q4 = kd_bits2d(0x3fb3b8c5b12e9282ull);
  double p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5)))));
  double q = 1.0 + t * (q1 + t * (q2 + t * (q3 + t * q4)));
  return p / q;
}

KD_HD double kd_asin(double x) {
  const double pio2_hi = kd_bits2d(0x3ff921fb54442d18ull);
  const double pio2_lo = kd_bits2d(0x3c91a62633145c07ull);
  const double pio4_hi = kd_bits2d(0x3fe921fb54442d18ull);
  double ax = x < 0.0 ? -x : x;
  if (x != x) return x;
  if (ax >= 1.0) {
    if (ax == 1.0) return x * pio2_hi + x * pio2_lo;
    return (x - x) / (x - x);
  }
  if (ax < 0.5) {
    if (ax < kd_bits2d(0x3e40000000000000ull)) return x; /* |x| < 2^-27 */
    return x + x * kd_asin_rat(x * x);
  }
  /* asin(|x|) = pi/2 - 2*asin(s), s = sqrt((1-|x|)/2). */
  double t = (1.0 - ax) * 0.5;
  double r = kd_asin_rat(t);
  double s = kd_sqrt(t);
  double res;
  if (ax >= 0.975) {
    res = pio2_hi - (2.0 * (s + s * r) - pio2_lo);
  } else {
    /* s = sh + c exactly-ish: sh is s with its low 32 bits cleared. */
    double sh = kd_bits2d(kd_d2bits(s) & 0xffffffff00000000ull);
    double c = (t - sh * sh) / (s + sh);
    double p = 2.0 * s * r - (pio2_lo - 2.0 * c);
    double q = pio4_hi - 2.0 * sh;
    res = pio4_hi - (p - q);
  }
  return x < 0.0 ? -res : res;
}

/* ---------------------------------------------------------------- exp ---- */
/* fdlibm e_exp.c method: x = k ln2 + r with |r| <= ln2/2 (ln2 split in a
 * high part with trailing zeros and a low part), exp(r) = 1 + r + r c/(2 - c)
 * with c = r - r^2 R(r^2) (Remez degree 5), then scaling by 2^k.  Used for
 * every decay factor e^(+-log_scale) on the host and on the device. */
KD_HD double kd_exp(double x) {
  const double invln2 = kd_bits2d(0x3ff71547652b82feull);
  const double ln2hi = kd_bits2d(0x3fe62e42fee00000ull);
  const double ln2lo = kd_bits2d(0x3dea39ef35793c76ull);
  const double p1 = kd_bits2d(0x3fc555555555553eull);
  const double p2 = kd_bits2d(0xbf66c16c16bebd93ull);
  const double p3 = kd_bits2d(0x3f11566aaf25de2cull);
  const double p4 = kd_bits2d(0xbebbbd41c5d26bf1ull);
  const double p5 = kd_bits2d(0x3e66376972bea4d0ull);
  const double o_thr = kd_bits2d(0x40862e42fefa39efull);
  const double u_thr = kd_bits2d(0xc0874910d52d3051ull);
  if (x != x) return x;
  if (x > o_thr) return 1.0 / 0.0;
  if (x < u_thr) return 0.0;
  double ax = x < 0.0 ? -x : x;
  if (ax < kd_bits2d(0x3e30000000000000ull)) return 1.0 + x; /* |x| < 2^-28 */
  int k;
  double hi, lo;
  if (ax > kd_bits2d(0x3fd62e42fefa39efull)) { /* |x| > ln2/2 */
    if (ax < kd_bits2d(0x3ff0a2b23f3bab73ull)) {                            /* |x| < 1.5 ln2 */
      k = x < 0.0 ? -1 : 1;
      hi = x < 0.0 ? x + ln2hi : x - ln2hi;
      lo = x < 0.0 ? -ln2lo : ln2lo;
    } else {
      k = (int)(invln2 * x + (x < 0.0 ? -0.5 : 0.5));
      double dk = (double)k;
      hi = x - dk * ln2hi;
      lo = dk * ln2lo;
    }
  } else {
    k = 0;
    hi = x;
    lo = 0.0;
  }
  double r = hi - lo;
  double t = r * r;
  double c = r - t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5))));
  double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
  if (k == 0) return y;
  if (k >= -1021) return kd_bits2d(kd_d2bits(y) + ((uint64_t)(int64_t)k << 52));
  /* Subnormal result: scale in two steps. */
  return kd_bits2d(kd_d2bits(y) + ((uint64_t)(int64_t)(k + 1000) << 52)) *
         kd_bits2d(0x0170000000000000ull);
}

/* ------------------------------------------------------ scale functions -- */
/* Scale ids: 0 = K0 (linear), 1 = K1 (asin), 2 = K2 (log-odds), 3 = K3. */
enum { KD_K0 = 0, KD_K1 = 1, KD_K2 = 2, KD_K3 = 3 };

KD_HD double kd_clamp_q(double q) {
  /* scale.mojo:32-47 (FP64 eps = 1e-15). */
  const double eps = 1e-15;
  if (q < eps) return eps;
  if (q > 1.0 - eps) return 1.0 - eps;
  return q;
}

KD_HD double kd_normalizer(int scale, double compression, double n) {
  /* scale.mojo:86-89, 109-112, 139-143, 171-175. */
  const double pi = kd_bits2d(0x400921fb54442d18ull);
  switch (scale) {
    case KD_K0: return compression / 2.0;
    case KD_K1: return compression / (2.0 * pi);
    case KD_K2: return compression / (4.0 * kd_log(n / compression) + 24.0);
    default: return compression / (4.0 * kd_log(n / compression) + 21.0);
  }
}

KD_HD double kd_k(int scale, double q, double nrm) {
  /* scale.mojo:91-94, 114-118, 145-150, 177-184. */
  switch (scale) {
    case KD_K0: return nrm * q;
    case KD_K1: return nrm * kd_asin(2.0 * q - 1.0);
    case KD_K2: {
      double qc = kd_clamp_q(q);
      return nrm * kd_log(qc / (1.0 - qc));
    }
    default: {
      double qc = kd_clamp_q(q);
      if (qc <= 0.5) return nrm * kd_log(2.0 * qc);
      return -nrm * kd_log(2.0 * (1.0 - qc));
    }
  }
}

#ifdef __cplusplus
}
#endif

#endif /* KD_MATH_H */
