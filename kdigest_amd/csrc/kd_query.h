// kd_query.h — quantile / cdf of one digest's centroid list on the device
// (kdigest/tdigest.mojo:485-587), shared by every query kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kd {

struct QState {
  double total, vmin, vmax, eff;
};

__device__ inline bool singleton(double stored, double eff) { return stored * eff <= 1.0 + 1e-9; }

// tdigest.mojo:485-538 for one query.
__device__ inline double quantile_one(const double* m, const double* w, const double* S, int64_t nc,
                               QState st, double q) {
  if (nc == 0) return __builtin_nan("");
  if (nc == 1) return m[0];
  double n = st.total;
  if (q <= 0.0) return st.vmin;
  if (q >= 1.0) return st.vmax;
  double target = q * n;
  double first_half = w[0] / 2.0;
  if (target < first_half) {
    if (singleton(w[0], st.eff)) return st.vmin;
    return st.vmin + (m[0] - st.vmin) * (target / first_half);
  }
  double lm = m[nc - 1], lw = w[nc - 1];
  double last_half = lw / 2.0;
  if (target > n - last_half) {
    if (singleton(lw, st.eff)) return st.vmax;
    double frac = (target - (n - last_half)) / last_half;
    return lm + (st.vmax - lm) * frac;
  }
  // First i in [0, nc-2] with target <= S[i+1].
  int64_t lo = 1, hi = nc;  // search j = i+1 in [1, nc-1]; hi == nc means none
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (target <= S[mid]) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= nc) return lm;
  int64_t i = lo - 1;
  double dw = (w[i] + w[i + 1]) / 2.0;
  double frac = (target - S[i]) / dw;
  return m[i] + (m[i + 1] - m[i]) * frac;
}

// tdigest.mojo:540-587 for one query.
__device__ inline double cdf_one(const double* m, const double* w, const double* S, int64_t nc,
                          QState st, double x) {
  if (nc == 0) return __builtin_nan("");
  if (x < st.vmin) return 0.0;
  if (x > st.vmax) return 1.0;
  if (nc == 1) {
    if (x < m[0]) return 0.0;
    if (x > m[0]) return 1.0;
    return 0.5;
  }
  double n = st.total;
  if (x < m[0]) {
    if (st.vmin == m[0]) return 0.0;
    double frac = (x - st.vmin) / (m[0] - st.vmin);
    return (w[0] / 2.0) * frac / n;
  }
  double lm = m[nc - 1], lw = w[nc - 1];
  if (x > lm) {
    if (st.vmax == lm) return 1.0;
    double frac = (x - lm) / (st.vmax - lm);
    return (n - lw / 2.0 + (lw / 2.0) * frac) / n;
  }
  // First i in [0, nc-2] with x <= m[i+1].
  int64_t lo = 1, hi = nc;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (x <= m[mid]) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= nc) return 1.0;
  int64_t i = lo - 1;
  double dw = (w[i] + w[i + 1]) / 2.0;
  if (m[i + 1] == m[i]) return (S[i] + dw / 2.0) / n;
  double frac = (x - m[i]) / (m[i + 1] - m[i]);
  return (S[i] + dw * frac) / n;
}

}  // namespace kd
