/*
 * kd_oracle.c — CPU restatement of the reference MergingDigest (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path.  Nothing under
 * kdigest_amd/ links or loads this file.
 *
 * Built with -O2 -ffp-contract=off so every expression rounds exactly as
 * written, in the operation order of the reference.  k() uses kd_math.h (the
 * same portable FP64 log/asin the device uses; its accuracy against glibc is
 * pinned by the tests).  Decay factors e^(+-log_scale) use kd_exp from the
 * same header (fdlibm method), as the host layer and the device do.
 *
 * Reference files restated (paths relative to the reference root):
 *   src/kdigest/centroid.mojo:25-40    incremental weighted mean
 *   src/kdigest/scale.mojo:32-192      K0..K3, clamp
 *   src/kdigest/decay.mojo:23-54       Decay
 *   src/kdigest/tdigest.mojo:39-726    MergingDigest
 *   src/kdigest/grouped.mojo:60-162    per-key batches
 */
#include "kd_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../kdigest_amd/csrc/kd_math.h"

#define BUFFER_FACTOR 5                 /* tdigest.mojo:39 */
#define RENORM_LOG_THRESHOLD 400.0      /* tdigest.mojo:43 */
#define PRUNE_EPSILON 1e-6              /* tdigest.mojo:47 */
#define GPU_CROSSOVER 60000             /* tdigest.mojo:55 */
#define GPU_CROSSOVER_SESSION 40000     /* tdigest.mojo:58 */
#define SERDE_VERSION 1                 /* tdigest.mojo:61 */

#ifdef KD_ORACLE_GLIBC_K
/* Sensitivity variant: the same scale formulas through glibc's log/asin. */
static double glibc_k(int scale, double q, double nrm) {
  switch (scale) {
    case KD_K0: return nrm * q;
    case KD_K1: return nrm * asin(2.0 * q - 1.0);
    case KD_K2: { double qc = kd_clamp_q(q); return nrm * log(qc / (1.0 - qc)); }
    default: {
      double qc = kd_clamp_q(q);
      if (qc <= 0.5) return nrm * log(2.0 * qc);
      return -nrm * log(2.0 * (1.0 - qc));
    }
  }
}
static double glibc_norm(int scale, double c, double n) {
  switch (scale) {
    case KD_K0: return c / 2.0;
    case KD_K1: return c / (2.0 * M_PI);
    case KD_K2: return c / (4.0 * log(n / c) + 24.0);
    default: return c / (4.0 * log(n / c) + 21.0);
  }
}
#define K_FN glibc_k
#define NORM_FN glibc_norm
#else
#define K_FN kd_k
#define NORM_FN kd_normalizer
#endif

double ko_k(int scale, double q, double nrm) { return K_FN(scale, q, nrm); }
double ko_normalizer(int scale, double c, double n) { return NORM_FN(scale, c, n); }
double ko_log(double x) { return kd_log(x); }
double ko_asin(double x) { return kd_asin(x); }
double ko_exp(double x) { return kd_exp(x); }

/* centroid.mojo:25-40 */
static inline void c_add_sample(ko_centroid *c, double value, double w) {
  double new_weight = c->weight + w;
  c->mean += (value - c->mean) * (w / new_weight);
  c->weight = new_weight;
}

static int reserve(ko_centroid **p, int64_t *cap, int64_t need) {
  if (need <= *cap) return 0;
  int64_t nc = *cap ? *cap : 16;
  while (nc < need) nc *= 2;
  ko_centroid *q = (ko_centroid *)realloc(*p, (size_t)nc * sizeof(ko_centroid));
  if (!q) return KO_ENOMEM;
  *p = q;
  *cap = nc;
  return 0;
}

/* tdigest.mojo:135-151 */
ko_digest *ko_new(int scale, double compression, double lambda) {
  ko_digest *d = (ko_digest *)calloc(1, sizeof(ko_digest));
  if (!d) return NULL;
  d->scale = scale;
  d->compression = compression;
  d->lambda = lambda;
  d->buffer_limit = (int64_t)compression * BUFFER_FACTOR;
  return d;
}

void ko_free(ko_digest *d) {
  if (!d) return;
  free(d->cent);
  free(d->buf);
  free(d);
}

ko_digest *ko_clone(const ko_digest *s) {
  ko_digest *d = ko_new(s->scale, s->compression, s->lambda);
  if (!d) return NULL;
  reserve(&d->cent, &d->cap_c, s->nc);
  reserve(&d->buf, &d->cap_b, s->nb);
  if (s->nc) memcpy(d->cent, s->cent, (size_t)s->nc * sizeof(ko_centroid));
  if (s->nb) memcpy(d->buf, s->buf, (size_t)s->nb * sizeof(ko_centroid));
  d->nc = s->nc;
  d->nb = s->nb;
  d->buffer_limit = s->buffer_limit;
  d->stored_total = s->stored_total;
  d->log_scale = s->log_scale;
  d->clock = s->clock;
  d->min = s->min;
  d->max = s->max;
  return d;
}

static inline int decay_active(const ko_digest *d) { return d->lambda > 0.0; } /* decay.mojo:49-51 */

/* tdigest.mojo:317-322 */
static inline int validate(double value, double weight) {
  if (isnan(value) || isinf(value)) return KO_EINVAL;
  if (weight <= 0.0 || isnan(weight) || isinf(weight)) return KO_EINVAL;
  return KO_OK;
}

/* tdigest.mojo:359-371 */
static void renormalize(ko_digest *d) {
  double factor = kd_exp(-d->log_scale);
  for (int64_t i = 0; i < d->nc; i++) d->cent[i].weight *= factor;
  for (int64_t i = 0; i < d->nb; i++) d->buf[i].weight *= factor;
  d->stored_total *= factor;
  d->log_scale = 0.0;
}

/* tdigest.mojo:324-332 */
static void advance_clock(ko_digest *d, double new_clock) {
  if (decay_active(d) && new_clock > d->clock) {
    d->log_scale += d->lambda * (new_clock - d->clock);
    if (d->log_scale > RENORM_LOG_THRESHOLD) renormalize(d);
  }
  d->clock = new_clock;
}

/* tdigest.mojo:352-357 */
static inline double scale_factor(const ko_digest *d) {
  if (!decay_active(d)) return 1.0;
  return kd_exp(d->log_scale);
}

/* tdigest.mojo:373-377 */
static inline double effective_count(const ko_digest *d) {
  if (!decay_active(d)) return d->stored_total;
  return d->stored_total * kd_exp(-d->log_scale);
}

/* Stable merge sort by mean: the permutation of the stable insertion sort at
 * tdigest.mojo:64-77 (a record moves left only past strictly larger means). */
static void msort(ko_centroid *a, ko_centroid *tmp, int64_t n) {
  if (n <= 16) {
    for (int64_t i = 1; i < n; i++) {
      ko_centroid key = a[i];
      int64_t j = i - 1;
      while (j >= 0 && a[j].mean > key.mean) {
        a[j + 1] = a[j];
        j--;
      }
      a[j + 1] = key;
    }
    return;
  }
  int64_t h = n / 2;
  msort(a, tmp, h);
  msort(a + h, tmp, n - h);
  if (!(a[h - 1].mean > a[h].mean)) return; /* already in order */
  memcpy(tmp, a, (size_t)h * sizeof(ko_centroid));
  int64_t i = 0, j = h, k = 0;
  while (i < h && j < n) {
    if (a[j].mean < tmp[i].mean) a[k++] = a[j++];
    else a[k++] = tmp[i++];
  }
  while (i < h) a[k++] = tmp[i++];
}

void ko_stable_sort(ko_centroid *cs, int64_t n) {
  if (n < 2) return;
  ko_centroid *tmp = (ko_centroid *)malloc((size_t)(n / 2 + 1) * sizeof(ko_centroid));
  msort(cs, tmp, n);
  free(tmp);
}

/* The greedy sweep, tdigest.mojo:414-435. */
int64_t ko_sweep(int scale, double nrm, double stored_n, const double *means,
                 const double *weights, int64_t n, double *om, double *ow) {
  if (n == 0) return 0;
  int64_t c = 0;
  ko_centroid cur = {means[0], weights[0]};
  double weight_so_far = 0.0;
  for (int64_t i = 1; i < n; i++) {
    double cm = means[i], cw = weights[i];
    double q_left = weight_so_far / stored_n;
    double q_right = (weight_so_far + cur.weight + cw) / stored_n;
    double k_span = K_FN(scale, q_right, nrm) - K_FN(scale, q_left, nrm);
    if (cm == cur.mean || k_span <= 1.0) {
      c_add_sample(&cur, cm, cw);
    } else {
      om[c] = cur.mean;
      ow[c] = cur.weight;
      c++;
      weight_so_far += cur.weight;
      cur.mean = cm;
      cur.weight = cw;
    }
  }
  om[c] = cur.mean;
  ow[c] = cur.weight;
  return c + 1;
}

/* tdigest.mojo:441-458 */
static void prune(ko_digest *d) {
  double stored_floor = PRUNE_EPSILON * kd_exp(d->log_scale);
  double removed = 0.0;
  int64_t k = 0;
  for (int64_t i = 0; i < d->nc; i++) {
    if (d->cent[i].weight >= stored_floor) d->cent[k++] = d->cent[i];
    else removed += d->cent[i].weight;
  }
  if (removed > 0.0) {
    d->stored_total -= removed;
    d->nc = k;
  }
}

/* Sort + sweep of a work list; installs the result as the centroids. */
static void compress_install(ko_digest *d, ko_centroid *work, int64_t n) {
  ko_stable_sort(work, n);
  double effective_n = effective_count(d);                              /* :410 */
  double stored_n = d->stored_total;                                    /* :411 */
  double nrm = NORM_FN(d->scale, d->compression, effective_n);         /* :412 */
  double *m = (double *)malloc((size_t)n * sizeof(double) * 4);
  double *w = m + n, *om = m + 2 * n, *ow = m + 3 * n;
  for (int64_t i = 0; i < n; i++) {
    m[i] = work[i].mean;
    w[i] = work[i].weight;
  }
  int64_t c = ko_sweep(d->scale, nrm, stored_n, m, w, n, om, ow);
  reserve(&d->cent, &d->cap_c, c);
  for (int64_t i = 0; i < c; i++) {
    d->cent[i].mean = om[i];
    d->cent[i].weight = ow[i];
  }
  d->nc = c;
  free(m);
}

/* tdigest.mojo:391-439 */
void ko_merge(ko_digest *d) {
  if (d->nb == 0) return;
  int64_t n = d->nc + d->nb;
  ko_centroid *work = (ko_centroid *)malloc((size_t)n * sizeof(ko_centroid));
  memcpy(work, d->cent, (size_t)d->nc * sizeof(ko_centroid));
  memcpy(work + d->nc, d->buf, (size_t)d->nb * sizeof(ko_centroid));
  d->nb = 0;
  compress_install(d, work, n);
  free(work);
  if (decay_active(d)) prune(d);
}

/* tdigest.mojo:334-350 */
static void insert(ko_digest *d, double value, double weight) {
  if (d->stored_total == 0.0 && d->nb == 0) {
    d->min = value;
    d->max = value;
  } else {
    if (value < d->min) d->min = value;
    if (value > d->max) d->max = value;
  }
  double scaled = weight * scale_factor(d);
  reserve(&d->buf, &d->cap_b, d->nb + 1);
  d->buf[d->nb].mean = value;
  d->buf[d->nb].weight = scaled;
  d->nb++;
  d->stored_total += scaled;
  if (d->nb >= d->buffer_limit) ko_merge(d);
}

/* tdigest.mojo:153-161 */
int ko_add(ko_digest *d, double value, double weight) {
  int rc = validate(value, weight);
  if (rc) return rc;
  advance_clock(d, d->clock + 1.0);
  insert(d, value, weight);
  return KO_OK;
}

/* tdigest.mojo:163-175 */
int ko_add_at(ko_digest *d, double value, double timestamp, double weight) {
  int rc = validate(value, weight);
  if (rc) return rc;
  if (timestamp > d->clock) advance_clock(d, timestamp);
  insert(d, value, weight);
  return KO_OK;
}

/* tdigest.mojo:199-202: the per-sample loop (validation per sample, so a bad
 * value mid-batch leaves the earlier samples applied, as in the reference). */
int ko_add_loop(ko_digest *d, const double *values, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    int rc = ko_add(d, values[i], 1.0);
    if (rc) return rc;
  }
  return KO_OK;
}

/* B semantics: _gpu_batch_prepare (tdigest.mojo:236-296), the compress as the
 * CPU greedy sweep (391-439) over the prepared records in FP64, and
 * _gpu_batch_finish (298-315): stored_total reconciled to the sum of the
 * compressed weights, then prune when decay is active. */
int ko_add_batch_b(ko_digest *d, const double *values, int64_t n) {
  if (n == 0) return KO_OK;
  for (int64_t i = 0; i < n; i++)
    if (validate(values[i], 1.0)) return KO_EINVAL;
  advance_clock(d, d->clock + (double)n);                  /* :258 */
  double sf = scale_factor(d);                              /* :259 */
  int64_t tot = d->nc + d->nb + n;
  ko_centroid *work = (ko_centroid *)malloc((size_t)tot * sizeof(ko_centroid));
  memcpy(work, d->cent, (size_t)d->nc * sizeof(ko_centroid));
  memcpy(work + d->nc, d->buf, (size_t)d->nb * sizeof(ko_centroid));
  int64_t had_centroids = d->nc;
  d->nb = 0;                                                /* :271 */
  double vmin = values[0], vmax = values[0];
  for (int64_t i = 0; i < n; i++) {
    double v = values[i];
    work[tot - n + i].mean = v;   /* centroids ++ buffer ++ values (:263-277) */
    work[tot - n + i].weight = sf;
    if (v < vmin) vmin = v;
    if (v > vmax) vmax = v;
  }
  d->stored_total += (double)n * sf;                        /* :282 */
  if (d->stored_total == (double)n * sf && had_centroids == 0) { /* :285-295 */
    d->min = vmin;
    d->max = vmax;
  } else {
    if (vmin < d->min) d->min = vmin;
    if (vmax > d->max) d->max = vmax;
  }
  compress_install(d, work, tot);
  free(work);
  double total = 0.0;                                       /* :304-313 */
  for (int64_t i = 0; i < d->nc; i++) total += d->cent[i].weight;
  d->stored_total = total;
  if (decay_active(d)) prune(d);                            /* :314-315 */
  return KO_OK;
}

/* tdigest.mojo:177-221: the dispatcher with a GPU present. */
int ko_add_batch(ko_digest *d, const double *values, int64_t n, int with_session) {
  if (n == 0) return KO_OK;
  int64_t cross = with_session ? GPU_CROSSOVER_SESSION : GPU_CROSSOVER;
  if (n < cross) return ko_add_loop(d, values, n);
  return ko_add_batch_b(d, values, n);
}

double ko_total_weight(const ko_digest *d) { return effective_count(d); } /* :465-467 */

/* tdigest.mojo:475-483 */
static inline int is_singleton(const ko_digest *d, double stored) {
  double eff = decay_active(d) ? stored * kd_exp(-d->log_scale) : stored;
  return eff <= 1.0 + 1e-9;
}

/* tdigest.mojo:485-538 */
int ko_quantile(ko_digest *d, double q, double *out) {
  if (q < 0.0 || q > 1.0) return KO_ERANGE;
  ko_merge(d);
  const ko_centroid *cs = d->cent;
  int64_t nc = d->nc;
  if (nc == 0) return KO_EEMPTY;
  if (nc == 1) { *out = cs[0].mean; return KO_OK; }
  double n = d->stored_total;
  if (q <= 0.0) { *out = d->min; return KO_OK; }
  if (q >= 1.0) { *out = d->max; return KO_OK; }
  double target = q * n;
  double first_half = cs[0].weight / 2.0;
  if (target < first_half) {
    if (is_singleton(d, cs[0].weight)) { *out = d->min; return KO_OK; }
    *out = d->min + (cs[0].mean - d->min) * (target / first_half);
    return KO_OK;
  }
  ko_centroid last = cs[nc - 1];
  double last_half = last.weight / 2.0;
  if (target > n - last_half) {
    if (is_singleton(d, last.weight)) { *out = d->max; return KO_OK; }
    double frac = (target - (n - last_half)) / last_half;
    *out = last.mean + (d->max - last.mean) * frac;
    return KO_OK;
  }
  double weight_so_far = cs[0].weight / 2.0;
  for (int64_t i = 0; i < nc - 1; i++) {
    ko_centroid left = cs[i], right = cs[i + 1];
    double dw = (left.weight + right.weight) / 2.0;
    if (target <= weight_so_far + dw) {
      double frac = (target - weight_so_far) / dw;
      *out = left.mean + (right.mean - left.mean) * frac;
      return KO_OK;
    }
    weight_so_far += dw;
  }
  *out = last.mean;
  return KO_OK;
}

/* tdigest.mojo:540-587 */
int ko_cdf(ko_digest *d, double x, double *out) {
  ko_merge(d);
  const ko_centroid *cs = d->cent;
  int64_t nc = d->nc;
  if (nc == 0) return KO_EEMPTY;
  if (x < d->min) { *out = 0.0; return KO_OK; }
  if (x > d->max) { *out = 1.0; return KO_OK; }
  if (nc == 1) {
    if (x < cs[0].mean) *out = 0.0;
    else if (x > cs[0].mean) *out = 1.0;
    else *out = 0.5;
    return KO_OK;
  }
  double n = d->stored_total;
  if (x < cs[0].mean) {
    if (d->min == cs[0].mean) { *out = 0.0; return KO_OK; }
    double frac = (x - d->min) / (cs[0].mean - d->min);
    *out = (cs[0].weight / 2.0) * frac / n;
    return KO_OK;
  }
  ko_centroid last = cs[nc - 1];
  if (x > last.mean) {
    if (d->max == last.mean) { *out = 1.0; return KO_OK; }
    double frac = (x - last.mean) / (d->max - last.mean);
    *out = (n - last.weight / 2.0 + (last.weight / 2.0) * frac) / n;
    return KO_OK;
  }
  double weight_so_far = cs[0].weight / 2.0;
  for (int64_t i = 0; i < nc - 1; i++) {
    ko_centroid left = cs[i], right = cs[i + 1];
    if (x <= right.mean) {
      double dw = (left.weight + right.weight) / 2.0;
      if (right.mean == left.mean) { *out = (weight_so_far + dw / 2.0) / n; return KO_OK; }
      double frac = (x - left.mean) / (right.mean - left.mean);
      *out = (weight_so_far + dw * frac) / n;
      return KO_OK;
    }
    weight_so_far += (left.weight + right.weight) / 2.0;
  }
  *out = 1.0;
  return KO_OK;
}

/* tdigest.mojo:589-623 */
int ko_merge_into(ko_digest *self, ko_digest *other) {
  ko_merge(other);
  if (effective_count(other) == 0.0) return KO_OK;
  ko_merge(self);
  if (decay_active(self)) renormalize(self);
  if (decay_active(other)) renormalize(other);
  int was_empty = self->stored_total == 0.0;
  reserve(&self->buf, &self->cap_b, self->nb + other->nc);
  for (int64_t i = 0; i < other->nc; i++) self->buf[self->nb++] = other->cent[i];
  self->stored_total += other->stored_total;
  if (was_empty) {
    self->min = other->min;
    self->max = other->max;
  } else {
    self->min = fmin(self->min, other->min);
    self->max = fmax(self->max, other->max);
  }
  self->clock = self->clock > other->clock ? self->clock : other->clock;
  ko_merge(self);
  return KO_OK;
}

/* tdigest.mojo:80-94 */
static void put_f64(uint8_t *p, double v) {
  uint64_t b = kd_d2bits(v);
  for (int i = 0; i < 8; i++) p[i] = (uint8_t)((b >> (8 * i)) & 0xff);
}
static int get_f64(const uint8_t *data, int64_t n, int64_t off, double *v) {
  if (off + 8 > n) return KO_ETRUNC;
  uint64_t b = 0;
  for (int i = 0; i < 8; i++) b |= (uint64_t)data[off + i] << (8 * i);
  *v = kd_bits2d(b);
  return KO_OK;
}

/* tdigest.mojo:627-652 */
int64_t ko_to_bytes(ko_digest *d, uint8_t *out, int64_t cap) {
  ko_merge(d);
  int64_t need = 1 + 8 * 8 + 16 * d->nc;
  if (!out || cap < need) return need;
  out[0] = SERDE_VERSION;
  double hdr[8] = {d->compression, d->lambda, d->log_scale, d->clock,
                   d->stored_total, d->min, d->max, (double)d->nc};
  for (int i = 0; i < 8; i++) put_f64(out + 1 + 8 * i, hdr[i]);
  for (int64_t i = 0; i < d->nc; i++) {
    put_f64(out + 65 + 16 * i, d->cent[i].mean);
    put_f64(out + 73 + 16 * i, d->cent[i].weight);
  }
  return need;
}

/* tdigest.mojo:654-698 */
int ko_from_bytes(const uint8_t *data, int64_t n, int scale, ko_digest **out) {
  *out = NULL;
  if (n == 0 || data[0] != SERDE_VERSION) return KO_ESERDE;
  double h[8];
  int64_t off = 1;
  for (int i = 0; i < 8; i++, off += 8)
    if (get_f64(data, n, off, &h[i])) return KO_ETRUNC;
  int64_t count = (int64_t)h[7];
  ko_digest *d = ko_new(scale, h[0], h[1]);
  if (count > 0 && reserve(&d->cent, &d->cap_c, count)) { ko_free(d); return KO_ENOMEM; }
  for (int64_t i = 0; i < count; i++) {
    double m, w;
    if (get_f64(data, n, off, &m) || get_f64(data, n, off + 8, &w)) {
      ko_free(d);
      return KO_ETRUNC;
    }
    off += 16;
    d->cent[i].mean = m;
    d->cent[i].weight = w;
    d->nc = i + 1;
  }
  d->log_scale = h[2];
  d->clock = h[3];
  d->stored_total = h[4];
  d->min = h[5];
  d->max = h[6];
  *out = d;
  return KO_OK;
}

/* tdigest.mojo:700-726 */
ko_digest *ko_from_centroids(int scale, double compression, const double *means,
                             const double *weights, int64_t n, double min_v, double max_v) {
  ko_digest *d = ko_new(scale, compression, 0.0);
  reserve(&d->cent, &d->cap_c, n);
  double total = 0.0;
  for (int64_t i = 0; i < n; i++) {
    d->cent[i].mean = means[i];
    d->cent[i].weight = weights[i];
    total += weights[i];
  }
  d->nc = n;
  d->stored_total = total;
  d->clock = total;
  d->min = min_v;
  d->max = max_v;
  return d;
}

int64_t ko_ncentroids(const ko_digest *d) { return d->nc; }
int64_t ko_nbuffer(const ko_digest *d) { return d->nb; }

void ko_get_centroids(const ko_digest *d, double *means, double *weights) {
  for (int64_t i = 0; i < d->nc; i++) {
    means[i] = d->cent[i].mean;
    weights[i] = d->cent[i].weight;
  }
}

void ko_get_state(const ko_digest *d, double *o) {
  o[0] = d->stored_total;
  o[1] = d->log_scale;
  o[2] = d->clock;
  o[3] = d->min;
  o[4] = d->max;
  o[5] = d->compression;
  o[6] = d->lambda;
}

void ko_set_state(ko_digest *d, const double *means, const double *weights, int64_t nc,
                  const double *bm, const double *bw, int64_t nb, const double *s) {
  reserve(&d->cent, &d->cap_c, nc);
  reserve(&d->buf, &d->cap_b, nb);
  for (int64_t i = 0; i < nc; i++) {
    d->cent[i].mean = means[i];
    d->cent[i].weight = weights[i];
  }
  for (int64_t i = 0; i < nb; i++) {
    d->buf[i].mean = bm[i];
    d->buf[i].weight = bw[i];
  }
  d->nc = nc;
  d->nb = nb;
  d->stored_total = s[0];
  d->log_scale = s[1];
  d->clock = s[2];
  d->min = s[3];
  d->max = s[4];
}

/* ---------------------------------------------------------------- grouped */
typedef struct {
  ko_digest **digests;
  const int64_t *off;
  const double *vals;
  int64_t k0, k1;
} grp_job;

static void *grp_worker(void *arg) {
  grp_job *j = (grp_job *)arg;
  for (int64_t k = j->k0; k < j->k1; k++) {
    int64_t a = j->off[k], b = j->off[k + 1];
    if (b > a) ko_add_batch_b(j->digests[k], j->vals + a, b - a);
  }
  return NULL;
}

/* grouped.mojo:60-162 with an accelerator: validate all (91-97), group per key
 * preserving batch order (99-104), then one B-semantics update per touched key
 * (113-162).  Keys are partitioned over `threads` host threads. */
int ko_grouped_add_batch(ko_digest **digests, int64_t num_keys, const uint32_t *keys,
                         const double *values, int64_t n, int threads) {
  if (n == 0) return KO_OK;
  for (int64_t i = 0; i < n; i++) {
    if ((int64_t)keys[i] >= num_keys) return KO_EINVAL;
    if (validate(values[i], 1.0)) return KO_EINVAL;
  }
  int64_t *off = (int64_t *)calloc((size_t)num_keys + 1, sizeof(int64_t));
  double *vals = (double *)malloc((size_t)n * sizeof(double));
  for (int64_t i = 0; i < n; i++) off[keys[i] + 1]++;
  for (int64_t k = 0; k < num_keys; k++) off[k + 1] += off[k];
  int64_t *cur = (int64_t *)malloc((size_t)num_keys * sizeof(int64_t));
  memcpy(cur, off, (size_t)num_keys * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) vals[cur[keys[i]]++] = values[i];
  free(cur);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  grp_job jobs[256];
  pthread_t th[256];
  /* Balance by sample count: contiguous key ranges of ~n/threads samples. */
  int64_t k = 0;
  for (int t = 0; t < threads; t++) {
    int64_t target = (int64_t)((double)n * (t + 1) / threads);
    jobs[t].digests = digests;
    jobs[t].off = off;
    jobs[t].vals = vals;
    jobs[t].k0 = k;
    while (k < num_keys && (off[k + 1] <= target || t == threads - 1)) k++;
    jobs[t].k1 = k;
  }
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, grp_worker, &jobs[t]);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(vals);
  free(off);
  return KO_OK;
}

/* test_tdigest.mojo:93-103: fraction of samples <= v, minus q. */
double ko_rank_error(const double *sorted, int64_t n, double v, double q) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (sorted[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  double r = (double)lo / (double)n - q;
  return r < 0 ? -r : r;
}

/* Vectors of queries: out[i] = quantile(qs[i]) (or cdf), stopping at the
 * first error. */
int ko_quantiles(ko_digest *d, const double *qs, int64_t n, double *out) {
  for (int64_t i = 0; i < n; i++) {
    int rc = ko_quantile(d, qs[i], &out[i]);
    if (rc) return rc;
  }
  return KO_OK;
}

int ko_cdfs(ko_digest *d, const double *xs, int64_t n, double *out) {
  for (int64_t i = 0; i < n; i++) {
    int rc = ko_cdf(d, xs[i], &out[i]);
    if (rc) return rc;
  }
  return KO_OK;
}

/* add_at per sample (tdigest.mojo:163-175), in order. */
int ko_add_at_many(ko_digest *d, const double *values, const double *times,
                   const double *weights, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    int rc = ko_add_at(d, values[i], times[i], weights[i]);
    if (rc) return rc;
  }
  return KO_OK;
}

/* ------------------------------------------------------- threaded replays */
typedef struct {
  ko_digest **ds;
  const int64_t *off;
  const double *v, *t, *w;
  int64_t d0, d1;
  int rc;
} at_job;

static void *at_worker(void *arg) {
  at_job *j = (at_job *)arg;
  for (int64_t d = j->d0; d < j->d1 && !j->rc; d++) {
    int64_t a = j->off[d], b = j->off[d + 1];
    j->rc = ko_add_at_many(j->ds[d], j->v + a, j->t + a, j->w + a, b - a);
  }
  return NULL;
}

/* Digest d gets add_at of samples [off[d], off[d+1]) in order
 * (tdigest.mojo:163-175 per sample); digests split over `threads` host
 * threads by sample count. */
int ko_batch_add_at(ko_digest **ds, int64_t nd, const int64_t *off, const double *v,
                    const double *t, const double *w, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  at_job jobs[256];
  pthread_t th[256];
  int64_t d = 0, n = off[nd];
  for (int i = 0; i < threads; i++) {
    int64_t target = (int64_t)((double)n * (i + 1) / threads);
    jobs[i] = (at_job){ds, off, v, t, w, d, d, 0};
    while (d < nd && (off[d + 1] <= target || i == threads - 1)) d++;
    jobs[i].d1 = d;
  }
  for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, at_worker, &jobs[i]);
  int rc = 0;
  for (int i = 0; i < threads; i++) {
    pthread_join(th[i], NULL);
    if (jobs[i].rc && !rc) rc = jobs[i].rc;
  }
  return rc;
}

typedef struct {
  ko_digest **ds;
  int64_t nreq;
  const int64_t *keys, *off;
  const double *data;
  int tid, threads;
  int rc;
} rp_job;

static void *rp_worker(void *arg) {
  rp_job *j = (rp_job *)arg;
  for (int64_t r = 0; r < j->nreq && !j->rc; r++) {
    if (j->keys[r] % j->threads != j->tid) continue;
    int64_t a = j->off[r], b = j->off[r + 1];
    j->rc = ko_add_batch_b(j->ds[j->keys[r]], j->data + a, b - a);
  }
  return NULL;
}

/* Ingest requests in arrival order, each one B-semantics batch merge of its
 * key (tdigest.mojo:223-315): request r adds data[off[r], off[r+1]) to
 * ds[keys[r]].  Keys are split over `threads` host threads (key mod
 * threads), so each key sees its requests in order. */
int ko_replay_ingests(ko_digest **ds, int64_t nreq, const int64_t *keys, const int64_t *off,
                      const double *data, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  rp_job jobs[256];
  pthread_t th[256];
  for (int i = 0; i < threads; i++) {
    jobs[i] = (rp_job){ds, nreq, keys, off, data, i, threads, 0};
    pthread_create(&th[i], NULL, rp_worker, &jobs[i]);
  }
  int rc = 0;
  for (int i = 0; i < threads; i++) {
    pthread_join(th[i], NULL);
    if (jobs[i].rc && !rc) rc = jobs[i].rc;
  }
  return rc;
}

/* L sequential additions W = W + x, one rounding each (the running weight of
 * centroid.mojo:34-36 over equal weights). */
double ko_add_loop_sum(double W, double x, int64_t L) {
  for (int64_t i = 0; i < L; i++) W = W + x;
  return W;
}
