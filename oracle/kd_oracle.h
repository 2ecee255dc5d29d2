/*
 * kd_oracle.h — CPU restatement of the reference MergingDigest.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU path is compared
 * against; nothing in the kdigest_amd package links, loads or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 */
#ifndef KD_ORACLE_H
#define KD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double mean;
  double weight;
} ko_centroid;

/* kdigest/tdigest.mojo:116-151 (state). */
typedef struct {
  int scale;             /* 0..3 = ScaleK0..K3 */
  double compression;
  double lambda;         /* Decay.rate_lambda (decay.mojo:23-37) */
  ko_centroid *cent;
  int64_t nc, cap_c;
  ko_centroid *buf;
  int64_t nb, cap_b;
  int64_t buffer_limit;  /* Int(compression) * 5 (tdigest.mojo:146) */
  double stored_total;
  double log_scale;
  double clock;
  double min, max;
} ko_digest;

enum {
  KO_OK = 0,
  KO_EINVAL = -1,   /* invalid value / weight / argument */
  KO_EEMPTY = -2,   /* query on an empty digest */
  KO_ERANGE = -3,   /* q outside [0, 1] */
  KO_ESERDE = -4,   /* bad version byte */
  KO_ETRUNC = -5,   /* truncated bytes */
  KO_ENOMEM = -6
};

ko_digest *ko_new(int scale, double compression, double lambda);
void ko_free(ko_digest *d);
ko_digest *ko_clone(const ko_digest *d);

int ko_add(ko_digest *d, double value, double weight);                    /* A */
int ko_add_at(ko_digest *d, double value, double timestamp, double weight);
int ko_add_loop(ko_digest *d, const double *values, int64_t n);           /* A loop */
int ko_add_batch_b(ko_digest *d, const double *values, int64_t n);        /* B */
int ko_add_batch(ko_digest *d, const double *values, int64_t n, int with_session);
void ko_merge(ko_digest *d);
int ko_quantile(ko_digest *d, double q, double *out);
int ko_cdf(ko_digest *d, double x, double *out);
int ko_merge_into(ko_digest *self, ko_digest *other);
int ko_quantiles(ko_digest *d, const double *qs, int64_t n, double *out);
int ko_cdfs(ko_digest *d, const double *xs, int64_t n, double *out);
int ko_add_at_many(ko_digest *d, const double *values, const double *times,
                   const double *weights, int64_t n);
double ko_total_weight(const ko_digest *d);

/* Serde v1 (tdigest.mojo:627-698).  ko_to_bytes returns the size; writes when
 * out != NULL and cap is large enough. */
int64_t ko_to_bytes(ko_digest *d, uint8_t *out, int64_t cap);
int ko_from_bytes(const uint8_t *data, int64_t n, int scale, ko_digest **out);
ko_digest *ko_from_centroids(int scale, double compression, const double *means,
                             const double *weights, int64_t n, double min_v,
                             double max_v);

/* State access for tests. */
int64_t ko_ncentroids(const ko_digest *d);
int64_t ko_nbuffer(const ko_digest *d);
void ko_get_centroids(const ko_digest *d, double *means, double *weights);
void ko_get_state(const ko_digest *d, double *out7); /* stored_total, log_scale, clock, min, max, compression, lambda */
void ko_set_state(ko_digest *d, const double *means, const double *weights, int64_t nc,
                  const double *bmeans, const double *bweights, int64_t nbuf,
                  const double *state5); /* stored_total, log_scale, clock, min, max */

/* The bare greedy sweep over an already mean-sorted record list
 * (tdigest.mojo:419-435): returns the number of output centroids. */
int64_t ko_sweep(int scale, double nrm, double stored_n, const double *means,
                 const double *weights, int64_t n, double *out_means,
                 double *out_weights);
/* Stable sort by mean (same permutation as tdigest.mojo:64-77). */
void ko_stable_sort(ko_centroid *cs, int64_t n);

/* Grouped per-key B semantics (grouped.mojo:60-162 with a GPU present):
 * keys are dense ids < num_keys.  Validates everything before mutating. */
int ko_grouped_add_batch(ko_digest **digests, int64_t num_keys, const uint32_t *keys,
                         const double *values, int64_t n, int threads);

/* add_at of samples [off[d], off[d+1]) into ds[d], digests over `threads`
 * host threads. */
int ko_batch_add_at(ko_digest **ds, int64_t nd, const int64_t *off, const double *v,
                    const double *t, const double *w, int threads);
/* Ingest request r (data[off[r], off[r+1])) as one B batch merge of
 * ds[keys[r]], in arrival order per key; keys over `threads` host threads. */
int ko_replay_ingests(ko_digest **ds, int64_t nreq, const int64_t *keys, const int64_t *off,
                      const double *data, int threads);

/* L sequential additions W = W + x (one rounding each). */
double ko_add_loop_sum(double W, double x, int64_t L);

/* Scalar math exposed for tests (the same kd_math.h the device uses). */
double ko_k(int scale, double q, double nrm);
double ko_normalizer(int scale, double compression, double n);
double ko_log(double x);
double ko_asin(double x);
double ko_exp(double x);

/* Exact rank error: |#(x <= v)/n - q| with x sorted ascending. */
double ko_rank_error(const double *sorted, int64_t n, double v, double q);

#ifdef __cplusplus
}
#endif
#endif
