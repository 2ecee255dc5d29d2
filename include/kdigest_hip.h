/*
 * kdigest_hip.h — C ABI of the MI355X (gfx950) digest path.
 *
 * Plain C types only (no torch, no C++ in signatures).  Every function returns
 * KD_OK (0) or a negative KD_E* code; kd_last_error() gives a thread-local
 * message.  Validation happens before any state changes, as in the reference
 * (kdigest/tdigest.mojo:252-254, kdigest/grouped.mojo:91-97).
 *
 * Scale ids: 0..3 = ScaleK0..ScaleK3 (kdigest/scale.mojo:78-200), a runtime
 * int in place of the reference's compile-time type parameter.
 *
 * Citations are of the reference interface each entry point replaces (paths
 * relative to the reference repository root).
 */
#ifndef KDIGEST_HIP_H
#define KDIGEST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KD_OK 0
#define KD_EINVAL -1   /* invalid value, weight, key, scale or argument */
#define KD_EEMPTY -2   /* query on an empty digest */
#define KD_ERANGE -3   /* quantile q outside [0, 1] */
#define KD_ESERDE -4   /* unrecognized serialization version */
#define KD_ETRUNC -5   /* truncated serialized digest */
#define KD_EDEVICE -6  /* HIP runtime / device failure */
#define KD_ENOMEM -7   /* out of device or host memory */

typedef struct kd_session kd_session;
typedef struct kd_tdigest kd_tdigest;
typedef struct kd_grouped kd_grouped;
typedef struct kd_batch kd_batch;
typedef struct kd_server kd_server;
typedef struct kd_keymap kd_keymap;

/* Thread-local description of the last error on this thread. */
const char *kd_last_error(void);

/* ---------------------------------------------------------- session ----- */
/* GpuSession (kdigest/gpu/session.mojo:19-27): one device, one in-order HIP
 * stream and workspaces that grow geometrically and are reused across calls
 * (CompressWorkspace.reserve, kdigest/gpu/add_batch.mojo:400-404). */
int kd_session_create(int device, kd_session **out);
void kd_session_destroy(kd_session *s);
int kd_session_sync(kd_session *s);
/* Per-phase device timing (HIP events on the session stream around each
 * kernel group).  enable != 0 turns it on and clears the totals. */
int kd_session_profile(kd_session *s, int enable);
/* Writes {"phase": [ms_total, launches, algorithmic_bytes], ...} as JSON;
 * returns the length (writes only when cap is large enough). */
int64_t kd_session_phase_stats(kd_session *s, char *out, int64_t cap);
/* Compress strategy for digests receiving more than 4096 new values in one
 * merge: enable != 0 (the default) compresses them by selection (bucket
 * counts, rank-space sweep, exact values only at cluster boundaries) and
 * falls back to a full sort where a boundary may tie; 0 always sorts.  Both
 * give the same clusters.  No counterpart in the reference (its GPU compress
 * always sorts, kdigest/gpu/add_batch.mojo:629-663). */
int kd_session_set_select(kd_session *s, int enable);
/* Digests of at most 4096 merged records are compressed in one block each
 * (one launch for all of them) unless enable == 0, which routes them through
 * the multi-kernel sort path.  Same clusters either way.  No counterpart in
 * the reference. */
int kd_session_set_fused(kd_session *s, int enable);
/* Launch shape of the grouped key partition (radix by key): items per thread
 * of a scatter sub-tile (4, 8 or 16) and the target number of chunks per
 * pass (1..2^20; one workgroup per chunk).  Results do not depend on it.  No
 * counterpart in the reference (it groups keys on the host,
 * kdigest/gpu/grouped.mojo:454-538). */
int kd_session_set_partition_shape(kd_session *s, int items, int64_t chunks);
/* Tiles of the selection compress (digests receiving more than 4096 new
 * values): each tile holds at least tile_min values (1024..2^24) and a digest
 * is split into at most tiles_max tiles (1..4096).  Results do not depend on
 * it.  No counterpart in the reference. */
int kd_session_set_select_shape(kd_session *s, int64_t tile_min, int64_t tiles_max);
/* Small digests (at most 4096 merged records): 0 (the default) compressed on
 * the session stream after the selection samples of large ones; 1 on a second
 * stream once the selection count is done, overlapping the rest of the
 * selection compress; 2 on a second stream from the start of the compress.
 * Other values: KD_EINVAL.  Same results every way.  No counterpart in the
 * reference. */
int kd_session_set_overlap(kd_session *s, int enable);
/* Grouped key partition of batches of 2^22 pairs or more (and keys below
 * 2^17): keys estimated from a sample to hold at least min_pairs of the
 * batch's pairs (at most 255 of them) are placed by the first radix pass
 * alone; 0 sends every key through both passes.  Same results either way.
 * No counterpart in the reference. */
int kd_session_set_heavy_keys(kd_session *s, int64_t min_pairs);
/* Grouped batches over at most 2^bits keys (bits 0..27, default 27) are
 * grouped by the radix partition (one to three passes, stable: each key's
 * values keep their batch order); more keys by the counting partition, whose
 * order within a key follows its atomics (the selection compress may then
 * differ in the last bit between runs).  No counterpart in the reference. */
int kd_session_set_radix_key_bits(kd_session *s, int bits);
/* Autotuner of the grouped ingest: times ingests of the device-resident batch
 * (d_keys, d_values, n pairs, keys below num_keys) into fresh digests of the
 * given scale and compression under candidate launch shapes (coordinate
 * descent over partition items and chunks, selection tile_min and tiles_max,
 * overlap; best of `reps` runs each) and leaves the fastest on the session.
 * out_shape5 = {items, chunks, tile_min, tiles_max, overlap}; out_ms2 = {ms
 * with the shape the session had, ms with the chosen one}.  Digests built
 * with any shape are identical.  No counterpart in the reference (its block
 * size and per-key path are fixed, kdigest/gpu/grouped.mojo:579-624). */
int kd_session_autotune_grouped(kd_session *s, int scale, int64_t num_keys, double compression,
                                const uint32_t *d_keys, const double *d_values, int64_t n,
                                int reps, int64_t *out_shape5, double *out_ms2);
/* The same with the batch in host memory (copied to the device once). */
int kd_session_autotune_grouped_host(kd_session *s, int scale, int64_t num_keys,
                                     double compression, const uint32_t *keys,
                                     const double *values, int64_t n, int reps,
                                     int64_t *out_shape5, double *out_ms2);
/* Every shape the autotuner chooses is recorded for the process, keyed by the
 * batch's class (device, scale, compression, and the powers of two at or
 * above the key count and the pair count); a later grouped ingest of that
 * class, in any session, runs with it unless the session's shape was set
 * explicitly (kd_session_set_partition_shape / _select_shape / _overlap).
 * kd_autotune_save writes the table to a text file, kd_autotune_load merges
 * one in (*loaded = entries read).  No counterpart in the reference. */
int kd_autotune_save(const char *path);
/* The session's current grouped launch shape: {items, chunks, tile_min,
 * tiles_max, overlap}. */
int kd_session_shape(kd_session *s, int64_t *out5);
int kd_autotune_load(const char *path, int64_t *loaded);
/* Digests compressed so far by selection (out3[0]), by the multi-kernel sort
 * path (out3[1]) and in one block each (out3[2]). */
int kd_session_path_counts(kd_session *s, int64_t *out3);

/* ------------------------------------------------ operator entry points -- */
/* gpu_compress_batch (kdigest/gpu/add_batch.mojo:629-663) as an FP64
 * greedy-sweep compress.  means[0..n_prior+n): the first n_prior records carry
 * weights[i] (prior centroids ++ buffer, in that order), the next n carry
 * w_uniform.  Result: the centroids of the reference CPU merge pass
 * (kdigest/tdigest.mojo:391-439) with q ratios over stored_n and the given
 * normalizer nrm.  Host pointers; out arrays hold n_prior + n entries. */
int kd_compress(kd_session *s, int scale, double nrm, double stored_n, const double *means,
                const double *weights, double w_uniform, int64_t n_prior, int64_t n,
                double *out_means, double *out_weights, int64_t *out_count);

/* gpu_quantile / gpu_cdf (kdigest/gpu/query.mojo:320-342), batched over many
 * digests.  Digest d has centroids [off[d], off[d+1]) of means/weights and
 * state[4*d .. 4*d+4) = {stored_total, min, max, eff_factor}, where
 * eff_factor = e^(-log_scale) with decay and 1 without (the singleton test of
 * tdigest.mojo:475-483).  Query i reads digest q_digest[i].  Host pointers. */
int kd_quantile(kd_session *s, int64_t num_digests, const int64_t *off, const double *means,
                const double *weights, const double *state, const int64_t *q_digest,
                const double *qs, int64_t nq, double *out);
int kd_cdf(kd_session *s, int64_t num_digests, const int64_t *off, const double *means,
           const double *weights, const double *state, const int64_t *q_digest,
           const double *xs, int64_t nq, double *out);

/* Device-level FP64 scale evaluation, the analogue of gpu_scale_eval
 * (kdigest/gpu/scale_device.mojo:53-98): out_k[i] = k(q[i], nrm).  Host
 * pointers; used to pin device math against the host. */
int kd_scale_eval(kd_session *s, int scale, double nrm, const double *q, int64_t n,
                  double *out_k);
/* Device log / asin / sqrt of kd_math.h over an array (fn 0 / 1 / 2; host
 * pointers). */
int kd_math_eval(kd_session *s, int fn, const double *x, int64_t n, double *out);

/* Host-side scale functions (kdigest/scale.mojo:78-192), the same code the
 * device evaluates: normalizer(compression, n), k(q, nrm), q(k, nrm). */
double kd_scale_normalizer(int scale, double compression, double n);
double kd_scale_k(int scale, double q, double nrm);
double kd_scale_q(int scale, double k, double nrm);
/* The running weight of a cluster over L equal weights, as the reference
 * sums it (centroid.mojo:34-36: L sequential W = W + x), in O(log) steps:
 * the host side of the helper the device sweeps use to check their
 * boundaries in the reference's summation order (tdigest.mojo:419-435). */
double kd_add_repeated(double W, double x, int64_t L);
/* The running sums S = S + x[i], i = 0..m-1 (x[i] >= 0), as the reference's
 * per-sample loops form them (the decay exponent, tdigest.mojo:324-332; the
 * stored total, :334-347), by the binade rounds the batched replay runs in
 * parallel on the device: bit-identical to the loop.  With stop_on the first
 * nonzero x[i] whose sum exceeds thr is not consumed (its sum goes to
 * *stop_val).  Returns the number consumed; *S is the sum after them and
 * out[i] (if out is given) the sum after x[i].  Host only, for tests. */
int64_t kd_add_sequence(double *S, const double *x, int64_t m, int stop_on, double thr,
                        double *out, double *stop_val);

/* ------------------------------------------------ TDigest (one digest) -- */
/* TDigest[S] (kdigest/tdigest.mojo:97-726).  decay_lambda is Decay.rate_lambda
 * (kdigest/decay.mojo:23-51); 0 disables decay. */
int kd_tdigest_create(kd_session *s, int scale, double compression, double decay_lambda,
                      kd_tdigest **out);
void kd_tdigest_destroy(kd_tdigest *d);
int kd_tdigest_add(kd_tdigest *d, double value, double weight);              /* :153-161 */
int kd_tdigest_add_at(kd_tdigest *d, double value, double timestamp, double weight); /* :163-175 */
/* add_batch (:177-221): below the crossover (60000, or 40000 when
 * with_session) the per-sample path, otherwise one batch merge (:223-315). */
int kd_tdigest_add_batch(kd_tdigest *d, const double *values, int64_t n, int with_session);
/* One batch merge of device-resident values (B semantics, :223-315). */
int kd_tdigest_add_batch_device(kd_tdigest *d, const double *d_values, int64_t n);
/* n device-resident values as consecutive batch merges of `batch` values
 * each (the last one shorter): the same digest as kd_tdigest_add_batch_device
 * called on each slice in order (B semantics, :223-315, per slice), except
 * that every value is validated before any slice is merged (a non-finite
 * value anywhere leaves the digest unchanged).  The work that does not
 * depend on the merges before it (validation, each slice's splitter sample)
 * runs for all slices up front.  No counterpart in the reference, whose
 * callers loop over add_batch. */
int kd_tdigest_add_batches_device(kd_tdigest *d, const double *d_values, int64_t n,
                                  int64_t batch);
int kd_tdigest_merge(kd_tdigest *d);                                          /* :391-439 */
int kd_tdigest_quantile(kd_tdigest *d, const double *qs, int64_t nq, double *out); /* :485-538 */
int kd_tdigest_cdf(kd_tdigest *d, const double *xs, int64_t nq, double *out);      /* :540-587 */
int kd_tdigest_merge_into(kd_tdigest *self, kd_tdigest *other);               /* :589-623 */
/* to_bytes (:627-652): returns the size; writes when cap is large enough. */
int64_t kd_tdigest_to_bytes(kd_tdigest *d, uint8_t *out, int64_t cap);
int kd_tdigest_from_bytes(kd_session *s, const uint8_t *data, int64_t n, int scale,
                          kd_tdigest **out);                                   /* :654-698 */
int kd_tdigest_from_centroids(kd_session *s, int scale, double compression, const double *means,
                              const double *weights, int64_t n, double min_v, double max_v,
                              kd_tdigest **out);                               /* :700-726 */
/* __len__ (:460-463): number of merged centroids. */
int64_t kd_tdigest_size(kd_tdigest *d);
/* Centroids (means, weights) of the merged digest; *count gets the number. */
int kd_tdigest_centroids(kd_tdigest *d, double *means, double *weights, int64_t cap,
                         int64_t *count);
/* {stored_total, log_scale, clock, min, max, compression, lambda,
 *  total_weight (:465-467), buffered samples}. */
int kd_tdigest_state(kd_tdigest *d, double *out9);

/* ------------------------------------------ GroupedTDigest (per key) ---- */
/* GroupedTDigest[S] (kdigest/grouped.mojo:30-190): num_keys digests with
 * dense ids (1 <= num_keys <= 2^31 - 1, the reference's Int32 id range), state
 * resident in device memory. */
int kd_grouped_create(kd_session *s, int scale, int64_t num_keys, double compression,
                      double decay_lambda, kd_grouped **out);
void kd_grouped_destroy(kd_grouped *g);
/* add_batch (grouped.mojo:60-162): keys uint32 < num_keys (the reference's
 * Int32 dense ids), host arrays; at most 2^32 - 2^16 pairs per call. */
int kd_grouped_add_batch(kd_grouped *g, const uint32_t *keys, const double *values, int64_t n);
/* Same with device-resident keys and values. */
int kd_grouped_add_batch_device(kd_grouped *g, const uint32_t *d_keys, const double *d_values,
                                int64_t n);
/* quantile / cdf (grouped.mojo:164-174), vectorized: query i on key keys[i]. */
int kd_grouped_quantile(kd_grouped *g, const int64_t *keys, const double *qs, int64_t nq,
                        double *out);
int kd_grouped_cdf(kd_grouped *g, const int64_t *keys, const double *xs, int64_t nq,
                   double *out);
/* Device-resident query arrays (keys int64, values, output).  Every query is
 * checked on the device: a key outside [0, num_keys) returns KD_EINVAL, a q
 * outside [0, 1] KD_ERANGE, an empty key KD_EEMPTY (those outputs are NaN, the
 * others are written).  Synchronous: returns after the results are in d_out. */
int kd_grouped_quantile_device(kd_grouped *g, const int64_t *d_keys, const double *d_qs,
                               int64_t nq, double *d_out);
/* No counterpart in the reference.  kd_grouped_quantile_device for callers
 * that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid it): same arguments, checks and results, and
 * no allocation of any kind.  Its scratch (every key's query prefix and
 * state, the error flag) is a fixed 64 MiB array per GPU that the library
 * holds until the process ends, so a digest may have at most
 * 256 + 32 * num_keys + 8 * centroids <= 64 MiB bytes of it (else KD_EINVAL,
 * nothing launched).  All calls on one GPU share that array: calls that
 * overlap on one GPU (several host threads or streams at once) race. */
int kd_grouped_quantile_device_noalloc(kd_grouped *g, const int64_t *d_keys, const double *d_qs,
                                       int64_t nq, double *d_out);
int64_t kd_grouped_to_bytes(kd_grouped *g, int64_t key, uint8_t *out, int64_t cap); /* :176-180 */
int kd_grouped_centroid_count(kd_grouped *g, int64_t key, int64_t *out);            /* :182-185 */
int kd_grouped_total_weight(kd_grouped *g, int64_t key, double *out);               /* :187-190 */
int kd_grouped_centroids(kd_grouped *g, int64_t key, double *means, double *weights,
                         int64_t cap, int64_t *count);
/* Per-key scalar state {stored_total, log_scale, clock, min, max}. */
int kd_grouped_key_state(kd_grouped *g, int64_t key, double *out5);
/* Wire format of every key at once: key k's digest in the per-digest layout
 * (to_bytes, tdigest.mojo:627-652; byte-identical to kd_grouped_to_bytes for
 * key k) occupies bytes [off[k], off[k+1]) of one buffer, keys in order.
 * kd_grouped_wire_size gives the total (off[num_keys]).  The restore builds a
 * grouped digest of num_keys keys from such blobs: every blob is checked
 * first (KD_ETRUNC for a short one, KD_ESERDE for a bad version byte or
 * count, or for a compression or decay rate that differs from key 0's) and
 * nothing is created on an error.  The reference serializes one digest at a
 * time and has no grouped restore (tdigest.mojo:627-698, grouped.mojo:
 * 176-180): these batch it for checkpoint and resume of device-resident
 * state.  _device: buffers and offsets in device memory; _all: host arrays. */
int kd_grouped_wire_size(kd_grouped *g, int64_t *out_bytes);
/* Compression and decay rate of a grouped digest (as created or restored). */
int kd_grouped_params(kd_grouped *g, double *compression, double *lambda);
int kd_grouped_to_bytes_device(kd_grouped *g, uint8_t *d_out, int64_t cap, int64_t *d_off);
int kd_grouped_to_bytes_all(kd_grouped *g, uint8_t *out, int64_t cap, int64_t *off);
int kd_grouped_from_bytes_device(kd_session *s, int scale, int64_t num_keys,
                                 const uint8_t *d_data, int64_t n, const int64_t *d_off,
                                 kd_grouped **out);
int kd_grouped_from_bytes_all(kd_session *s, int scale, int64_t num_keys, const uint8_t *data,
                              int64_t n, const int64_t *off, kd_grouped **out);

/* ------------------------------------------- batched per-sample digests -- */
/* num_digests independent TDigest[S] fed sample by sample with add_at
 * (kdigest/tdigest.mojo:163-175, 324-350) and reduced by merge_into
 * (tdigest.mojo:589-623); state resident in device memory, one workgroup per
 * digest.  Capacity: 5 * compression <= 1536 buffered samples and at most 1536
 * centroids per digest. */
int kd_batch_create(kd_session *s, int scale, int64_t num_digests, double compression,
                    double decay_lambda, kd_batch **out);
void kd_batch_destroy(kd_batch *b);
/* add_at(values[i], times[i], weights[i]) for i in [off[d], off[d+1]) of each
 * digest d, in order.  Every sample is validated before any digest changes.
 * Device pointers. */
int kd_batch_add_at_device(kd_batch *b, const int64_t *d_off, const double *d_values,
                           const double *d_times, const double *d_weights);
/* Same with host arrays. */
int kd_batch_add_at(kd_batch *b, const int64_t *off, const double *values, const double *times,
                    const double *weights);
/* merge() of every digest (tdigest.mojo:391-439). */
int kd_batch_flush(kd_batch *b);
/* merge() of every digest, then level by level digest i*2^(l+1) absorbs
 * digest i*2^(l+1) + 2^l with merge_into; the result is digest 0. */
int kd_batch_merge_tree(kd_batch *b);
/* Whether kd_batch_merge_tree runs as a HIP graph (enable != 0, the default:
 * the flush and every level are captured on the first call and the graph is
 * replayed afterwards) or as eager launches.  Same results either way.  No
 * counterpart in the reference. */
int kd_batch_set_graph(kd_batch *b, int enable);
/* {stored_total, log_scale, clock, min, max, buffered samples, error flag}. */
int kd_batch_state(kd_batch *b, int64_t d, double *out7);
int kd_batch_centroids(kd_batch *b, int64_t d, double *means, double *weights, int64_t cap,
                       int64_t *count);
/* A TDigest holding a copy of digest d (centroids, buffer and state). */
int kd_batch_to_tdigest(kd_batch *b, int64_t d, kd_tdigest **out);
/* Compression and decay rate of a batch (as created or loaded). */
int kd_batch_params(kd_batch *b, double *compression, double *lambda);
/* Map-reduce finalize (test_edge.mojo:183-215, query.mojo:345-369): a batch
 * of num_digests digests loaded from their serialized bytes (the per-digest
 * layout of to_bytes, tdigest.mojo:627-652; digest d at bytes [off[d],
 * off[d+1]) of one buffer): centroids and state, empty buffers, as from_bytes
 * gives them (tdigest.mojo:656-698).  kd_batch_merge_tree then reduces them
 * into digest 0.  Every blob is checked before the batch is created
 * (KD_ETRUNC for a short one, KD_ESERDE for a bad version byte or count or a
 * compression or decay rate other than digest 0's, KD_EINVAL for more
 * centroids than a batched digest holds or a compression outside the
 * batched range).  _device: device buffers; _all: host arrays. */
int kd_batch_from_bytes_device(kd_session *s, int scale, int64_t num_digests,
                               const uint8_t *d_data, int64_t n, const int64_t *d_off,
                               kd_batch **out);
int kd_batch_from_bytes_all(kd_session *s, int scale, int64_t num_digests, const uint8_t *data,
                            int64_t n, const int64_t *off, kd_batch **out);

/* --------------------------------------------------------------- key map -- */
/* Dense ids for sparse uint32 keys, held on the device, for grouped digests
 * over keys that are not 0..num_keys-1 (the reference takes dense ids,
 * grouped.mojo:85-104, and assembles its per-key blocks on the host,
 * gpu/grouped.mojo:454-538).  At most capacity keys; a key gets its id the
 * first time it is mapped with insert != 0, ids 0, 1, 2, ... in order of
 * first occurrence within that batch.  A batch that would exceed the
 * capacity fails with KD_ENOMEM and leaves the map unchanged.  Without
 * insert, an unknown key maps to 0xffffffff. */
int kd_keymap_create(kd_session *s, int64_t capacity, kd_keymap **out);
void kd_keymap_destroy(kd_keymap *m);
int kd_keymap_map_device(kd_keymap *m, const uint32_t *d_keys, int64_t n, uint32_t *d_ids,
                         int insert);
int kd_keymap_map(kd_keymap *m, const uint32_t *keys, int64_t n, uint32_t *ids, int insert);
int kd_keymap_size(kd_keymap *m, int64_t *out);
/* The key of every id 0..size-1 (cap >= size). */
int kd_keymap_keys(kd_keymap *m, uint32_t *out, int64_t cap);

/* --------------------------------------------------------------- serving -- */
/* Continuous-batching server over num_keys digests resident in device
 * memory (GroupedTDigest semantics per key, grouped.mojo:60-190).  Each key
 * holds at most slot_capacity centroids; requests are admitted in arrival
 * order while the device memory their batch occupies fits mem_budget bytes
 * (a query 28 B per value: key, x, result; an ingest 8 B per value staged
 * plus its merge's workspace, 24 B per record over its values and the key's
 * slot_capacity prior centroids; one request larger than the budget runs
 * alone).  kind: 0 ingest batch (one B
 * merge, tdigest.mojo:223-315), 1 quantile batch, 2 cdf batch
 * (tdigest.mojo:485-587).  Requests of one key run in arrival order. */
int kd_server_create(kd_session *s, int scale, int64_t num_keys, double compression,
                     double decay_lambda, int64_t slot_capacity, int64_t mem_budget,
                     kd_server **out);
void kd_server_destroy(kd_server *v);
/* Copies the request data; validates it before queueing. */
int kd_server_submit(kd_server *v, int kind, int64_t key, const double *data, int64_t n,
                     int64_t *id);
/* n requests at once: request r has kind kinds[r], key keys[r] and data
 * data[off[r] .. off[r+1]); all are validated before any is queued. */
int kd_server_submit_many(kd_server *v, int64_t n, const int *kinds, const int64_t *keys,
                          const int64_t *off, const double *data, int64_t *ids);
/* Ids of the requests completed since the last call (at most cap) and the
 * server clock (kd_server_clock) when each completed. */
int kd_server_take_completed(kd_server *v, int64_t *ids, double *t, int64_t cap, int64_t *count);
/* Seconds since the server was created (steady clock). */
int kd_server_clock(kd_server *v, double *now);
/* Admits and runs one continuous batch: all admitted queries in one launch,
 * then up to max_rounds rounds of ingests (one per key per round). */
int kd_server_step(kd_server *v, int max_rounds, int64_t *completed);
/* Worker mode: two host threads serve requests as they are submitted, so a
 * request that arrives while others are on the device is admitted without
 * waiting for them.  Queries whose key has no earlier ingest outstanding run
 * in batched launches on a stream of their own, beside the ingest rounds
 * (up to max_rounds per key per batch) on the session's stream; each key
 * still sees its requests in submission order, so every result equals the
 * one kd_server_step gives.  Completions are read with
 * kd_server_take_completed / kd_server_result as before; kd_server_step,
 * kd_server_key_state and kd_server_key_centroids fail (KD_EINVAL) while the
 * workers run, and the session must not be used for other work meanwhile.
 * kd_server_stop waits until every submitted request completed, then joins
 * the workers (a device error in a worker fails its requests with
 * KD_EDEVICE and is returned by kd_server_stop).  kd_server_destroy stops
 * them too.  No counterpart in the reference, whose digests are
 * single-threaded objects (the serving loop is the caller's). */
int kd_server_start(kd_server *v, int max_rounds);
int kd_server_stop(kd_server *v);
/* *done = 1 when the request completed; query results are copied to out
 * (cap entries) and the request is then forgotten.  A query that found its
 * key empty completes with KD_EEMPTY (tdigest.mojo:485-487). */
int kd_server_result(kd_server *v, int64_t id, double *out, int64_t cap, int *done);
int kd_server_pending(kd_server *v, int64_t *out);
/* {steps, requests admitted, bytes admitted, largest step in bytes, ingest
 * rounds, requests still queued}. */
int kd_server_stats(kd_server *v, int64_t *out6);
/* {stored_total, log_scale, clock, min, max, centroids} of one key. */
int kd_server_key_state(kd_server *v, int64_t key, double *out6);
int kd_server_key_centroids(kd_server *v, int64_t key, double *means, double *weights,
                            int64_t cap, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
