// kd_api.hip — the C ABI (include/kdigest_hip.h): session, operator entry
// points, query kernels and the host layer of TDigest / GroupedTDigest.
//
// The host layer keeps only scalar digest state (clock, decay exponent,
// stored total, min, max) and the per-sample ingest buffer on the host, as the
// reference does (kdigest/tdigest.mojo:116-151).  Centroids live in device
// memory; every merge, batch compress and query runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kdigest_hip.h"
#include "kd_batch.h"
#include "kd_engine.h"
#include "kd_exact.h"
#include "kd_math.h"
#include "kd_query.h"
#include "kd_serve.h"

static_assert(kd::KD_EEMPTY_SERVE == KD_EEMPTY && kd::KD_ENOMEM_SERVE == KD_ENOMEM,
              "server failure statuses are the C ABI's codes");

using kd::DevBuf;
using kd::SegDesc;
using kd::SegResult;
using kd::launched;
using kd::QState;
using kd::cdf_one;
using kd::quantile_one;

namespace {

constexpr double kRenormThreshold = 400.0;  // tdigest.mojo:43
constexpr double kPruneEpsilon = 1e-6;      // tdigest.mojo:47
constexpr int64_t kCrossover = 60000;       // tdigest.mojo:55
constexpr int64_t kCrossoverSession = 40000;// tdigest.mojo:58
constexpr uint8_t kSerdeVersion = 1;        // tdigest.mojo:61

thread_local std::string g_err;

struct KdError : std::runtime_error {
  int code;
  KdError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <typename F>
int guarded(F&& f) {
  // hipGetLastError is sticky per thread: drop whatever another library in
  // the process (or a caller's own HIP use) left behind, so each launch check
  // below reports only our own launches.
  (void)hipGetLastError();
  try {
    return f();
  } catch (const KdError& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(KD_ENOMEM, "out of memory");
  } catch (const std::exception& e) {
    return fail(KD_EDEVICE, e.what());
  }
}

void hip_ok(hipError_t e, const char* what) { kd::check_hip(e, what); }

using clk = std::chrono::steady_clock;
// Milliseconds since t; resets t (consecutive host phases).
// Host threads for per-key loops (the box's CPU share; at most 16).
int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(n, 16u));
}

// fn(begin, end, thread) over [0, n) split into nthr contiguous chunks.
template <class Fn>
void parallel_for(int64_t n, int nthr, Fn fn) {
  if (nthr <= 1 || n < 2) {
    fn(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthr - 1) / nthr;
  for (int t = 0; t < nthr; t++) {
    int64_t a = std::min(n, t * chunk), b = std::min(n, a + chunk);
    th.emplace_back([=, &fn] { fn(a, b, t); });
  }
  for (auto& x : th) x.join();
}

double ms_since(clk::time_point& t) {
  clk::time_point now = clk::now();
  double ms = std::chrono::duration<double, std::milli>(now - t).count();
  t = now;
  return ms;
}

bool is_finite_val(double v) { return std::isfinite(v); }

// ------------------------------------------------------------ kernels
__global__ void k_scale_eval(int scale, double nrm, const double* q, int64_t n, double* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = kd_k(scale, q[i], nrm);
}

__global__ void k_math_eval(int fn, const double* x, int64_t n, double* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fn == 0 ? kd_log(x[i]) : (fn == 1 ? kd_asin(x[i]) : kd_sqrt(x[i]));
}

// Interior walk prefix of quantile/cdf (tdigest.mojo:528-536, 575-585):
// S[0] = w0/2, S[i+1] = S[i] + (w[i] + w[i+1])/2, accumulated in order so the
// binary search below sees exactly the sums of the sequential walk.
__global__ void k_query_prefix(int64_t nd, const int64_t* off, const double* w, double* S) {
  int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  int64_t a = off[d], b = off[d + 1];
  if (b <= a) return;
  double acc = w[a] / 2.0;
  S[a] = acc;
  for (int64_t i = a; i + 1 < b; i++) {
    acc += (w[i] + w[i + 1]) / 2.0;
    S[i + 1] = acc;
  }
}

__global__ void k_query(int is_cdf, const int64_t* off, const double* m, const double* w,
                        const double* S, const double* state, const int64_t* qd,
                        const double* qs, int64_t nq, double* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int64_t d = qd[i];
  int64_t a = off[d], nc = off[d + 1] - a;
  QState st{state[4 * d], state[4 * d + 1], state[4 * d + 2], state[4 * d + 3]};
  out[i] = is_cdf ? cdf_one(m + a, w + a, S + a, nc, st, qs[i])
                  : quantile_one(m + a, w + a, S + a, nc, st, qs[i]);
}

// Grouped queries: the state arrays are separate per-key vectors.  A query
// whose key is out of range, whose quantile q is outside [0, 1] or whose key
// is empty gets NaN and sets its bit in *err (the reference raises for each,
// grouped.mojo:164-174 via tdigest.mojo:492-498, 545-546).
// Queries of many keys.  When every query of a block names the same key (a
// batch of quantiles per key, in key order), the block stages that key's
// centroids, prefix and state in LDS once and every search runs there; the
// arithmetic is the same either way.
constexpr int kQStage = 512;  // centroids a block stages

__global__ __launch_bounds__(256) void k_query_grouped(int is_cdf, int64_t nk, const int64_t* off,
                                                       const double* m, const double* w,
                                                       const double* S, const double* total,
                                                       const double* vmin, const double* vmax,
                                                       const double* eff, const int64_t* qk,
                                                       const double* qs, int64_t nq, double* out,
                                                       int* err) {
  __shared__ double lm[kQStage], lw[kQStage], lS[kQStage];
  __shared__ int64_t bkey;
  __shared__ int mixed;
  __shared__ QState bst;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t i = i0 + threadIdx.x;
  const bool live = i < nq;
  const int64_t d = live ? qk[i] : -1;
  const double q = live ? qs[i] : 0.0;  // issued before the key's data is needed
  if (threadIdx.x == 0) {
    bkey = qk[i0];
    mixed = 0;
  }
  __syncthreads();
  if (live && d != bkey) mixed = 1;
  __syncthreads();
  const int64_t kb = bkey;
  bool staged = false;
  int64_t a = 0, nc = 0;
  if (!mixed && kb >= 0 && kb < nk) {
    a = off[kb];
    nc = off[kb + 1] - a;
    if (nc > 0 && nc <= kQStage) {
      for (int64_t c = threadIdx.x; c < nc; c += blockDim.x) {
        lm[c] = m[a + c];
        lw[c] = w[a + c];
        lS[c] = S[a + c];
      }
      if (threadIdx.x == 0) bst = QState{total[kb], vmin[kb], vmax[kb], eff[kb]};
      staged = true;
    }
  }
  __syncthreads();
  if (!live) return;
  int bad = 0;
  if (!staged) {
    if (d < 0 || d >= nk) {
      bad = 1;
    } else {
      a = off[d];
      nc = off[d + 1] - a;
    }
  }
  if (!bad && !is_cdf && !(q >= 0.0 && q <= 1.0)) bad = 2;
  else if (!bad && nc == 0) bad = 4;
  if (bad) {
    out[i] = __builtin_nan("");
    if (err) atomicOr(err, bad);
    return;
  }
  const QState st = staged ? bst : QState{total[d], vmin[d], vmax[d], eff[d]};
  const double* pm = staged ? lm : m + a;
  const double* pw = staged ? lw : w + a;
  const double* pS = staged ? lS : S + a;
  out[i] = is_cdf ? cdf_one(pm, pw, pS, nc, st, q) : quantile_one(pm, pw, pS, nc, st, q);
}

// A digest's new centroids (means and weights) copied out of the engine's
// output in one launch.
__global__ void k_copy_pair(double* a, const double* sa, double* b, const double* sb, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a[i] = sa[i];
    b[i] = sb[i];
  }
}

__global__ void k_scale_weights(double* w, int64_t n, double factor) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] *= factor;
}

// Scale the centroid weights of a list of keys (grouped renormalize).
// Flags non-finite values and out-of-range keys (validate-then-commit).
__global__ void k_validate(const uint32_t* keys, const double* vals, int64_t n, int64_t nk,
                           int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = vals[i];
  bool bad = !(v - v == 0.0);  // NaN or inf
  if (keys && (int64_t)keys[i] >= nk) bad = true;
  if (bad) atomicOr(flag, 1);
}


unsigned blocks_for(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }
// Grid of 256-thread blocks for kernels that give each item one wave and walk
// the items with a wave stride: at most 2^20 blocks (2^22 waves), so the grid
// stays far inside the 32-bit launch limits for any item count.
// One-thread-per-pair launches stay below 2^32 work items per dimension.
constexpr int64_t kMaxGroupedBatch = ((int64_t)1 << 32) - ((int64_t)1 << 16);
unsigned wave_grid(int64_t items) {
  const int64_t b = (items + 3) / 4;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, (int64_t)1 << 20));
}

}  // namespace

// =================================================================== session
struct kd_session {
  kd::Engine eng;
  DevBuf q_off, q_m, q_w, q_s, q_state, q_qd, q_q, q_out, e_m, e_w, v_buf, tmp, tmp2, presort;
  // Grouped-ingest scratch shared by every grouped digest of the session.
  DevBuf g_counts, g_koff, g_vals, g_keys, g_valsin, g_segof, g_xoff;
  // Page-locked segment tables and results, reused by every compress call.
  kd::SegVec segv;
  kd::ResVec resv;
  // Set once the caller fixes a launch shape itself: the shapes the
  // autotuner recorded are then no longer applied to this session.
  bool shape_pinned = false;
  // Centroid total after the last grouped batch (keys, pairs) of the
  // session: the first guess of the next CSR's size for a batch of that
  // shape.
  int64_t csr_keys = -1, csr_pairs = -1, csr_total = 0;
  explicit kd_session(int dev) : eng(dev) {}
  hipStream_t st() { return eng.stream(); }
  template <typename T>
  T* buf(DevBuf& b, int64_t n) {
    return static_cast<T*>(b.reserve(std::max<int64_t>(n, 1) * sizeof(T)));
  }
  void h2d(void* d, const void* h, size_t bytes) {
    if (bytes) hip_ok(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st()), "H2D");
  }
  void d2h(void* h, const void* d, size_t bytes) {
    if (bytes) hip_ok(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st()), "D2H");
  }
  void d2d(void* d, const void* s, size_t bytes) {
    if (bytes) hip_ok(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, st()), "D2D");
  }
  void sync() { hip_ok(hipStreamSynchronize(st()), "stream sync"); }
};

extern "C" const char* kd_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------- allocation-free scratch
// Device memory of kd_grouped_quantile_device_noalloc: one fixed array per
// GPU, held until the process ends, shared by every call on that GPU (calls
// that overlap on one GPU race).  In host code `scratch` is only a stand-in
// name; the device address of each GPU's copy is resolved once, with that GPU
// current, when a session is created on it, and kept in g_scratch.
namespace {
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch[kScratchDevices] = {};

void resolve_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kScratchDevices)
    throw KdError(KD_EINVAL, "device id beyond the scratch table");
  if (g_scratch[device]) return;
  hip_ok(hipSetDevice(device), "hipSetDevice");
  void* p = nullptr;
  hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw KdError(KD_EDEVICE, "no device address for the scratch array");
  g_scratch[device] = p;
}

void* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kScratchDevices || !g_scratch[device])
    throw KdError(KD_EDEVICE, "scratch not set up on this device (create a session on it first)");
  return g_scratch[device];
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace

extern "C" int kd_session_create(int device, kd_session** out) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
      throw KdError(KD_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) throw KdError(KD_EINVAL, "device id out of range");
    std::unique_ptr<kd_session> s(new kd_session(device));
    resolve_scratch(device);
    *out = s.release();
    return KD_OK;
  });
}

extern "C" void kd_session_destroy(kd_session* s) { delete s; }

extern "C" int kd_session_sync(kd_session* s) {
  return guarded([&] {
    s->sync();
    return KD_OK;
  });
}

extern "C" int kd_session_profile(kd_session* s, int enable) {
  return guarded([&] {
    s->eng.set_profiling(enable != 0);
    return KD_OK;
  });
}

extern "C" int kd_session_set_select(kd_session* s, int enable) {
  return guarded([&] {
    s->eng.set_select(enable != 0);
    return KD_OK;
  });
}

extern "C" int kd_session_set_fused(kd_session* s, int enable) {
  return guarded([&] {
    s->eng.set_fused(enable != 0);
    return KD_OK;
  });
}

extern "C" int kd_session_set_partition_shape(kd_session* s, int items, int64_t chunks) {
  return guarded([&] {
    if (items != 4 && items != 8 && items != 16) throw KdError(KD_EINVAL, "items must be 4, 8 or 16");
    if (chunks < 1 || chunks > (1 << 20)) throw KdError(KD_EINVAL, "chunks must be in [1, 2^20]");
    s->eng.set_partition_shape(items, chunks);
    s->shape_pinned = true;
    return KD_OK;
  });
}

extern "C" int kd_session_set_select_shape(kd_session* s, int64_t tile_min, int64_t tiles_max) {
  return guarded([&] {
    if (tile_min < 1024 || tile_min > (1 << 24)) throw KdError(KD_EINVAL, "tile_min must be in [1024, 2^24]");
    if (tiles_max < 1 || tiles_max > 4096) throw KdError(KD_EINVAL, "tiles_max must be in [1, 4096]");
    s->eng.set_select_shape(tile_min, tiles_max);
    s->shape_pinned = true;
    return KD_OK;
  });
}

extern "C" int kd_session_set_overlap(kd_session* s, int enable) {
  return guarded([&] {
    if (enable < 0 || enable > 2) return KD_EINVAL;
    s->eng.set_overlap(enable);
    s->shape_pinned = true;
    return KD_OK;
  });
}

extern "C" int kd_session_set_heavy_keys(kd_session* s, int64_t min_pairs) {
  return guarded([&] {
    if (min_pairs < 0) throw KdError(KD_EINVAL, "min_pairs must be >= 0");
    s->eng.set_heavy_keys(min_pairs);
    return KD_OK;
  });
}

extern "C" int kd_session_set_radix_key_bits(kd_session* s, int bits) {
  return guarded([&] {
    if (bits < 0 || bits > 27) throw KdError(KD_EINVAL, "bits must be in [0, 27]");
    s->eng.set_radix_key_bits(bits);
    return KD_OK;
  });
}

extern "C" int kd_session_path_counts(kd_session* s, int64_t* out3) {
  for (int i = 0; i < 3; i++) out3[i] = s->eng.path_counts()[i];
  return KD_OK;
}

extern "C" int64_t kd_session_phase_stats(kd_session* s, char* out, int64_t cap) {
  std::string js = "{";
  try {
    bool first = true;
    for (const auto& kv : s->eng.phase_stats()) {
      char buf[160];
      snprintf(buf, sizeof buf, "%s\"%s\": [%.6f, %lld, %.1f]", first ? "" : ", ",
               kv.first.c_str(), kv.second.ms, (long long)kv.second.launches, kv.second.bytes);
      js += buf;
      first = false;
    }
  } catch (const std::exception& e) {
    return fail(KD_EDEVICE, e.what());
  }
  js += "}";
  int64_t need = (int64_t)js.size() + 1;
  if (out && cap >= need) memcpy(out, js.c_str(), need);
  return need;
}

// ======================================================= operator entries
extern "C" int kd_compress(kd_session* s, int scale, double nrm, double stored_n,
                           const double* means, const double* weights, double w_uniform,
                           int64_t n_prior, int64_t n, double* out_means, double* out_weights,
                           int64_t* out_count) {
  return guarded([&] {
    if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
    if (n_prior < 0 || n < 0) throw KdError(KD_EINVAL, "negative size");
    for (int64_t i = 0; i < n_prior + n; i++)
      if (!is_finite_val(means[i])) throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
    for (int64_t i = 0; i < n_prior; i++)
      if (!(weights[i] > 0.0) || !is_finite_val(weights[i]))
        throw KdError(KD_EINVAL, "weight must be positive and finite");
    *out_count = 0;
    if (n_prior + n == 0) return KD_OK;
    double* em = s->buf<double>(s->e_m, n_prior);
    double* ew = s->buf<double>(s->e_w, n_prior);
    double* v = s->buf<double>(s->v_buf, n);
    s->h2d(em, means, n_prior * sizeof(double));
    s->h2d(ew, weights, n_prior * sizeof(double));
    s->h2d(v, means + n_prior, n * sizeof(double));
    kd::SegVec& segs = s->segv;
    segs.resize(1);
    segs[0] = SegDesc{0, n, 0, n_prior, 0, 0, 0, w_uniform, stored_n, nrm, 0.0};
    kd::ResVec& res = s->resv;
    s->eng.compress(scale, segs, v, em, ew, res);
    int64_t c = res[0].count;
    s->d2h(out_means, s->eng.cent_means() + segs[0].x_off, c * sizeof(double));
    s->d2h(out_weights, s->eng.cent_weights() + segs[0].x_off, c * sizeof(double));
    s->sync();
    *out_count = c;
    return KD_OK;
  });
}

static int run_query(kd_session* s, int is_cdf, int64_t nd, const int64_t* off,
                     const double* means, const double* weights, const double* state,
                     const int64_t* qd, const double* qs, int64_t nq, double* out) {
  return guarded([&] {
    if (nq == 0) return KD_OK;
    int64_t nc = off[nd];
    for (int64_t i = 0; i < nq; i++) {
      if (qd[i] < 0 || qd[i] >= nd) throw KdError(KD_EINVAL, "query digest out of range");
      if (off[qd[i] + 1] == off[qd[i]])
        throw KdError(KD_EEMPTY, is_cdf ? "cdf of an empty digest is undefined"
                                        : "quantile of an empty digest is undefined");
      if (!is_cdf && (qs[i] < 0.0 || qs[i] > 1.0))
        throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
    }
    int64_t* doff = s->buf<int64_t>(s->q_off, nd + 1);
    double* dm = s->buf<double>(s->q_m, nc);
    double* dw = s->buf<double>(s->q_w, nc);
    double* dS = s->buf<double>(s->q_s, nc);
    double* dst = s->buf<double>(s->q_state, 4 * nd);
    int64_t* dqd = s->buf<int64_t>(s->q_qd, nq);
    double* dq = s->buf<double>(s->q_q, nq);
    double* dout = s->buf<double>(s->q_out, nq);
    s->h2d(doff, off, (nd + 1) * sizeof(int64_t));
    s->h2d(dm, means, nc * sizeof(double));
    s->h2d(dw, weights, nc * sizeof(double));
    s->h2d(dst, state, 4 * nd * sizeof(double));
    s->h2d(dqd, qd, nq * sizeof(int64_t));
    s->h2d(dq, qs, nq * sizeof(double));
    k_query_prefix<<<blocks_for(nd, 64), 64, 0, s->st()>>>(nd, doff, dw, dS);
    launched("k_query_prefix");
    k_query<<<blocks_for(nq), 256, 0, s->st()>>>(is_cdf, doff, dm, dw, dS, dst, dqd, dq, nq, dout);
    launched("k_query");
    hip_ok(hipGetLastError(), "query launch");
    s->d2h(out, dout, nq * sizeof(double));
    s->sync();
    return KD_OK;
  });
}

extern "C" int kd_quantile(kd_session* s, int64_t nd, const int64_t* off, const double* means,
                           const double* weights, const double* state, const int64_t* qd,
                           const double* qs, int64_t nq, double* out) {
  return run_query(s, 0, nd, off, means, weights, state, qd, qs, nq, out);
}

extern "C" int kd_cdf(kd_session* s, int64_t nd, const int64_t* off, const double* means,
                      const double* weights, const double* state, const int64_t* qd,
                      const double* xs, int64_t nq, double* out) {
  return run_query(s, 1, nd, off, means, weights, state, qd, xs, nq, out);
}

extern "C" int kd_scale_eval(kd_session* s, int scale, double nrm, const double* q, int64_t n,
                             double* out_k) {
  return guarded([&] {
    if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
    if (n == 0) return KD_OK;
    double* dq = s->buf<double>(s->tmp, n);
    double* dk = s->buf<double>(s->tmp2, n);
    s->h2d(dq, q, n * sizeof(double));
    k_scale_eval<<<blocks_for(n), 256, 0, s->st()>>>(scale, nrm, dq, n, dk);
    launched("k_scale_eval");
    hip_ok(hipGetLastError(), "scale launch");
    s->d2h(out_k, dk, n * sizeof(double));
    s->sync();
    return KD_OK;
  });
}

extern "C" int kd_math_eval(kd_session* s, int fn, const double* x, int64_t n, double* out) {
  return guarded([&] {
    if (fn < 0 || fn > 2) throw KdError(KD_EINVAL, "fn must be 0 (log), 1 (asin) or 2 (sqrt)");
    if (n == 0) return KD_OK;
    double* dx = s->buf<double>(s->tmp, n);
    double* dy = s->buf<double>(s->tmp2, n);
    s->h2d(dx, x, n * sizeof(double));
    k_math_eval<<<blocks_for(n), 256, 0, s->st()>>>(fn, dx, n, dy);
    launched("k_math_eval");
    hip_ok(hipGetLastError(), "math launch");
    s->d2h(out, dy, n * sizeof(double));
    s->sync();
    return KD_OK;
  });
}

// ================================================================ TDigest
struct kd_tdigest {
  kd_session* s;
  int scale;
  double compression, lambda;
  int64_t buffer_limit;
  double stored_total = 0.0, log_scale = 0.0, clock = 0.0, vmin = 0.0, vmax = 0.0;
  std::vector<double> bm, bw;  // ingest buffer (stored weights)
  DevBuf cm, cw;               // centroids (device)
  int64_t nc = 0;

  bool decay() const { return lambda > 0.0; }                       // decay.mojo:49-51
  double scale_factor() const { return decay() ? kd_exp(log_scale) : 1.0; }  // :352-357
  double effective_count() const {                                   // :373-377
    return decay() ? stored_total * kd_exp(-log_scale) : stored_total;
  }
  double* dm() { return static_cast<double*>(cm.p); }
  double* dw() { return static_cast<double*>(cw.p); }

  void renormalize() {  // tdigest.mojo:359-371
    double factor = kd_exp(-log_scale);
    if (nc > 0) {
      k_scale_weights<<<blocks_for(nc), 256, 0, s->st()>>>(dw(), nc, factor);
      launched("k_scale_weights");
      hip_ok(hipGetLastError(), "renormalize launch");
    }
    for (double& w : bw) w *= factor;
    stored_total *= factor;
    log_scale = 0.0;
  }

  void advance_clock(double new_clock) {  // tdigest.mojo:324-332
    if (decay() && new_clock > clock) {
      log_scale += lambda * (new_clock - clock);
      if (log_scale > kRenormThreshold) renormalize();
    }
    clock = new_clock;
  }

  // Compress E = centroids ++ buffer (++ device values with weight sf).
  SegResult run(const double* d_values, int64_t n, double sf) {
    int64_t ne = nc + (int64_t)bm.size();
    // The engine only reads its explicit records: with nothing buffered on
    // the host they are the digest's own centroids, read in place.
    const double* em = dm();
    const double* ew = dw();
    if (!bm.empty()) {
      double* tm = s->buf<double>(s->e_m, ne);
      double* tw = s->buf<double>(s->e_w, ne);
      s->d2d(tm, dm(), nc * sizeof(double));
      s->d2d(tw, dw(), nc * sizeof(double));
      s->h2d(tm + nc, bm.data(), bm.size() * sizeof(double));
      s->h2d(tw + nc, bw.data(), bw.size() * sizeof(double));
      em = tm;
      ew = tw;
    }
    bm.clear();
    bw.clear();
    double eff = effective_count();
    double nrm = kd_normalizer(scale, compression, eff);
    double floor = decay() ? kPruneEpsilon * kd_exp(log_scale) : 0.0;
    kd::SegVec& segs = s->segv;
    segs.resize(1);
    segs[0] = SegDesc{0, n, 0, ne, 0, 0, 0, sf, stored_total, nrm, floor};
    kd::ResVec& res = s->resv;
    s->eng.compress(scale, segs, d_values, em, ew, res);
    int64_t c = res[0].count;
    cm.reserve(std::max<int64_t>(c, 1) * sizeof(double));
    cw.reserve(std::max<int64_t>(c, 1) * sizeof(double));
    if (c > 0) {
      k_copy_pair<<<blocks_for(c), 256, 0, s->st()>>>(dm(), s->eng.cent_means() + segs[0].x_off, dw(),
                                                       s->eng.cent_weights() + segs[0].x_off, c);
      launched("k_copy_pair");
    }
    nc = c;
    return res[0];
  }

  void merge() {  // tdigest.mojo:391-439 (+ prune 438-439)
    if (bm.empty()) return;
    SegResult r = run(nullptr, 0, 1.0);
    if (decay() && r.removed > 0.0) stored_total -= r.removed;
  }

  void insert(double value, double weight) {  // tdigest.mojo:334-350
    if (stored_total == 0.0 && bm.empty()) {
      vmin = value;
      vmax = value;
    } else {
      if (value < vmin) vmin = value;
      if (value > vmax) vmax = value;
    }
    double scaled = weight * scale_factor();
    bm.push_back(value);
    bw.push_back(scaled);
    stored_total += scaled;
    if ((int64_t)bm.size() >= buffer_limit) merge();
  }

  static void validate(double value, double weight) {  // tdigest.mojo:317-322
    if (!is_finite_val(value)) throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
    if (weight <= 0.0 || !is_finite_val(weight)) throw KdError(KD_EINVAL, "weight must be positive and finite");
  }

  // B semantics (tdigest.mojo:223-315) of n values already in device memory.
  // vmin_in/vmax_in: the batch range when the caller knows it (host input),
  // otherwise taken from the sorted batch.
  void batch(const double* d_values, int64_t n, bool have_range, double bmin, double bmax) {
    if (n == 0) return;
    advance_clock(clock + (double)n);  // :258
    double sf = scale_factor();        // :259
    int64_t had_centroids = nc;
    stored_total += (double)n * sf;    // :282
    SegResult r = run(d_values, n, sf);
    if (!have_range) {
      bmin = r.vmin;
      bmax = r.vmax;
    }
    if (stored_total == (double)n * sf && had_centroids == 0) {  // :285-295
      vmin = bmin;
      vmax = bmax;
    } else {
      if (bmin < vmin) vmin = bmin;
      if (bmax > vmax) vmax = bmax;
    }
    stored_total = r.total;  // :304-315 (reconcile, then prune)
  }

  void query(int is_cdf, const double* qs, int64_t nq, double* out) {
    if (nq == 0) return;
    double* dq = s->buf<double>(s->q_q, nq);
    double* dout = s->buf<double>(s->q_out, nq);
    double* dS = s->buf<double>(s->q_s, nc);
    int64_t* doff = s->buf<int64_t>(s->q_off, 2);
    double* dst = s->buf<double>(s->q_state, 4);
    int64_t* dqd = s->buf<int64_t>(s->q_qd, nq);
    int64_t hoff[2] = {0, nc};
    double hst[4] = {stored_total, vmin, vmax, decay() ? kd_exp(-log_scale) : 1.0};
    s->h2d(doff, hoff, sizeof hoff);
    s->h2d(dst, hst, sizeof hst);
    s->h2d(dq, qs, nq * sizeof(double));
    hip_ok(hipMemsetAsync(dqd, 0, nq * sizeof(int64_t), s->st()), "memset");
    k_query_prefix<<<1, 64, 0, s->st()>>>(1, doff, dw(), dS);
    launched("k_query_prefix");
    k_query<<<blocks_for(nq), 256, 0, s->st()>>>(is_cdf, doff, dm(), dw(), dS, dst, dqd, dq, nq, dout);
    launched("k_query");
    hip_ok(hipGetLastError(), "query launch");
    s->d2h(out, dout, nq * sizeof(double));
    s->sync();
  }

  void download(std::vector<double>& m, std::vector<double>& w) {
    m.resize(nc);
    w.resize(nc);
    s->d2h(m.data(), dm(), nc * sizeof(double));
    s->d2h(w.data(), dw(), nc * sizeof(double));
    s->sync();
  }

  void upload(const double* m, const double* w, int64_t n) {
    cm.reserve(std::max<int64_t>(n, 1) * sizeof(double));
    cw.reserve(std::max<int64_t>(n, 1) * sizeof(double));
    s->h2d(dm(), m, n * sizeof(double));
    s->h2d(dw(), w, n * sizeof(double));
    s->sync();
    nc = n;
  }
};

static kd_tdigest* new_digest(kd_session* s, int scale, double compression, double lambda) {
  if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
  auto* d = new kd_tdigest();
  d->s = s;
  d->scale = scale;
  d->compression = compression;
  d->lambda = lambda;
  d->buffer_limit = (int64_t)compression * 5;  // tdigest.mojo:146
  return d;
}

extern "C" int kd_tdigest_create(kd_session* s, int scale, double compression, double lambda,
                                 kd_tdigest** out) {
  return guarded([&] {
    *out = new_digest(s, scale, compression, lambda);
    return KD_OK;
  });
}

extern "C" void kd_tdigest_destroy(kd_tdigest* d) { delete d; }

extern "C" int kd_tdigest_add(kd_tdigest* d, double value, double weight) {
  return guarded([&] {
    kd_tdigest::validate(value, weight);
    d->advance_clock(d->clock + 1.0);
    d->insert(value, weight);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_add_at(kd_tdigest* d, double value, double timestamp, double weight) {
  return guarded([&] {
    kd_tdigest::validate(value, weight);
    if (timestamp > d->clock) d->advance_clock(timestamp);
    d->insert(value, weight);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_add_batch(kd_tdigest* d, const double* values, int64_t n,
                                    int with_session) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    int64_t cross = with_session ? kCrossoverSession : kCrossover;
    if (n < cross) {  // tdigest.mojo:199-202: the per-sample path
      for (int64_t i = 0; i < n; i++) {
        kd_tdigest::validate(values[i], 1.0);
        d->advance_clock(d->clock + 1.0);
        d->insert(values[i], 1.0);
      }
      return KD_OK;
    }
    double bmin = values[0], bmax = values[0];
    for (int64_t i = 0; i < n; i++) {  // :253-254 validate everything first
      kd_tdigest::validate(values[i], 1.0);
      if (values[i] < bmin) bmin = values[i];
      if (values[i] > bmax) bmax = values[i];
    }
    double* v = d->s->buf<double>(d->s->v_buf, n);
    d->s->h2d(v, values, n * sizeof(double));
    d->batch(v, n, true, bmin, bmax);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_add_batch_device(kd_tdigest* d, const double* d_values, int64_t n) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    int* flag = d->s->buf<int>(d->s->tmp, 1);
    hip_ok(hipMemsetAsync(flag, 0, sizeof(int), d->s->st()), "memset");
    k_validate<<<blocks_for(n), 256, 0, d->s->st()>>>(nullptr, d_values, n, 0, flag);
    launched("k_validate");
    int hflag = 0;
    d->s->d2h(&hflag, flag, sizeof(int));
    d->s->sync();
    if (hflag) throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
    d->batch(d_values, n, false, 0.0, 0.0);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_add_batches_device(kd_tdigest* d, const double* d_values, int64_t n,
                                             int64_t batch) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    if (batch <= 0) throw KdError(KD_EINVAL, "batch size must be positive");
    int* flag = d->s->buf<int>(d->s->tmp, 1);
    hip_ok(hipMemsetAsync(flag, 0, sizeof(int), d->s->st()), "memset");
    k_validate<<<blocks_for(n), 256, 0, d->s->st()>>>(nullptr, d_values, n, 0, flag);
    launched("k_validate");
    // Slices that can take the selection path get their splitter samples
    // sorted up front, all in one launch (they do not depend on the merges).
    const int64_t nb = (n + batch - 1) / batch;
    double* ps = nullptr;
    if (batch > kd::kCap) {
      ps = d->s->buf<double>(d->s->presort, nb * kd::Engine::kSelSampleMax);
      d->s->eng.presort_samples(d_values, n, batch, ps);
    }
    int hflag = 0;
    d->s->d2h(&hflag, flag, sizeof(int));
    d->s->sync();
    if (hflag) throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
    struct Clear {
      kd::Engine& e;
      ~Clear() { e.set_presorted_sample(nullptr); }
    } clear{d->s->eng};
    for (int64_t b = 0; b < nb; b++) {
      const int64_t v0 = b * batch, nv = std::min<int64_t>(batch, n - v0);
      d->s->eng.set_presorted_sample(ps ? ps + b * kd::Engine::kSelSampleMax : nullptr);
      d->batch(d_values + v0, nv, false, 0.0, 0.0);
    }
    return KD_OK;
  });
}

extern "C" int kd_tdigest_merge(kd_tdigest* d) {
  return guarded([&] {
    d->merge();
    d->s->sync();
    return KD_OK;
  });
}

extern "C" int kd_tdigest_quantile(kd_tdigest* d, const double* qs, int64_t nq, double* out) {
  return guarded([&] {
    for (int64_t i = 0; i < nq; i++)  // :492-493 before the merge
      if (qs[i] < 0.0 || qs[i] > 1.0) throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
    d->merge();
    if (d->nc == 0) throw KdError(KD_EEMPTY, "quantile of an empty digest is undefined");
    d->query(0, qs, nq, out);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_cdf(kd_tdigest* d, const double* xs, int64_t nq, double* out) {
  return guarded([&] {
    d->merge();
    if (d->nc == 0) throw KdError(KD_EEMPTY, "cdf of an empty digest is undefined");
    d->query(1, xs, nq, out);
    return KD_OK;
  });
}

extern "C" int kd_tdigest_merge_into(kd_tdigest* self, kd_tdigest* other) {
  return guarded([&] {
    // tdigest.mojo:589-623
    other->merge();
    if (other->effective_count() == 0.0) return KD_OK;
    self->merge();
    if (self->decay()) self->renormalize();
    if (other->decay()) other->renormalize();
    bool was_empty = self->stored_total == 0.0;
    std::vector<double> om, ow;
    other->download(om, ow);
    self->bm.insert(self->bm.end(), om.begin(), om.end());
    self->bw.insert(self->bw.end(), ow.begin(), ow.end());
    self->stored_total += other->stored_total;
    if (was_empty) {
      self->vmin = other->vmin;
      self->vmax = other->vmax;
    } else {
      self->vmin = std::min(self->vmin, other->vmin);
      self->vmax = std::max(self->vmax, other->vmax);
    }
    self->clock = std::max(self->clock, other->clock);
    self->merge();
    self->s->sync();
    return KD_OK;
  });
}

static void put_f64(uint8_t* p, double v) {  // tdigest.mojo:80-84
  uint64_t b = kd_d2bits(v);
  for (int i = 0; i < 8; i++) p[i] = (uint8_t)((b >> (8 * i)) & 0xff);
}

static int64_t serialize(const double* hdr7, const std::vector<double>& m,
                         const std::vector<double>& w, uint8_t* out, int64_t cap) {
  int64_t need = 1 + 8 * 8 + 16 * (int64_t)m.size();
  if (!out || cap < need) return need;
  out[0] = kSerdeVersion;
  for (int i = 0; i < 7; i++) put_f64(out + 1 + 8 * i, hdr7[i]);
  put_f64(out + 57, (double)m.size());
  for (size_t i = 0; i < m.size(); i++) {
    put_f64(out + 65 + 16 * i, m[i]);
    put_f64(out + 73 + 16 * i, w[i]);
  }
  return need;
}

extern "C" int64_t kd_tdigest_to_bytes(kd_tdigest* d, uint8_t* out, int64_t cap) {
  int64_t r = 0;
  int rc = guarded([&] {
    d->merge();  // :638
    std::vector<double> m, w;
    d->download(m, w);
    double hdr[7] = {d->compression, d->lambda, d->log_scale, d->clock, d->stored_total, d->vmin, d->vmax};
    r = serialize(hdr, m, w, out, cap);
    return KD_OK;
  });
  return rc == KD_OK ? r : rc;
}

static int get_f64(const uint8_t* data, int64_t n, int64_t off, double* v) {  // :87-94
  if (off + 8 > n) return KD_ETRUNC;
  uint64_t b = 0;
  for (int i = 0; i < 8; i++) b |= (uint64_t)data[off + i] << (8 * i);
  *v = kd_bits2d(b);
  return KD_OK;
}

extern "C" int kd_tdigest_from_bytes(kd_session* s, const uint8_t* data, int64_t n, int scale,
                                     kd_tdigest** out) {
  return guarded([&] {
    if (n == 0 || data[0] != kSerdeVersion)
      throw KdError(KD_ESERDE, "unrecognized digest serialization version");
    double h[8];
    for (int i = 0; i < 8; i++)
      if (get_f64(data, n, 1 + 8 * i, &h[i]))
        throw KdError(KD_ETRUNC, "truncated digest: not enough bytes for a Float64");
    // The count is stored as a double: a whole number of centroids, checked
    // before anything is sized by it, and the blob must hold that many pairs.
    if (!(h[7] >= 0.0) || !(h[7] < 4503599627370496.0) || h[7] != std::floor(h[7]))
      throw KdError(KD_ESERDE, "digest centroid count is not a whole number >= 0");
    int64_t count = (int64_t)h[7];
    if (count > (n - 65) / 16)
      throw KdError(KD_ETRUNC, "truncated digest: not enough bytes for a Float64");
    std::vector<double> m(count), w(count);
    int64_t off = 65;
    for (int64_t i = 0; i < count; i++, off += 16)
      if (get_f64(data, n, off, &m[i]) || get_f64(data, n, off + 8, &w[i]))
        throw KdError(KD_ETRUNC, "truncated digest: not enough bytes for a Float64");
    std::unique_ptr<kd_tdigest> d(new_digest(s, scale, h[0], h[1]));
    d->upload(m.data(), w.data(), (int64_t)m.size());
    d->log_scale = h[2];
    d->clock = h[3];
    d->stored_total = h[4];
    d->vmin = h[5];
    d->vmax = h[6];
    *out = d.release();
    return KD_OK;
  });
}

extern "C" int kd_tdigest_from_centroids(kd_session* s, int scale, double compression,
                                         const double* means, const double* weights, int64_t n,
                                         double min_v, double max_v, kd_tdigest** out) {
  return guarded([&] {
    std::unique_ptr<kd_tdigest> d(new_digest(s, scale, compression, 0.0));
    double total = 0.0;
    for (int64_t i = 0; i < n; i++) total += weights[i];  // :716-722
    d->upload(means, weights, n);
    d->stored_total = total;
    d->clock = total;
    d->vmin = min_v;
    d->vmax = max_v;
    *out = d.release();
    return KD_OK;
  });
}

extern "C" int64_t kd_tdigest_size(kd_tdigest* d) { return d->nc; }

extern "C" int kd_tdigest_centroids(kd_tdigest* d, double* means, double* weights, int64_t cap,
                                    int64_t* count) {
  return guarded([&] {
    *count = d->nc;
    if (cap < d->nc) return KD_OK;
    d->s->d2h(means, d->dm(), d->nc * sizeof(double));
    d->s->d2h(weights, d->dw(), d->nc * sizeof(double));
    d->s->sync();
    return KD_OK;
  });
}

extern "C" int kd_tdigest_state(kd_tdigest* d, double* o) {
  o[0] = d->stored_total;
  o[1] = d->log_scale;
  o[2] = d->clock;
  o[3] = d->vmin;
  o[4] = d->vmax;
  o[5] = d->compression;
  o[6] = d->lambda;
  o[7] = d->effective_count();
  o[8] = (double)d->bm.size();
  return KD_OK;
}

// ========================================================= GroupedTDigest
// Per-key state and centroids live on the device; add_batch plans, compresses
// and rebuilds the centroid CSR there (one host sync for the plan summary,
// one at the end).  Host copies of the offsets and per-key scalars are
// pulled on demand for the host-side accessors.
struct kd_grouped {
  kd_session* s;
  int scale;
  int64_t num_keys;
  double compression, lambda;
  // Host mirrors (valid when !dirty).
  std::vector<double> stored_total, log_scale, clock, vmin, vmax;
  std::vector<int64_t> hoff;      // host copy of the CSR offsets
  bool dirty = false;
  int64_t total_cent = 0;         // centroids over all keys
  DevBuf off, cm, cw, S;          // device CSR of centroids + query prefix
  DevBuf noff, ncm, ncw;          // next CSR (swapped in after a batch)
  DevBuf state;                   // 5 x num_keys: stored total, log scale, clock, min, max
  DevBuf qstate, seg_key, seg_of, ncnt, tot, qerr;
  int64_t* h_tot = nullptr;       // page-locked new centroid total
  bool prefix_valid = false;

  ~kd_grouped() {
    if (h_tot) (void)hipHostFree(h_tot);
  }
  bool decay() const { return lambda > 0.0; }
  int64_t* doff() { return static_cast<int64_t*>(off.p); }
  double* dm() { return static_cast<double*>(cm.p); }
  double* dw() { return static_cast<double*>(cw.p); }
  double* dstate(int i) { return static_cast<double*>(state.p) + (size_t)i * num_keys; }
  void pull() {
    if (!dirty) return;
    s->d2h(hoff.data(), doff(), (num_keys + 1) * sizeof(int64_t));
    s->d2h(stored_total.data(), dstate(0), num_keys * sizeof(double));
    s->d2h(log_scale.data(), dstate(1), num_keys * sizeof(double));
    s->d2h(clock.data(), dstate(2), num_keys * sizeof(double));
    s->d2h(vmin.data(), dstate(3), num_keys * sizeof(double));
    s->d2h(vmax.data(), dstate(4), num_keys * sizeof(double));
    s->sync();
    dirty = false;
  }
  int64_t count(int64_t k) {
    pull();
    return hoff[k + 1] - hoff[k];
  }
  double eff(int64_t k) {
    pull();
    return decay() ? stored_total[k] * kd_exp(-log_scale[k]) : stored_total[k];
  }

  void check_key(int64_t k) const {
    if (k < 0 || k >= num_keys) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
  }

  void add_device(const uint32_t* d_keys, const double* d_values, int64_t n);
  // d_err (may be null): bit 0 a key outside [0, num_keys), bit 1 a quantile
  // q outside [0, 1], bit 2 a key with an empty digest; those outputs are NaN.
  void query(int is_cdf, const int64_t* d_keys, const double* d_q, int64_t nq, double* d_out,
             int* d_err = nullptr);
};

namespace {
// Per segment of a grouped batch: min / max (grouped.mojo via
// tdigest.mojo:285-295) and the reconciled stored total (:298-315).
// only: -1 every segment, 0 those on the sort path (mode 0), 1 the others.
__global__ void k_grouped_keys(int64_t nseg, const kd::SegDesc* segs, const kd::SegResult* res,
                               const int64_t* seg_key, double* total, double* vmin, double* vmax,
                               int only) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  const kd::SegDesc sd = segs[i];
  if (only >= 0 && (sd.mode == 0) != (only == 0)) return;
  const kd::SegResult r = res[i];
  const int64_t k = seg_key[i];
  if (sd.stored_n == (double)sd.v_cnt * sd.sf && sd.e_cnt == 0) {
    vmin[k] = r.vmin;
    vmax[k] = r.vmax;
  } else {
    if (r.vmin < vmin[k]) vmin[k] = r.vmin;
    if (r.vmax > vmax[k]) vmax[k] = r.vmax;
  }
  total[k] = r.total;
}

// Whether segment sg is left out under only (-1 none, 0 all but the sort
// path's, 1 the sort path's).
__device__ inline bool seg_skipped(const kd::SegDesc* segs, int64_t sg, int only) {
  return only >= 0 && (segs[sg].mode == 0) != (only == 0);
}

// Centroid count of every key after the batch (0 for a key whose segment is
// left out under only).
__global__ void k_grouped_counts(int64_t nk, const int64_t* off, const int64_t* seg_of,
                                 const kd::SegDesc* segs, const kd::SegResult* res, int only,
                                 int64_t* cnt) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  int64_t sg = seg_of[k];
  cnt[k] = sg >= 0 ? (seg_skipped(segs, sg, only) ? 0 : res[sg].count) : off[k + 1] - off[k];
}

// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total) in three
// launches: block sums of 1024-entry chunks, one block scanning those (at
// most 1024 * 1024 entries), then every chunk scanned from its base.
constexpr int kOffChunk = 1024;

__device__ inline int64_t block_scan_1024(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[17];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    int64_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int i = 0; i < 16; i++) {
      int64_t t = wsum[i];
      wsum[i] = run;
      run += t;
    }
    wsum[16] = run;
  }
  __syncthreads();
  int64_t r = inc - v + wsum[w];
  *total = wsum[16];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(1024) void k_off_sums(const int64_t* cnt, int64_t n, int64_t* bsum) {
  const int64_t i = (int64_t)blockIdx.x * kOffChunk + threadIdx.x;
  int64_t tot;
  block_scan_1024(i < n ? cnt[i] : 0, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Block sums scanned in place; each thread owns a contiguous run of them.
__global__ __launch_bounds__(1024) void k_off_top(int64_t* bsum, int64_t nb, int64_t* total) {
  const int64_t per = (nb + 1023) / 1024;
  const int64_t lo = std::min<int64_t>(nb, threadIdx.x * per), hi = std::min<int64_t>(nb, lo + per);
  int64_t v = 0;
  for (int64_t i = lo; i < hi; i++) v += bsum[i];
  int64_t tot;
  int64_t run = block_scan_1024(v, &tot);
  for (int64_t i = lo; i < hi; i++) {
    int64_t x = bsum[i];
    bsum[i] = run;
    run += x;
  }
  if (threadIdx.x == 0) *total = tot;
}

__global__ __launch_bounds__(1024) void k_off_down(const int64_t* cnt, int64_t n, const int64_t* bsum,
                                                   int64_t* off) {
  const int64_t i = (int64_t)blockIdx.x * kOffChunk + threadIdx.x;
  int64_t tot;
  int64_t e = block_scan_1024(i < n ? cnt[i] : 0, &tot);
  if (i < n) off[i] = bsum[blockIdx.x] + e;
}

// New CSR: one wave per key copies its centroids (the engine's output for a
// touched key, the old ones otherwise).  Nothing is written when the new
// total exceeds cap, the capacity of nm and nw, nor for a key whose segment
// is left out under only (its count is 0 then).
__global__ void k_gather_keys(int64_t nk, const int64_t* old_off, const double* om,
                              const double* ow, const int64_t* seg_of, const kd::SegDesc* segs,
                              const double* em, const double* ew, const int64_t* new_off,
                              double* nm, double* nw, int64_t cap, int only) {
  if (new_off[nk] > cap) return;
  const int lane = threadIdx.x & 63;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nk; k += stride) {
    const int64_t dst = new_off[k], cnt = new_off[k + 1] - dst;
    const int64_t sg = seg_of[k];
    if (sg >= 0 && seg_skipped(segs, sg, only)) continue;
    const double* sm = sg >= 0 ? em + segs[sg].x_off : om + old_off[k];
    const double* sw = sg >= 0 ? ew + segs[sg].x_off : ow + old_off[k];
    for (int64_t i = lane; i < cnt; i += 64) {
      nm[dst + i] = sm[i];
      nw[dst + i] = sw[i];
    }
  }
}

// Query-side state of every key: stored total, min, max, e^-log_scale.
__global__ void k_grouped_qstate(int64_t nk, const double* total, const double* logs,
                                 const double* vmin, const double* vmax, int decay, double* q) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  q[k] = total[k];
  q[nk + k] = vmin[k];
  q[2 * nk + k] = vmax[k];
  q[3 * nk + k] = decay ? kd_exp(-logs[k]) : 1.0;
}
}  // namespace

extern "C" int kd_grouped_create(kd_session* s, int scale, int64_t num_keys, double compression,
                                 double lambda, kd_grouped** out) {
  return guarded([&] {
    if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
    if (num_keys <= 0 || num_keys > 0x7fffffffLL) throw KdError(KD_EINVAL, "num_keys must be in [1, 2^31 - 1]");
    std::unique_ptr<kd_grouped> g(new kd_grouped());
    g->s = s;
    g->scale = scale;
    g->num_keys = num_keys;
    g->compression = compression;
    g->lambda = lambda;
    g->stored_total.assign(num_keys, 0.0);
    g->log_scale.assign(num_keys, 0.0);
    g->clock.assign(num_keys, 0.0);
    g->vmin.assign(num_keys, 0.0);
    g->vmax.assign(num_keys, 0.0);
    g->hoff.assign(num_keys + 1, 0);
    g->off.reserve((num_keys + 1) * sizeof(int64_t));
    hip_ok(hipMemsetAsync(g->off.p, 0, (num_keys + 1) * sizeof(int64_t), s->st()), "memset");
    g->state.reserve(5 * num_keys * sizeof(double));
    hip_ok(hipMemsetAsync(g->state.p, 0, 5 * num_keys * sizeof(double), s->st()), "memset");
    g->cm.reserve(sizeof(double));
    g->cw.reserve(sizeof(double));
    void* p = nullptr;
    hip_ok(hipHostMalloc(&p, sizeof(int64_t), hipHostMallocDefault), "hipHostMalloc");
    g->h_tot = static_cast<int64_t*>(p);
    s->sync();
    *out = g.release();
    return KD_OK;
  });
}

extern "C" void kd_grouped_destroy(kd_grouped* g) { delete g; }

static void apply_recorded_shape(kd_session* s, int scale, double compression, int64_t num_keys,
                                 int64_t n);

void kd_grouped::add_device(const uint32_t* d_keys, const double* d_values, int64_t n) {
  apply_recorded_shape(s, scale, compression, num_keys, n);
  // grouped.mojo:60-162.  The whole batch is validated (:91-97) before any
  // key changes: the partition (or k_validate) sets the flag, and the device
  // plan changes nothing when it is set.
  if (n > kMaxGroupedBatch)
    throw KdError(KD_EINVAL, "grouped batch above 2^32 - 2^16 pairs: split it into several calls");
  clk::time_point tp = clk::now();
  int* flag = s->buf<int>(s->tmp, 1);
  hip_ok(hipMemsetAsync(flag, 0, sizeof(int), s->st()), "memset");
  const bool fused_check = s->eng.key_partition_validates(n, num_keys);
  if (!fused_check) {
    k_validate<<<blocks_for(n), 256, 0, s->st()>>>(d_keys, d_values, n, num_keys, flag);
    launched("k_validate");
  }
  // Group per key (:99-104) on the device.
  int64_t* dcnt = static_cast<int64_t*>(s->g_counts.reserve(num_keys * sizeof(int64_t)));
  int64_t* dkoff = static_cast<int64_t*>(s->g_koff.reserve((num_keys + 1) * sizeof(int64_t)));
  double* dv = static_cast<double*>(s->g_vals.reserve(std::max<int64_t>(n, 1) * sizeof(double)));
  s->eng.key_partition(d_keys, d_values, n, num_keys, dcnt, dkoff, dv, fused_check ? flag : nullptr);
  s->eng.host_phase("key_partition", ms_since(tp));

  // Per touched key: B-semantics prepare (tdigest.mojo:236-296), segment
  // table and compress, all on the device.
  int64_t* dsk = static_cast<int64_t*>(seg_key.reserve(num_keys * sizeof(int64_t)));
  int64_t* dso = static_cast<int64_t*>(seg_of.reserve(num_keys * sizeof(int64_t)));
  kd::GroupedPlanArgs a{};
  a.scale = scale;
  a.compression = compression;
  a.lambda = lambda;
  a.num_keys = num_keys;
  a.kcnt = dcnt;
  a.kof = dkoff;
  a.off = doff();
  a.w = dw();
  a.total = dstate(0);
  a.logs = dstate(1);
  a.clock = dstate(2);
  a.bad = flag;
  a.seg_key = dsk;
  a.seg_of = dso;
  // Selection fallbacks (rare) are finished after the sync below, not with
  // a sync of their own.
  const kd::PlanSummary P = s->eng.compress_grouped(a, dv, dm(), dw(), true);
  s->eng.host_phase("compress", ms_since(tp));
  if (P.bad == 1) throw KdError(KD_EINVAL, "key out of range [0, num_keys) or non-finite value");
  if (P.bad) throw std::runtime_error("too many explicit records in one digest");
  if (P.nseg == 0) return;

  // Per-key min / max / stored total, then the new CSR (untouched keys keep
  // their centroids).
  const kd::SegDesc* segs = s->eng.dev_segments();
  const kd::SegResult* res = s->eng.dev_results();
  int64_t* dc = static_cast<int64_t*>(ncnt.reserve(num_keys * sizeof(int64_t)));
  int64_t* dno = static_cast<int64_t*>(noff.reserve((num_keys + 1) * sizeof(int64_t)));
  const int64_t nb = (num_keys + kOffChunk - 1) / kOffChunk;
  int64_t* bsum = static_cast<int64_t*>(tot.reserve((nb + 1) * sizeof(int64_t)));
  // The new total is known on the device only.  The next CSR is sized from
  // a guess (the current total, or the session's last batch of this shape,
  // at most the bound total + n: a key ends with at most its old centroids
  // plus its new values) and gathered at once; the gather writes nothing
  // when the total exceeds the guess, and then runs again after the sync,
  // with the exact size.
  int64_t guess = total_cent;
  if (s->csr_keys == num_keys && s->csr_pairs == n) guess = std::max(guess, s->csr_total);
  guess = std::max<int64_t>(std::min<int64_t>(guess, total_cent + n), 1);
  // The key state of the segments under only_keys, and the CSR with the
  // segments under only_csr (-1 every one; with fallen-back segments
  // pending, 1 all but those, which count as empty; 0 those).
  auto gather = [&](int64_t cap_items, int only_csr) {
    double* dnm = static_cast<double*>(ncm.reserve(cap_items * sizeof(double)));
    double* dnw = static_cast<double*>(ncw.reserve(cap_items * sizeof(double)));
    const int64_t cap = (int64_t)(std::min(ncm.cap, ncw.cap) / sizeof(double));
    k_gather_keys<<<wave_grid(num_keys), 256, 0, s->st()>>>(
        num_keys, doff(), dm(), dw(), dso, segs, s->eng.cent_means(), s->eng.cent_weights(), dno,
        dnm, dnw, cap, only_csr);
    launched("k_gather_keys");
    return cap;
  };
  auto rebuild = [&](int only_keys, int only_csr, int64_t cap_items) {
    k_grouped_keys<<<blocks_for(P.nseg), 256, 0, s->st()>>>(P.nseg, segs, res, dsk, dstate(0),
                                                            dstate(3), dstate(4), only_keys);
    launched("k_grouped_keys");
    k_grouped_counts<<<blocks_for(num_keys), 256, 0, s->st()>>>(num_keys, doff(), dso, segs, res,
                                                                only_csr, dc);
    launched("k_grouped_counts");
    k_off_sums<<<(unsigned)nb, kOffChunk, 0, s->st()>>>(dc, num_keys, bsum);
    launched("k_off_sums");
    k_off_top<<<1, kOffChunk, 0, s->st()>>>(bsum, nb, dno + num_keys);
    launched("k_off_top");
    k_off_down<<<(unsigned)nb, kOffChunk, 0, s->st()>>>(dc, num_keys, bsum, dno);
    launched("k_off_down");
    const int64_t cap = gather(cap_items, only_csr);
    s->d2h(h_tot, dno + num_keys, sizeof(int64_t));
    s->sync();
    return cap;
  };
  const int first = s->eng.grouped_deferred() ? 1 : -1;
  int64_t cap = rebuild(first, first, guess);
  // Fallen-back segments finished now: their keys' state, then every count
  // and the gather again.
  if (s->eng.finish_grouped(scale, P, dv, dm(), dw())) cap = rebuild(0, -1, guess);
  const int64_t new_total = *h_tot;
  if (new_total > cap) {
    gather(new_total, -1);
    s->sync();
  }
  s->csr_keys = num_keys;
  s->csr_pairs = n;
  s->csr_total = new_total;
  off.swap(noff);
  cm.swap(ncm);
  cw.swap(ncw);
  total_cent = new_total;
  dirty = true;
  prefix_valid = false;
  s->eng.host_phase("csr", ms_since(tp));
}

void kd_grouped::query(int is_cdf, const int64_t* d_keys, const double* d_q, int64_t nq,
                       double* d_out, int* d_err) {
  double* dS = static_cast<double*>(S.reserve(std::max<int64_t>(total_cent, 1) * sizeof(double)));
  double* st = static_cast<double*>(qstate.reserve(4 * num_keys * sizeof(double)));
  if (!prefix_valid) {
    k_query_prefix<<<blocks_for(num_keys, 64), 64, 0, s->st()>>>(num_keys, doff(), dw(), dS);
    launched("k_query_prefix");
    k_grouped_qstate<<<blocks_for(num_keys), 256, 0, s->st()>>>(
        num_keys, dstate(0), dstate(1), dstate(3), dstate(4), decay() ? 1 : 0, st);
    launched("k_grouped_qstate");
    prefix_valid = true;
  }
  k_query_grouped<<<blocks_for(nq), 256, 0, s->st()>>>(is_cdf, num_keys, doff(), dm(), dw(), dS,
                                                       st, st + num_keys, st + 2 * num_keys,
                                                       st + 3 * num_keys, d_keys, d_q, nq, d_out,
                                                       d_err);
  launched("k_query_grouped");
  hip_ok(hipGetLastError(), "grouped query launch");
}

// ------------------------------------------------------------- autotuner
// Picks the launch shape of the grouped ingest for a representative batch by
// timing it: the key partition (items per sub-tile, chunks per pass), the
// selection tiling (tile_min, tiles_max) and whether the small-digest
// compress overlaps the selection path on a second stream.  Coordinate
// descent from the session's current shape: every value of one knob is timed
// with the others fixed (best of `reps` ingests into fresh digests), the
// fastest is kept, then the next knob.  Every shape gives the same digests,
// so the choice only changes speed.
namespace {
struct GroupedShape {
  int items;
  int64_t chunks, tile_min, tiles_max;
  int overlap;
};

GroupedShape current_shape(kd_session* s) {
  GroupedShape c;
  c.items = s->eng.partition_items();
  c.chunks = s->eng.partition_chunks();
  c.tile_min = s->eng.select_tile_min();
  c.tiles_max = s->eng.select_tiles_max();
  c.overlap = s->eng.overlap();
  return c;
}

void apply_shape(kd_session* s, const GroupedShape& c) {
  s->eng.set_partition_shape(c.items, c.chunks);
  s->eng.set_select_shape(c.tile_min, c.tiles_max);
  s->eng.set_overlap(c.overlap);
}

// Shapes the autotuner chose, for the whole process, keyed by the batch's
// class: device, scale, compression, and the powers of two at or above the
// key count and the pair count.  A grouped ingest whose class has an entry
// runs with that shape (unless its session's shape was fixed by the caller);
// kd_autotune_save / kd_autotune_load keep the table across processes.
using TuneKey = std::tuple<int, int, uint64_t, int, int>;
std::mutex g_tune_mu;
std::map<TuneKey, GroupedShape> g_tune;

int pow2_class(int64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll((uint64_t)(x - 1)); }

TuneKey tune_key(int device, int scale, double compression, int64_t num_keys, int64_t n) {
  return TuneKey{device, scale, kd_d2bits(compression), pow2_class(num_keys), pow2_class(n)};
}

void record_shape(const TuneKey& k, const GroupedShape& c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[k] = c;
}

bool recorded_shape(const TuneKey& k, GroupedShape* c) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  *c = it->second;
  return true;
}
}  // namespace

static void apply_recorded_shape(kd_session* s, int scale, double compression, int64_t num_keys,
                                 int64_t n) {
  if (s->shape_pinned) return;
  GroupedShape c;
  if (!recorded_shape(tune_key(s->eng.device(), scale, compression, num_keys, n), &c)) return;
  const GroupedShape cur = current_shape(s);
  if (memcmp(&c, &cur, sizeof c) != 0) apply_shape(s, c);
}

extern "C" int kd_session_shape(kd_session* s, int64_t* out5) {
  return guarded([&] {
    const GroupedShape c = current_shape(s);
    out5[0] = c.items;
    out5[1] = c.chunks;
    out5[2] = c.tile_min;
    out5[3] = c.tiles_max;
    out5[4] = c.overlap;
    return KD_OK;
  });
}

extern "C" int kd_autotune_save(const char* path) {
  return guarded([&] {
    std::ofstream f(path);
    if (!f) throw KdError(KD_EINVAL, std::string("cannot write ") + path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    f << "# kdigest grouped autotune: device scale compression_bits keys_class pairs_class "
         "items chunks tile_min tiles_max overlap\n";
    for (const auto& e : g_tune) {
      const TuneKey& k = e.first;
      const GroupedShape& c = e.second;
      f << std::get<0>(k) << ' ' << std::get<1>(k) << ' ' << std::get<2>(k) << ' ' << std::get<3>(k)
        << ' ' << std::get<4>(k) << ' ' << c.items << ' ' << c.chunks << ' ' << c.tile_min << ' '
        << c.tiles_max << ' ' << c.overlap << '\n';
    }
    if (!f) throw KdError(KD_EINVAL, std::string("cannot write ") + path);
    return KD_OK;
  });
}

extern "C" int kd_autotune_load(const char* path, int64_t* loaded) {
  return guarded([&] {
    std::ifstream f(path);
    if (!f) throw KdError(KD_EINVAL, std::string("cannot read ") + path);
    std::string line;
    int64_t cnt = 0;
    std::map<TuneKey, GroupedShape> add;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream is(line);
      int dev, scale, kc, nc, items, ov;
      uint64_t cb;
      int64_t chunks, tmin, tmax;
      if (!(is >> dev >> scale >> cb >> kc >> nc >> items >> chunks >> tmin >> tmax >> ov))
        throw KdError(KD_EINVAL, "malformed autotune line: " + line);
      if ((items != 4 && items != 8 && items != 16) || chunks < 1 || chunks > (1 << 20) ||
          tmin < 1024 || tmin > (1 << 24) || tmax < 1 || tmax > 4096 || scale < 0 || scale > 3)
        throw KdError(KD_EINVAL, "autotune line out of range: " + line);
      add[TuneKey{dev, scale, cb, kc, nc}] = GroupedShape{items, chunks, tmin, tmax, ov != 0};
      cnt++;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& e : add) g_tune[e.first] = e.second;
    if (loaded) *loaded = cnt;
    return KD_OK;
  });
}

extern "C" int kd_session_autotune_grouped(kd_session* s, int scale, int64_t num_keys,
                                           double compression, const uint32_t* d_keys,
                                           const double* d_values, int64_t n, int reps,
                                           int64_t* out_shape5, double* out_ms2) {
  return guarded([&] {
    if (n <= 0) throw KdError(KD_EINVAL, "autotune needs a non-empty batch");
    if (reps < 1) reps = 1;
    const bool prof = s->eng.profiling();
    s->eng.set_profiling(false);
    s->shape_pinned = true;  // the trials run exactly the shapes they set
    auto trial = [&](const GroupedShape& c) {
      apply_shape(s, c);
      double best = 1e300;
      for (int r = 0; r < reps; r++) {
        kd_grouped* g = nullptr;
        int rc = kd_grouped_create(s, scale, num_keys, compression, 0.0, &g);
        if (rc != KD_OK) throw KdError(rc, g_err);
        std::unique_ptr<kd_grouped> hold(g);
        s->sync();
        auto t0 = clk::now();
        g->add_device(d_keys, d_values, n);
        s->sync();
        best = std::min(best, ms_since(t0));
      }
      return best;
    };
    GroupedShape best = current_shape(s);
    trial(best);  // warm: buffers at their final sizes, kernels loaded
    double best_ms = trial(best);
    const double first_ms = best_ms;
    const int items[] = {4, 8, 16};
    const int64_t chunks[] = {1024, 2048, 4096, 8192, 16384};
    const int64_t tmins[] = {16384, 32768, 65536, 131072};
    const int64_t tmaxs[] = {16, 32, 64, 128, 256, 512, 1024};
    auto knob = [&](const auto& values, auto set) {
      for (auto v : values) {
        GroupedShape c = best;
        set(c, v);
        if (memcmp(&c, &best, sizeof c) == 0) continue;
        double ms = trial(c);
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    };
    knob(items, [](GroupedShape& c, int v) { c.items = v; });
    knob(chunks, [](GroupedShape& c, int64_t v) { c.chunks = v; });
    knob(tmins, [](GroupedShape& c, int64_t v) { c.tile_min = v; });
    knob(tmaxs, [](GroupedShape& c, int64_t v) { c.tiles_max = v; });
    const int ov[] = {0, 1, 2};
    knob(ov, [](GroupedShape& c, int v) { c.overlap = v; });
    apply_shape(s, best);
    s->shape_pinned = false;
    record_shape(tune_key(s->eng.device(), scale, compression, num_keys, n), best);
    s->eng.set_profiling(prof);
    if (out_shape5) {
      out_shape5[0] = best.items;
      out_shape5[1] = best.chunks;
      out_shape5[2] = best.tile_min;
      out_shape5[3] = best.tiles_max;
      out_shape5[4] = best.overlap;
    }
    if (out_ms2) {
      out_ms2[0] = first_ms;
      out_ms2[1] = best_ms;
    }
    return KD_OK;
  });
}

extern "C" int kd_session_autotune_grouped_host(kd_session* s, int scale, int64_t num_keys,
                                                double compression, const uint32_t* keys,
                                                const double* values, int64_t n, int reps,
                                                int64_t* out_shape5, double* out_ms2) {
  return guarded([&] {
    if (n <= 0) throw KdError(KD_EINVAL, "autotune needs a non-empty batch");
    DevBuf dk, dv;
    dk.reserve(n * sizeof(uint32_t));
    dv.reserve(n * sizeof(double));
    s->h2d(dk.p, keys, n * sizeof(uint32_t));
    s->h2d(dv.p, values, n * sizeof(double));
    s->sync();
    int rc = kd_session_autotune_grouped(s, scale, num_keys, compression,
                                         static_cast<const uint32_t*>(dk.p),
                                         static_cast<const double*>(dv.p), n, reps, out_shape5,
                                         out_ms2);
    if (rc != KD_OK) throw KdError(rc, g_err);
    return KD_OK;
  });
}

extern "C" int kd_grouped_add_batch(kd_grouped* g, const uint32_t* keys, const double* values,
                                    int64_t n) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    for (int64_t i = 0; i < n; i++) {  // grouped.mojo:91-97
      if ((int64_t)keys[i] >= g->num_keys) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
      if (!is_finite_val(values[i])) throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
    }
    uint32_t* dk = static_cast<uint32_t*>(g->s->g_keys.reserve(n * sizeof(uint32_t)));
    double* dv = static_cast<double*>(g->s->g_valsin.reserve(n * sizeof(double)));
    g->s->h2d(dk, keys, n * sizeof(uint32_t));
    g->s->h2d(dv, values, n * sizeof(double));
    g->add_device(dk, dv, n);
    return KD_OK;
  });
}

extern "C" int kd_grouped_add_batch_device(kd_grouped* g, const uint32_t* d_keys,
                                           const double* d_values, int64_t n) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    g->add_device(d_keys, d_values, n);
    return KD_OK;
  });
}

static int grouped_query_host(kd_grouped* g, int is_cdf, const int64_t* keys, const double* qs,
                              int64_t nq, double* out) {
  return guarded([&] {
    if (nq <= 0) return KD_OK;
    for (int64_t i = 0; i < nq; i++) {
      g->check_key(keys[i]);
      if (!is_cdf && (qs[i] < 0.0 || qs[i] > 1.0))
        throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
      if (g->count(keys[i]) == 0)
        throw KdError(KD_EEMPTY, is_cdf ? "cdf of an empty digest is undefined"
                                        : "quantile of an empty digest is undefined");
    }
    kd_session* s = g->s;
    int64_t* dk = s->buf<int64_t>(s->q_qd, nq);
    double* dq = s->buf<double>(s->q_q, nq);
    double* dout = s->buf<double>(s->q_out, nq);
    s->h2d(dk, keys, nq * sizeof(int64_t));
    s->h2d(dq, qs, nq * sizeof(double));
    g->query(is_cdf, dk, dq, nq, dout);
    s->d2h(out, dout, nq * sizeof(double));
    s->sync();
    return KD_OK;
  });
}

extern "C" int kd_grouped_quantile(kd_grouped* g, const int64_t* keys, const double* qs,
                                   int64_t nq, double* out) {
  return grouped_query_host(g, 0, keys, qs, nq, out);
}

extern "C" int kd_grouped_cdf(kd_grouped* g, const int64_t* keys, const double* xs, int64_t nq,
                              double* out) {
  return grouped_query_host(g, 1, keys, xs, nq, out);
}

extern "C" int kd_grouped_quantile_device(kd_grouped* g, const int64_t* d_keys, const double* d_qs,
                                          int64_t nq, double* d_out) {
  return guarded([&] {
    if (nq <= 0) return KD_OK;
    if (nq > kMaxGroupedBatch) throw KdError(KD_EINVAL, "too many queries in one call");
    int* derr = static_cast<int*>(g->qerr.reserve(sizeof(int)));
    hip_ok(hipMemsetAsync(derr, 0, sizeof(int), g->s->st()), "memset");
    g->query(0, d_keys, d_qs, nq, d_out, derr);
    int herr = 0;
    g->s->d2h(&herr, derr, sizeof(int));
    g->s->sync();
    // Checked in the reference's order: key, then q range, then emptiness.
    if (herr & 1) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
    if (herr & 2) throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
    if (herr & 4) throw KdError(KD_EEMPTY, "quantile of an empty digest is undefined");
    return KD_OK;
  });
}

// The same queries with no device allocation: the prefix of every key, the
// per-key query state and the error flag live in the GPU's `scratch` array
// (rebuilt by every call), so a digest needs 256 + 32 * num_keys + 8 * centroids
// bytes of it at most.
extern "C" int kd_grouped_quantile_device_noalloc(kd_grouped* g, const int64_t* d_keys,
                                                  const double* d_qs, int64_t nq, double* d_out) {
  return guarded([&] {
    if (nq <= 0) return KD_OK;
    if (nq > kMaxGroupedBatch) throw KdError(KD_EINVAL, "too many queries in one call");
    const int64_t nk = g->num_keys, nc = g->total_cent;
    const size_t need = 256 + (size_t)32 * (size_t)nk + (size_t)8 * (size_t)std::max<int64_t>(nc, 1);
    if (need > kScratchBytes)
      throw KdError(KD_EINVAL, "digest needs " + std::to_string(need) +
                                   " bytes of query scratch; the allocation-free entry point has " +
                                   std::to_string(kScratchBytes));
    char* base = static_cast<char*>(scratch_of(g->s->eng.device()));
    int* derr = reinterpret_cast<int*>(base);
    double* st = reinterpret_cast<double*>(base + 256);
    double* dS = st + 4 * nk;
    hipStream_t sm = g->s->st();
    k_fill_int<<<1, 64, 0, sm>>>(derr, 1, 0);
    launched("k_fill_int");
    k_query_prefix<<<blocks_for(nk, 64), 64, 0, sm>>>(nk, g->doff(), g->dw(), dS);
    launched("k_query_prefix");
    k_grouped_qstate<<<blocks_for(nk), 256, 0, sm>>>(nk, g->dstate(0), g->dstate(1), g->dstate(3),
                                                     g->dstate(4), g->decay() ? 1 : 0, st);
    launched("k_grouped_qstate");
    k_query_grouped<<<blocks_for(nq), 256, 0, sm>>>(0, nk, g->doff(), g->dm(), g->dw(), dS, st,
                                                    st + nk, st + 2 * nk, st + 3 * nk, d_keys, d_qs,
                                                    nq, d_out, derr);
    launched("k_query_grouped");
    int herr = 0;
    hip_ok(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, sm), "D2H");
    hip_ok(hipStreamSynchronize(sm), "stream sync");
    if (herr & 1) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
    if (herr & 2) throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
    if (herr & 4) throw KdError(KD_EEMPTY, "quantile of an empty digest is undefined");
    return KD_OK;
  });
}

extern "C" int64_t kd_grouped_to_bytes(kd_grouped* g, int64_t key, uint8_t* out, int64_t cap) {
  int64_t r = 0;
  int rc = guarded([&] {
    g->check_key(key);
    int64_t c = g->count(key);  // pulls the host mirrors
    std::vector<double> m(c), w(c);
    g->s->d2h(m.data(), g->dm() + g->hoff[key], c * sizeof(double));
    g->s->d2h(w.data(), g->dw() + g->hoff[key], c * sizeof(double));
    g->s->sync();
    double hdr[7] = {g->compression, g->lambda, g->log_scale[key], g->clock[key],
                     g->stored_total[key], g->vmin[key], g->vmax[key]};
    r = serialize(hdr, m, w, out, cap);
    return KD_OK;
  });
  return rc == KD_OK ? r : rc;
}

// ------------------------------------------- wire format of every key at once
// Every key's digest in the per-digest layout (tdigest.mojo:627-652, the
// bytes kd_grouped_to_bytes gives for one key), concatenated in key order on
// the device, and the inverse: a grouped digest restored from such blobs.
// Checkpoint and resume of device-resident grouped state without a host
// round trip per key.
namespace {
constexpr int kWireHeader = 65;  // version byte + 8 f64 (count as f64)

__device__ inline uint64_t wire_u64(const uint8_t* p) {
  uint64_t b = 0;
  for (int i = 0; i < 8; i++) b |= (uint64_t)p[i] << (8 * i);
  return b;
}

// Blob sizes: 65 + 16 * centroids.
__global__ void k_wire_sizes(int64_t nk, const int64_t* off, int64_t* size) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nk) size[k] = kWireHeader + 16 * (off[k + 1] - off[k]);
}

// One key's blob, written byte by byte by one wave (a blob starts at any byte).
__device__ void wire_write_key(int64_t k, int lane, int64_t nk, const int64_t* off,
                               const double* cm, const double* cw, const double* state,
                               double compression, double lambda, const int64_t* woff,
                               uint8_t* out) {
  const int64_t c0 = off[k], c = off[k + 1] - c0, base = woff[k];
  const int64_t size = kWireHeader + 16 * c;
  // Header order: compression, lambda, log scale, clock, stored total, min,
  // max, count (tdigest.mojo:640-651).
  for (int64_t i = lane; i < size; i += 64) {
    uint8_t b;
    if (i == 0) {
      b = kSerdeVersion;
    } else {
      double v;
      int byte;
      if (i < kWireHeader) {
        const int f = (int)((i - 1) >> 3);
        byte = (int)((i - 1) & 7);
        switch (f) {
          case 0: v = compression; break;
          case 1: v = lambda; break;
          case 2: v = state[1 * nk + k]; break;
          case 3: v = state[2 * nk + k]; break;
          case 4: v = state[0 * nk + k]; break;
          case 5: v = state[3 * nk + k]; break;
          case 6: v = state[4 * nk + k]; break;
          default: v = (double)c; break;
        }
      } else {
        const int64_t q = i - kWireHeader;
        const int64_t pidx = q >> 4;
        byte = (int)(q & 7);
        v = (q & 8) ? cw[c0 + pidx] : cm[c0 + pidx];
      }
      b = (uint8_t)(kd_d2bits(v) >> (8 * byte));
    }
    out[base + i] = b;
  }
}

__global__ void k_wire_write(int64_t nk, const int64_t* off, const double* cm, const double* cw,
                             const double* state, double compression, double lambda,
                             const int64_t* woff, uint8_t* out) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nk; k += stride)
    wire_write_key(k, lane, nk, off, cm, cw, state, compression, lambda, woff, out);
}

// Per key: header checks and scalars.  err[0] the first truncated key,
// err[1] the first with a bad version byte or count, err[2] the first whose
// compression or decay differs from key 0's (all nk when none).
__global__ void k_wire_parse(int64_t nk, const uint8_t* data, int64_t n, const int64_t* woff,
                             int64_t* cnt, double* state, unsigned long long* err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const int64_t a = woff[k], b = woff[k + 1];
  cnt[k] = 0;
  if (a < 0 || b < a || b > n) {
    atomicMin(&err[0], (unsigned long long)k);
    return;
  }
  if (b == a || data[a] != kSerdeVersion) {
    atomicMin(&err[1], (unsigned long long)k);
    return;
  }
  if (b - a < kWireHeader) {
    atomicMin(&err[0], (unsigned long long)k);
    return;
  }
  double h[8];
  for (int i = 0; i < 8; i++) h[i] = kd_bits2d(wire_u64(data + a + 1 + 8 * i));
  // The count is stored as a double: a whole number of centroids.
  if (!(h[7] >= 0.0) || !(h[7] < 4503599627370496.0) || h[7] != floor(h[7])) {
    atomicMin(&err[1], (unsigned long long)k);
    return;
  }
  const int64_t c = (int64_t)h[7];
  if (b - a < kWireHeader + 16 * c) {
    atomicMin(&err[0], (unsigned long long)k);
    return;
  }
  // Key 0's compression and decay rate (read only when its own header is
  // in bounds; otherwise its thread reports it).
  const int64_t a0 = woff[0], b0 = woff[1];
  if (a0 >= 0 && b0 <= n && b0 - a0 >= kWireHeader) {
    const uint64_t c0 = wire_u64(data + a0 + 1), l0 = wire_u64(data + a0 + 9);
    if (kd_d2bits(h[0]) != c0 || kd_d2bits(h[1]) != l0) atomicMin(&err[2], (unsigned long long)k);
  }
  cnt[k] = c;
  state[0 * nk + k] = h[4];  // stored total
  state[1 * nk + k] = h[2];  // log scale
  state[2 * nk + k] = h[3];  // clock
  state[3 * nk + k] = h[5];  // min
  state[4 * nk + k] = h[6];  // max
}

// One wave per key: its centroid pairs into the CSR.
__global__ void k_wire_pairs(int64_t nk, const uint8_t* data, const int64_t* woff,
                             const int64_t* off, double* cm, double* cw) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nk; k += stride) {
    const int64_t c0 = off[k], c = off[k + 1] - c0;
    const uint8_t* p = data + woff[k] + kWireHeader;
    for (int64_t i = lane; i < c; i += 64) {
      cm[c0 + i] = kd_bits2d(wire_u64(p + 16 * i));
      cw[c0 + i] = kd_bits2d(wire_u64(p + 16 * i + 8));
    }
  }
}
}  // namespace

// Exclusive scan of cnt[0..n) into off[0..n] (off[n] = total).
static void scan_offsets(kd_session* s, const int64_t* cnt, int64_t n, int64_t* off, DevBuf& scratch) {
  const int64_t nb = (n + kOffChunk - 1) / kOffChunk;
  int64_t* bsum = static_cast<int64_t*>(scratch.reserve((nb + 1) * sizeof(int64_t)));
  k_off_sums<<<(unsigned)nb, kOffChunk, 0, s->st()>>>(cnt, n, bsum);
  launched("k_off_sums");
  k_off_top<<<1, kOffChunk, 0, s->st()>>>(bsum, nb, off + n);
  launched("k_off_top");
  k_off_down<<<(unsigned)nb, kOffChunk, 0, s->st()>>>(cnt, n, bsum, off);
  launched("k_off_down");
}

extern "C" int kd_grouped_params(kd_grouped* g, double* compression, double* lambda) {
  *compression = g->compression;
  *lambda = g->lambda;
  return KD_OK;
}

extern "C" int kd_grouped_wire_size(kd_grouped* g, int64_t* out_bytes) {
  return guarded([&] {
    *out_bytes = kWireHeader * g->num_keys + 16 * g->total_cent;
    return KD_OK;
  });
}

static void grouped_to_wire(kd_grouped* g, uint8_t* d_out, int64_t cap, int64_t* d_off) {
  const int64_t nk = g->num_keys;
  const int64_t need = kWireHeader * nk + 16 * g->total_cent;
  if (cap < need) throw KdError(KD_EINVAL, "output buffer smaller than kd_grouped_wire_size");
  kd_session* s = g->s;
  int64_t* sizes = s->buf<int64_t>(s->tmp, nk);
  k_wire_sizes<<<blocks_for(nk), 256, 0, s->st()>>>(nk, g->doff(), sizes);
  launched("k_wire_sizes");
  scan_offsets(s, sizes, nk, d_off, s->tmp2);
  k_wire_write<<<wave_grid(nk), 256, 0, s->st()>>>(nk, g->doff(), g->dm(), g->dw(),
                                                         static_cast<double*>(g->state.p),
                                                         g->compression, g->lambda, d_off, d_out);
  launched("k_wire_write");
}

extern "C" int kd_grouped_to_bytes_device(kd_grouped* g, uint8_t* d_out, int64_t cap, int64_t* d_off) {
  return guarded([&] {
    grouped_to_wire(g, d_out, cap, d_off);
    g->s->sync();
    return KD_OK;
  });
}

extern "C" int kd_grouped_to_bytes_all(kd_grouped* g, uint8_t* out, int64_t cap, int64_t* off) {
  return guarded([&] {
    const int64_t nk = g->num_keys;
    const int64_t need = kWireHeader * nk + 16 * g->total_cent;
    if (cap < need) throw KdError(KD_EINVAL, "output buffer smaller than kd_grouped_wire_size");
    kd_session* s = g->s;
    uint8_t* dout = s->buf<uint8_t>(s->v_buf, need);
    int64_t* doff = s->buf<int64_t>(s->q_off, nk + 1);
    grouped_to_wire(g, dout, need, doff);
    s->d2h(out, dout, need);
    s->d2h(off, doff, (nk + 1) * sizeof(int64_t));
    s->sync();
    return KD_OK;
  });
}

static kd_grouped* grouped_from_wire(kd_session* s, int scale, int64_t nk, const uint8_t* d_data,
                                     int64_t n, const int64_t* d_off) {
  if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
  if (nk <= 0 || nk > 0xffffffffLL) throw KdError(KD_EINVAL, "num_keys out of range");
  std::unique_ptr<kd_grouped> g(new kd_grouped());
  g->s = s;
  g->scale = scale;
  g->num_keys = nk;
  g->stored_total.assign(nk, 0.0);
  g->log_scale.assign(nk, 0.0);
  g->clock.assign(nk, 0.0);
  g->vmin.assign(nk, 0.0);
  g->vmax.assign(nk, 0.0);
  g->hoff.assign(nk + 1, 0);
  g->off.reserve((nk + 1) * sizeof(int64_t));
  g->state.reserve(5 * nk * sizeof(double));
  void* p = nullptr;
  hip_ok(hipHostMalloc(&p, 4 * sizeof(int64_t), hipHostMallocDefault), "hipHostMalloc");
  g->h_tot = static_cast<int64_t*>(p);
  int64_t* cnt = s->buf<int64_t>(s->tmp, nk + 3);
  auto* err = reinterpret_cast<unsigned long long*>(cnt + nk);
  const unsigned long long none[3] = {(unsigned long long)nk, (unsigned long long)nk,
                                      (unsigned long long)nk};
  s->h2d(err, none, sizeof(none));
  k_wire_parse<<<blocks_for(nk), 256, 0, s->st()>>>(nk, d_data, n, d_off, cnt,
                                                    static_cast<double*>(g->state.p), err);
  launched("k_wire_parse");
  unsigned long long herr[3];
  s->d2h(herr, err, sizeof(herr));
  s->sync();
  if (herr[0] < (unsigned long long)nk)
    throw KdError(KD_ETRUNC, "truncated digest for key " + std::to_string(herr[0]));
  if (herr[1] < (unsigned long long)nk)
    throw KdError(KD_ESERDE, "unrecognized digest serialization (version or count) for key " +
                                 std::to_string(herr[1]));
  if (herr[2] < (unsigned long long)nk)
    throw KdError(KD_ESERDE, "key " + std::to_string(herr[2]) +
                                 " has another compression or decay rate than key 0");
  uint8_t h0[17];
  int64_t w0 = 0;
  s->d2h(&w0, d_off, sizeof(int64_t));
  s->sync();
  s->d2h(h0, d_data + w0, sizeof(h0));
  scan_offsets(s, cnt, nk, g->doff(), s->tmp2);
  s->d2h(g->h_tot, g->doff() + nk, sizeof(int64_t));
  s->sync();
  uint64_t cb = 0, lb = 0;
  for (int i = 0; i < 8; i++) {
    cb |= (uint64_t)h0[1 + i] << (8 * i);
    lb |= (uint64_t)h0[9 + i] << (8 * i);
  }
  g->compression = kd_bits2d(cb);
  g->lambda = kd_bits2d(lb);
  g->total_cent = *g->h_tot;
  g->cm.reserve(std::max<int64_t>(g->total_cent, 1) * sizeof(double));
  g->cw.reserve(std::max<int64_t>(g->total_cent, 1) * sizeof(double));
  k_wire_pairs<<<wave_grid(nk), 256, 0, s->st()>>>(nk, d_data, d_off, g->doff(), g->dm(),
                                                         g->dw());
  launched("k_wire_pairs");
  s->sync();
  g->dirty = true;
  g->prefix_valid = false;
  return g.release();
}

extern "C" int kd_grouped_from_bytes_device(kd_session* s, int scale, int64_t num_keys,
                                            const uint8_t* d_data, int64_t n, const int64_t* d_off,
                                            kd_grouped** out) {
  return guarded([&] {
    *out = grouped_from_wire(s, scale, num_keys, d_data, n, d_off);
    return KD_OK;
  });
}

extern "C" int kd_grouped_from_bytes_all(kd_session* s, int scale, int64_t num_keys,
                                         const uint8_t* data, int64_t n, const int64_t* off,
                                         kd_grouped** out) {
  return guarded([&] {
    if (num_keys <= 0 || num_keys > 0x7fffffffLL) throw KdError(KD_EINVAL, "num_keys must be in [1, 2^31 - 1]");
    uint8_t* dd = s->buf<uint8_t>(s->v_buf, std::max<int64_t>(n, 1));
    int64_t* doff = s->buf<int64_t>(s->q_off, num_keys + 1);
    s->h2d(dd, data, n);
    s->h2d(doff, off, (num_keys + 1) * sizeof(int64_t));
    *out = grouped_from_wire(s, scale, num_keys, dd, n, doff);
    return KD_OK;
  });
}

extern "C" int kd_grouped_centroid_count(kd_grouped* g, int64_t key, int64_t* out) {
  return guarded([&] {
    g->check_key(key);
    *out = g->count(key);
    return KD_OK;
  });
}

extern "C" int kd_grouped_total_weight(kd_grouped* g, int64_t key, double* out) {
  return guarded([&] {
    g->check_key(key);
    *out = g->eff(key);
    return KD_OK;
  });
}

extern "C" int kd_grouped_centroids(kd_grouped* g, int64_t key, double* means, double* weights,
                                    int64_t cap, int64_t* count) {
  return guarded([&] {
    g->check_key(key);
    int64_t c = g->count(key);
    *count = c;
    if (cap < c) return KD_OK;
    g->s->d2h(means, g->dm() + g->hoff[key], c * sizeof(double));
    g->s->d2h(weights, g->dw() + g->hoff[key], c * sizeof(double));
    g->s->sync();
    return KD_OK;
  });
}

extern "C" int kd_grouped_key_state(kd_grouped* g, int64_t key, double* o) {
  return guarded([&] {
    g->check_key(key);
    g->pull();
    o[0] = g->stored_total[key];
    o[1] = g->log_scale[key];
    o[2] = g->clock[key];
    o[3] = g->vmin[key];
    o[4] = g->vmax[key];
    return KD_OK;
  });
}

// ================================================ host scale functions
extern "C" double kd_scale_normalizer(int scale, double compression, double n) {
  return kd_normalizer(scale, compression, n);
}

extern "C" double kd_scale_k(int scale, double q, double nrm) { return kd_k(scale, q, nrm); }

extern "C" double kd_add_repeated(double W, double x, int64_t L) { return kd_repeat_add(W, x, L); }

extern "C" int64_t kd_add_sequence(double* S, const double* x, int64_t m, int stop_on, double thr,
                                   double* out, double* stop_val) {
  return kd_running_sum(S, x, m, stop_on, thr, out, stop_val);
}

// Inverse maps of scale.mojo:96-99, 120-124, 152-157, 186-192.
extern "C" double kd_scale_q(int scale, double k, double nrm) {
  switch (scale) {
    case KD_K0: return k / nrm;
    case KD_K1: return (std::sin(k / nrm) + 1.0) / 2.0;
    case KD_K2: {
      double w = std::exp(k / nrm);
      return w / (1.0 + w);
    }
    default:
      if (k <= 0.0) return std::exp(k / nrm) / 2.0;
      return 1.0 - std::exp(-k / nrm) / 2.0;
  }
}

// ======================================================= batched digests
// Many independent digests fed sample by sample (add_at) and reduced by a
// merge_into tree; state resident in device memory (kd_batch.hip).
struct kd_batch {
  kd_session* s;
  int scale;
  int64_t nd;
  double compression, lambda;
  int limit;
  DevBuf st, cm, cw, bm, bw, order, flag;
  int64_t* tmp_order = nullptr;
  // The merge_into tree (flush + one launch per level) captured once into a
  // HIP graph over this object's fixed buffers and replayed on later calls.
  bool use_graph = true;
  hipGraphExec_t tree_exec = nullptr;
  ~kd_batch() {
    if (tree_exec) (void)hipGraphExecDestroy(tree_exec);
  }
};

static std::unique_ptr<kd_batch> new_batch(kd_session* s, int scale, int64_t num_digests,
                                           double compression, double lambda) {
  if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
  if (num_digests <= 0) throw KdError(KD_EINVAL, "num_digests must be positive");
  int64_t limit = (int64_t)compression * 5;  // tdigest.mojo:146
  if (limit < 1 || limit > kd::batch_capacity_buffer())
    throw KdError(KD_EINVAL, "compression out of range for the batched path (5*compression <= " +
                                 std::to_string(kd::batch_capacity_buffer()) + ")");
  std::unique_ptr<kd_batch> b(new kd_batch());
  b->s = s;
  b->scale = scale;
  b->nd = num_digests;
  b->compression = compression;
  b->lambda = lambda;
  b->limit = (int)limit;
  size_t cc = (size_t)num_digests * kd::batch_capacity_centroids();
  size_t bc = (size_t)num_digests * kd::batch_capacity_buffer();
  b->st.reserve(num_digests * sizeof(kd::BatchState));
  b->cm.reserve(cc * sizeof(double));
  b->cw.reserve(cc * sizeof(double));
  b->bm.reserve(bc * sizeof(double));
  b->bw.reserve(bc * sizeof(double));
  hip_ok(hipMemsetAsync(b->st.p, 0, num_digests * sizeof(kd::BatchState), s->st()), "memset");
  s->sync();
  return b;
}

extern "C" int kd_batch_create(kd_session* s, int scale, int64_t num_digests, double compression,
                               double lambda, kd_batch** out) {
  return guarded([&] {
    *out = new_batch(s, scale, num_digests, compression, lambda).release();
    return KD_OK;
  });
}

extern "C" void kd_batch_destroy(kd_batch* b) { delete b; }

extern "C" int kd_batch_params(kd_batch* b, double* compression, double* lambda) {
  *compression = b->compression;
  *lambda = b->lambda;
  return KD_OK;
}

// Map-reduce finalize (test_edge.mojo:183-215, query.mojo:345-369): many
// serialized digests loaded into a batch, for kd_batch_merge_tree to reduce.
namespace {
__device__ void batch_load_one(int64_t d, int lane, int64_t nd, const uint8_t* data,
                               const int64_t* woff, const int64_t* cnt, const double* st5,
                               int64_t capc, kd::BatchState* st, double* cm, double* cw,
                               unsigned long long* over) {
  const int64_t c = cnt[d];
  if (c > capc) {
    if (lane == 0) atomicMin(over, (unsigned long long)d);
    return;
  }
  if (lane == 0) {
    kd::BatchState b;
    b.stored_total = st5[d];
    b.log_scale = st5[nd + d];
    b.clock = st5[2 * nd + d];
    b.vmin = st5[3 * nd + d];
    b.vmax = st5[4 * nd + d];
    b.nc = (int)c;
    b.nb = 0;
    b.err = 0;
    b.pad = 0;
    st[d] = b;
  }
  const uint8_t* p = data + woff[d] + kWireHeader;
  for (int64_t i = lane; i < c; i += 64) {
    cm[d * capc + i] = kd_bits2d(wire_u64(p + 16 * i));
    cw[d * capc + i] = kd_bits2d(wire_u64(p + 16 * i + 8));
  }
}

__global__ void k_batch_load(int64_t nd, const uint8_t* data, const int64_t* woff,
                             const int64_t* cnt, const double* st5, int64_t capc,
                             kd::BatchState* st, double* cm, double* cw,
                             unsigned long long* over) {
  const int lane = threadIdx.x & 63;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t d = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; d < nd; d += stride)
    batch_load_one(d, lane, nd, data, woff, cnt, st5, capc, st, cm, cw, over);
}
}  // namespace

static kd_batch* batch_from_wire(kd_session* s, int scale, int64_t nd, const uint8_t* d_data,
                                 int64_t n, const int64_t* d_off) {
  if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
  if (nd <= 0) throw KdError(KD_EINVAL, "num_digests must be positive");
  int64_t* cnt = s->buf<int64_t>(s->tmp, nd + 4);
  auto* err = reinterpret_cast<unsigned long long*>(cnt + nd);
  double* st5 = s->buf<double>(s->tmp2, 5 * nd);
  const unsigned long long none[4] = {(unsigned long long)nd, (unsigned long long)nd,
                                      (unsigned long long)nd, (unsigned long long)nd};
  s->h2d(err, none, sizeof(none));
  k_wire_parse<<<blocks_for(nd), 256, 0, s->st()>>>(nd, d_data, n, d_off, cnt, st5, err);
  launched("k_wire_parse");
  unsigned long long herr[3];
  s->d2h(herr, err, sizeof(herr));
  s->sync();
  if (herr[0] < (unsigned long long)nd)
    throw KdError(KD_ETRUNC, "truncated digest " + std::to_string(herr[0]));
  if (herr[1] < (unsigned long long)nd)
    throw KdError(KD_ESERDE, "unrecognized digest serialization (version or count) for digest " +
                                 std::to_string(herr[1]));
  if (herr[2] < (unsigned long long)nd)
    throw KdError(KD_ESERDE, "digest " + std::to_string(herr[2]) +
                                 " has another compression or decay rate than digest 0");
  int64_t w0 = 0;
  s->d2h(&w0, d_off, sizeof(int64_t));
  s->sync();
  uint8_t h0[17];
  s->d2h(h0, d_data + w0, sizeof(h0));
  s->sync();
  uint64_t cb = 0, lb = 0;
  for (int i = 0; i < 8; i++) {
    cb |= (uint64_t)h0[1 + i] << (8 * i);
    lb |= (uint64_t)h0[9 + i] << (8 * i);
  }
  std::unique_ptr<kd_batch> b = new_batch(s, scale, nd, kd_bits2d(cb), kd_bits2d(lb));
  const int64_t capc = kd::batch_capacity_centroids();
  k_batch_load<<<wave_grid(nd), 256, 0, s->st()>>>(
      nd, d_data, d_off, cnt, st5, capc, static_cast<kd::BatchState*>(b->st.p),
      static_cast<double*>(b->cm.p), static_cast<double*>(b->cw.p), err + 3);
  launched("k_batch_load");
  unsigned long long hover = 0;
  s->d2h(&hover, err + 3, sizeof(hover));
  s->sync();
  if (hover < (unsigned long long)nd)
    throw KdError(KD_EINVAL, "digest " + std::to_string(hover) + " has more than " +
                                 std::to_string(capc) + " centroids (the batch capacity)");
  return b.release();
}

extern "C" int kd_batch_from_bytes_device(kd_session* s, int scale, int64_t num_digests,
                                          const uint8_t* d_data, int64_t n, const int64_t* d_off,
                                          kd_batch** out) {
  return guarded([&] {
    *out = batch_from_wire(s, scale, num_digests, d_data, n, d_off);
    return KD_OK;
  });
}

extern "C" int kd_batch_from_bytes_all(kd_session* s, int scale, int64_t num_digests,
                                       const uint8_t* data, int64_t n, const int64_t* off,
                                       kd_batch** out) {
  return guarded([&] {
    if (num_digests <= 0) throw KdError(KD_EINVAL, "num_digests must be positive");
    uint8_t* dd = s->buf<uint8_t>(s->v_buf, std::max<int64_t>(n, 1));
    int64_t* doff = s->buf<int64_t>(s->q_off, num_digests + 1);
    s->h2d(dd, data, n);
    s->h2d(doff, off, (num_digests + 1) * sizeof(int64_t));
    *out = batch_from_wire(s, scale, num_digests, dd, n, doff);
    return KD_OK;
  });
}

namespace {
__global__ void k_validate_samples(const double* v, const double* w, int64_t n, int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool bad = !(v[i] - v[i] == 0.0) || !(w[i] > 0.0) || !(w[i] - w[i] == 0.0);
  if (bad) atomicOr(flag, 1);
}
}  // namespace

static void batch_add(kd_batch* b, const int64_t* d_off, const int64_t* h_off, const double* d_v,
                      const double* d_t, const double* d_w) {
  kd_session* s = b->s;
  int64_t n = h_off[b->nd] - h_off[0];
  if (n <= 0) return;
  // Validate every sample before any digest changes (tdigest.mojo:317-322).
  int* flag = static_cast<int*>(b->flag.reserve(sizeof(int)));
  hip_ok(hipMemsetAsync(flag, 0, sizeof(int), s->st()), "memset");
  k_validate_samples<<<blocks_for(n), 256, 0, s->st()>>>(d_v + h_off[0], d_w + h_off[0], n, flag);
  launched("k_validate_samples");
  int hflag = 0;
  s->d2h(&hflag, flag, sizeof(int));
  s->sync();
  if (hflag) throw KdError(KD_EINVAL, "values must be finite and weights positive and finite");
  // Largest streams first (LPT): they carry the longest sequential chains.
  std::vector<int64_t> ord(b->nd);
  for (int64_t d = 0; d < b->nd; d++) ord[d] = d;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
    return (h_off[x + 1] - h_off[x]) > (h_off[y + 1] - h_off[y]);
  });
  int64_t* dord = static_cast<int64_t*>(b->order.reserve(b->nd * sizeof(int64_t)));
  s->h2d(dord, ord.data(), b->nd * sizeof(int64_t));
  int ph = s->eng.phase_begin("replay", 24.0 * (double)n);
  kd::launch_replay(s->st(), b->scale, b->compression, b->lambda, b->limit, dord, b->nd, d_off, d_v,
                    d_t, d_w, static_cast<kd::BatchState*>(b->st.p), static_cast<double*>(b->cm.p),
                    static_cast<double*>(b->cw.p), static_cast<double*>(b->bm.p),
                    static_cast<double*>(b->bw.p));
  launched("k_replay");
  s->eng.phase_end(ph);
  s->sync();
}

extern "C" int kd_batch_add_at_device(kd_batch* b, const int64_t* d_off, const double* d_values,
                                      const double* d_times, const double* d_weights) {
  return guarded([&] {
    std::vector<int64_t> h(b->nd + 1);
    b->s->d2h(h.data(), d_off, (b->nd + 1) * sizeof(int64_t));
    b->s->sync();
    batch_add(b, d_off, h.data(), d_values, d_times, d_weights);
    return KD_OK;
  });
}

extern "C" int kd_batch_add_at(kd_batch* b, const int64_t* off, const double* values,
                               const double* times, const double* weights) {
  return guarded([&] {
    kd_session* s = b->s;
    int64_t n = off[b->nd];
    int64_t* doff = s->buf<int64_t>(s->tmp, b->nd + 1);
    double* dv = s->buf<double>(s->g_vals, n);
    double* dt = s->buf<double>(s->g_valsin, n);
    double* dw = s->buf<double>(s->tmp2, n);
    s->h2d(doff, off, (b->nd + 1) * sizeof(int64_t));
    s->h2d(dv, values, n * sizeof(double));
    s->h2d(dt, times, n * sizeof(double));
    s->h2d(dw, weights, n * sizeof(double));
    batch_add(b, doff, off, dv, dt, dw);
    return KD_OK;
  });
}

static void batch_check_err(kd_batch* b) {
  std::vector<kd::BatchState> h(b->nd);
  b->s->d2h(h.data(), b->st.p, b->nd * sizeof(kd::BatchState));
  b->s->sync();
  for (int64_t d = 0; d < b->nd; d++)
    if (h[d].err)
      throw KdError(KD_EINVAL, "digest " + std::to_string(d) + " exceeds " +
                                   std::to_string(kd::batch_capacity_centroids()) +
                                   " centroids (batched path capacity)");
}

extern "C" int kd_batch_flush(kd_batch* b) {
  return guarded([&] {
    int ph = b->s->eng.phase_begin("flush", 0.0);
    kd::launch_flush(b->s->st(), b->scale, b->compression, b->lambda, b->nd,
                     static_cast<kd::BatchState*>(b->st.p), static_cast<double*>(b->cm.p),
                     static_cast<double*>(b->cw.p), static_cast<double*>(b->bm.p),
                     static_cast<double*>(b->bw.p));
    launched("k_flush");
    b->s->eng.phase_end(ph);
    batch_check_err(b);
    return KD_OK;
  });
}

// Flush of every digest, then the pairwise levels (stride 1, 2, 4, ...).
static void enqueue_tree(kd_batch* b, hipStream_t st) {
  auto* bs = static_cast<kd::BatchState*>(b->st.p);
  auto* cm = static_cast<double*>(b->cm.p);
  auto* cw = static_cast<double*>(b->cw.p);
  auto* bm = static_cast<double*>(b->bm.p);
  auto* bw = static_cast<double*>(b->bw.p);
  kd::launch_flush(st, b->scale, b->compression, b->lambda, b->nd, bs, cm, cw, bm, bw);
  launched("k_flush");
  for (int64_t stride = 1; stride < b->nd; stride *= 2) {
    kd::launch_merge_pairs(st, b->scale, b->compression, b->lambda, b->nd, stride, bs, cm, cw, bm,
                           bw);
    launched("k_merge_pairs");
  }
}

extern "C" int kd_batch_merge_tree(kd_batch* b) {
  return guarded([&] {
    kd_session* s = b->s;
    if (b->use_graph && !b->tree_exec) {
      hip_ok(hipStreamBeginCapture(s->st(), hipStreamCaptureModeThreadLocal), "begin capture");
      hipGraph_t g = nullptr;
      try {
        enqueue_tree(b, s->st());
      } catch (...) {
        (void)hipStreamEndCapture(s->st(), &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_ok(hipStreamEndCapture(s->st(), &g), "end capture");
      hipError_t e = hipGraphInstantiate(&b->tree_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_ok(e, "graph instantiate");
    }
    int ph = s->eng.phase_begin("merge_tree", 0.0);
    if (b->use_graph) hip_ok(hipGraphLaunch(b->tree_exec, s->st()), "graph launch");
    else enqueue_tree(b, s->st());
    s->eng.phase_end(ph);
    batch_check_err(b);
    return KD_OK;
  });
}

extern "C" int kd_batch_set_graph(kd_batch* b, int enable) {
  return guarded([&] {
    b->use_graph = enable != 0;
    return KD_OK;
  });
}

extern "C" int kd_batch_state(kd_batch* b, int64_t d, double* o) {
  return guarded([&] {
    if (d < 0 || d >= b->nd) throw KdError(KD_EINVAL, "digest out of range");
    kd::BatchState h;
    b->s->d2h(&h, static_cast<kd::BatchState*>(b->st.p) + d, sizeof h);
    b->s->sync();
    o[0] = h.stored_total;
    o[1] = h.log_scale;
    o[2] = h.clock;
    o[3] = h.vmin;
    o[4] = h.vmax;
    o[5] = (double)h.nb;
    o[6] = (double)h.err;
    return KD_OK;
  });
}

extern "C" int kd_batch_centroids(kd_batch* b, int64_t d, double* means, double* weights,
                                  int64_t cap, int64_t* count) {
  return guarded([&] {
    if (d < 0 || d >= b->nd) throw KdError(KD_EINVAL, "digest out of range");
    kd::BatchState h;
    b->s->d2h(&h, static_cast<kd::BatchState*>(b->st.p) + d, sizeof h);
    b->s->sync();
    *count = h.nc;
    if (cap < h.nc) return KD_OK;
    int64_t base = d * kd::batch_capacity_centroids();
    b->s->d2h(means, static_cast<double*>(b->cm.p) + base, h.nc * sizeof(double));
    b->s->d2h(weights, static_cast<double*>(b->cw.p) + base, h.nc * sizeof(double));
    b->s->sync();
    return KD_OK;
  });
}

extern "C" int kd_batch_to_tdigest(kd_batch* b, int64_t d, kd_tdigest** out) {
  return guarded([&] {
    if (d < 0 || d >= b->nd) throw KdError(KD_EINVAL, "digest out of range");
    kd::BatchState h;
    b->s->d2h(&h, static_cast<kd::BatchState*>(b->st.p) + d, sizeof h);
    b->s->sync();
    std::unique_ptr<kd_tdigest> t(new_digest(b->s, b->scale, b->compression, b->lambda));
    std::vector<double> m(h.nc), w(h.nc), bmv(h.nb), bwv(h.nb);
    int64_t cb = d * kd::batch_capacity_centroids(), bb = d * kd::batch_capacity_buffer();
    b->s->d2h(m.data(), static_cast<double*>(b->cm.p) + cb, h.nc * sizeof(double));
    b->s->d2h(w.data(), static_cast<double*>(b->cw.p) + cb, h.nc * sizeof(double));
    b->s->d2h(bmv.data(), static_cast<double*>(b->bm.p) + bb, h.nb * sizeof(double));
    b->s->d2h(bwv.data(), static_cast<double*>(b->bw.p) + bb, h.nb * sizeof(double));
    b->s->sync();
    t->upload(m.data(), w.data(), h.nc);
    t->bm = bmv;
    t->bw = bwv;
    t->stored_total = h.stored_total;
    t->log_scale = h.log_scale;
    t->clock = h.clock;
    t->vmin = h.vmin;
    t->vmax = h.vmax;
    *out = t.release();
    return KD_OK;
  });
}

// ================================================================ key map
// Dense ids for sparse uint32 keys, on the device (upstream of the grouped
// ingest: the reference takes dense ids, grouped.mojo:85-104, and builds its
// per-key blocks on the host, gpu/grouped.mojo:454-538).  An open-addressing
// table of 64-bit words (key << 1 | 1; 0 is empty) with at least twice the
// capacity in slots; a key's id is assigned when it is first mapped with
// insert, in order of first occurrence in that batch (a flag per position,
// scanned), so the ids do not depend on thread timing.  A batch that would
// exceed the capacity changes nothing: its pending slots are cleared again
// (slots inserted by one batch never sit inside an older key's probe chain,
// so clearing them keeps every older key reachable).
struct kd_keymap {
  kd_session* s;
  int64_t capacity = 0, size = 0, slots = 0;
  DevBuf word, id, first, rev, slot_of, flag, scan, scratch;
  uint64_t* h_res = nullptr;  // pinned: probe overflow, new keys
  ~kd_keymap() {
    if (h_res) (void)hipHostFree(h_res);
  }
};

namespace {
constexpr int32_t kKmPending = -1;
constexpr uint32_t kKmUnknown = 0xffffffffu;

__device__ inline uint64_t km_hash(uint32_t k, uint64_t mask) {
  return ((uint64_t)k * 0x9E3779B97F4A7C15ull >> 17) & mask;
}

__global__ void k_km_init(uint64_t* word, int32_t* id, unsigned long long* first, int64_t slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  word[i] = 0ull;
  id[i] = kKmPending;
  first[i] = ~0ull;
}

// Finds (or, with insert, claims) each key's slot; a pending slot records
// the first position that mapped it.  No slot is claimed once the map would
// hold more than `capacity` keys (res[1] counts the claims; res[0] is set
// instead), so the table (at least twice the capacity) never fills beyond
// the claims in flight; and at most `slots` probes per key.
__global__ void k_km_probe(const uint32_t* keys, int64_t n, unsigned long long* word,
                           const int32_t* id, unsigned long long* first, int64_t slots, int insert,
                           int64_t size, int64_t capacity, int64_t* slot_of,
                           unsigned long long* res) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  const unsigned long long w = ((unsigned long long)k << 1) | 1ull;
  const uint64_t mask = (uint64_t)slots - 1;
  uint64_t h = km_hash(k, mask);
  int64_t found = -1;
  bool over = false;
  // Table words and counters are read from the L2 (other workgroups claim
  // slots during the launch; a stale cached 0 would only cost an atomic).
  for (int64_t p = 0; p < slots; p++) {
    unsigned long long cur = __hip_atomic_load(&word[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0ull) {
      if (!insert) break;
      const unsigned long long claimed =
          __hip_atomic_load(&res[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (size + (int64_t)claimed >= capacity) {
        // Full: unless this key has just been claimed here by another
        // position (the slot is read again), the batch does not fit.
        cur = atomicCAS(&word[h], 0ull, 0ull);
        if (cur == 0ull) {
          over = true;
          break;
        }
      } else {
        cur = atomicCAS(&word[h], 0ull, w);
        if (cur == 0ull) {
          atomicAdd(&res[1], 1ull);
          cur = w;
        }
      }
    }
    if (cur == w) {
      found = (int64_t)h;
      break;
    }
    h = (h + 1) & mask;
  }
  if (insert && (over || found < 0)) atomicMax(&res[0], 1ull);
  slot_of[i] = found;
  // The first position of a new key: most positions come after the one
  // already recorded and skip the atomic (a hot key's occurrences would
  // otherwise all serialize on one word).
  if (found >= 0 && id[found] == kKmPending) {
    const unsigned long long f =
        __hip_atomic_load(&first[found], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned long long)i < f) atomicMin(&first[found], (unsigned long long)i);
  }
}

// flag[i] = 1 where position i is the first to map a new key.
__global__ void k_km_first(int64_t n, const int64_t* slot_of, const int32_t* id,
                           const unsigned long long* first, int64_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t h = slot_of[i];
  flag[i] = (h >= 0 && id[h] == kKmPending && first[h] == (unsigned long long)i) ? 1 : 0;
}

__global__ void k_km_assign(const uint32_t* keys, int64_t n, const int64_t* slot_of,
                            const int64_t* flag, const int64_t* rank, int64_t base, int32_t* id,
                            unsigned long long* first, uint32_t* rev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int64_t h = slot_of[i];
  id[h] = (int32_t)(base + rank[i]);
  first[h] = ~0ull;
  rev[base + rank[i]] = keys[i];
}

// Clears the slots this batch claimed (still pending).
__global__ void k_km_rollback(int64_t slots, unsigned long long* word, const int32_t* id,
                              unsigned long long* first) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= slots || id[h] != kKmPending) return;
  word[h] = 0ull;
  first[h] = ~0ull;
}

__global__ void k_km_lookup(int64_t n, const int64_t* slot_of, const int32_t* id, uint32_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t h = slot_of[i];
  out[i] = (h >= 0 && id[h] != kKmPending) ? (uint32_t)id[h] : kKmUnknown;
}
}  // namespace

extern "C" int kd_keymap_create(kd_session* s, int64_t capacity, kd_keymap** out) {
  return guarded([&] {
    if (capacity <= 0 || capacity > 0x7fffffffLL)
      throw KdError(KD_EINVAL, "key map capacity must be in [1, 2^31)");
    std::unique_ptr<kd_keymap> m(new kd_keymap());
    m->s = s;
    m->capacity = capacity;
    int64_t slots = 64;
    while (slots < 2 * capacity) slots <<= 1;
    m->slots = slots;
    m->word.reserve(slots * sizeof(uint64_t));
    m->id.reserve(slots * sizeof(int32_t));
    m->first.reserve(slots * sizeof(uint64_t));
    m->rev.reserve(capacity * sizeof(uint32_t));
    k_km_init<<<blocks_for(slots), 256, 0, s->st()>>>(static_cast<uint64_t*>(m->word.p),
                                                      static_cast<int32_t*>(m->id.p),
                                                      static_cast<unsigned long long*>(m->first.p),
                                                      slots);
    launched("k_km_init");
    void* p = nullptr;
    hip_ok(hipHostMalloc(&p, 2 * sizeof(uint64_t), hipHostMallocDefault), "hipHostMalloc");
    m->h_res = static_cast<uint64_t*>(p);
    s->sync();
    *out = m.release();
    return KD_OK;
  });
}

extern "C" void kd_keymap_destroy(kd_keymap* m) { delete m; }

static void keymap_map(kd_keymap* m, const uint32_t* d_keys, int64_t n, uint32_t* d_ids, int insert) {
  kd_session* s = m->s;
  if (n <= 0) return;
  auto* word = static_cast<unsigned long long*>(m->word.p);
  auto* id = static_cast<int32_t*>(m->id.p);
  auto* first = static_cast<unsigned long long*>(m->first.p);
  int64_t* slot_of = static_cast<int64_t*>(m->slot_of.reserve(n * sizeof(int64_t)));
  auto* res = static_cast<unsigned long long*>(m->scratch.reserve(2 * sizeof(uint64_t)));
  hip_ok(hipMemsetAsync(res, 0, 2 * sizeof(uint64_t), s->st()), "memset");
  k_km_probe<<<blocks_for(n), 256, 0, s->st()>>>(d_keys, n, word, id, first, m->slots, insert,
                                                 m->size, m->capacity, slot_of, res);
  launched("k_km_probe");
  if (insert) {
    int64_t* flag = static_cast<int64_t*>(m->flag.reserve(n * sizeof(int64_t)));
    int64_t* rank = static_cast<int64_t*>(m->scan.reserve((n + 1) * sizeof(int64_t)));
    k_km_first<<<blocks_for(n), 256, 0, s->st()>>>(n, slot_of, id, first, flag);
    launched("k_km_first");
    DevBuf bsum;
    scan_offsets(s, flag, n, rank, bsum);
    s->d2h(m->h_res, res, sizeof(uint64_t));
    s->d2h(m->h_res + 1, rank + n, sizeof(int64_t));
    s->sync();
    const int64_t fresh = (int64_t)m->h_res[1];
    if (m->h_res[0] || m->size + fresh > m->capacity) {
      k_km_rollback<<<blocks_for(m->slots), 256, 0, s->st()>>>(m->slots, word, id, first);
      launched("k_km_rollback");
      s->sync();
      throw KdError(KD_ENOMEM, "key map capacity exceeded (" + std::to_string(m->capacity) +
                                   " keys; " + std::to_string(m->size) + " held)");
    }
    k_km_assign<<<blocks_for(n), 256, 0, s->st()>>>(d_keys, n, slot_of, flag, rank, m->size, id,
                                                    first, static_cast<uint32_t*>(m->rev.p));
    launched("k_km_assign");
    m->size += fresh;
  }
  k_km_lookup<<<blocks_for(n), 256, 0, s->st()>>>(n, slot_of, id, d_ids);
  launched("k_km_lookup");
  s->sync();
}

extern "C" int kd_keymap_map_device(kd_keymap* m, const uint32_t* d_keys, int64_t n,
                                    uint32_t* d_ids, int insert) {
  return guarded([&] {
    keymap_map(m, d_keys, n, d_ids, insert);
    return KD_OK;
  });
}

extern "C" int kd_keymap_map(kd_keymap* m, const uint32_t* keys, int64_t n, uint32_t* ids,
                             int insert) {
  return guarded([&] {
    if (n <= 0) return KD_OK;
    kd_session* s = m->s;
    uint32_t* dk = s->buf<uint32_t>(s->tmp, n);
    uint32_t* di = s->buf<uint32_t>(s->tmp2, n);
    s->h2d(dk, keys, n * sizeof(uint32_t));
    keymap_map(m, dk, n, di, insert);
    s->d2h(ids, di, n * sizeof(uint32_t));
    s->sync();
    return KD_OK;
  });
}

extern "C" int kd_keymap_size(kd_keymap* m, int64_t* out) {
  *out = m->size;
  return KD_OK;
}

extern "C" int kd_keymap_keys(kd_keymap* m, uint32_t* out, int64_t cap) {
  return guarded([&] {
    if (cap < m->size) throw KdError(KD_EINVAL, "output smaller than the key map's size");
    m->s->d2h(out, m->rev.p, m->size * sizeof(uint32_t));
    m->s->sync();
    return KD_OK;
  });
}

// ================================================================ serving
struct kd_server {
  kd_session* s;
  std::unique_ptr<kd::Server> srv;
};

extern "C" int kd_server_create(kd_session* s, int scale, int64_t num_keys, double compression,
                                double lambda, int64_t slot_capacity, int64_t mem_budget,
                                kd_server** out) {
  return guarded([&] {
    if (scale < 0 || scale > 3) throw KdError(KD_EINVAL, "scale must be 0..3");
    if (num_keys <= 0) throw KdError(KD_EINVAL, "num_keys must be positive");
    if (slot_capacity <= 0 || slot_capacity > kd::kECap)
      throw KdError(KD_EINVAL, "slot capacity must be in 1.." + std::to_string(kd::kECap));
    if (mem_budget <= 0) throw KdError(KD_EINVAL, "memory budget must be positive");
    std::unique_ptr<kd_server> v(new kd_server());
    v->s = s;
    v->srv.reset(new kd::Server(s->eng, scale, num_keys, compression, lambda, slot_capacity,
                                mem_budget));
    *out = v.release();
    return KD_OK;
  });
}

extern "C" void kd_server_destroy(kd_server* v) { delete v; }

extern "C" int kd_server_submit(kd_server* v, int kind, int64_t key, const double* data, int64_t n,
                                int64_t* id) {
  return guarded([&] {
    if (kind < 0 || kind > 2) throw KdError(KD_EINVAL, "kind must be 0 (ingest), 1 (quantile) or 2 (cdf)");
    if (key < 0 || key >= v->srv->num_keys()) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
    if (n <= 0) throw KdError(KD_EINVAL, "empty request");
    for (int64_t i = 0; i < n; i++) {
      if (kind == 0 && !is_finite_val(data[i]))
        throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
      if (kind == 1 && !(data[i] >= 0.0 && data[i] <= 1.0))
        throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
    }
    *id = v->srv->submit(kind, key, data, n);
    return KD_OK;
  });
}

extern "C" int kd_server_submit_many(kd_server* v, int64_t n, const int* kinds, const int64_t* keys,
                                     const int64_t* off, const double* data, int64_t* ids) {
  return guarded([&] {
    for (int64_t r = 0; r < n; r++) {  // validate every request before queueing any
      int kind = kinds[r];
      int64_t key = keys[r];
      if (kind < 0 || kind > 2) throw KdError(KD_EINVAL, "kind must be 0 (ingest), 1 (quantile) or 2 (cdf)");
      if (key < 0 || key >= v->srv->num_keys()) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
      if (off[r + 1] <= off[r]) throw KdError(KD_EINVAL, "empty request");
      for (int64_t i = off[r]; i < off[r + 1]; i++) {
        if (kind == 0 && !is_finite_val(data[i]))
          throw KdError(KD_EINVAL, "value must be finite (got NaN or inf)");
        if (kind == 1 && !(data[i] >= 0.0 && data[i] <= 1.0))
          throw KdError(KD_ERANGE, "quantile q must be in [0, 1]");
      }
    }
    for (int64_t r = 0; r < n; r++)
      ids[r] = v->srv->submit(kinds[r], keys[r], data + off[r], off[r + 1] - off[r]);
    return KD_OK;
  });
}

extern "C" int kd_server_take_completed(kd_server* v, int64_t* ids, double* t, int64_t cap,
                                        int64_t* count) {
  return guarded([&] {
    *count = v->srv->take_completed(ids, t, cap);
    return KD_OK;
  });
}

extern "C" int kd_server_clock(kd_server* v, double* now) {
  *now = v->srv->now();
  return KD_OK;
}

// Calls that need the server's state at rest.
static void check_not_running(kd_server* v) {
  if (v->srv->running()) throw KdError(KD_EINVAL, "the server's workers are running; stop them first");
}

extern "C" int kd_server_step(kd_server* v, int max_rounds, int64_t* completed) {
  return guarded([&] {
    check_not_running(v);
    *completed = v->srv->step(max_rounds);
    return KD_OK;
  });
}

extern "C" int kd_server_start(kd_server* v, int max_rounds) {
  return guarded([&] {
    if (max_rounds < 1) throw KdError(KD_EINVAL, "max_rounds must be at least 1");
    v->srv->start(max_rounds);
    return KD_OK;
  });
}

extern "C" int kd_server_stop(kd_server* v) {
  return guarded([&] {
    v->srv->stop();
    return KD_OK;
  });
}

extern "C" int kd_server_result(kd_server* v, int64_t id, double* out, int64_t cap, int* done) {
  return guarded([&] {
    std::vector<double> r;
    int err = 0;
    std::string msg;
    *done = v->srv->result(id, r, err, msg, false) ? 1 : 0;
    if (*done) {
      if (err != 0) {
        v->srv->result(id, r, err, msg, true);
        throw KdError(err, msg);
      }
      if (cap < (int64_t)r.size()) throw KdError(KD_EINVAL, "output capacity below the request size");
      if (!r.empty()) memcpy(out, r.data(), r.size() * sizeof(double));
      v->srv->result(id, r, err, msg, true);
    }
    return KD_OK;
  });
}

extern "C" int kd_server_pending(kd_server* v, int64_t* out) {
  *out = v->srv->pending();
  return KD_OK;
}

extern "C" int kd_server_stats(kd_server* v, int64_t* out6) {
  v->srv->stats(out6);
  return KD_OK;
}

extern "C" int kd_server_key_state(kd_server* v, int64_t key, double* out6) {
  return guarded([&] {
    if (key < 0 || key >= v->srv->num_keys()) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
    check_not_running(v);
    v->srv->key_state(key, out6);
    return KD_OK;
  });
}

extern "C" int kd_server_key_centroids(kd_server* v, int64_t key, double* means, double* weights,
                                       int64_t cap, int64_t* count) {
  return guarded([&] {
    if (key < 0 || key >= v->srv->num_keys()) throw KdError(KD_EINVAL, "key out of range [0, num_keys)");
    check_not_running(v);
    std::vector<double> m, w;
    v->srv->key_centroids(key, m, w);
    *count = (int64_t)m.size();
    if (cap >= *count && *count > 0) {
      memcpy(means, m.data(), m.size() * sizeof(double));
      memcpy(weights, w.data(), w.size() * sizeof(double));
    }
    return KD_OK;
  });
}
