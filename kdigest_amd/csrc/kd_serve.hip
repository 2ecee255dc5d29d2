// kd_serve.hip — serving loop over device-resident per-key digests.
//
// Requests (ingest batches and quantile/cdf query batches, each for one key)
// are queued in arrival order.  kd_server_step admits requests in that order
// while their data fits the memory budget, and runs them as one continuous
// batch:
//   * every admitted query runs in one launch against the state the key had
//     before this step's ingests (only queries that precede the key's next
//     ingest in arrival order are admitted);
//   * ingests run in rounds, at most one per key per round, so each request
//     is its own batch merge (B semantics of kdigest/tdigest.mojo:223-315) in
//     its key's arrival order; every round is one batched compress over all
//     its keys.
// Per-key results therefore do not depend on what else shares the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "kd_engine.h"
#include "kd_math.h"
#include "kd_query.h"
#include "kd_serve.h"

namespace kd {

namespace {

// Prefix of the interior walk of quantile/cdf (tdigest.mojo:528-536), per key slot.
__global__ void k_slot_prefix(const int64_t* keys, int64_t nk, int64_t cap, const int64_t* count,
                              const double* w, double* S) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk) return;
  int64_t k = keys[t];
  int64_t a = k * cap, nc = count[k];
  if (nc <= 0) return;
  double acc = w[a] / 2.0;
  S[a] = acc;
  for (int64_t i = 0; i + 1 < nc; i++) {
    acc += (w[a + i] + w[a + i + 1]) / 2.0;
    S[a + i + 1] = acc;
  }
}

// One query (kind 1: quantile tdigest.mojo:485-538, kind 2: cdf :540-587).
__global__ void k_slot_query(const int* kinds, const int64_t* qkeys, const double* xs, int64_t nq,
                             int64_t cap, const int64_t* count, const double* m, const double* w,
                             const double* S, const double* state /* 4 per key */, double* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int64_t k = qkeys[i];
  int64_t a = k * cap, nc = count[k];
  QState st{state[4 * k], state[4 * k + 1], state[4 * k + 2], state[4 * k + 3]};
  out[i] = kinds[i] == 1 ? quantile_one(m + a, w + a, S + a, nc, st, xs[i])
                         : cdf_one(m + a, w + a, S + a, nc, st, xs[i]);
}

// Scatters packed per-key query state (stored total, min, max, eff factor).
__global__ void k_scatter_state(const int64_t* keys, const double* packed, int64_t nk,
                                double* state) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk) return;
  int64_t k = keys[t];
  for (int j = 0; j < 4; j++) state[4 * k + j] = packed[4 * t + j];
}

// Installs each segment's compressed centroids into its key slot.
__global__ void k_install(const int64_t* keys, const int64_t* xoff, const int64_t* counts,
                          int64_t nseg, int64_t cap, const double* em, const double* ew,
                          double* m, double* w, int64_t* count) {
  int64_t sidx = blockIdx.x;
  if (sidx >= nseg) return;
  int64_t k = keys[sidx], c = counts[sidx];
  for (int64_t i = threadIdx.x; i < c; i += blockDim.x) {
    m[k * cap + i] = em[xoff[sidx] + i];
    w[k * cap + i] = ew[xoff[sidx] + i];
  }
  if (threadIdx.x == 0) count[k] = c;
}

// _renormalize of one key per block (tdigest.mojo:366-367): its centroid
// weights times e^(-log_scale).  The weights before it are kept in
// backup[r * cap ...], so a request whose merge then fails can restore them.
__global__ void k_renorm_slots(const int64_t* keys, const double* factor, int64_t nr, int64_t cap,
                               const int64_t* count, double* w, double* backup) {
  const int64_t r = blockIdx.x;
  if (r >= nr) return;
  const int64_t k = keys[r], c = count[k];
  const double f = factor[r];
  for (int64_t i = threadIdx.x; i < c; i += blockDim.x) {
    const double x = w[k * cap + i];
    backup[r * cap + i] = x;
    w[k * cap + i] = x * f;
  }
}

// Puts back the weights k_renorm_slots saved for the renormalized keys whose
// merge failed: list[j] is the index r of such a key in that launch.
__global__ void k_restore_slots(const int64_t* keys, const int64_t* list, int64_t nl, int64_t cap,
                                const int64_t* count, const double* backup, double* w) {
  const int64_t j = blockIdx.x;
  if (j >= nl) return;
  const int64_t r = list[j], k = keys[r], c = count[k];
  for (int64_t i = threadIdx.x; i < c; i += blockDim.x) w[k * cap + i] = backup[r * cap + i];
}

unsigned nblocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

Server::Server(Engine& eng, int scale, int64_t num_keys, double compression, double lambda,
               int64_t slot_cap, int64_t mem_budget)
    : eng_(eng), scale_(scale), nk_(num_keys), compression_(compression), lambda_(lambda),
      cap_(slot_cap), budget_(mem_budget) {
  stored_.assign(nk_, 0.0);
  logs_.assign(nk_, 0.0);
  clock_.assign(nk_, 0.0);
  vmin_.assign(nk_, 0.0);
  vmax_.assign(nk_, 0.0);
  count_.assign(nk_, 0);
  size_t slots = (size_t)nk_ * cap_;
  m_.reserve(slots * sizeof(double));
  w_.reserve(slots * sizeof(double));
  S_.reserve(slots * sizeof(double));
  dcount_.reserve(nk_ * sizeof(int64_t));
  dstate_.reserve(4 * nk_ * sizeof(double));
  check_hip(hipMemsetAsync(dcount_.p, 0, nk_ * sizeof(int64_t), eng_.stream()), "memset");
  std::vector<double> st(4 * nk_);
  for (int64_t k = 0; k < nk_; k++) {
    st[4 * k] = 0.0;
    st[4 * k + 1] = 0.0;
    st[4 * k + 2] = 0.0;
    st[4 * k + 3] = 1.0;
  }
  check_hip(hipMemcpyAsync(dstate_.p, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                           eng_.stream()),
            "H2D state");
  check_hip(hipStreamSynchronize(eng_.stream()), "sync");
}

// Device memory a request occupies while its batch runs: a query's key, x
// and result (plus its kind); an ingest's staged values, and the compress
// workspace of its merge: the merged records and centroid outputs of its
// new values and of the key's prior centroids (at most the slot capacity).
int64_t Server::request_bytes(const Request& r) const {
  const int64_t n = (int64_t)r.data.size();
  return r.kind == 0 ? 8 * n + 24 * (n + cap_) : 28 * n;
}

int64_t Server::submit(int kind, int64_t key, const double* data, int64_t n) {
  Request r;
  r.kind = kind;
  r.key = key;
  r.data.assign(data, data + n);
  int64_t id;
  bool workers;
  {
    std::lock_guard<std::mutex> lk(mu_);
    id = r.id = next_id_++;
    Request& q = reqs_.emplace(id, std::move(r)).first->second;
    workers = running_;
    if (workers) enqueue(q);
    else pending_.push_back(id);
  }
  if (workers) cv_.notify_all();
  return id;
}

// Pushes the query state of `keys` (host scalars, centroid prefix sums) to
// the device.  Asynchronous: the staging vectors are rewritten only after a
// later synchronization (every compress ends with one, and so does step()).
void Server::refresh_state(const std::vector<int64_t>& keys) {
  if (keys.empty()) return;
  std::vector<double>& st = h_state_;
  st.resize(4 * keys.size());
  h_keys_ = keys;
  for (size_t i = 0; i < keys.size(); i++) {
    int64_t k = keys[i];
    st[4 * i] = stored_[k];
    st[4 * i + 1] = vmin_[k];
    st[4 * i + 2] = vmax_[k];
    st[4 * i + 3] = lambda_ > 0.0 ? kd_exp(-logs_[k]) : 1.0;
  }
  int64_t* dk = static_cast<int64_t*>(tmpk_.reserve(keys.size() * sizeof(int64_t)));
  double* dp = static_cast<double*>(tmps_.reserve(st.size() * sizeof(double)));
  check_hip(hipMemcpyAsync(dk, h_keys_.data(), keys.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           eng_.stream()),
            "H2D keys");
  check_hip(hipMemcpyAsync(dp, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                           eng_.stream()),
            "H2D key state");
  k_scatter_state<<<nblocks((int64_t)keys.size()), 256, 0, eng_.stream()>>>(
      dk, dp, (int64_t)keys.size(), static_cast<double*>(dstate_.p));
  launched("k_scatter_state");
  k_slot_prefix<<<nblocks((int64_t)keys.size(), 64), 64, 0, eng_.stream()>>>(
      dk, (int64_t)keys.size(), cap_, static_cast<int64_t*>(dcount_.p),
      static_cast<double*>(w_.p), static_cast<double*>(S_.p));
  launched("k_slot_prefix");
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

void Server::finish(Request& r) {
  r.done = true;
  r.data.clear();
  r.data.shrink_to_fit();
  completed_.push_back(r.id);
  completed_t_.push_back(now());
}

int64_t Server::take_completed(int64_t* ids, double* t, int64_t cap) {
  std::lock_guard<std::mutex> lk(mu_);
  int64_t n = std::min<int64_t>(cap, (int64_t)completed_.size());
  std::copy(completed_.begin(), completed_.begin() + n, ids);
  std::copy(completed_t_.begin(), completed_t_.begin() + n, t);
  completed_.erase(completed_.begin(), completed_.begin() + n);
  completed_t_.erase(completed_t_.begin(), completed_t_.begin() + n);
  return n;
}

static double ms_since(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

int64_t Server::step(int max_rounds) {
  if (running()) throw std::runtime_error("the server's workers are running (kd_server_start)");
  const auto t_step = std::chrono::steady_clock::now();
  // ---- admission (arrival order, memory budget, per-key Q* I{0..R}) ----
  std::vector<Request*> queries;
  std::vector<std::vector<Request*>> rounds(std::max(max_rounds, 1));
  std::unordered_map<int64_t, int> key_ingests;  // admitted ingests per key this step
  std::unordered_map<int64_t, bool> key_blocked;  // a later request must wait for the next step
  int64_t bytes = 0;
  std::deque<int64_t> keep;
  bool budget_full = false;
  int64_t n_empty = 0;
  std::unique_lock<std::mutex> lk(mu_);  // submit() may run on other threads
  for (int64_t id : pending_) {
    Request& r = reqs_.at(id);
    bool admit = false;
    const int64_t need = request_bytes(r);
    if (!budget_full && !key_blocked[r.key]) {
      if (bytes + need > budget_ && bytes > 0) {
        budget_full = true;  // FIFO: nothing behind it jumps the queue
      } else if (r.kind == 0) {
        int& c = key_ingests[r.key];
        if (c < (int)rounds.size()) {
          rounds[c].push_back(&r);
          c++;
          admit = true;
        } else {
          key_blocked[r.key] = true;
        }
      } else if (key_ingests[r.key] == 0) {
        if (count_[r.key] == 0) {
          r.err = KD_EEMPTY_SERVE;  // tdigest.mojo:485-487 raises on an empty digest
          r.msg = "query on a key whose digest is empty";
          finish(r);
          n_empty++;
        } else {
          queries.push_back(&r);
        }
        admit = true;
      } else {
        key_blocked[r.key] = true;  // would need the post-ingest state
      }
    }
    if (admit) {
      bytes += need;
      admitted_++;
    }
    else keep.push_back(id);
  }
  pending_.swap(keep);
  last_admitted_bytes_ = bytes;
  steps_++;
  admitted_bytes_ += bytes;
  max_step_bytes_ = std::max(max_step_bytes_, bytes);
  lk.unlock();
  int64_t done = n_empty;
  try {
    if (!queries.empty()) {
      run_queries(queries, eng_.stream());
      std::lock_guard<std::mutex> g(mu_);
      for (Request* r : queries) finish(*r);
      done += (int64_t)queries.size();
    }
    for (auto& round : rounds) {
      if (round.empty()) continue;
      run_ingest_round(round);
      std::lock_guard<std::mutex> g(mu_);
      for (Request* r : round) finish(*r);
      done += (int64_t)round.size();
    }
    check_hip(hipStreamSynchronize(eng_.stream()), "sync");
  } catch (const std::exception& e) {
    // The step's requests that did not complete fail with the error.
    std::lock_guard<std::mutex> g(mu_);
    auto fail = [&](Request* r) {
      if (r->done) return;
      r->err = KD_EDEVICE_SERVE;
      r->msg = e.what();
      finish(*r);
    };
    for (Request* r : queries) fail(r);
    for (auto& round : rounds)
      for (Request* r : round) fail(r);
    throw;
  }
  if (eng_.profiling()) eng_.host_phase("serve_step", ms_since(t_step));
  return done;
}

// All the given queries in one launch on stream st, results into each
// request (the caller completes them).  Uses the query staging buffers.
void Server::run_queries(const std::vector<Request*>& queries, hipStream_t st) {
  std::vector<int>& kinds = q_kinds_;
  std::vector<int64_t>& qk = q_keys_;
  std::vector<double>& xs = q_xs_;
  kinds.clear();
  qk.clear();
  xs.clear();
  for (const Request* r : queries) {
    for (double x : r->data) {
      kinds.push_back(r->kind);
      qk.push_back(r->key);
      xs.push_back(x);
    }
  }
  int64_t nq = (int64_t)xs.size();
  int* dkind = static_cast<int*>(qkind_.reserve(std::max<int64_t>(nq, 1) * sizeof(int)));
  int64_t* dqk = static_cast<int64_t*>(qkey_.reserve(std::max<int64_t>(nq, 1) * sizeof(int64_t)));
  double* dx = static_cast<double*>(qx_.reserve(std::max<int64_t>(nq, 1) * sizeof(double)));
  double* dout = static_cast<double*>(qout_.reserve(std::max<int64_t>(nq, 1) * sizeof(double)));
  check_hip(hipMemcpyAsync(dkind, kinds.data(), nq * sizeof(int), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(dqk, qk.data(), nq * sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(dx, xs.data(), nq * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
  const bool main = st == eng_.stream();
  int ph = main ? eng_.phase_begin("serve_query", 24.0 * (double)nq) : -1;
  k_slot_query<<<nblocks(nq), 256, 0, st>>>(dkind, dqk, dx, nq, cap_,
                                            static_cast<int64_t*>(dcount_.p),
                                            static_cast<double*>(m_.p), static_cast<double*>(w_.p),
                                            static_cast<double*>(S_.p),
                                            static_cast<double*>(dstate_.p), dout);
  launched("k_slot_query");
  if (main) eng_.phase_end(ph);
  std::vector<double> out(nq);
  check_hip(hipMemcpyAsync(out.data(), dout, nq * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipStreamSynchronize(st), "sync");
  int64_t pos = 0;
  for (Request* r : queries) {
    r->result.assign(out.begin() + pos, out.begin() + pos + (int64_t)r->data.size());
    pos += (int64_t)r->data.size();
  }
}

// One round of ingests (at most one per key): each its own batch merge
// (_gpu_batch_prepare, the compress, _gpu_batch_finish), all in one batched
// compress.  Sets err / msg of a request whose merge does not fit its slot.
// Ends with the round's device work complete.
void Server::run_ingest_round(std::vector<Request*>& round) {
  rounds_++;
  std::sort(round.begin(), round.end(), [](const Request* a, const Request* b) { return a->key < b->key; });
  SegVec& segs = segv_;
  segs.clear();
  std::vector<double> vals;
  std::vector<int64_t> keys;
  std::vector<char> had;
  std::vector<double> sfs;
  struct Saved {
    double stored, logs, clock;
  };
  std::vector<Saved> saved;
  std::vector<int64_t>& rkeys = h_rkeys_;
  std::vector<double>& rfac = h_rfac_;
  std::vector<int64_t> ridx;  // per segment: its index among the renormalized keys, or -1
  rkeys.clear();
  rfac.clear();
  for (const Request* r : round) {
    int64_t k = r->key, n = (int64_t)r->data.size();
    saved.push_back(Saved{stored_[k], logs_[k], clock_[k]});
    // _gpu_batch_prepare (tdigest.mojo:236-296).
    double nclock = clock_[k] + (double)n;
    if (lambda_ > 0.0 && nclock > clock_[k]) {
      logs_[k] += lambda_ * (nclock - clock_[k]);
      if (logs_[k] > 400.0) {  // _renormalize (:359-371)
        double f = kd_exp(-logs_[k]);
        ridx.resize(keys.size(), -1);
        ridx.push_back((int64_t)rkeys.size());
        rkeys.push_back(k);  // the slot's weights are rescaled on the device below
        rfac.push_back(f);
        stored_[k] *= f;
        logs_[k] = 0.0;
      }
    }
    clock_[k] = nclock;
    double sf = lambda_ > 0.0 ? kd_exp(logs_[k]) : 1.0;
    had.push_back(count_[k] > 0);
    stored_[k] += (double)n * sf;
    double effn = lambda_ > 0.0 ? stored_[k] * kd_exp(-logs_[k]) : stored_[k];
    SegDesc sd{};
    sd.v_off = (int64_t)vals.size();
    sd.v_cnt = n;
    sd.e_off = k * cap_;
    sd.e_cnt = count_[k];
    sd.sf = sf;
    sd.stored_n = stored_[k];
    sd.nrm = kd_normalizer(scale_, compression_, effn);
    sd.floor = lambda_ > 0.0 ? 1e-6 * kd_exp(logs_[k]) : 0.0;
    segs.push_back(sd);
    keys.push_back(k);
    sfs.push_back(sf);
    vals.insert(vals.end(), r->data.begin(), r->data.end());
  }
  hipStream_t st = eng_.stream();
  if (!rkeys.empty()) {
    const int64_t nr = (int64_t)rkeys.size();
    int64_t* drk = static_cast<int64_t*>(rkey_.reserve(nr * sizeof(int64_t)));
    double* drf = static_cast<double*>(rfac_.reserve(nr * sizeof(double)));
    check_hip(hipMemcpyAsync(drk, rkeys.data(), nr * sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
    check_hip(hipMemcpyAsync(drf, rfac.data(), nr * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
    double* dbk = static_cast<double*>(rbak_.reserve(nr * cap_ * sizeof(double)));
    k_renorm_slots<<<(unsigned)nr, 64, 0, st>>>(drk, drf, nr, cap_, static_cast<int64_t*>(dcount_.p),
                                                static_cast<double*>(w_.p), dbk);
    launched("k_renorm_slots");
  }
  ridx.resize(keys.size(), -1);
  double* dv = static_cast<double*>(vbuf_.reserve(std::max<size_t>(vals.size(), 1) * sizeof(double)));
  check_hip(hipMemcpyAsync(dv, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
  ResVec& res = resv_;
  eng_.compress(scale_, segs, dv, static_cast<double*>(m_.p), static_cast<double*>(w_.p), res);
  // A merge that would not fit its key's slot fails that request alone:
  // the key keeps its state (a renormalized slot gets its weights back)
  // and nothing is installed for it.
  std::vector<int64_t> ok, restore;
  for (size_t i = 0; i < segs.size(); i++) {
    if (res[i].count > cap_) {
      int64_t k = keys[i];
      stored_[k] = saved[i].stored;
      logs_[k] = saved[i].logs;
      clock_[k] = saved[i].clock;
      if (ridx[i] >= 0) restore.push_back(ridx[i]);
      Request* r = round[i];
      r->err = KD_ENOMEM_SERVE;
      r->msg = "key " + std::to_string(k) + " needs " + std::to_string(res[i].count) +
               " centroids; the slot capacity is " + std::to_string(cap_);
    } else {
      ok.push_back((int64_t)i);
    }
  }
  if (!restore.empty()) {
    const int64_t nl = (int64_t)restore.size();
    int64_t* dl = static_cast<int64_t*>(rlist_.reserve(nl * sizeof(int64_t)));
    check_hip(hipMemcpyAsync(dl, restore.data(), nl * sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
    k_restore_slots<<<(unsigned)nl, 64, 0, st>>>(static_cast<int64_t*>(rkey_.p), dl, nl, cap_,
                                                 static_cast<int64_t*>(dcount_.p),
                                                 static_cast<double*>(rbak_.p), static_cast<double*>(w_.p));
    launched("k_restore_slots");
    check_hip(hipStreamSynchronize(st), "sync");  // `restore` is a local staging vector
  }
  int64_t ns = (int64_t)ok.size();
  std::vector<int64_t>& meta = h_meta_;
  meta.resize(3 * std::max<int64_t>(ns, 1));
  for (int64_t q = 0; q < ns; q++) {
    int64_t i = ok[q];
    meta[q] = keys[i];
    meta[ns + q] = segs[i].x_off;
    meta[2 * ns + q] = res[i].count;
  }
  if (ns > 0) {
    int64_t* dmeta = static_cast<int64_t*>(meta_.reserve(3 * ns * sizeof(int64_t)));
    check_hip(hipMemcpyAsync(dmeta, meta.data(), 3 * ns * sizeof(int64_t), hipMemcpyHostToDevice, st),
              "H2D");
    k_install<<<(unsigned)ns, 64, 0, st>>>(dmeta, dmeta + ns, dmeta + 2 * ns, ns, cap_,
                                           eng_.cent_means(), eng_.cent_weights(),
                                           static_cast<double*>(m_.p), static_cast<double*>(w_.p),
                                           static_cast<int64_t*>(dcount_.p));
    launched("k_install");
  }
  {
    std::lock_guard<std::mutex> lk(mu_);  // count_ and the extremes are read by other threads
    for (int64_t i : ok) {
      int64_t k = keys[i];
      double m = (double)segs[i].v_cnt;
      if (stored_[k] == m * sfs[i] && !had[i]) {  // :285-295
        vmin_[k] = res[i].vmin;
        vmax_[k] = res[i].vmax;
      } else {
        if (res[i].vmin < vmin_[k]) vmin_[k] = res[i].vmin;
        if (res[i].vmax > vmax_[k]) vmax_[k] = res[i].vmax;
      }
      stored_[k] = res[i].total;  // :304-315
      count_[k] = res[i].count;
    }
  }
  // Every key of the round gets its query state rewritten (a failed key's
  // is unchanged, rewriting it keeps the code path single).
  refresh_state(keys);
  check_hip(hipStreamSynchronize(st), "sync");
}

// ------------------------------------------------------------ workers
// Two host threads serve the queue while the caller keeps submitting: the
// query worker answers every query whose key has no earlier ingest still
// outstanding, in batches, on a stream of its own; the ingest worker runs
// ingest rounds (at most max_rounds per key per batch) on the engine's
// stream for every ingest whose key has no earlier request outstanding.  A
// request's per-key predecessors are counted when it is submitted, so each
// key still sees its requests in arrival order (results equal those of the
// synchronous step), while queries of other keys run beside the ingests.
void Server::start(int max_rounds) {
  std::lock_guard<std::mutex> lk(mu_);
  if (running_) return;
  check_hip(hipStreamCreateWithFlags(&qstream_, hipStreamNonBlocking), "hipStreamCreate");
  max_rounds_ = std::max(max_rounds, 1);
  // Requests queued before the start join the workers' queues in order
  // (every request counted in key_seq_ before has completed or is among them).
  key_seq_.assign(4 * nk_, 0);
  for (int64_t id : pending_) enqueue(reqs_.at(id));
  pending_.clear();
  stop_ = false;
  running_ = true;
  worker_err_.clear();
  qthr_ = std::thread([this] { query_loop(); });
  ithr_ = std::thread([this] { ingest_loop(); });
}

Server::~Server() {
  try {
    stop();
  } catch (...) {
  }
}

void Server::stop() {
  {
    std::unique_lock<std::mutex> lk(mu_);
    if (!running_) return;
    idle_cv_.wait(lk, [&] { return qq_.empty() && iq_.empty() && busy_ == 0; });
    stop_ = true;
  }
  cv_.notify_all();
  qthr_.join();
  ithr_.join();
  std::lock_guard<std::mutex> lk(mu_);
  running_ = false;
  // Requests submitted while stop() was under way stay queued for step() or
  // the next start, in submission order.
  for (Request* r : qq_) pending_.push_back(r->id);
  for (Request* r : iq_) pending_.push_back(r->id);
  qq_.clear();
  iq_.clear();
  std::sort(pending_.begin(), pending_.end());
  check_hip(hipStreamDestroy(qstream_), "hipStreamDestroy");
  qstream_ = nullptr;
  if (!worker_err_.empty()) {
    std::string e;
    e.swap(worker_err_);
    throw std::runtime_error(e);
  }
}

// Sequencing of a new request among its key's requests (mu_ held).
void Server::enqueue(Request& r) {
  int64_t* ks = key_seq_.data() + 4 * r.key;  // ingests submitted, done; queries submitted, done
  r.need_ing = ks[0];
  r.need_qry = ks[2];
  if (r.kind == 0) {
    ks[0]++;
    iq_.push_back(&r);
  } else {
    ks[2]++;
    qq_.push_back(&r);
  }
}

// A worker's batch that threw: each request fails with the error (the
// server's first such error is also raised by stop()).
void Server::fail_batch(const std::vector<Request*>& batch, const char* what) {
  std::lock_guard<std::mutex> lk(mu_);
  if (worker_err_.empty()) worker_err_ = what;
  for (Request* r : batch) {
    r->err = KD_EDEVICE_SERVE;
    r->msg = what;
  }
}

void Server::query_loop() {
  check_hip(hipSetDevice(eng_.device()), "hipSetDevice");
  std::vector<Request*> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (stop_) return;
        int64_t bytes = 0;
        for (auto it = qq_.begin(); it != qq_.end();) {
          Request* r = *it;
          const int64_t* ks = key_seq_.data() + 4 * r->key;
          const int64_t need = request_bytes(*r);
          if (ks[1] >= r->need_ing && (bytes == 0 || bytes + need <= budget_)) {
            it = qq_.erase(it);
            if (count_[r->key] == 0) {
              r->err = KD_EEMPTY_SERVE;  // tdigest.mojo:485-487 raises on an empty digest
              r->msg = "query on a key whose digest is empty";
              finish(*r);
              key_seq_[4 * r->key + 3]++;
              admitted_++;
              continue;
            }
            batch.push_back(r);
            bytes += need;
          } else {
            ++it;
          }
        }
        if (!batch.empty()) {
          busy_++;
          admitted_ += (int64_t)batch.size();
          break;
        }
        idle_cv_.notify_all();
        cv_.wait(lk);
      }
    }
    try {
      run_queries(batch, qstream_);
    } catch (const std::exception& e) {
      fail_batch(batch, e.what());
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (Request* r : batch) {
        finish(*r);
        key_seq_[4 * r->key + 3]++;
      }
      busy_--;
    }
    cv_.notify_all();
    idle_cv_.notify_all();
  }
}

void Server::ingest_loop() {
  check_hip(hipSetDevice(eng_.device()), "hipSetDevice");
  std::vector<std::vector<Request*>> rounds;
  for (;;) {
    rounds.assign(max_rounds_, {});
    {
      std::unique_lock<std::mutex> lk(mu_);
      for (;;) {
        if (stop_) return;
        // Per key, the ingests taken this batch (they run in consecutive
        // rounds): each must follow every earlier request of its key.
        std::unordered_map<int64_t, int> taken;
        int64_t bytes = 0;
        int64_t n = 0;
        for (auto it = iq_.begin(); it != iq_.end();) {
          Request* r = *it;
          const int64_t* ks = key_seq_.data() + 4 * r->key;
          int& c = taken[r->key];
          const int64_t need = request_bytes(*r);
          const bool ready = ks[1] + c >= r->need_ing && ks[3] >= r->need_qry;
          if (ready && c < max_rounds_ && (bytes == 0 || bytes + need <= budget_)) {
            it = iq_.erase(it);
            rounds[c].push_back(r);
            c++;
            bytes += need;
            n++;
          } else {
            c = max_rounds_;  // later ingests of this key wait for the next batch
            ++it;
          }
        }
        if (n > 0) {
          busy_++;
          steps_++;
          admitted_ += n;
          admitted_bytes_ += bytes;
          last_admitted_bytes_ = bytes;
          max_step_bytes_ = std::max(max_step_bytes_, bytes);
          break;
        }
        idle_cv_.notify_all();
        cv_.wait(lk);
      }
    }
    for (auto& round : rounds) {
      if (round.empty()) continue;
      try {
        run_ingest_round(round);
      } catch (const std::exception& e) {
        fail_batch(round, e.what());
      }
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (Request* r : round) {
          finish(*r);
          key_seq_[4 * r->key + 1]++;
        }
      }
      cv_.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      busy_--;
    }
    cv_.notify_all();
    idle_cv_.notify_all();
  }
}

bool Server::result(int64_t id, std::vector<double>& out, int& err, std::string& msg, bool erase) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) throw std::runtime_error("unknown request id");
  if (!it->second.done) return false;
  out = it->second.result;
  err = it->second.err;
  msg = it->second.msg;
  if (erase) reqs_.erase(it);
  return true;
}

void Server::key_state(int64_t key, double* o) {
  std::lock_guard<std::mutex> lk(mu_);
  if (running_) throw std::runtime_error("the server's workers are running; stop them first");
  o[0] = stored_[key];
  o[1] = logs_[key];
  o[2] = clock_[key];
  o[3] = vmin_[key];
  o[4] = vmax_[key];
  o[5] = (double)count_[key];
}

void Server::key_centroids(int64_t key, std::vector<double>& m, std::vector<double>& w) {
  if (running()) throw std::runtime_error("the server's workers are running; stop them first");
  m.resize(count_[key]);
  w.resize(count_[key]);
  check_hip(hipMemcpyAsync(m.data(), static_cast<double*>(m_.p) + key * cap_, m.size() * sizeof(double),
                           hipMemcpyDeviceToHost, eng_.stream()), "D2H");
  check_hip(hipMemcpyAsync(w.data(), static_cast<double*>(w_.p) + key * cap_, w.size() * sizeof(double),
                           hipMemcpyDeviceToHost, eng_.stream()), "D2H");
  check_hip(hipStreamSynchronize(eng_.stream()), "sync");
}

}  // namespace kd
