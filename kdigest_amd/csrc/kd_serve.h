// kd_serve.h — continuous-batching server over device-resident per-key digests.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kd_engine.h"

namespace kd {

// Failure statuses of a request (the C ABI's KD_EEMPTY and KD_ENOMEM).
constexpr int KD_EEMPTY_SERVE = -2;
constexpr int KD_EDEVICE_SERVE = -6;
constexpr int KD_ENOMEM_SERVE = -7;

class Server {
 public:
  // kind of a request: 0 ingest batch, 1 quantile batch, 2 cdf batch.
  Server(Engine& eng, int scale, int64_t num_keys, double compression, double lambda,
         int64_t slot_cap, int64_t mem_budget);
  ~Server();
  int64_t submit(int kind, int64_t key, const double* data, int64_t n);
  // Admits and runs one continuous batch; returns the number of completed
  // requests.  Not while the workers run.
  int64_t step(int max_rounds);
  // Starts the two worker threads (queries on a stream of their own, ingest
  // rounds of at most max_rounds per key per batch on the engine's stream)
  // that serve requests as they are submitted; stop() waits until every
  // submitted request is complete and joins them.  While they run the
  // session's engine belongs to the server.
  void start(int max_rounds);
  void stop();
  bool running() {
    std::lock_guard<std::mutex> lk(mu_);
    return running_;
  }
  // Copies a completed request's result (query values); false while pending.
  // err: 0, or the status of a request that failed (a query on a key
  // without centroids, a merge beyond the slot capacity, a device error in a
  // worker), with a message.
  bool result(int64_t id, std::vector<double>& out, int& err, std::string& msg, bool erase);
  int64_t pending() {
    std::lock_guard<std::mutex> lk(mu_);
    return (int64_t)(pending_.size() + qq_.size() + iq_.size());
  }
  // Seconds since the server was created (the clock of take_completed).
  double now() const;
  // Ids of requests completed since the last call, with completion times.
  int64_t take_completed(int64_t* ids, double* t, int64_t cap);
  // {steps (worker batches), requests admitted, bytes admitted, largest step
  //  in bytes, ingest rounds, requests still queued}
  void stats(int64_t* out6) {
    std::lock_guard<std::mutex> lk(mu_);
    out6[0] = steps_;
    out6[1] = admitted_;
    out6[2] = admitted_bytes_;
    out6[3] = max_step_bytes_;
    out6[4] = rounds_;
    out6[5] = (int64_t)(pending_.size() + qq_.size() + iq_.size());
  }
  int64_t num_keys() const { return nk_; }
  // A key's state and centroids; not while the workers run.
  void key_state(int64_t key, double* out6);
  void key_centroids(int64_t key, std::vector<double>& m, std::vector<double>& w);

 private:
  struct Request {
    int64_t id = 0;
    int kind = 0;
    int64_t key = 0;
    std::vector<double> data, result;
    bool done = false;
    int err = 0;  // failure status (no result), see result()
    std::string msg;
    // Requests of the same key submitted before this one while the workers
    // run: ingests, queries.
    int64_t need_ing = 0, need_qry = 0;
  };
  void refresh_state(const std::vector<int64_t>& keys);
  void finish(Request& r);
  int64_t request_bytes(const Request& r) const;
  void run_queries(const std::vector<Request*>& queries, hipStream_t st);
  void run_ingest_round(std::vector<Request*>& round);
  void enqueue(Request& r);
  void query_loop();
  void ingest_loop();
  void fail_batch(const std::vector<Request*>& batch, const char* what);

  Engine& eng_;
  int scale_;
  int64_t nk_;
  double compression_, lambda_;
  int64_t cap_, budget_;
  std::vector<double> stored_, logs_, clock_, vmin_, vmax_;
  std::vector<int64_t> count_;
  DevBuf m_, w_, S_, dcount_, dstate_, tmpk_, tmps_, qkind_, qkey_, qx_, qout_, vbuf_, meta_,
      rkey_, rfac_, rbak_, rlist_;
  std::deque<int64_t> pending_;
  std::unordered_map<int64_t, Request> reqs_;
  int64_t next_id_ = 1;
  int64_t last_admitted_bytes_ = 0;
  int64_t steps_ = 0, admitted_ = 0, admitted_bytes_ = 0, max_step_bytes_ = 0, rounds_ = 0;
  std::vector<int64_t> completed_;
  std::vector<double> completed_t_;
  std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
  // Host staging of the asynchronous state and install copies.
  std::vector<int64_t> h_keys_, h_meta_, h_rkeys_;
  std::vector<double> h_state_, h_rfac_;
  std::vector<int> q_kinds_;
  std::vector<int64_t> q_keys_;
  std::vector<double> q_xs_;
  SegVec segv_;
  ResVec resv_;

  // Worker mode.  mu_ guards the request table and queues, the completion
  // list, the counters, key_seq_ and the per-key host state that queries
  // read (count_, extremes).
  std::mutex mu_;
  std::condition_variable cv_, idle_cv_;
  std::thread qthr_, ithr_;
  hipStream_t qstream_ = nullptr;
  bool running_ = false, stop_ = false;
  int max_rounds_ = 1;
  int busy_ = 0;
  std::string worker_err_;
  std::list<Request*> qq_, iq_;  // queued queries, ingests (submission order)
  // Per key: ingests submitted, ingests done, queries submitted, queries done.
  std::vector<int64_t> key_seq_;
};

}  // namespace kd
