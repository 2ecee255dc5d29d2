// kd_engine.h — internal C++ interface of the gfx950 digest engine.
//
// The engine runs the B-semantics compress of many segments at once.  A
// segment is one digest (or one key of a grouped digest) and holds
//   * explicit records E (means + weights, in input order: centroids ++ buffer),
//   * new values V (contiguous, unsorted) that all carry the weight sf.
// For each segment it produces the greedy-sweep clustering of the stable,
// mean-sorted union E ++ V (kdigest/tdigest.mojo:391-439), the prune of
// tdigest.mojo:441-458 and the stored-total reconcile of tdigest.mojo:298-315.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace kd {

// LDS capacities (doubles) of the bucket sort and of the E sort.
constexpr int kBlock = 256;
constexpr int kCap = 4096;      // one LDS sort tile of new values
constexpr int kECap = 4096;     // explicit records per segment
constexpr int kMaxSpl = 511;    // splitters per partition part
constexpr int kMaxBkt = 2 * kMaxSpl + 1;
constexpr int kTile = 1024;     // records per checkpoint tile of the cluster sums

struct SegDesc {
  int64_t v_off, v_cnt;   // new values in the V buffer
  int64_t e_off, e_cnt;   // explicit records (input order)
  int64_t x_off;          // merged region [x_off, x_off + e_cnt + v_cnt), also centroid output
  int64_t t_off;          // first checkpoint tile of the segment
  int64_t es_off;         // sorted E records (engine-assigned, compact)
  double sf, stored_n, nrm, floor;
  // Engine-assigned path: 0 multi-kernel sort, 1 selection (no full sort;
  // the device resets it to 0 when the segment falls back to sorting),
  // 2 one block per digest (at most kCap merged records).
  int32_t mode = 0;
  int32_t sel = -1;       // index among the selection-path segments
};

struct SegResult {
  int64_t count;    // centroids after prune
  double total;     // stored total after reconcile and prune (B semantics)
  double removed;   // weight dropped by the prune (A semantics subtracts it)
  double vmin, vmax;// min / max of the new values (unused if v_cnt == 0)
  int64_t replays;  // tie decisions resolved by an exact replay of the running mean
  int64_t path;     // 0: multi-kernel sort path, 1: selection, 2: one block per digest
};

// What a compress plan holds, however it was made (on the host from a
// segment table, or on the device for a grouped batch): segment, class and
// selection-path counts and the sizes of the buffers the launches need.
struct PlanSummary {
  int64_t nseg, nsel, n_mode0, n_eseg, max_e_items;
  int64_t n_small[5];  // one-block classes: merged records <= 256, 512, 1024, 2048, 4096
  int64_t t_small, t_mid, sel_t;  // selection tiles: end of class A, of class B, total
  int64_t sel_b, sel_c, sel_m;    // selection bucket / cluster / mark array sizes
  int64_t total_e, total_x, total_t, total_v;
  double sel_v;                   // new values on the selection path
  int64_t bad;                    // 1: invalid batch; 2: too many prior centroids in a digest
  int64_t sample_cls;             // 8 | bit c: some segment samples in class c; 16 / 32: some
                                  // segment counts with the small / large table (0: unknown)
};

// Inputs of the device-side plan of a grouped batch (per-key B-semantics
// prepare of tdigest.mojo:236-296 for every key the batch touches, then the
// segment table, size classes and selection layout of the compress).
struct GroupedPlanArgs {
  int scale;
  double compression, lambda;
  int64_t num_keys;
  const int64_t* kcnt;   // new values per key
  const int64_t* kof;    // first new value of each key (grouped values)
  const int64_t* off;    // prior centroid CSR (num_keys + 1)
  double* w;             // prior centroid weights (scaled in place on renormalize)
  double* total;         // per-key stored total, decay exponent, clock (updated)
  double* logs;
  double* clock;
  const int* bad;        // batch validation flag (no change when set)
  int64_t* seg_key;      // out: key of each segment
  int64_t* seg_of;       // out: segment of each key, or -1
};

// Page-locked host allocator (hipHostMalloc): copies from and to such memory
// run at link speed without staging.  Allocation is slow, so vectors using it
// are kept and reused.
template <class T>
struct PinnedAlloc {
  using value_type = T;
  PinnedAlloc() = default;
  template <class U>
  PinnedAlloc(const PinnedAlloc<U>&) {}
  T* allocate(size_t n) {
    void* p = nullptr;
    if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess || !p)
      throw std::bad_alloc();
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t) { (void)hipHostFree(p); }
  template <class U>
  bool operator==(const PinnedAlloc<U>&) const { return true; }
  template <class U>
  bool operator!=(const PinnedAlloc<U>&) const { return false; }
};

using SegVec = std::vector<SegDesc, PinnedAlloc<SegDesc>>;
using ResVec = std::vector<SegResult, PinnedAlloc<SegResult>>;

// Owning device allocation, grown on demand (contents are not preserved when
// it grows).  Move-only: a copy would free the same pointer twice.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  void swap(DevBuf& o) noexcept {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
  void* reserve(size_t bytes);
  void release();
  ~DevBuf() { release(); }
};

// Checks the launch that was just enqueued (configuration errors surface here,
// named after the kernel).
void launched(const char* kernel);

struct SelSeg;  // selection-path segment (kd_engine.hip)

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

  // Compress segments.  d_v: new values, concatenated in segment order at
  // SegDesc::v_off.  d_em/d_ew: explicit records at SegDesc::e_off.  The
  // caller fills v_cnt/e_cnt/sf/stored_n/nrm/floor; the engine assigns the
  // offsets.  Results stay on the device (cent_means()/cent_weights() at
  // SegDesc::x_off) and per-segment summaries are returned in `res`.
  void compress(int scale, SegVec& segs, const double* d_v,
                const double* d_em, const double* d_ew, ResVec& res,
                bool sync_results = true);

  // The same compress for a grouped batch planned on the device: the plan
  // (per-key prepare, segment table, classes) runs in one kernel and only
  // its summary comes back to the host; the results stay on the device
  // (dev_results(), dev_segments()).  Returns the summary; summary.bad != 0
  // means nothing was changed or launched.  With defer, segments of the
  // selection path that fall back to sorting are left unfinished (the host
  // learns of them only after a sync): grouped_deferred() tells the caller
  // that their results are not final yet, and after its own sync the caller
  // runs finish_grouped(), which finishes them (true) or finds none (false).
  PlanSummary compress_grouped(const GroupedPlanArgs& a, const double* d_v, const double* d_em,
                               const double* d_ew, bool defer = false);
  bool grouped_deferred() const { return defer_pending_; }
  bool finish_grouped(int scale, const PlanSummary& P, const double* d_v, const double* d_em,
                      const double* d_ew);
  const SegDesc* dev_segments() const { return dsegs_; }

  const double* cent_means() const { return static_cast<const double*>(cm_.p); }
  const double* cent_weights() const { return static_cast<const double*>(cw_.p); }
  const double* sorted_records() const { return static_cast<const double*>(x_.p); }
  const SegResult* dev_results() const { return static_cast<const SegResult*>(res_.p); }

  // Key partition of (key, value) pairs: counts per key and values grouped by
  // key.  The radix partition (up to 2^radix_key_bits() keys) keeps every
  // key's values in batch order; the counting partition beyond it does not,
  // and the selection compress of a key's values can differ in the last bit
  // with their order.
  // With d_bad != nullptr and key_partition_validates(n, num_keys), the
  // batch is also validated on the way (a key >= num_keys or a non-finite
  // value sets *d_bad; the grouping is then meaningless but stays in bounds).
  void key_partition(const uint32_t* d_keys, const double* d_vals, int64_t n,
                     int64_t num_keys, int64_t* d_counts, int64_t* d_offsets, double* d_out,
                     int* d_bad = nullptr);
  bool key_partition_validates(int64_t n, int64_t num_keys) const;
  // Launch shape of the radix key partition: items per thread of a scatter
  // sub-tile (4, 8 or 16) and the target number of chunks per pass.
  void set_partition_shape(int items, int64_t chunks) {
    rx_items_ = items;
    rx_chunks_ = chunks;
  }
  int partition_items() const { return rx_items_; }
  // Keys of a large batch estimated (from a sample) to hold at least
  // min_pairs pairs are placed by the first partition pass alone; 0 turns
  // this off.  Results are the same either way.
  void set_heavy_keys(int64_t min_pairs) { heavy_min_ = std::max<int64_t>(min_pairs, 0); }
  int64_t heavy_keys() const { return heavy_min_; }
  // Grouped batches over at most 2^bits keys (bits <= 27) take the radix
  // partition (stable: every key's values keep their batch order); more
  // keys take the counting partition, whose order within a key follows its
  // atomics.
  void set_radix_key_bits(int bits) { rx_max_bits_ = bits; }
  int radix_key_bits() const { return rx_max_bits_; }
  // Tiles of the selection count / gather passes: at least tile_min values,
  // at most tiles_max tiles per segment (one workgroup per tile).
  // One-block compress of small segments on a second stream, overlapping the
  // selection path (off by default: both are bound by the same CU resources
  // on the grouped workloads measured so far).
  void set_overlap(int mode) { overlap_ = mode; }
  int overlap() const { return overlap_; }
  void set_select_shape(int64_t tile_min, int64_t tiles_max) {
    sel_tile_min_ = tile_min;
    sel_tiles_max_ = tiles_max;
  }
  int64_t select_tile_min() const { return sel_tile_min_; }
  int64_t select_tiles_max() const { return sel_tiles_max_; }
  int64_t partition_chunks() const { return rx_chunks_; }

  DevBuf& scratch(int i) { return extra_[i]; }

  // Segments of more than kCap new values are compressed by selection
  // (bucket counts, rank-space chain, gathered boundary buckets) unless this
  // is switched off; both paths give the same clusters.
  void set_select(bool on) { sel_on_ = on; }
  bool select_enabled() const { return sel_on_; }
  // Segments compressed so far by selection ([0]), by the multi-kernel
  // sorting path ([1]) and in one block each ([2]).
  const int64_t* path_counts() const { return path_counts_; }
  // Segments of at most kCap records are compressed in one block each unless
  // this is switched off (same results either way).
  void set_fused(bool on) { fused_on_ = on; }

  // Per-phase device time from HIP events recorded on the engine stream
  // around each kernel group (enabled on demand; off by default).
  struct PhaseStat {
    double ms = 0.0;
    int64_t launches = 0;
    double bytes = 0.0;  // algorithmic HBM bytes of the launches
  };
  void set_profiling(bool on);
  const std::map<std::string, PhaseStat>& phase_stats() { collect(); return stats_; }
  int phase_begin(const char* name, double bytes, hipStream_t on = nullptr);
  void phase_end(int id);
  void collect();
  // Host wall time of a host-side step (recorded as "host:<name>").
  void host_phase(const char* name, double ms) {
    if (!prof_) return;
    PhaseStat& st = stats_[std::string("host:") + name];
    st.ms += ms;
    st.launches += 1;
  }
  bool profiling() const { return prof_; }

  // Consecutive batches of one digest's values (batch values each, the last
  // one shorter), merged one after the other: their splitter samples do not
  // depend on the merges before them, so all of them are sorted up front in
  // one launch (out: kSelSampleMax doubles per batch).  A later compress of
  // one segment with exactly batch b's values takes batch b's sorted sample
  // when set_presorted_sample() names it (null: none).
  static constexpr int kSelSampleMax = 8192;
  void presort_samples(const double* d_v, int64_t n, int64_t batch, double* out);
  void set_presorted_sample(const double* p) { presorted_ = p; }

 private:
  void sort_values(int nseg, const double* d_v, int64_t total_v);
  // Launches of a planned compress: the segment table, small-class lists,
  // selection segments and E-segment list are on the device already.
  // part 0: every launch, deciding on the sort path for selection
  // fallbacks with a sync; 1: the same without that sync (fallen-back
  // segments are left unfinished and defer_pending_ is set, for the caller to
  // run part 2 once it has synchronized); 2: the sort path and finish of the
  // fallen-back segments only.
  void execute(int scale, const PlanSummary& P, const double* d_v, const double* d_em,
               const double* d_ew, int part = 0);
  bool defer_pending_ = false;
  bool fb_count_pending_ = false;  // the fallback count of a deferred grouped batch is unread
  DevBuf plan_sum_, plan_part_;
  PlanSummary* h_plan_ = nullptr;  // page-locked copy of the device plan summary
  const double* presorted_ = nullptr;

  int device_;
  hipStream_t stream_ = nullptr;
  // Second stream (default priority): the two smaller splitter-sample
  // classes run on it beside the largest, and with set_overlap the
  // one-block compress of small segments, overlapping the selection path of
  // large segments; fork_/join_ order it after the shared inputs and before
  // the work that reads its results.
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  int overlap_ = 0;  // small segments: 0 session stream, 1 side stream after the count, 2 side stream from the start
  DevBuf seg_, res_, em_s_, ew_s_, se_, pose_, vs_, x_, cm_, cw_, cs_, cnt_, t1_, t2_;
  DevBuf parts_, parts2_, spl_, nspl_, bcnt_, bcur_, tiles_, tcnt_, fb_, fbl_, route_, tsum_, long_,
      order_, clist_;
  DevBuf extra_[4];
  // Selection path (large segments compressed without a full sort).
  DevBuf sel_seg_, sel_st_, sel_spl_, sel_cnt_, sel_bst_, sel_info_, sel_bnd_, sel_ce_, sel_csl_, sel_cwv_,
      sel_mk_, sel_ml_, sel_mlc_, sel_bix_, sel_mlx_, sel_stm_, sel_tmap_, sel_wv_,
      sel_lcnt_, sel_slist_, sel_clist_;
  unsigned int* h_fallbacks_ = nullptr;  // pinned: selection fallbacks of the last batch
  // compress(): the segment table, E-segment list, small-class lists and
  // selection segments go up in one copy from page-locked staging (a copy
  // from pageable memory holds the host for tens of microseconds), and the
  // results come back through page-locked staging.
  DevBuf up_;
  void* h_up_ = nullptr;
  size_t h_up_cap_ = 0;
  void* h_res_ = nullptr;
  size_t h_res_cap_ = 0;
  // The device plan arrays of the current batch (in up_ after compress(), in
  // seg_ / order_ / small_list_ / sel_seg_ after compress_grouped()).
  SegDesc* dsegs_ = nullptr;
  int* deseg_ = nullptr;
  int* dsmall_ = nullptr;
  SelSeg* dsel_ = nullptr;
  bool sel_on_ = true;
  int64_t sel_tile_min_ = 65536, sel_tiles_max_ = 64;
  int64_t path_counts_[3] = {0, 0, 0};
  bool fused_on_ = true;
  // One-block compress size classes (merged records up to 256, 512, 1024,
  // 2048, 4096): LDS and threads per block grow with the class.
  static constexpr int kSmallClasses = 5;
  std::vector<int> small_[kSmallClasses], h_small_;
  DevBuf small_list_;
  // Radix key partition: pass-1 (and, over 2^18 keys, pass-2) output, tile
  // count matrix, scan partials, segment and tile offsets.
  DevBuf rx_keys_, rx_vals_, rx_mat_, rx_bsum_, rx_seg1_, rx_seg2_, rx_tile_;
  DevBuf rx_keys32_, rx_vals2_;
  int rx_items_ = 8;       // items per thread in a scatter sub-tile (4, 8 or 16)
  int rx_max_bits_ = 27;       // radix partition up to 2^27 keys (three passes)
  int64_t heavy_min_ = 32768;  // pairs that make a key heavy (0: no heavy split)
  DevBuf heavy_cnt_, heavy_list_, heavy_dig_;
  int64_t rx_chunks_ = 4096;  // target number of chunks per pass
  void key_partition_radix(const uint32_t* d_keys, const double* d_vals, int64_t n, int bits,
                           int64_t num_keys, int64_t* d_counts, int64_t* d_offsets, double* d_out,
                           int* d_bad);
  // Host staging that must outlive the asynchronous copies reading it.
  std::vector<char> h_parts_, h_tiles_, h_finals_;
  bool prof_ = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    double bytes;
    hipStream_t on;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, PhaseStat> stats_;
  hipEvent_t get_event();
};

void check_hip(hipError_t e, const char* what);

}  // namespace kd
