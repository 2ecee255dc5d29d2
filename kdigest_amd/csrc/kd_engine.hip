// kd_engine.hip — gfx950 kernels for the FP64 digest path.
//
// Pipeline of one compress (all segments of a call in the same launches):
//   1. k_sort_e         stable LDS sort of each segment's explicit records E
//   2. value sort       segments <= kCap: one LDS bitonic tile; larger ones:
//                       sample-sort partition rounds (splitters from a sorted
//                       sample, equality buckets for repeated splitters) until
//                       every bucket fits one LDS tile, then k_bucket_sort
//   3. k_merge_e        merge-path of sorted E into sorted V (E first on ties)
//   4. k_chain          greedy-sweep cluster boundaries, one lane per segment
//   5. k_tile_sums /    double-double sums of w*x per 1024-record tile, per
//      k_tile_scan      segment checkpoint prefixes, then
//      k_cluster_mean   per-cluster mean = sum / weight
//   6. k_finish         stored-total reconcile and prune per segment
//
// All arithmetic that decides a cluster boundary runs in the same order and
// with the same rounding as the CPU sweep (kdigest/tdigest.mojo:414-435);
// the device is compiled with -ffp-contract=off.
#include "kd_engine.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <type_traits>

#include "kd_exact.h"
#include "kd_math.h"

namespace kd {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

void launched(const char* kernel) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("launch of ") + kernel + ": " + hipGetErrorString(e));
}

void* DevBuf::reserve(size_t bytes) {
  if (bytes <= cap) return p;
  // Growth is geometric (a buffer that has grown grows by half again), so
  // a stream of slowly growing requests reallocates, and with hipFree
  // synchronizes the device, only a few times.
  const size_t old = cap;
  release();
  size_t want = std::max<size_t>(bytes, 256);
  if (old > 0) want = std::max<size_t>(want, old + old / 2);
  check_hip(hipMalloc(&p, want), "hipMalloc");
  cap = want;
  return p;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

// ----------------------------------------------------------------- helpers
// v of lane l (l the same in every lane), read through the scalar unit: a
// lane-serial loop over it is a chain of scalar reads rather than of LDS
// permutes.
__device__ inline double lane_value(double v, int l) {
  const uint64_t b = kd_d2bits(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return kd_bits2d(((uint64_t)hi << 32) | lo);
}
__device__ inline int64_t lane_value(int64_t v, int l) {
  const uint64_t b = (uint64_t)v;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

struct dd {
  double hi, lo;
};

__device__ inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

__device__ inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  double lo = s.lo + t.hi;
  dd u = two_sum(s.hi, lo);
  lo = u.lo + t.lo;
  return two_sum(u.hi, lo);
}

__device__ inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }

__device__ inline dd two_prod(double a, double b) {
  double p = a * b;
  double e = __builtin_fma(a, b, -p);
  return {p, e};
}

// (a.hi + a.lo) / b rounded to double (one correction step).
__device__ inline double dd_div(dd a, double b) {
  double q = a.hi / b;
  dd qb = two_prod(q, b);
  double r = ((a.hi - qb.hi) - qb.lo) + a.lo;
  return q + r / b;
}

// Total order key of a finite double with -0.0 and +0.0 mapped together.
__device__ inline uint64_t ord_key(double x) {
  uint64_t b = (x == 0.0) ? 0ull : kd_d2bits(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline void bitonic_lds(double* s, int npad) {
  for (int k = 2; k <= npad; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < (npad >> 1); t += blockDim.x) {
        int i = 2 * j * (t / j) + (t % j);
        int l = i + j;
        double a = s[i], b = s[l];
        double lo = a < b ? a : b;
        double hi = a < b ? b : a;
        bool asc = (i & k) == 0;
        s[i] = asc ? lo : hi;
        s[l] = asc ? hi : lo;
      }
      __syncthreads();
    }
  }
}

__device__ inline int pow2_at_least(int64_t n) {
  int p = 2;
  while (p < n) p <<= 1;
  return p;
}

// In-place exclusive scan of v[0..m) (m <= blockDim.x, blockDim.x a multiple
// of 64 and at most 1024); returns the total.  All threads must call it.
__device__ inline uint32_t block_excl_scan(uint32_t* v, int m) {
  __shared__ uint32_t wsum[17];
  int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  uint32_t x = tid < m ? v[tid] : 0u;
  uint32_t inc = x;
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; w++) {
      uint32_t t = wsum[w];
      wsum[w] = run;
      run += t;
    }
    wsum[16] = run;
  }
  __syncthreads();
  uint32_t total = wsum[16];
  if (tid < m) v[tid] = inc - x + wsum[wid];
  __syncthreads();
  return total;
}

// ------------------------------------------------------------ LDS sort
// Each thread sorts kPer values in registers (bitonic network), writes them
// as a run, then log2(n / kPer) merge-path rounds double the run length.
// LDS layout is padded by one slot per 16 so a thread's run of 16 doubles
// spreads over the banks.

constexpr int kPadCap = kCap + kCap / 16;

__device__ inline int padi(int e) { return e + (e >> 4); }

template <int PER>
__device__ inline void sort_regs(double (&v)[PER]) {
#pragma unroll
  for (int k = 2; k <= PER; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < PER; i++) {
        int l = i ^ j;
        if (l > i) {
          double a = v[i], b = v[l];
          double lo = __builtin_fmin(a, b), hi = __builtin_fmax(a, b);
          bool asc = (i & k) == 0;
          v[i] = asc ? lo : hi;
          v[l] = asc ? hi : lo;
        }
      }
    }
  }
}

// Sorts n finite doubles held in `a` at padded indices (n <= capacity of a
// and b); returns the buffer (a or b) that holds the sorted values.  PER
// values per thread; all threads of the block must call it.
template <int PER>
__device__ double* lds_sort_t(double* a, double* b, int n) {
  int npad = pow2_at_least(n);
  if (npad < PER) npad = PER;
  const int nthr = npad / PER;
  const int t = threadIdx.x;
  const double inf = __builtin_inf();
  for (int t0 = t; t0 < nthr; t0 += blockDim.x) {
    double v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      int e = t0 * PER + i;
      v[i] = e < n ? a[padi(e)] : inf;
    }
    sort_regs<PER>(v);
#pragma unroll
    for (int i = 0; i < PER; i++) a[padi(t0 * PER + i)] = v[i];
  }
  __syncthreads();
  double* src = a;
  double* dst = b;
  for (int L = PER; L < npad; L <<= 1) {
    for (int t0 = t; t0 < nthr; t0 += blockDim.x) {
      int d0 = t0 * PER;
      int base = d0 & ~(2 * L - 1);
      int d = d0 - base;
      int A = base, B = base + L;
      int lo = d - L > 0 ? d - L : 0;
      int hi = d < L ? d : L;
      while (lo < hi) {  // smallest i with A[i] > B[d-1-i]: ties go to A
        int mid = (lo + hi) >> 1;
        if (src[padi(A + mid)] <= src[padi(B + d - mid - 1)]) lo = mid + 1;
        else hi = mid;
      }
      int i = lo, j = d - lo;
#pragma unroll
      for (int k = 0; k < PER; k++) {
        double x = i < L ? src[padi(A + i)] : inf;
        double y = j < L ? src[padi(B + j)] : inf;
        bool take_a = (j >= L) || (i < L && x <= y);
        dst[padi(d0 + k)] = take_a ? x : y;
        i += take_a ? 1 : 0;
        j += take_a ? 0 : 1;
      }
    }
    __syncthreads();
    double* tmp = src;
    src = dst;
    dst = tmp;
  }
  return src;
}

// Picks the values-per-thread so that every thread of the block has a run
// (at most MAXPER, so a small class keeps its register budget small).
template <int MAXPER = 16>
__device__ inline double* lds_sort(double* a, double* b, int n) {
  int per = pow2_at_least(n) / (int)blockDim.x;
  if (MAXPER <= 2 || per <= 2) return lds_sort_t<2>(a, b, n);
  if (MAXPER <= 4 || per <= 4) return lds_sort_t<(MAXPER < 4 ? MAXPER : 4)>(a, b, n);
  if (MAXPER <= 8 || per <= 8) return lds_sort_t<(MAXPER < 8 ? MAXPER : 8)>(a, b, n);
  return lds_sort_t<16>(a, b, n);
}

// ---------------------------------------------------- 1. explicit records
// Stable sort of n <= 64 PER explicit records by (mean, input index) in one
// wave: PER slots per lane in registers, a bitonic network whose partners
// closer than PER are in the same lane and the others one shuffle away.
// Writes the sorted order to key / idx (slots 0 .. 64 PER - 1).
template <int PER>
__device__ inline void sort_e_wave(const double* em, int n, double* key, int* idx) {
  const int lane = threadIdx.x & 63, base = lane * PER;
  double kk[PER];
  int ii[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int i = base + r;
    kk[r] = i < n ? em[i] : __builtin_inf();
    ii[r] = i < n ? i : 0x7fffffff;
  }
  auto first = [](double a, int ia, double b, int ib) { return (a < b) || (!(b < a) && ia < ib); };
  for (int k = 2; k <= 64 * PER; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < PER) {
#pragma unroll
        for (int r = 0; r < PER; r++) {
#pragma unroll
          for (int jj = 1; jj < PER; jj <<= 1) {
            if (jj == j && (r & jj) == 0) {
              const int r2 = r | jj;
              const bool asc = ((base + r) & k) == 0;
              const bool sw = asc ? !first(kk[r], ii[r], kk[r2], ii[r2]) : first(kk[r], ii[r], kk[r2], ii[r2]);
              if (sw) {
                const double tk = kk[r];
                kk[r] = kk[r2];
                kk[r2] = tk;
                const int ti = ii[r];
                ii[r] = ii[r2];
                ii[r2] = ti;
              }
            }
          }
        }
        continue;
      }
#pragma unroll
      for (int r = 0; r < PER; r++) {
        const double pk = __shfl_xor(kk[r], j / PER, 64);
        const int pi = __shfl_xor(ii[r], j / PER, 64);
        const int p = base + r;
        const bool lower = (p & j) == 0, asc = (p & k) == 0;
        const bool mine = first(kk[r], ii[r], pk, pi);
        if ((lower == asc) ? !mine : mine) {
          kk[r] = pk;
          ii[r] = pi;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < PER; r++) {
    key[base + r] = kk[r];
    idx[base + r] = ii[r];
  }
}

__global__ __launch_bounds__(kBlock) void k_sort_e(const SegDesc* segs, int nseg,
                                                    const double* em, const double* ew,
                                                    double* esm, double* esw, double* se) {
  __shared__ double key[kECap];
  __shared__ int idx[kECap];
  int s = blockIdx.x;
  if (s >= nseg) return;
  SegDesc sd = segs[s];
  int n = (int)sd.e_cnt;
  if (n == 0) {
    if (threadIdx.x == 0) se[sd.es_off + s] = 0.0;
    return;
  }
  int npad = pow2_at_least(n);
  if (n <= 512) {
    // One wave, 2-8 slots per lane in registers: the same network over 64 PER
    // slots, partners across lanes exchanged by shuffles (no barriers).
    if (threadIdx.x < 64) {
      if (n <= 128) sort_e_wave<2>(em + sd.e_off, n, key, idx);
      else if (n <= 256) sort_e_wave<4>(em + sd.e_off, n, key, idx);
      else sort_e_wave<8>(em + sd.e_off, n, key, idx);
    }
    __syncthreads();
    npad = 0;  // sorted: the network below does nothing
  } else {
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
      key[i] = i < n ? em[sd.e_off + i] : __builtin_inf();
      idx[i] = i < n ? i : 0x7fffffff;
    }
    __syncthreads();
  }
  // Bitonic on (mean, input index): the stable order of the reference's
  // insertion sort (tdigest.mojo:64-77).
  for (int k = 2; k <= npad; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < (npad >> 1); t += blockDim.x) {
        int i = 2 * j * (t / j) + (t % j);
        int l = i + j;
        double a = key[i], b = key[l];
        int ia = idx[i], ib = idx[l];
        bool a_first = (a < b) || (!(b < a) && ia < ib);
        bool asc = (i & k) == 0;
        bool swap = asc ? !a_first : a_first;
        if (swap) {
          key[i] = b;
          key[l] = a;
          idx[i] = ib;
          idx[l] = ia;
        }
      }
      __syncthreads();
    }
  }
  // The sorted weights replace the keys in LDS for the prefix below.
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double wv = ew[sd.e_off + idx[i]];
    esm[sd.es_off + i] = em[sd.e_off + idx[i]];
    esw[sd.es_off + i] = wv;
    key[i] = wv;
  }
  __syncthreads();
  // Weight prefix in sorted order, summed one record at a time (as the
  // sweep does) by the first wave through lane broadcasts.
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    double acc = 0.0;
    double* p = se + sd.es_off + s;
    if (lane == 0) p[0] = 0.0;
    for (int c0 = 0; c0 < n; c0 += 64) {
      const double wl = c0 + lane < n ? key[c0 + lane] : 0.0;
      const int cnt = n - c0 < 64 ? n - c0 : 64;
      double mine = 0.0;
      for (int l = 0; l < cnt; l++) {
        acc += lane_value(wl, l);
        if (lane == l) mine = acc;
      }
      if (lane < cnt) p[c0 + lane + 1] = mine;
    }
  }
}

// ------------------------------------------------------ 2. value sorting
struct Part {
  int32_t seg, src;
  int64_t off, cnt;
};
struct PTile {
  int32_t part, len;
  int64_t start;
};
struct FBucket {
  int32_t seg, src, is_eq, pad;
  int64_t off, cnt, rank;
};

constexpr int kPTile = 4096;       // elements per partition tile
constexpr int kPerThread = kPTile / kBlock;

__device__ inline const double* pick(int id, const double* a, const double* b, const double* c) {
  return id == 0 ? a : (id == 1 ? b : c);
}

__global__ __launch_bounds__(kBlock) void k_splitters(const Part* parts, const double* s0,
                                                       const double* s1, const double* s2,
                                                       double* spl, int* nspl) {
  __shared__ double sa[kPadCap], sb[kPadCap];
  Part p = parts[blockIdx.x];
  const double* src = pick(p.src, s0, s1, s2) + p.off;
  int S = (int)std::min<int64_t>(p.cnt, kCap);
  for (int i = threadIdx.x; i < S; i += blockDim.x) sa[padi(i)] = src[((int64_t)i * p.cnt) / S];
  __syncthreads();
  const double* s = lds_sort(sa, sb, S);
  if (threadIdx.x == 0) {
    int64_t want = p.cnt / (kCap / 4);
    int ns = (int)std::max<int64_t>(1, std::min<int64_t>(want, kMaxSpl));
    int k = 0;
    double* out = spl + (int64_t)blockIdx.x * kMaxSpl;
    for (int j = 0; j < ns; j++) {
      double v = s[padi((int)(((int64_t)(j + 1) * S) / (ns + 1)))];
      if (k == 0 || out[k - 1] < v) out[k++] = v;
    }
    nspl[blockIdx.x] = k;
  }
}

// Bucket of x: 2r for the open interval (spl[r-1], spl[r]), 2r+1 for x == spl[r].
__device__ inline int classify(const double* spl, int ns, double x) {
  int lo = 0, hi = ns;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (spl[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < ns && spl[lo] == x) ? 2 * lo + 1 : 2 * lo;
}

__global__ __launch_bounds__(kBlock) void k_bucket_hist(const PTile* tiles, const Part* parts,
                                                         const double* s0, const double* s1,
                                                         const double* s2, const double* spl,
                                                         const int* nspl,
                                                         unsigned long long* bcnt,
                                                         const int64_t* ntiles) {
  __shared__ double lspl[kMaxSpl];
  __shared__ unsigned int lcnt[kMaxBkt];
  if ((int64_t)blockIdx.x >= *ntiles) return;
  PTile t = tiles[blockIdx.x];
  Part p = parts[t.part];
  int ns = nspl[t.part];
  for (int i = threadIdx.x; i < ns; i += blockDim.x) lspl[i] = spl[(int64_t)t.part * kMaxSpl + i];
  for (int i = threadIdx.x; i < kMaxBkt; i += blockDim.x) lcnt[i] = 0;
  __syncthreads();
  const double* src = pick(p.src, s0, s1, s2) + t.start;
  for (int i = threadIdx.x; i < t.len; i += blockDim.x) {
    int b = classify(lspl, ns, src[i]);
    atomicAdd(&lcnt[b], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 2 * ns + 1; b += blockDim.x)
    if (lcnt[b]) atomicAdd(&bcnt[(int64_t)t.part * kMaxBkt + b], (unsigned long long)lcnt[b]);
}

__global__ void k_bucket_scan(const Part* parts, int nparts, const int* nspl,
                              const unsigned long long* bcnt, unsigned long long* bcur) {
  int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= nparts) return;
  int nb = 2 * nspl[pi] + 1;
  unsigned long long run = (unsigned long long)parts[pi].off;
  for (int b = 0; b < nb; b++) {
    bcur[(int64_t)pi * kMaxBkt + b] = run;
    run += bcnt[(int64_t)pi * kMaxBkt + b];
  }
}

__global__ __launch_bounds__(kBlock) void k_bucket_scatter(const PTile* tiles, const Part* parts,
                                                            const double* s0, const double* s1,
                                                            const double* s2, double* dst,
                                                            const double* spl, const int* nspl,
                                                            unsigned long long* bcur,
                                                            const int64_t* ntiles) {
  __shared__ double lspl[kMaxSpl];
  __shared__ unsigned int lcnt[kMaxBkt];
  __shared__ unsigned long long lbase[kMaxBkt];
  if ((int64_t)blockIdx.x >= *ntiles) return;
  PTile t = tiles[blockIdx.x];
  Part p = parts[t.part];
  int ns = nspl[t.part];
  int nb = 2 * ns + 1;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) lspl[i] = spl[(int64_t)t.part * kMaxSpl + i];
  for (int i = threadIdx.x; i < kMaxBkt; i += blockDim.x) lcnt[i] = 0;
  __syncthreads();
  const double* src = pick(p.src, s0, s1, s2) + t.start;
  double xv[kPerThread];
  int bk[kPerThread];
#pragma unroll
  for (int r = 0; r < kPerThread; r++) {
    int i = threadIdx.x + r * kBlock;
    bk[r] = -1;
    if (i < t.len) {
      xv[r] = src[i];
      bk[r] = classify(lspl, ns, xv[r]);
      atomicAdd(&lcnt[bk[r]], 1u);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    unsigned int c = lcnt[b];
    lbase[b] = c ? atomicAdd(&bcur[(int64_t)t.part * kMaxBkt + b], (unsigned long long)c) : 0ull;
    lcnt[b] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPerThread; r++) {
    if (bk[r] >= 0) {
      unsigned int slot = atomicAdd(&lcnt[bk[r]], 1u);
      dst[lbase[bk[r]] + slot] = xv[r];
    }
  }
}

// One final bucket (<= kCap values, or all equal) -> sorted values at the
// segment's destination.
// Device-side routing of the partition rounds.  Counters: [0] small finals,
// [1] large finals, [2] next-round parts, [3] elements in next-round parts.
struct RouteCounters {
  unsigned long long n_small, n_large, n_next, next_elems;
};

constexpr int kSmallCap = 1024;

__device__ inline void push_final(const FBucket& f, FBucket* fs, FBucket* fl, RouteCounters* rc,
                                  int64_t cap) {
  bool small = !f.is_eq && f.cnt <= kSmallCap;
  unsigned long long slot = atomicAdd(small ? &rc->n_small : &rc->n_large, 1ull);
  if ((int64_t)slot < cap) (small ? fs : fl)[slot] = f;
}

// Segments of at most kCap new values are final buckets of their own; larger
// ones become the parts of the first partition round.
__global__ void k_seed(const SegDesc* segs, int nseg, FBucket* fs, FBucket* fl, Part* parts,
                       RouteCounters* rc, int64_t cap) {
  int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nseg) return;
  SegDesc sd = segs[si];
  if (sd.v_cnt == 0 || sd.mode != 0) return;
  if (sd.v_cnt <= kCap) {
    push_final(FBucket{si, 0, 0, 0, sd.v_off, sd.v_cnt, 0}, fs, fl, rc, cap);
  } else {
    unsigned long long slot = atomicAdd(&rc->n_next, 1ull);
    if ((int64_t)slot < cap) parts[slot] = Part{si, 0, sd.v_off, sd.v_cnt};
    atomicAdd(&rc->next_elems, (unsigned long long)sd.v_cnt);
  }
}

__global__ void k_part_tiles(const Part* parts, int np, int64_t* tcnt) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) tcnt[p] = (parts[p].cnt + kPTile - 1) / kPTile;
}

__global__ void k_make_tiles(const Part* parts, const int64_t* toff, PTile* tiles) {
  int p = blockIdx.x;
  Part q = parts[p];
  int64_t nt = (q.cnt + kPTile - 1) / kPTile;
  for (int64_t t = threadIdx.x; t < nt; t += blockDim.x)
    tiles[toff[p] + t] = PTile{p, (int32_t)std::min<int64_t>(kPTile, q.cnt - t * kPTile),
                               q.off + t * kPTile};
}

// After the scatter each bucket cursor points at the bucket's end.  Non-empty
// buckets become final buckets (equality buckets, or at most kCap values) or
// parts of the next round.
__global__ void k_route(const Part* parts, const int* nspl, const unsigned long long* bcnt,
                        const unsigned long long* bcur, const SegDesc* segs, int dst_id,
                        FBucket* fs, FBucket* fl, Part* next, RouteCounters* rc, int64_t cap) {
  int p = blockIdx.x;
  Part q = parts[p];
  int nb = 2 * nspl[p] + 1;
  int64_t v_off = segs[q.seg].v_off;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int64_t c = (int64_t)bcnt[(int64_t)p * kMaxBkt + b];
    if (c == 0) continue;
    int64_t off = (int64_t)bcur[(int64_t)p * kMaxBkt + b] - c;
    if ((b & 1) || c <= kCap) {
      push_final(FBucket{q.seg, dst_id, (b & 1), 0, off, c, off - v_off}, fs, fl, rc, cap);
    } else {
      unsigned long long slot = atomicAdd(&rc->n_next, 1ull);
      if ((int64_t)slot < cap) next[slot] = Part{q.seg, dst_id, off, c};
      atomicAdd(&rc->next_elems, (unsigned long long)c);
    }
  }
}

// Size classes: buckets of up to kSmallCap values sort in one wave with a
// 17 KB LDS footprint (many resident per CU); larger ones in a 256-thread
// block with the full kCap tile.
template <int CAP, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bucket_sort(const FBucket* fbs, const SegDesc* segs,
                                                       const double* s0, const double* s1,
                                                       const double* s2, double* x,
                                                       double* vs) {
  __shared__ double sa[CAP + CAP / 16], sb[CAP + CAP / 16];
  FBucket fb = fbs[blockIdx.x];
  SegDesc sd = segs[fb.seg];
  const double* src = pick(fb.src, s0, s1, s2) + fb.off;
  double* dst = (sd.e_cnt == 0 ? x + sd.x_off : vs + sd.v_off) + fb.rank;
  if (fb.is_eq) {
    for (int64_t i = threadIdx.x; i < fb.cnt; i += blockDim.x) dst[i] = src[i];
    return;
  }
  int n = (int)fb.cnt;
  for (int i = threadIdx.x; i < n; i += blockDim.x) sa[padi(i)] = src[i];
  __syncthreads();
  const double* s = lds_sort(sa, sb, n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = s[padi(i)];
}

// ------------------------------------------------------------ 3. E merge
// Merged order: E records precede new values with an equal mean, which is
// the stable order of centroids ++ buffer ++ values (tdigest.mojo:263-281).
__global__ __launch_bounds__(kBlock) void k_merge_e(const SegDesc* segs, const int* eseg, int neseg,
                                                     const double* vs, const double* esm,
                                                     double* x, int64_t* pose) {
  int si = blockIdx.y;
  if (si >= neseg) return;
  SegDesc sd = segs[eseg[si]];
  if (sd.mode != 0) return;  // compressed by selection or in one block
  int64_t tot = sd.v_cnt + sd.e_cnt;
  const double* e = esm + sd.es_off;
  const double* v = vs + sd.v_off;
  for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < tot;
       it += (int64_t)gridDim.x * blockDim.x) {
    if (it < sd.v_cnt) {
      double xv = v[it];
      int64_t lo = 0, hi = sd.e_cnt;  // # E <= xv
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (!(xv < e[mid])) lo = mid + 1;
        else hi = mid;
      }
      x[sd.x_off + it + lo] = xv;
    } else {
      int64_t i = it - sd.v_cnt;
      double xe = e[i];
      int64_t lo = 0, hi = sd.v_cnt;  // # V < xe
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (v[mid] < xe) lo = mid + 1;
        else hi = mid;
      }
      x[sd.x_off + i + lo] = xe;
      pose[sd.es_off + i] = i + lo;
    }
  }
}

// ------------------------------------------------------------ 4. chain
// Weight of the first j merged records: sf * (#V among them) + sum of the E
// weights among them.  `c` is a cursor (= # E records before j) that moves
// with the queries, so a monotone walk costs O(1) amortized.
struct PCur {
  const int64_t* pos;
  const double* se;
  int64_t ne;
  double sf;
  int64_t c;
  __device__ double P(int64_t j) {
    while (c > 0 && pos[c - 1] >= j) c--;
    while (c < ne && pos[c] < j) c++;
    return sf * (double)(j - c) + se[c];
  }
};

// First E record at or after merged index j (binary search over pos).
__device__ inline int64_t pcur_seek(const PCur& pc, int64_t j) {
  int64_t lo = 0, hi = pc.ne;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (pc.pos[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// A positive normal power of two: its multiples by integers below 2^53 are
// exact, so a running sum of n such weights is exactly n * w.
__device__ inline bool pow2_weight(double w) {
  const uint64_t b = kd_d2bits(w);
  return (b & 0x000fffffffffffffull) == 0 && ((b >> 52) & 0x7ff) != 0 && ((b >> 52) & 0x7ff) != 0x7ff;
}

// Running weight of the merged records [a, b) in the CPU's order (new values
// weigh sf, the prior record at merged position pos[c] weighs ew[c]); ce is
// the first prior record at or after a.  O(prior records + binades).
__device__ inline double run_weight(int64_t a, int64_t b, double sf, const int64_t* pos,
                                    const double* ew, int64_t ne, int64_t ce) {
  double W;
  if (ce < ne && pos[ce] == a) {
    W = ew[ce];
    ce++;
  } else {
    W = sf;
  }
  int64_t j = a + 1;
  while (j < b) {
    int64_t e = ce < ne ? pos[ce] : b;
    if (e > b) e = b;
    if (e > j) {
      W = kd_repeat_add(W, sf, e - j);
      j = e;
    }
    if (j < b) {
      W = W + ew[ce];
      ce++;
      j++;
    }
  }
  return W;
}

// Weight of merged record j.
__device__ inline double rec_weight(int64_t j, double sf, const int64_t* pos, const double* ew,
                                    int64_t ne) {
  if (ne == 0) return sf;
  int64_t lo = 0, hi = ne;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (pos[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return (lo < ne && pos[lo] == j) ? ew[lo] : sf;
}

__device__ inline double kinv_est(int scale, double k, double nrm) {
  double a = k / nrm;
  switch (scale) {
    case KD_K0: return a;
    case KD_K1: {
      const double h = 1.5707963267948966;
      if (a >= h) return 1.0;
      if (a <= -h) return 0.0;
      return (sin(a) + 1.0) * 0.5;
    }
    case KD_K2: return 1.0 / (1.0 + exp(-a));
    default: return a <= 0.0 ? exp(a) * 0.5 : 1.0 - exp(-a) * 0.5;
  }
}

// Out-of-line copies of the scale function, its inverse estimate and the
// sweep's test for the long kernels (chain walk, one-block compress).
// Inlined at each call site (every branch of the scale switch, with its
// logarithms and divisions) they made such a kernel's code 50-100 KB, near or
// above the size of the instruction cache.
__device__ __noinline__ double chain_k(int scale, double q, double nrm) { return kd_k(scale, q, nrm); }
__device__ __noinline__ double chain_kinv(int scale, double k, double nrm) {
  return kinv_est(scale, k, nrm);
}
__device__ __noinline__ int chain_closes(int scale, double nrm, double stored_n, double wsf, double cur,
                                         double w) {
  return kd_sweep_closes(scale, nrm, stored_n, wsf, cur, w);
}

// Smallest j with P(j) >= T (an estimate: the exact predicate search below
// corrects any rounding slack).
__device__ int64_t index_for_weight(PCur& pc, int64_t n, double T) {
  if (pc.ne == 0) {
    double j = ceil(T / pc.sf);
    if (!(j > 0.0)) return 0;
    if (j > (double)n) return n;
    return (int64_t)j;
  }
  // Largest E record index c with weight-before(c) < T, found by galloping
  // from the cursor.
  auto before = [&](int64_t c) { return pc.sf * (double)(pc.pos[c] - c) + pc.se[c]; };
  int64_t lo = -1, hi = pc.ne;  // before(lo) < T <= before(hi) (virtual ends)
  int64_t c0 = std::min<int64_t>(std::max<int64_t>(pc.c, 0), pc.ne - 1);
  if (before(c0) < T) {
    lo = c0;
    int64_t step = 1;
    while (lo + step < pc.ne && before(lo + step) < T) {
      lo += step;
      step <<= 1;
    }
    hi = std::min<int64_t>(lo + step, pc.ne);
  } else {
    hi = c0;
    int64_t step = 1;
    while (hi - step >= 0 && !(before(hi - step) < T)) {
      hi -= step;
      step <<= 1;
    }
    lo = std::max<int64_t>(hi - step, -1);
  }
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (before(mid) < T) lo = mid;
    else hi = mid;
  }
  // Records before E[lo+1]: V run after E[lo] (or from the start).
  int64_t base_j = lo < 0 ? 0 : pc.pos[lo] + 1;
  double base_w = lo < 0 ? 0.0 : pc.sf * (double)(pc.pos[lo] - lo) + pc.se[lo + 1];
  if (base_w >= T) return base_j;
  double j = (double)base_j + ceil((T - base_w) / pc.sf);
  int64_t lim = lo + 1 < pc.ne ? pc.pos[lo + 1] : n;
  if (j > (double)lim) return lim;
  return (int64_t)j;
}

// First index b > s whose record closes the cluster started at s in the
// greedy sweep (tdigest.mojo:424-435): the smallest b with
// NOT (k(P(b+1)/N) - k(P(s)/N) <= 1), or n when none does.  Found from the
// inverse scale function, then an exact gallop and binary search on the
// monotone predicate (P grows with the index).  Ties are not handled here.
__device__ inline int64_t chain_next(int scale, PCur& pc, int64_t s, int64_t n, double N,
                                     double nrm) {
  double ks = kd_k(scale, pc.P(s) / N, nrm);
  // pred(i): candidate i closes the cluster started at s, i.e. NOT
  // (k(q_right) - k(q_left) <= 1) with q_right = P(i+1)/N.
  auto pred = [&](int64_t i) -> bool {
    double span = kd_k(scale, pc.P(i + 1) / N, nrm) - ks;
    return !(span <= 1.0);
  };
  // Estimate from the inverse scale function, then an exact gallop.
  int64_t g;
  double qt = kinv_est(scale, ks + 1.0, nrm);
  if (!(qt < 1.0)) g = n - 1;
  else if (!(qt > 0.0)) g = s + 1;
  else g = index_for_weight(pc, n, qt * N) - 1;
  if (g < s + 1) g = s + 1;
  if (g > n - 1) g = n - 1;
  int64_t lo, hi;  // invariant: pred(lo) false (or lo == s), pred(hi) true (or hi == n)
  if (pred(g)) {
    hi = g;
    int64_t step = 1;
    lo = g - step;
    while (lo >= s + 1 && pred(lo)) {
      hi = lo;
      step <<= 1;
      lo = hi - step;
    }
    if (lo < s) lo = s;
  } else {
    lo = g;
    int64_t step = 1;
    hi = g + step;
    while (hi <= n - 1 && !pred(hi)) {
      lo = hi;
      step <<= 1;
      hi = lo + step;
    }
    if (hi > n) hi = n;  // hi == n is a virtual "true"
  }
  while (hi - lo > 1) {
    int64_t mid = lo + ((hi - lo) >> 1);
    if (mid >= n) {
      hi = mid;
      continue;
    }
    if (pred(mid)) hi = mid;
    else lo = mid;
  }
  return hi;
}

// End of the cluster started at s whose first closing candidate b < n has a
// value within 2 ulps of the record before it: the tie rule
// (tdigest.mojo:429, candidate.mean == current.mean absorbs) may extend it.
__device__ inline int64_t tie_extend(const double* x, int64_t n, int64_t s, int64_t b,
                                     const PCur& pc, const double* esw_seg, int64_t& replays) {
  const double xb = x[b];
  int64_t e = b;
  if (x[s] == xb) {
    // Every record in [s, b] equals xb: the running mean is exactly xb and
    // stays so; absorb the whole run of xb.
    int64_t l2 = b, h2 = n;
    while (l2 < h2) {
      int64_t mid = (l2 + h2) >> 1;
      if (x[mid] == xb) l2 = mid + 1;
      else h2 = mid;
    }
    e = l2;
  } else {
    // Replay the incremental mean of centroid.mojo:34-36 exactly.
    replays++;
    int64_t ce = 0;
    while (ce < pc.ne && pc.pos[ce] < s) ce++;
    double mean = 0.0, wt = 0.0;
    for (int64_t j = s; j < b; j++) {
      double w = pc.sf;
      if (ce < pc.ne && pc.pos[ce] == j) {
        w = esw_seg[ce];
        ce++;
      }
      if (j == s) {
        mean = x[j];
        wt = w;
      } else {
        double nw = wt + w;
        mean += (x[j] - mean) * (w / nw);
        wt = nw;
      }
    }
    while (e < n && x[e] == mean) e++;
  }
  return e;
}

// The greedy sweep of tdigest.mojo:414-435 over one segment's merged records
// x[0..n) (sorted; any address space).  next(s) gives the first candidate
// closing the cluster started at s (or n).  Writes each cluster's first
// record index to cs[c], calls done(a, b, c) for every cluster [a, b) but the
// last single record, and returns the cluster count.
template <class Next, class Done>
__device__ inline int64_t sweep_segment(const double* x, int64_t n, PCur pc,
                                        const double* esw_seg, int64_t* cs, int64_t& replays,
                                        Next next, Done done) {
  int64_t s = 0, c = 0;
  while (s < n) {
    cs[c] = s;
    c++;
    if (s == n - 1) break;
    int64_t b = next(s);  // first closing candidate, or n
    int64_t e = b;
    if (b < n && ord_key(x[b]) - ord_key(x[b - 1]) <= 2) e = tie_extend(x, n, s, b, pc, esw_seg, replays);
    done(s, e, c - 1);
    s = e;
  }
  return c;
}

__global__ __launch_bounds__(64) void k_chain(const SegDesc* segs, int nseg, int scale,
                                               const double* X, const int64_t* pose,
                                               const double* se, const double* esw,
                                               int64_t* cs, int64_t* ccount,
                                               SegResult* res, int64_t* long_list,
                                               unsigned long long* long_cnt) {
  int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nseg) return;
  SegDesc sd = segs[si];
  if (sd.mode != 0) return;  // compressed by selection or in one block
  int64_t n = sd.v_cnt + sd.e_cnt;
  if (n == 0) {
    ccount[si] = 0;
    res[si].replays = 0;
    return;
  }
  PCur pc{pose + sd.es_off, se + sd.es_off + si, sd.e_cnt, sd.sf, 0};
  int64_t replays = 0;
  const int64_t* pos = pose + sd.es_off;
  const double* ew = esw + sd.es_off;
  const int64_t ne = sd.e_cnt;
  // The closed-form prefix decides exactly when every weight is an integer
  // (or a power of two with no prior records); otherwise each boundary is
  // moved to where the CPU's running sums put it (kd_exact.h), with the
  // finished clusters' weights summed in order in wsf.
  bool closed_exact = ne == 0 ? pow2_weight(sd.sf) : (sd.sf == 1.0 && sd.stored_n < 0x1p53);
  for (int64_t i = 0; closed_exact && i < ne; i++) closed_exact = ew[i] == floor(ew[i]);
  double wsf = 0.0;
  auto next = [&](int64_t s) {
    int64_t b = chain_next(scale, pc, s, n, sd.stored_n, sd.nrm);
    if (closed_exact) return b;
    const int64_t ce = pcur_seek(pc, s);
    auto closes = [&](int64_t i) {
      const double cur = run_weight(s, i, sd.sf, pos, ew, ne, ce);
      return kd_sweep_closes(scale, sd.nrm, sd.stored_n, wsf, cur, rec_weight(i, sd.sf, pos, ew, ne)) != 0;
    };
    while (b < n && !closes(b)) b++;
    while (b - 1 > s && closes(b - 1)) b--;
    return b;
  };
  // Clusters spanning three or more checkpoint tiles are summed by
  // k_long_sum (one block each); the rest by k_cluster_mean.
  int64_t c = sweep_segment(X + sd.x_off, n, pc, esw + sd.es_off, cs + sd.x_off, replays, next,
                            [&](int64_t a, int64_t b, int64_t ci) {
                              if (!closed_exact)
                                wsf = wsf + run_weight(a, b, sd.sf, pos, ew, ne, pcur_seek(pc, a));
                              if (b / kTile - a / kTile >= 2) {
                                unsigned long long slot = atomicAdd(long_cnt, 1ull);
                                long_list[2 * slot] = si;
                                long_list[2 * slot + 1] = ci;
                              }
                            });
  ccount[si] = c;
  res[si].replays = replays;
}

// ----------------------------------------------------------- 5. sums
__device__ inline int find_seg_by_tile(const SegDesc* segs, int nseg, int64_t t) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (segs[mid].t_off <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_tile_sums(const SegDesc* segs, int nseg,
                                                       const double* X, const int64_t* pose,
                                                       const double* esw, double* tsum) {
  __shared__ double wl[kTile];
  __shared__ double rh[kBlock], rl[kBlock];
  int64_t t = blockIdx.x;
  int si = find_seg_by_tile(segs, nseg, t);
  SegDesc sd = segs[si];
  if (sd.mode != 0) return;  // compressed by selection or in one block
  int64_t n = sd.v_cnt + sd.e_cnt;
  int64_t r0 = (t - sd.t_off) * kTile;
  int len = (int)std::min<int64_t>(kTile, n - r0);
  for (int i = threadIdx.x; i < kTile; i += blockDim.x) wl[i] = sd.sf;
  __syncthreads();
  if (sd.e_cnt > 0) {
    const int64_t* pos = pose + sd.es_off;
    // E records inside [r0, r0 + len).
    int64_t lo = 0, hi = sd.e_cnt;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (pos[mid] < r0) lo = mid + 1;
      else hi = mid;
    }
    for (int64_t c = lo + threadIdx.x; c < sd.e_cnt && pos[c] < r0 + len; c += blockDim.x)
      wl[pos[c] - r0] = esw[sd.es_off + c];
  }
  __syncthreads();
  dd acc{0.0, 0.0};
  for (int i = threadIdx.x; i < len; i += blockDim.x)
    acc = dd_add(acc, two_prod(X[sd.x_off + r0 + i], wl[i]));
  rh[threadIdx.x] = acc.hi;
  rl[threadIdx.x] = acc.lo;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      dd a{rh[threadIdx.x], rl[threadIdx.x]};
      dd b{rh[threadIdx.x + w], rl[threadIdx.x + w]};
      dd c = dd_add(a, b);
      rh[threadIdx.x] = c.hi;
      rl[threadIdx.x] = c.lo;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    tsum[2 * t] = rh[0];
    tsum[2 * t + 1] = rl[0];
  }
}


__global__ void k_scan_counts(const int64_t* cnt, int n, int64_t* base, int64_t* total) {
  // Single block, chunked exclusive scan.
  __shared__ int64_t part[1024];
  int tid = threadIdx.x;
  int chunk = (n + blockDim.x - 1) / blockDim.x;
  int lo = std::min(n, tid * chunk), hi = std::min(n, lo + chunk);
  int64_t s = 0;
  for (int i = lo; i < hi; i++) s += cnt[i];
  part[tid] = s;
  __syncthreads();
  if (tid == 0) {
    int64_t run = 0;
    for (int i = 0; i < (int)blockDim.x; i++) {
      int64_t v = part[i];
      part[i] = run;
      run += v;
    }
    *total = run;
  }
  __syncthreads();
  int64_t run = part[tid];
  for (int i = lo; i < hi; i++) {
    base[i] = run;
    run += cnt[i];
  }
}

// Sum over [a, b) of w*x (double-double) and of w (plain, in record order:
// the running cluster weight of centroid.mojo:34-36); E positions through the
// cursor.
// When `run_mean` is given, the running mean of the cluster is replayed with
// the exact recurrence of centroid.mojo:34-36 as well.
// Records per chunk of the canonical double-double cluster sum: the sum of
// w*x over [a, b) is the in-order sum of the chunk sums over [a + 16k,
// a + 16k + 16), each summed in order, so one lane or many lanes (one chunk
// each) produce the same bits.
constexpr int64_t kSumChunk = 16;

__device__ inline dd direct_sum(const double* x, PCur& pc, const double* esw_seg, int64_t a,
                                int64_t b, double* wsum, double* run_mean = nullptr) {
  dd acc{0.0, 0.0}, part{0.0, 0.0};
  double ws = 0.0, mean = 0.0;
  int64_t c = pc.c;
  while (c > 0 && pc.pos[c - 1] >= a) c--;
  while (c < pc.ne && pc.pos[c] < a) c++;
  for (int64_t j = a; j < b; j++) {
    double w = pc.sf;
    if (c < pc.ne && pc.pos[c] == j) {
      w = esw_seg[c];
      c++;
    }
    double xj = x[j];
    if (run_mean) {
      if (j == a) {
        mean = xj;
      } else {
        double nw = ws + w;
        mean += (xj - mean) * (w / nw);
      }
    } else {
      part = dd_add(part, two_prod(xj, w));
      if ((j - a) % kSumChunk == kSumChunk - 1 || j == b - 1) {
        acc = dd_add(acc, part);
        part = dd{0.0, 0.0};
      }
    }
    ws = (j == a) ? w : ws + w;
  }
  pc.c = c;
  *wsum = ws;
  if (run_mean) *run_mean = mean;
  return acc;
}

// Clusters up to this many records replay the running centroid of the CPU
// sweep record by record (bit-identical means); longer ones use an exact
// double-double sum (means within an ulp or two).
constexpr int64_t kReplayMax = 256;

// The running mean of centroid.mojo:34-36 never moves off a value that every
// record of the cluster shares (x - mean is 0), so a run of equal records
// keeps that value exactly; a double-double quotient with fractional weights
// may land an ulp away, and the tie rule of the next sweep (tdigest.mojo:429)
// compares means for equality.  Zeros: after the second record the running
// mean is +0 whatever the signs.  Sorted input: x[a] == x[b-1] means all equal.
__device__ inline bool equal_run_mean(const double* x, int64_t a, int64_t b, double* mean) {
  double xa = x[a];
  if (!(xa == x[b - 1])) return false;
  *mean = (xa == 0.0 && b - a >= 2) ? 0.0 : xa;
  return true;
}

// One thread per cluster (grid-stride over the device-side cluster total);
// clusters spanning three or more tiles are left to k_long_sum.
__global__ void k_cluster_mean(const SegDesc* segs, int nseg, const int64_t* cbase,
                               const int64_t* ccount, const int64_t* cs, const double* X,
                               const int64_t* pose, const double* se, const double* esw,
                               double* cm, double* cw, const int64_t* total_p) {
  const int64_t total = *total_p;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {  // last segment whose first cluster index is <= g
      int mid = (lo + hi + 1) >> 1;
      if (cbase[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    int si = lo;
    SegDesc sd = segs[si];
    if (sd.mode != 0) continue;  // compressed by selection or in one block
    int64_t c = g - cbase[si];
    int64_t C = ccount[si];
    int64_t n = sd.v_cnt + sd.e_cnt;
    int64_t a = cs[sd.x_off + c];
    int64_t b = (c + 1 < C) ? cs[sd.x_off + c + 1] : n;
    if (b / kTile - a / kTile >= 2) continue;
    const double* x = X + sd.x_off;
    PCur pc{pose + sd.es_off, se + sd.es_off + si, sd.e_cnt, sd.sf, 0};
    const double* esw_seg = esw + sd.es_off;
    double w, mean;
    if (b - a <= kReplayMax) {
      direct_sum(x, pc, esw_seg, a, b, &w, &mean);
    } else {
      dd sum = direct_sum(x, pc, esw_seg, a, b, &w);
      if (!equal_run_mean(x, a, b, &mean)) mean = dd_div(sum, w);
    }
    cm[sd.x_off + c] = mean;
    cw[sd.x_off + c] = w;
  }
}

// Segments of at most CAP merged records compressed in one block each: the
// new values sorted in LDS, merged with the sorted prior records (prior
// first on equal means), the sweep walked by one lane, then the cluster
// means by all lanes with the arithmetic of k_cluster_mean.  Replaces the
// sort / merge / chain / sums launches for small digests.
template <int CAP, int BLOCK>
__global__ __launch_bounds__(BLOCK, (CAP <= 512 ? 4 : 1)) void k_small_compress(int scale, const SegDesc* segs,
                                                          const int* list, int nlist,
                                                          const double* vin, const double* esm,
                                                          const double* esw, const double* se,
                                                          int64_t* pose, double* cm, double* cw,
                                                          int64_t* ccount, SegResult* res) {
  // A cluster of more than kReplayMax records: at most CAP / kReplayMax.
  constexpr int kBigMax = CAP / (int)kReplayMax + 1;
  __shared__ double sa[CAP + CAP / 16], sb[CAP + CAP / 16];
  __shared__ uint16_t nxt[CAP];
  __shared__ int64_t s_C;
  __shared__ int s_nbig, s_frac, s_bad;
  __shared__ int s_big[kBigMax];
  __shared__ double s_bigw[kBigMax];
  __shared__ int s_bigc[kBigMax + 1];
  // Values per thread when a segment is loaded: all of them in flight at
  // once (one memory round trip rather than one per value).
  constexpr int kPer = CAP / BLOCK;
  for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int si = list[li];
    const SegDesc sd = segs[si];
    const int nv = (int)sd.v_cnt, ne = (int)sd.e_cnt, n = nv + ne;
    const double* v = vin + sd.v_off;
    {
      double r[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        r[u] = i < nv ? v[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        if (i < nv) sa[padi(i)] = r[u];
      }
    }
    __syncthreads();
    double* s = nv > 0 ? lds_sort<(CAP / BLOCK < 16 ? CAP / BLOCK : 16)>(sa, sb, nv) : sa;
    double* x = (s == sa) ? sb : sa;
    // The sorted prior means staged behind the sorted new values (nv + ne
    // <= CAP, so they fit in the padded buffer), for the merge's searches.
    double* e = s + padi(nv);
    if (ne > 0) {
      const double* eg = esm + sd.es_off;
      double r[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        r[u] = i < ne ? eg[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        if (i < ne) e[i] = r[u];
      }
      __syncthreads();
    }
    for (int i = threadIdx.x; i < nv; i += BLOCK) {
      double xv = s[padi(i)];
      int lo = 0, hi = ne;  // prior records <= xv
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (!(xv < e[mid])) lo = mid + 1;
        else hi = mid;
      }
      x[i + lo] = xv;
    }
    for (int i = threadIdx.x; i < ne; i += BLOCK) {
      double xe = e[i];
      int lo = 0, hi = nv;  // new values < xe
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s[padi(mid)] < xe) lo = mid + 1;
        else hi = mid;
      }
      x[i + lo] = xe;
      pose[sd.es_off + i] = i + lo;
    }
    __syncthreads();
    const int64_t* pos = pose + sd.es_off;
    const double* sw = se + sd.es_off + si;
    const double* ew = esw + sd.es_off;
    if (threadIdx.x == 0) {
      SegResult& r = res[si];
      r.vmin = nv > 0 ? s[padi(0)] : 0.0;
      r.vmax = nv > 0 ? s[padi(nv - 1)] : 0.0;
    }
    __syncthreads();
    // kv[j] = k(P(j) / N): the left and right ends of every sweep test
    // (P(j) = sf * (new values before j) + prior weight before j).
    double* kv = s;
    if (ne == 0) {
      // No prior records: P(j) = sf * j + (prior weight before 0).
      const double w0 = sw[0];
      for (int j = threadIdx.x; j <= n; j += BLOCK)
        kv[j] = chain_k(scale, (sd.sf * (double)j + w0) / sd.stored_n, sd.nrm);
    } else {
      for (int j = threadIdx.x; j <= n; j += BLOCK) {
        int lo = 0, hi = ne;  // prior records before merged index j
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (pos[mid] < j) lo = mid + 1;
          else hi = mid;
        }
        double P = sd.sf * (double)(j - lo) + sw[lo];
        kv[j] = chain_k(scale, P / sd.stored_n, sd.nrm);
      }
    }
    __syncthreads();
    // nxt[s]: first i > s with NOT (kv[i + 1] - kv[s] <= 1), or n (the
    // test is monotone in i, so a binary search finds the sweep's candidate).
    for (int st = threadIdx.x; st < n; st += BLOCK) {
      const double ks = kv[st];
      int lo = st + 1, hi = n;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (!(kv[mid + 1] - ks <= 1.0)) hi = mid;
        else lo = mid + 1;
      }
      // Bit 15 flags a candidate within 2 ulps of its predecessor, where the
      // tie rule may extend the cluster (n <= 4096 fits in the low bits).
      bool tie = lo < n && ord_key(x[lo]) - ord_key(x[lo - 1]) <= 2;
      nxt[st] = (uint16_t)(lo | (tie ? 0x8000 : 0));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // Cluster starts go to nxt itself (entries below the current start
      // are never read again; n <= CAP fits 16 bits).
      int64_t replays = 0;
      const PCur pc{pos, sw, ne, sd.sf, 0};
      int64_t C = 0, st = 0;
      while (st < n) {  // the sweep of sweep_segment, walking the table
        const uint16_t v = st < n - 1 ? nxt[st] : (uint16_t)0;
        nxt[C++] = (uint16_t)st;
        if (st == n - 1) break;
        const int64_t b = v & 0x7fff;
        st = (v & 0x8000) ? tie_extend(x, n, st, b, pc, ew, replays) : b;
      }
      res[si].replays = replays;
      s_C = C;
      s_nbig = 0;
      s_frac = 0;
    }
    __syncthreads();
    // The closed-form prefix is exact when every weight is an integer (or a
    // power of two with no prior records): the running sums of the CPU
    // sweep are then the same numbers.  Otherwise each boundary is checked
    // in the CPU's summation order (kd_exact.h): every cluster's running
    // weight (and the one before its last record), the finished-cluster
    // sums in order, then per cluster the last absorbed candidate must not
    // close and the first one after it must.  A disagreement (rare: mostly
    // the last candidate, where q_right is about 1) restarts the sweep at
    // that cluster, record by record as the CPU runs it.
    for (int i = threadIdx.x; i < ne; i += BLOCK)
      if (ew[i] != floor(ew[i])) s_frac = 1;
    __syncthreads();
    const bool closed_exact = (ne == 0 && pow2_weight(sd.sf)) ||
                              (sd.sf == 1.0 && s_frac == 0 && sd.stored_n < 0x1p53);
    if (!closed_exact && n > 1) {
      // kv is dead after the walk: T0[c] takes each cluster's running
      // weight, then (lane 0, in order) the sum of the clusters before it.
      double* T0 = s;
      const int64_t C0 = s_C;
      for (int64_t c = threadIdx.x; c < C0; c += BLOCK) {
        const int64_t a = nxt[c];
        const int64_t b = (c + 1 < C0) ? (int64_t)nxt[c + 1] : n;
        const int64_t ce = ne > 0 ? pcur_seek(PCur{pos, sw, ne, sd.sf, 0}, a) : 0;
        T0[c] = run_weight(a, b, sd.sf, pos, ew, ne, ce);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int64_t c = 0; c < C0; c++) {
          const double w = T0[c];
          T0[c] = acc;
          acc = acc + w;
        }
        s_bad = (int)C0;
      }
      __syncthreads();
      for (int64_t c = threadIdx.x; c < C0; c += BLOCK) {
        const int64_t a = nxt[c];
        const int64_t b = (c + 1 < C0) ? (int64_t)nxt[c + 1] : n;
        const int64_t ce = ne > 0 ? pcur_seek(PCur{pos, sw, ne, sd.sf, 0}, a) : 0;
        // Running weight before the last record, and with it (one more add,
        // as the sweep takes it) the cluster's weight.
        const double wsf = T0[c];
        const double kl = chain_k(scale, wsf / sd.stored_n, sd.nrm);
        // The sweep's test with the reference's operands (kd_sweep_closes),
        // k(q_left) shared by both candidates.
        auto closes = [&](double cur, double w) {
          return !(chain_k(scale, (wsf + cur + w) / sd.stored_n, sd.nrm) - kl <= 1.0);
        };
        double W = rec_weight(a, sd.sf, pos, ew, ne);
        bool bad = false;
        if (b - a >= 2) {
          const double wl = rec_weight(b - 1, sd.sf, pos, ew, ne);
          const double Wp = run_weight(a, b - 1, sd.sf, pos, ew, ne, ce);
          bad = closes(Wp, wl);
          W = Wp + wl;
        }
        if (b < n) bad = bad || !closes(W, rec_weight(b, sd.sf, pos, ew, ne));
        if (bad) atomicMin(&s_bad, (int)c);
      }
      __syncthreads();
      if (threadIdx.x == 0 && s_bad < (int)C0) {
        // The CPU sweep (tdigest.mojo:419-435) from the first cluster in
        // question: tie rule, running mean and weight, boundary test.
        int64_t c = s_bad;
        int64_t st = nxt[c];
        double wsf = T0[c];
        int64_t ce = ne > 0 ? pcur_seek(PCur{pos, sw, ne, sd.sf, 0}, st) : 0;
        double cur_w = sd.sf;
        if (ce < ne && pos[ce] == st) cur_w = ew[ce++];
        double cur_m = x[st];
        nxt[c++] = (uint16_t)st;
        for (int64_t i = st + 1; i < n; i++) {
          double wi = sd.sf;
          if (ce < ne && pos[ce] == i) wi = ew[ce++];
          const double xi = x[i];
          if (xi == cur_m || !chain_closes(scale, sd.nrm, sd.stored_n, wsf, cur_w, wi)) {
            const double nw = cur_w + wi;
            cur_m += (xi - cur_m) * (wi / nw);
            cur_w = nw;
          } else {
            wsf = wsf + cur_w;
            nxt[c++] = (uint16_t)i;
            cur_w = wi;
            cur_m = xi;
          }
        }
        s_C = c;
        res[si].replays += 1;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) ccount[si] = s_C;
    const int64_t C = s_C;
    const uint16_t* cs = nxt;
    double* T = s;  // kv is no longer needed: per-record scratch of the means
    // Clusters of at most kReplayMax records replay the running mean of
    // centroid.mojo:34-36: the ratios w / W' depend only on the running
    // weight, so each cluster's weight chain (one lane each) stores W' per
    // record, all lanes then form the ratios (independent divisions), and
    // the mean recurrence runs on them (one lane each).  Longer clusters get
    // their weight here and their double-double sum from all lanes below.
    // With no prior records and a power-of-two weight per new value, every
    // running weight is exactly (j - a + 1) * sf: no chain is needed.
    const bool exact_w = ne == 0 && pow2_weight(sd.sf);
    for (int64_t c = threadIdx.x; c < C; c += BLOCK) {
      const int64_t a = cs[c];
      const int64_t b = (c + 1 < C) ? (int64_t)cs[c + 1] : n;
      const bool replay = b - a <= kReplayMax;
      double ws = 0.0;
      if (exact_w) {
        ws = (double)(b - a) * sd.sf;
      } else if (ne == 0) {
        ws = sd.sf;
        if (replay) T[a] = ws;
        for (int64_t j = a + 1; j < b; j++) {
          ws += sd.sf;
          if (replay) T[j] = ws;
        }
      } else {
        PCur pc{pos, sw, ne, sd.sf, 0};
        int64_t ce = pcur_seek(pc, a);
        for (int64_t j = a; j < b; j++) {
          double w = sd.sf;
          if (ce < ne && pos[ce] == j) {
            w = ew[ce];
            ce++;
          }
          ws = (j == a) ? w : ws + w;
          if (replay) T[j] = ws;
        }
      }
      double mean;
      if (!replay) {
        if (equal_run_mean(x, a, b, &mean)) {
          cm[sd.x_off + c] = mean;
        } else {
          int k = atomicAdd(&s_nbig, 1);
          s_big[k] = (int)c;
          s_bigw[k] = ws;
        }
      }
      cw[sd.x_off + c] = ws;
    }
    __syncthreads();
    // w / W' for every record (also where unused: cluster starts, long
    // clusters, whose T entries are rewritten below).
    for (int j = threadIdx.x; j < n; j += BLOCK) {
      double w = sd.sf;
      double W;
      if (exact_w) {
        int lo = 0, hi = (int)C - 1;  // the cluster holding j: last start <= j
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if ((int)cs[mid] <= j) lo = mid;
          else hi = mid - 1;
        }
        W = (double)(j - (int)cs[lo] + 1) * sd.sf;
      } else {
        W = T[j];
        if (ne > 0) {
          int lo = 0, hi = ne;  // the prior record at merged index j, if any
          while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (pos[mid] < j) lo = mid + 1;
            else hi = mid;
          }
          if (lo < ne && pos[lo] == j) w = ew[lo];
        }
      }
      T[j] = w / W;
    }
    __syncthreads();
    for (int64_t c = threadIdx.x; c < C; c += BLOCK) {
      const int a = (int)cs[c];
      const int b = (c + 1 < C) ? (int)cs[c + 1] : n;
      if (b - a > kReplayMax) continue;
      double mean = x[a];
      int j = a + 1;
      // Four records at a time: their loads do not wait for the chain.
      for (; j + 4 <= b; j += 4) {
        const double x0 = x[j], x1 = x[j + 1], x2 = x[j + 2], x3 = x[j + 3];
        const double t0 = T[j], t1 = T[j + 1], t2 = T[j + 2], t3 = T[j + 3];
        mean += (x0 - mean) * t0;
        mean += (x1 - mean) * t1;
        mean += (x2 - mean) * t2;
        mean += (x3 - mean) * t3;
      }
      for (; j < b; j++) mean += (x[j] - mean) * T[j];
      cm[sd.x_off + c] = mean;
    }
    __syncthreads();
    const int nbig = s_nbig;
    if (nbig > 0) {
      // Chunk sums of the long clusters (chunk k of cluster [a, b) stored at
      // T[a + 2k], T[a + 2k + 1]: inside the cluster's own records), the
      // chunks of all of them spread over the block at once.
      if (threadIdx.x == 0) {
        int acc = 0;
        for (int q = 0; q < nbig; q++) {
          const int c = s_big[q];
          const int a = (int)cs[c];
          const int b = (c + 1 < C) ? (int)cs[c + 1] : n;
          s_bigc[q] = acc;
          acc += (b - a + (int)kSumChunk - 1) / (int)kSumChunk;
        }
        s_bigc[nbig] = acc;
      }
      __syncthreads();
      const int nflat = s_bigc[nbig];
      for (int f = threadIdx.x; f < nflat; f += BLOCK) {
        int q = 0;
        while (s_bigc[q + 1] <= f) q++;
        const int c = s_big[q];
        const int64_t a = cs[c];
        const int64_t b = (c + 1 < C) ? (int64_t)cs[c + 1] : n;
        const int64_t k = f - s_bigc[q];
        const int64_t j0 = a + k * kSumChunk, j1 = min(b, j0 + kSumChunk);
        PCur pc{pos, sw, ne, sd.sf, 0};
        int64_t ce = ne > 0 ? pcur_seek(pc, j0) : 0;
        dd part{0.0, 0.0};
        for (int64_t j = j0; j < j1; j++) {
          double w = sd.sf;
          if (ce < ne && pos[ce] == j) {
            w = ew[ce];
            ce++;
          }
          part = dd_add(part, two_prod(x[j], w));
        }
        T[a + 2 * k] = part.hi;
        T[a + 2 * k + 1] = part.lo;
      }
      __syncthreads();
      for (int q = threadIdx.x; q < nbig; q += BLOCK) {
        const int c = s_big[q];
        const int64_t a = cs[c];
        const int64_t b = (c + 1 < C) ? (int64_t)cs[c + 1] : n;
        const int64_t nch = (b - a + kSumChunk - 1) / kSumChunk;
        dd acc{0.0, 0.0};
        for (int64_t k = 0; k < nch; k++) acc = dd_add(acc, dd{T[a + 2 * k], T[a + 2 * k + 1]});
        cm[sd.x_off + c] = dd_div(acc, s_bigw[q]);
      }
    }
    __syncthreads();
  }
}

// One block per long cluster [a, b): head [a, end of tile ta), the full tiles
// ta+1 .. tb-1 from their tile sums, tail [start of tile tb, b); E records in
// head and tail are corrected from sf to their own weight.
__global__ __launch_bounds__(kBlock) void k_long_sum(const SegDesc* segs, const int64_t* ccount,
                                                     const int64_t* cs, const double* X,
                                                     const int64_t* pose, const double* se,
                                                     const double* esw, const double* tsum,
                                                     const int64_t* long_list,
                                                     const unsigned long long* long_cnt,
                                                     double* cm, double* cw) {
  __shared__ double rh[kBlock], rl[kBlock];
  const int64_t nl = (int64_t)*long_cnt;
  for (int64_t li = blockIdx.x; li < nl; li += gridDim.x) {
    int si = (int)long_list[2 * li];
    int64_t c = long_list[2 * li + 1];
    SegDesc sd = segs[si];
    int64_t n = sd.v_cnt + sd.e_cnt;
    int64_t C = ccount[si];
    int64_t a = cs[sd.x_off + c];
    int64_t b = (c + 1 < C) ? cs[sd.x_off + c + 1] : n;
    int64_t ta = a / kTile, tb = b / kTile;
    const double* x = X + sd.x_off;
    dd acc{0.0, 0.0};
    for (int64_t j = a + threadIdx.x; j < (ta + 1) * kTile; j += blockDim.x)
      acc = dd_add(acc, two_prod(x[j], sd.sf));
    for (int64_t j = tb * kTile + threadIdx.x; j < b; j += blockDim.x)
      acc = dd_add(acc, two_prod(x[j], sd.sf));
    for (int64_t t = ta + 1 + threadIdx.x; t < tb; t += blockDim.x)
      acc = dd_add(acc, dd{tsum[2 * (sd.t_off + t)], tsum[2 * (sd.t_off + t) + 1]});
    // Weight: sf per new value plus the E weights inside the cluster, summed
    // directly (a difference of prefix sums would lose relative precision).
    dd ewt{0.0, 0.0};
    double ecount = 0.0;
    if (sd.e_cnt > 0) {
      const int64_t* pos = pose + sd.es_off;
      const double* ew = esw + sd.es_off;
      int64_t lo = 0, hi = sd.e_cnt;
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (pos[mid] < a) lo = mid + 1;
        else hi = mid;
      }
      for (int64_t e = lo + threadIdx.x; e < sd.e_cnt && pos[e] < b; e += blockDim.x) {
        int64_t pe = pos[e];
        ewt = dd_add(ewt, dd{ew[e], 0.0});
        ecount += 1.0;
        // Inside head or tail the record was summed with weight sf above;
        // inside a full tile its tile sum already used the E weight.
        if (pe < (ta + 1) * kTile || pe >= tb * kTile) {
          double xe = x[pe];
          acc = dd_add(acc, dd_add(two_prod(xe, ew[e]), dd_neg(two_prod(xe, sd.sf))));
        }
      }
    }
    __shared__ double re[kBlock], rc[kBlock], rf[kBlock];
    rh[threadIdx.x] = acc.hi;
    rl[threadIdx.x] = acc.lo;
    re[threadIdx.x] = ewt.hi;
    rf[threadIdx.x] = ewt.lo;
    rc[threadIdx.x] = ecount;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        dd s2 = dd_add(dd{rh[threadIdx.x], rl[threadIdx.x]},
                       dd{rh[threadIdx.x + w], rl[threadIdx.x + w]});
        rh[threadIdx.x] = s2.hi;
        rl[threadIdx.x] = s2.lo;
        dd e2 = dd_add(dd{re[threadIdx.x], rf[threadIdx.x]},
                       dd{re[threadIdx.x + w], rf[threadIdx.x + w]});
        re[threadIdx.x] = e2.hi;
        rf[threadIdx.x] = e2.lo;
        rc[threadIdx.x] += rc[threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      dd wsum = dd_add(two_prod(sd.sf, (double)(b - a) - rc[0]), dd{re[0], rf[0]});
      double w = wsum.hi + wsum.lo;
      double mean;
      if (!equal_run_mean(x, a, b, &mean)) mean = dd_div(dd{rh[0], rl[0]}, w);
      cm[sd.x_off + c] = mean;
      // The stored weight is the cluster's running weight, summed in record
      // order as the CPU sums it (the exact sum above is for the mean).
      const int64_t* pos = pose + sd.es_off;
      const int64_t ce = pcur_seek(PCur{pos, se, sd.e_cnt, sd.sf, 0}, a);
      cw[sd.x_off + c] = run_weight(a, b, sd.sf, pos, esw + sd.es_off, sd.e_cnt, ce);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------ selection path
// Segments with more than kCap new values are compressed without sorting
// them.  The new values are bucketed by a sorted splitter set (a regular
// sample of the values plus every prior centroid mean); the bucket counts fix
// the rank range of every bucket and the merged position of every prior
// centroid, and the greedy sweep's cluster boundaries depend on ranks only
// (P(j), the weight of the first j merged records, is a function of j and of
// those positions), so the chain is walked in rank space.  Only buckets that
// hold a record next to a boundary are gathered; a radix select inside them
// gives the exact values and partial sums at the boundary ranks.  Every other
// value is added to its cluster as three fixed-point slices whose sums are
// exact, so the cluster sums do not depend on the order of the additions.
// A boundary whose two records lie within two ulps of each other (where the
// sweep's tie rule, tdigest.mojo:429, may apply), or any capacity limit,
// sends the segment back to the sorting path, which handles ties exactly.

constexpr int kSelSplMax = 8191;     // splitters per segment (buckets: one more)
constexpr int kSelSample = 8192;     // regular sample sorted in LDS
constexpr int kSelPad = kSelSample + kSelSample / 16;
constexpr int kSelCmax = 1024;       // clusters per segment on this path
constexpr int kSelW = 24;            // bits per slice: exact for 2^29 values per cluster
constexpr int kSelThreads = 1024;
constexpr int64_t kSelMaxV = (int64_t)1 << 29;
constexpr int kSelLdsVals = 16384;   // marked buckets up to this size are staged in LDS
constexpr int kSelSmallS = 1023;     // small LDS class of the count and gather passes
constexpr int kSelSmallC = 256;

struct SelSeg {
  int32_t si;        // segment index
  int32_t cap_s;     // splitter capacity
  int32_t cap_c;     // cluster capacity
  int32_t pad;
  int64_t b_off;     // bucket arrays: cap_s + 2 entries
  int64_t c_off;     // boundary and cluster arrays: cap_c + 1 entries
  int64_t m_off;     // marked buckets: 2 * cap_c + 2 entries
  int64_t tile0;     // first tile of the count and gather passes
  int64_t tile_len;  // values per tile
};

struct SelState {
  int32_t nspl, ncl, nmk, fail;
  unsigned long long kmin, kmax;  // order keys of the smallest / largest new value
  int32_t wdone;                  // 1: the chain left every cluster's running weight in cwv
  int32_t blast;                  // bucket of the last new-value rank (nv - 1), from the plan
};

struct SelMark {
  int64_t bucket, goff, cnt, vstart;  // gathered at x_off + goff; first new-value rank
  unsigned int cursor;
  int first_bound;                    // first boundary i >= 1 with vr >= vstart
  double sum_hi, sum_lo, vmin, vmax;
};

// A gathered bucket of the smallest size class, as its wave reads it: where
// its values are, its first new-value rank, its boundaries [b0, b1) (indices
// into the segment's boundary table at c_off) and its mark.
struct SelWave {
  int64_t src;     // index of the bucket's first gathered value in X
  int64_t vstart;  // new-value rank of the bucket's first value
  int64_t c_off;   // the segment's boundary table
  int64_t mark;    // index of the bucket's SelMark
  int32_t n, b0, b1, pad;
};

struct SelBound {
  int64_t b;           // merged rank of the cluster's first record
  int64_t vr;          // new values before it
  int32_t v_hi, v_lo;  // record b / record b - 1 is a new value
  int32_t dup, pad;    // the value at new-value rank vr repeats below that rank
  double x_hi, x_lo;   // values at new-value ranks vr and vr - 1 (same bucket only)
  double sl_hi, sl_lo; // sum of the bucket's values below rank vr
  int32_t bk_hi, bk_lo;  // buckets of new-value ranks vr and vr - 1 (-1: no such rank)
};

// Selection-path sizing of one digest (shared by the host and device plans).
__host__ __device__ inline int sel_cap_s(int64_t v_cnt, int64_t e_cnt) {
  int64_t a = v_cnt / 32 > 63 ? v_cnt / 32 : 63;
  a += e_cnt;
  return (int)(a < kSelSplMax ? a : kSelSplMax);
}

// Two consecutive clusters span more than one unit of k, and k is evaluated
// on q clamped to [1e-15, 1 - 1e-15].
__host__ __device__ inline int sel_cap_c(int scale, double nrm) {
  double kr = kd_k(scale, 1.0 - 1e-15, nrm) - kd_k(scale, 1e-15, nrm);
  int64_t cc = (kr >= 0.0 && kr < 1e6) ? (int64_t)(2.0 * kr) + 4 : kSelCmax;
  if (cc > kSelCmax) cc = kSelCmax;
  return (int)(cc < 2 ? 2 : cc);
}

// Values per tile of the count and gather passes: at most tiles_max tiles,
// at least tile_min values per tile, except that a digest of more than
// 2^20 / 64 values gets at least 64 tiles (at most 16384 values each) so a
// batch holding one large digest still spreads over the GPU.
__host__ __device__ inline int64_t sel_tile_len(int64_t v_cnt, int64_t tile_min, int64_t tiles_max) {
  int64_t lo = v_cnt / 64 > 16384 ? v_cnt / 64 : 16384;
  lo = lo < tile_min ? lo : tile_min;
  int64_t t = (v_cnt + tiles_max - 1) / tiles_max;
  return t > lo ? t : lo;
}

// Tiles counted by one block of the count pass: enough for at least
// kSelCountMin values (building a block's search table costs about as much
// as counting tens of thousands of values).
constexpr int64_t kSelCountMin = 65536;
constexpr int64_t kSelCountGroupTiles = 2048;  // tiles of a launch from which blocks group them
__host__ __device__ inline int64_t sel_count_group(int64_t tile_len) {
  return tile_len >= kSelCountMin ? 1 : (kSelCountMin + tile_len - 1) / tile_len;
}

// LDS footprint class of the count and gather passes: 0 few splitters and
// clusters, 1 many splitters and few clusters, 2 the rest.
__host__ __device__ inline int sel_class(int cap_s, int cap_c) {
  return cap_c <= kSelSmallC ? (cap_s <= kSelSmallS ? 0 : 1) : 2;
}

__device__ inline double key_value(unsigned long long k) {
  uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return kd_bits2d(b);
}

// Bucket of v: the number of splitters <= v.
__device__ inline int sel_bucket(const double* spl, int S, double v) {
  int lo = 0, hi = S;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (spl[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Bucket holding new-value rank r: the last b with bst[b] <= r.
__device__ inline int sel_bucket_of_rank(const unsigned int* bst, int S, int64_t r) {
  int lo = 0, hi = S;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if ((int64_t)bst[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// 1.5 * 2^k: adding and subtracting it rounds to a multiple of 2^(k - 52).
__device__ inline double sel_magic(int k) {
  return kd_bits2d(((uint64_t)(k + 1023) << 52) | 0x0008000000000000ull);
}

// v = s1 + s2 + r with s_i multiples of 2^(e - i W), |v| < 2^e and
// |r| <= 2^(e - 2W - 1): the dropped remainder is below 2^-48 of the
// cluster's magnitude bound, far inside the 1e-12 parity bar for means,
// while the slice sums stay exact (order-independent) for 2^29 values.
__device__ inline void sel_slices(double v, int e, double& s1, double& s2) {
  double m1 = sel_magic(e - kSelW + 52), m2 = sel_magic(e - 2 * kSelW + 52);
  s1 = (v + m1) - m1;
  double r1 = v - s1;
  s2 = (r1 + m2) - m2;
}

// Smallest e with max(|lo|, |hi|) < 2^e (-900 for zeros).
__device__ inline int sel_exponent(double lo, double hi) {
  double m = fmax(fabs(lo), fabs(hi));
  if (m == 0.0) return -900;
  return (int)((kd_d2bits(m) >> 52) & 0x7ff) - 1022;
}

__device__ inline dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  return two_sum(p.hi, p.lo + a.lo * b);
}

// Sample size for a segment of nv new values (about two per splitter).
__host__ __device__ inline int sel_sample_size(int64_t nv) {
  int64_t m = nv / 16;
  if (m < 512) m = 512;
  if (m > kSelSample) m = kSelSample;
  if (m > nv) m = nv;
  return (int)m;
}

// Per-segment selection state, the bucket counters and list sizes cleared,
// the segment of every tile of the count and gather passes, and the work
// lists (grid-stride): per splitter-sample size class its segments
// (slist[c * nsel ...], lcnt[c]), per count instance its blocks' first tiles
// (clist[i * sel_t ...], lcnt[3 + i]), so that each launch's working blocks
// come first.  lcnt is zero on entry.
__global__ void k_sel_init(const SegDesc* segs, const SelSeg* sel, SelState* st, int nsel,
                           unsigned int* cnt, int64_t ncnt, unsigned int* mlc, int nmlc,
                           int* tmap, int64_t sel_t, unsigned int* lcnt, int* slist, int* clist) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = j; i < ncnt; i += stride) cnt[i] = 0u;
  if (j < nmlc) mlc[j] = 0u;
  const int64_t grp_min = sel_t >= kSelCountGroupTiles ? kSelCountMin : 0;
  // Segments in whole waves: list places are claimed with one atomic per
  // wave and list (the counters are shared by every block).
  const int lane = (int)(threadIdx.x & 63);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t ib = j - lane; ib < nsel; ib += stride) {
    const int64_t i = ib + lane;
    const bool act = i < nsel;
    int c = -1, ci = -1;
    int64_t nblk = 0, grp = 1, tile0 = 0;
    if (act) {
      st[i] = SelState{0, 0, 0, 0, ~0ull, 0ull, 0, 0};
      const SelSeg g = sel[i];
      const int64_t nv = segs[g.si].v_cnt;
      const int64_t nt = (nv + g.tile_len - 1) / g.tile_len;
      for (int64_t k = 0; k < nt; k++) tmap[g.tile0 + k] = (int)i;
      const int need = std::max(sel_sample_size(nv), g.cap_s);
      c = need <= 1024 ? 0 : need <= 4096 ? 1 : 2;
      grp = grp_min > 0 ? sel_count_group(g.tile_len) : 1;
      nblk = (nt + grp - 1) / grp;
      ci = g.cap_s <= kSelSmallS ? 0 : 1;
      tile0 = g.tile0;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const uint64_t m = __ballot(c == q);
      if (m == 0ull) continue;
      const int leader = __ffsll((long long)m) - 1;
      unsigned int base = 0u;
      if (lane == leader) base = atomicAdd(&lcnt[q], (unsigned int)__popcll(m));
      base = __shfl(base, leader, 64);
      if (c == q) slist[(int64_t)q * nsel + base + (unsigned int)__popcll(m & below)] = (int)i;
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      // Exclusive prefix of the blocks of this wave's segments of instance q.
      const unsigned int mine = ci == q ? (unsigned int)nblk : 0u;
      unsigned int inc = mine;
      for (int d = 1; d < 64; d <<= 1) {
        const unsigned int y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      const unsigned int tot = __shfl(inc, 63, 64);
      if (tot == 0u) continue;
      unsigned int base = 0u;
      if (lane == 0) base = atomicAdd(&lcnt[3 + q], tot);
      base = __shfl(base, 0, 64) + inc - mine;
      for (int64_t k = 0; k < (int64_t)mine; k++) clist[q * sel_t + base + k] = (int)(tile0 + k * grp);
    }
  }
}

// Splitters: a regular sample of the new values, sorted in LDS, merged with
// the sorted prior means (each kept exactly), repeats dropped.  Size classes:
// this instance takes the segments whose sample size and splitter capacity
// are at most CAP (and not both at most LO).
// The regular sample of a segment's m new values (sample i at
// ((2i + 1) nv) / (2m)), loaded into sa at padded indices, eight loads in
// flight per thread (each is a separate cache line).
template <int BLOCK>
__device__ inline void sel_load_sample(const double* v, int64_t nv, int m, double* sa) {
  for (int i0 = 0; i0 < m; i0 += 8 * BLOCK) {
    double xs[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u * BLOCK + (int)threadIdx.x;
      xs[u] = i < m ? v[((2 * (int64_t)i + 1) * nv) / (2 * (int64_t)m)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u * BLOCK + (int)threadIdx.x;
      if (i < m) sa[padi(i)] = xs[u];
    }
  }
}

// The sorted regular samples of consecutive batches of one digest's values
// (batch values each, the last one shorter), ahead of their merges: block b
// sorts batch b's sample as k_sel_sample's largest class would (same block
// size, same sort) into out + b * kSelSample.
__global__ __launch_bounds__(kSelThreads) void k_sel_presort(const double* vin, int64_t n,
                                                             int64_t batch, double* out) {
  __shared__ double sa[kSelSample + kSelSample / 16], sb[kSelSample + kSelSample / 16];
  const int64_t v0 = (int64_t)blockIdx.x * batch;
  const int64_t nv = std::min<int64_t>(batch, n - v0);
  const int m = sel_sample_size(nv);
  sel_load_sample<kSelThreads>(vin + v0, nv, m, sa);
  __syncthreads();
  const double* s = lds_sort(sa, sb, m);
  double* o = out + (int64_t)blockIdx.x * kSelSample;
  for (int i = threadIdx.x; i < m; i += blockDim.x) o[i] = s[padi(i)];
}

// presorted: the segment's sorted sample from k_sel_presort, or null (only
// the largest class, whose block size and sort k_sel_presort shares, and only
// for a lone segment: j == 0).
template <int CAP, int BLOCK, int LO>
__global__ __launch_bounds__(BLOCK) void k_sel_sample(const SegDesc* segs, const SelSeg* sel,
                                                      const double* vin, const double* esm,
                                                      double* spl, SelState* st,
                                                      const double* presorted, const int* list,
                                                      const unsigned int* lcnt) {
  __shared__ double sa[CAP + CAP / 16], sb[CAP + CAP / 16];
  __shared__ uint32_t part[BLOCK];
  if (blockIdx.x >= *lcnt) return;
  const int j = list[blockIdx.x];
  const SelSeg g = sel[j];
  const SegDesc sd = segs[g.si];
  const int64_t nv = sd.v_cnt;
  const int m = sel_sample_size(nv);
  const int need = m > g.cap_s ? m : g.cap_s;
  if (need <= LO || need > CAP) return;
  double* s;
  if (presorted) {
    // Eight loads in flight per thread before their LDS stores.
    for (int i0 = 0; i0 < m; i0 += 8 * BLOCK) {
      double xs[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = i0 + u * BLOCK + (int)threadIdx.x;
        xs[u] = i < m ? presorted[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = i0 + u * BLOCK + (int)threadIdx.x;
        if (i < m) sa[padi(i)] = xs[u];
      }
    }
    __syncthreads();
    s = sa;
  } else {
    sel_load_sample<BLOCK>(vin + sd.v_off, nv, m, sa);
    __syncthreads();
    s = lds_sort(sa, sb, m);
  }
  double* f = (s == sa) ? sb : sa;
  const int ne = (int)sd.e_cnt;
  int T = (int)std::min<int64_t>(std::max<int64_t>(nv / 32, 63), (int64_t)(g.cap_s - ne));
  if (T > m - 1) T = m - 1;
  // Sample quantiles t/(T+1), t = 1..T, into f.
  for (int t = threadIdx.x; t < T; t += BLOCK) f[t] = s[padi((int)(((t + 1) * (int64_t)m) / (T + 1)))];
  __syncthreads();
  // Merge with the prior means into s: a quantile goes before equal means.
  // The prior means are staged after the quantiles in f (T + ne <= cap_s <
  // CAP), so the searches below stay in LDS.
  double* e = f + T;
  for (int b = threadIdx.x; b < ne; b += BLOCK) e[b] = esm[sd.es_off + b];
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += BLOCK) {
    double x = f[t];
    int l = 0, h = ne;
    while (l < h) {
      int mid = (l + h) >> 1;
      if (e[mid] < x) l = mid + 1;
      else h = mid;
    }
    s[t + l] = x;
  }
  for (int b = threadIdx.x; b < ne; b += BLOCK) {
    double y = e[b];
    int l = 0, h = T;
    while (l < h) {
      int mid = (l + h) >> 1;
      if (f[mid] <= y) l = mid + 1;
      else h = mid;
    }
    s[b + l] = y;
  }
  __syncthreads();
  // Drop repeats: keep element k when it exceeds element k - 1.
  const int L = T + ne;
  const int per = (L + BLOCK - 1) / BLOCK;
  const int k0 = (int)threadIdx.x * per < L ? (int)threadIdx.x * per : L;
  const int k1 = k0 + per < L ? k0 + per : L;
  uint32_t kept = 0;
  for (int k = k0; k < k1; k++) kept += (k == 0 || s[k - 1] < s[k]) ? 1u : 0u;
  part[threadIdx.x] = kept;
  __syncthreads();
  uint32_t total = block_excl_scan(part, BLOCK);
  uint32_t o = part[threadIdx.x];
  double* out = spl + g.b_off;
  for (int k = k0; k < k1; k++)
    if (k == 0 || s[k - 1] < s[k]) out[o++] = s[k];
  if (threadIdx.x == 0) st[j].nspl = (int)total;
}

// Two-level lookup of a value's bucket.  The splitters' order keys are
// searched in full (64 bits, so no comparison needs the doubles), guided by
// a table on 32 bits of each key: the bits right after the prefix that every
// splitter shares (its window).  The window range [wA, wB] of the splitters
// is cut into cells of 2^csh window values (at most 2^log_cells cells);
// cell c is cut again into 2^p fine slots, p chosen from the number n of
// splitters in the cell (n < 2^p <= 2n, so a slot holds one splitter or
// none on average), and fs[s] counts the splitters whose slot is below s.
// A value's bucket then lies among the splitters of its fine slot, fs[s] ..
// fs[s + 1] (every splitter with the value's window is there), found by a
// bisection of as many rounds as the fullest slot that the wave's values hit
// needs.
__host__ __device__ inline int sel_cell_shift(uint32_t span, int log_cells) {
  int bits = 0;
  while (bits < 32 && (span >> bits) != 0u) bits++;
  return bits > log_cells ? bits - log_cells : 0;
}

// log2 of the fine slots of a cell of n splitters and 2^csh window values.
__host__ __device__ inline int sel_fine_bits(uint32_t n, int csh) {
  int p = 0;
  while (p < 30 && (n >> p) != 0u) p++;  // n < 2^p
  return p < csh ? p : csh;
}

__host__ __device__ inline uint32_t sel_window(uint64_t key, int sh) {
  return (uint32_t)((sh >= 64 ? 0ull : (key << sh)) >> 32);
}

// Fine slot of window offset d (= window - wA) under packed cell entries
// (first slot | (csh - p) << 16).
__device__ inline int sel_slot(const uint32_t* cell, int csh, uint32_t cmask, uint32_t d) {
  const uint32_t ce = cell[d >> csh];
  return (int)(ce & 0xffffu) + (int)((d & cmask) >> (ce >> 16));
}

// x[q] = min(x[q], ..., x[n]) for q <= n, by the whole block: entries
// still at their initial maximum take the value of the next set entry
// (x[n] must be set).  Each thread takes a run of at most PER consecutive
// entries; the runs' minima are combined right to left across the block.
template <int BLOCK, int PER, typename T>
__device__ inline void block_suffix_min(T* x, int n, unsigned int* wmin) {
  const int lane = (int)(threadIdx.x & 63), wid = (int)(threadIdx.x >> 6);
  const int per = (n + BLOCK) / BLOCK;  // n + 1 entries
  const int q0 = (int)threadIdx.x * per;
  const unsigned int kMax = (unsigned int)(T)~(T)0;
  unsigned int v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) v[k] = (k < per && q0 + k <= n) ? (unsigned int)x[q0 + k] : kMax;
  unsigned int m = kMax;
#pragma unroll
  for (int k = PER - 1; k >= 0; k--) {
    m = min(m, v[k]);
    v[k] = m;
  }
  // c = minimum of the runs of this lane and the lanes above it in the wave.
  unsigned int c = m;
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned int y = __shfl_down(c, d, 64);
    if (lane + d < 64) c = min(c, y);
  }
  if (lane == 0) wmin[wid] = c;
  __syncthreads();
  unsigned int carry = kMax;
  for (int w = wid + 1; w < BLOCK / 64; w++) carry = min(carry, wmin[w]);
  const unsigned int above = __shfl_down(c, 1, 64);
  if (lane < 63) carry = min(carry, above);
#pragma unroll
  for (int k = 0; k < PER; k++)
    if (k < per && q0 + k <= n) x[q0 + k] = (T)min(v[k], carry);
}

// SMAX: the largest splitter count of the segments this instance serves;
// the LDS footprint follows it.  An instance serves the segments of tiles
// [t0, t1) whose splitter capacity is in (SMIN, SMAX]; the others' blocks
// return at once.
template <int SMAX, int SMIN, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sel_count(const SegDesc* segs, const SelSeg* sel,
                                                     int nsel, int64_t t0, int64_t t1,
                                                     const double* vin, const double* spl,
                                                     SelState* st, unsigned int* cnt,
                                                     uint16_t* bix, const int* tmap,
                                                     const int* list, const unsigned int* lcnt) {
  constexpr int kLogCells = SMAX >= 4096 ? 9 : 7;
  constexpr int kCells = 1 << kLogCells;
  constexpr int kFine = 2 * SMAX + kCells + 1;  // fine slots at most, plus the end
  static_assert(kCells + 1 <= SMAX + 1, "cell counts live in the counters");
  static_assert(kCells % 64 == 0 && kFine < 65536, "table layout");
  __shared__ uint64_t sk[SMAX + 1];
  __shared__ unsigned int lc[SMAX + 1];
  __shared__ uint16_t fs[kFine + 1];
  __shared__ uint32_t cell[kCells];
  __shared__ unsigned long long smin, smax;
  __shared__ int s_nslot;
  __shared__ unsigned int wmin[BLOCK / 64];
  // The blocks of this instance's work list (first tiles of its segments'
  // tile groups) come first; the rest return.
  if (blockIdx.x >= *lcnt) return;
  const int64_t t = list[blockIdx.x];
  const int j = tmap[t];
  const SelSeg g = sel[j];
  if (g.cap_s <= SMIN || g.cap_s > SMAX) return;
  // With many tiles (the GPU is full anyway), a block counts
  // sel_count_group consecutive tiles of its segment (the blocks of the
  // other tiles return), so that building the table is paid for at least
  // kSelCountMin values; a few tiles keep one block each.
  const int64_t lt = t - g.tile0;
  const int64_t grp = t1 - t0 >= kSelCountGroupTiles ? sel_count_group(g.tile_len) : 1;
  if (lt % grp != 0) return;
  const SegDesc sd = segs[g.si];
  const int S = st[j].nspl;
  const double* sg = spl + g.b_off;
  const int64_t a = lt * g.tile_len;
  const int64_t b = std::min<int64_t>(sd.v_cnt, a + grp * g.tile_len);
  const double* v = vin + sd.v_off;
  // The tile's first values are loaded before the table is built.
  constexpr int U = 8;
  double xn[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int64_t i = a + (int64_t)u * BLOCK + threadIdx.x;
    xn[u] = v[i < b ? i : b - 1];
  }
  // Setup loops unrolled over a thread's kPer entries, so that their loads
  // are all in flight together.
  constexpr int kPer = (SMAX + 1 + BLOCK - 1) / BLOCK;
  {
    double tv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int i = u * BLOCK + (int)threadIdx.x;
      tv[u] = i < S ? sg[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int i = u * BLOCK + (int)threadIdx.x;
      if (i < S) sk[i] = ord_key(tv[u]);
    }
  }
  for (int c = threadIdx.x; c <= kCells; c += BLOCK) lc[c] = ~0u;
  if (threadIdx.x == 0) {
    smin = ~0ull;
    smax = 0ull;
  }
  __syncthreads();
  const uint64_t kf = S > 0 ? sk[0] : 0ull, kl = S > 0 ? sk[S - 1] : 0ull;
  const int sh = kf == kl ? 64 : __clzll(kf ^ kl);
  const uint32_t wA = sel_window(kf, sh);
  const int csh = sel_cell_shift(sel_window(kl, sh) - wA, kLogCells);
  const uint32_t cmask = (1u << csh) - 1u;
  // cf[c] (in lc) = splitters in cells below c: splitter i writes it for
  // the cells after splitter i - 1's cell up to its own (the keys are
  // sorted, so no counting is needed); then each cell's fine slots and first
  // slot, scanned by the first wave.
  unsigned int* cf = lc;
  uint32_t wd[kPer], wq[kPer];  // window offsets of splitters i and i - 1
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int i = u * BLOCK + (int)threadIdx.x;
    wd[u] = i < S ? sel_window(sk[i], sh) - wA : 0u;
    wq[u] = (i > 0 && i <= S) ? sel_window(sk[i - 1], sh) - wA : 0u;
  }
  // The first splitter of each cell marks it; cells without one take the
  // next cell's mark (a suffix minimum).
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int i = u * BLOCK + (int)threadIdx.x;
    if (i <= S) {
      const int cp = i > 0 ? (int)(wq[u] >> csh) : -1;
      const int ci = i < S ? (int)(wd[u] >> csh) : kCells;
      if (ci != cp) cf[ci] = (unsigned)i;
    }
  }
  __syncthreads();
  block_suffix_min<BLOCK, (kCells + BLOCK) / BLOCK>(cf, kCells, wmin);
  __syncthreads();
  if (threadIdx.x < 64) {
    constexpr int per = kCells / 64;
    const int c0 = (int)threadIdx.x * per;
    uint32_t m[per], sum = 0;
#pragma unroll
    for (int u = 0; u < per; u++) {
      m[u] = 1u << sel_fine_bits(cf[c0 + u + 1] - cf[c0 + u], csh);
      sum += m[u];
    }
    uint32_t inc = sum;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, 64);
      if ((int)threadIdx.x >= d) inc += y;
    }
    uint32_t base = inc - sum;
#pragma unroll
    for (int u = 0; u < per; u++) {
      const int p = 31 - __clz((int)m[u]);
      cell[c0 + u] = base | ((uint32_t)(csh - p) << 16);
      base += m[u];
    }
    if (threadIdx.x == 63) s_nslot = (int)inc;
  }
  __syncthreads();
  const int nslot = s_nslot;
  for (int q = threadIdx.x; q <= nslot; q += BLOCK) fs[q] = 0xffffu;
  __syncthreads();
  int sp[kPer], si[kPer];
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int i = u * BLOCK + (int)threadIdx.x;
    sp[u] = (i > 0 && i <= S) ? sel_slot(cell, csh, cmask, wq[u]) : -1;
    si[u] = i < S ? sel_slot(cell, csh, cmask, wd[u]) : nslot;
  }
  // fs[s] = i for the first splitter i of slot s; slots without one take
  // the next slot's (a suffix minimum); fs[nslot] = S.
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int i = u * BLOCK + (int)threadIdx.x;
    if (i <= S && si[u] != sp[u]) fs[si[u]] = (uint16_t)i;
  }
#pragma unroll
  for (int u = 0; u < kPer; u++) {
    const int i = u * BLOCK + (int)threadIdx.x;
    if (i <= S) lc[i] = 0u;
  }
  __syncthreads();
  block_suffix_min<BLOCK, (kFine + BLOCK) / BLOCK>(fs, nslot, wmin);
  __syncthreads();
  double mn = __builtin_inf(), mx = -__builtin_inf();
  uint16_t* bx = bix + sd.v_off;
  // U values per thread searched together (independent LDS reads in
  // flight), the next U loaded meanwhile: bucket = number of splitters <=
  // x, as sel_bucket (the order key maps -0.0 and +0.0 together, as the
  // double comparison does).
  for (int64_t i0 = a; i0 < b; i0 += (int64_t)U * BLOCK) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; u++) x[u] = xn[u];
    // The next loads go out before this round's stores (an index past the
    // tile reads its last value, then is ignored).
    const int64_t i1 = i0 + (int64_t)U * BLOCK;
    if (i1 < b) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int64_t i = i1 + (int64_t)u * BLOCK + threadIdx.x;
        xn[u] = v[i < b ? i : b - 1];
      }
    }
    uint64_t kx[U];
    int lo[U], len[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      kx[u] = ord_key(x[u]);
      const bool below = kx[u] < kf, above = kx[u] >= kl;
      const uint32_t d = (below || above) ? 0u : sel_window(kx[u], sh) - wA;
      const int q = sel_slot(cell, csh, cmask, d);
      const int f0 = fs[q], f1 = fs[q + 1];
      lo[u] = below ? 0 : (above ? S : f0);
      len[u] = (below || above) ? 0 : f1 - f0;
    }
    // Upper bound of the value's key among the slot's splitters, in as many
    // rounds as the fullest slot this wave's values hit needs.
    int ml = len[0];
#pragma unroll
    for (int u = 1; u < U; u++) ml = max(ml, len[u]);
    for (int d = 32; d > 0; d >>= 1) ml = max(ml, __shfl_xor(ml, d, 64));
    int rounds = 0;
    while ((ml >> rounds) != 0) rounds++;
    for (int r = 0; r < rounds; r++) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int half = len[u] >> 1;
        const bool right = len[u] > 0 && sk[lo[u] + half] <= kx[u];
        lo[u] = right ? lo[u] + half + 1 : lo[u];
        len[u] = right ? len[u] - half - 1 : half;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = i0 + (int64_t)u * BLOCK + threadIdx.x;
      if (i < b) {
        const int k = lo[u];
        atomicAdd(&lc[k], 1u);
        bx[i] = (uint16_t)k;  // the gather pass reads it back instead of searching again
        mn = fmin(mn, x[u]);
        mx = fmax(mx, x[u]);
      }
    }
  }
  // Zero extremes keep no sign here (the order key maps -0.0 and +0.0
  // together); the batch extremes' zero signs are set afterwards.
  for (int d = 32; d > 0; d >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, d, 64));
    mx = fmax(mx, __shfl_xor(mx, d, 64));
  }
  if ((threadIdx.x & 63) == 0 && mn <= mx) {
    atomicMin(&smin, ord_key(mn));
    atomicMax(&smax, ord_key(mx));
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= S; i += BLOCK)
    if (lc[i]) atomicAdd(&cnt[g.b_off + i], lc[i]);
  if (threadIdx.x == 0) {
    atomicMin(&st[j].kmin, smin);
    atomicMax(&st[j].kmax, smax);
  }
}

// Per segment: bucket ranks and the merged positions of the prior centroids.
__global__ __launch_bounds__(kSelThreads) void k_sel_ranks(const SegDesc* segs, const SelSeg* sel,
                                                           const double* spl,
                                                           const unsigned int* cnt,
                                                           unsigned int* bst, int* binfo,
                                                           const double* esm, int64_t* pose,
                                                           SelState* st) {
  __shared__ uint32_t part[kSelThreads];
  __shared__ int s_fail;
  // Every 16th splitter (the last of each run of 16): a prior centroid's
  // search runs there, then over its run's 16 splitters loaded at once.
  constexpr int kFence = 16;
  __shared__ double sfence[(kSelSplMax + kFence - 1) / kFence];
  const int j = blockIdx.x;
  const SelSeg g = sel[j];
  const SegDesc sd = segs[g.si];
  const int tid = threadIdx.x;
  const int S = st[j].nspl;
  const int64_t nv = sd.v_cnt;
  const int ne = (int)sd.e_cnt;
  const double* sp = spl + g.b_off;
  const unsigned int* c = cnt + g.b_off;
  unsigned int* B = bst + g.b_off;
  int* info = binfo + g.b_off;
  if (tid == 0) s_fail = 0;
  // 1. Exclusive scan of the S + 1 bucket counts; buckets start unmarked.
  const int nb = S + 1;
  const int per = (nb + kSelThreads - 1) / kSelThreads;
  const int lo0 = std::min(nb, tid * per), hi0 = std::min(nb, lo0 + per);
  // A thread's counts are loaded together (at most kPerB of them).
  constexpr int kPerB = (kSelSplMax + 1 + kSelThreads - 1) / kSelThreads;
  uint32_t cv[kPerB];
#pragma unroll
  for (int u = 0; u < kPerB; u++) cv[u] = lo0 + u < hi0 ? c[lo0 + u] : 0u;
  // This thread's first prior mean, read beside the counts.
  const double e_own = tid < ne ? esm[sd.es_off + tid] : 0.0;
  const int nf = (S + kFence - 1) / kFence;
  if (ne > 0)
    for (int k = tid; k < nf; k += blockDim.x) sfence[k] = sp[std::min(k * kFence + kFence - 1, S - 1)];
  uint32_t acc = 0;
#pragma unroll
  for (int u = 0; u < kPerB; u++) {
    acc += cv[u];
    if (lo0 + u < hi0) info[lo0 + u] = 0;
  }
  part[tid] = acc;
  __syncthreads();
  uint32_t total = block_excl_scan(part, kSelThreads);
  uint32_t run = part[tid];
#pragma unroll
  for (int u = 0; u < kPerB; u++) {
    if (lo0 + u < hi0) B[lo0 + u] = run;
    run += cv[u];
  }
  if (tid == 0) {
    B[nb] = total;
    if ((int64_t)total != nv) s_fail = 1;
  }
  __syncthreads();
  // 2. Merged positions of the prior centroids: E precede equal new values.
  for (int i = tid; i < ne; i += blockDim.x) {
    const double e = i == tid ? e_own : esm[sd.es_off + i];
    int l = 0, h = nf;  // first run whose last splitter is >= e
    while (l < h) {
      int mid = (l + h) >> 1;
      if (sfence[mid] < e) l = mid + 1;
      else h = mid;
    }
    // First splitter >= e (e itself is a splitter): inside run l, whose
    // last splitter is >= e; S when there is no such run.
    if (l < nf) {
      const int r0 = l * kFence;
      double v[kFence];
#pragma unroll
      for (int u = 0; u < kFence; u++) v[u] = r0 + u < S ? sp[r0 + u] : __builtin_inf();
      int f = kFence - 1;
#pragma unroll
      for (int u = kFence - 1; u >= 0; u--)
        if (!(v[u] < e)) f = u;
      l = r0 + f;
    } else {
      l = S;
    }
    if (l >= S || !(sp[l] == e)) s_fail = 1;
    else pose[sd.es_off + i] = i + (int64_t)B[l + 1];
  }
  __syncthreads();
  if (tid == 0 && s_fail) st[j].fail = 1;
}

// The greedy sweep in rank space (tdigest.mojo:414-435), one wave per
// segment: the boundary of each cluster is the first candidate b > s with
// NOT (k(P(b+1)/N) - k(P(s)/N) <= 1) (n when none), found by the 64 lanes
// testing 64 candidates at once, first around the estimate from the inverse
// scale function and then evenly spread over the bracket that is left, so a
// cluster costs about one scale-function evaluation of latency.  The prior
// records' merged positions and weight prefixes are staged in LDS when they
// fit.  Same boundaries as chain_next (the predicate is evaluated with the
// same arithmetic; only the order of the probes differs).
constexpr int kChainWaves = 4;  // segments per block
constexpr int kChainGroups = 6;  // clusters tested per speculative pass
// First lane of each speculative group (and 64).
// Lane ranges of the groups (a table the compiler folds into the unrolled
// group loop).
__device__ constexpr int kChainGroup[kChainGroups + 1] = {0, 4, 12, 22, 34, 48, 64};
constexpr int kChainLds = 512;  // prior records staged per segment

__global__ __launch_bounds__(64 * kChainWaves) void k_sel_chain(int scale, const SegDesc* segs,
                                                               const SelSeg* sel, int nsel,
                                                               const double* se, const double* esw,
                                                               const int64_t* pose, SelState* st,
                                                               SelBound* bnd, double* csl,
                                                               double* cwv) {
  __shared__ int64_t lpos[kChainWaves][kChainLds];
  __shared__ double lse[kChainWaves][kChainLds + 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = blockIdx.x * kChainWaves + wv;
  const bool active = j < nsel && !st[j].fail;
  SelSeg g{};
  SegDesc sd{};
  int64_t ne = 0;
  const int64_t* pos = nullptr;
  const double* sep = nullptr;
  if (active) {
    g = sel[j];
    sd = segs[g.si];
    ne = sd.e_cnt;
    pos = pose + sd.es_off;
    sep = se + sd.es_off + g.si;
    if (ne <= kChainLds) {
      for (int i = lane; i < ne; i += 64) lpos[wv][i] = pos[i];
      for (int i = lane; i <= ne; i += 64) lse[wv][i] = sep[i];
    }
  }
  __syncthreads();
  if (!active) return;
  if (ne <= kChainLds) {
    pos = lpos[wv];
    sep = lse[wv];
  }
  const int64_t n = sd.v_cnt + sd.e_cnt;
  const double N = sd.stored_n, nrm = sd.nrm, sf = sd.sf;
  // Prior records before the current start s (the walk only moves up).
  int64_t cur = 0;
  // First index c >= cur whose predicate fails (monotone over [cur, ne):
  // true then false), galloping up from the cursor: O(log distance).
  auto gallop = [&](auto below) -> int64_t {
    if (cur >= ne || !below(cur)) return cur;
    int64_t a = cur, step = 1, b = cur + 1;  // below(a); b: unknown or ne
    while (b < ne && below(b)) {
      a = b;
      step <<= 1;
      b = a + step;
    }
    if (b > ne) b = ne;
    while (b - a > 1) {
      const int64_t mid = a + ((b - a) >> 1);
      if (below(mid)) a = mid;
      else b = mid;
    }
    return b;
  };
  // P(j) for j >= s: weight of the merged records before index j (PCur::P).
  // *cp: the prior records before jj.
  auto Pc = [&](int64_t jj, int64_t* cp) -> double {
    const int64_t c = gallop([&](int64_t i) { return pos[i] < jj; });
    *cp = c;
    return sf * (double)(jj - c) + sep[c];
  };
  auto kvc = [&](int64_t jj, int64_t* cp) -> double { return chain_k(scale, Pc(jj, cp) / N, nrm); };
  auto kv = [&](int64_t jj) -> double {
    int64_t c;
    return kvc(jj, &c);
  };
  // Smallest index with P >= T, estimated (index_for_weight), for T at or
  // above the weight before s.
  auto index_est = [&](double T) -> int64_t {
    // Largest prior record c with weight-before(c) < T.
    const int64_t lo = gallop([&](int64_t i) { return sf * (double)(pos[i] - i) + sep[i] < T; }) - 1;
    const int64_t base_j = lo < 0 ? 0 : pos[lo] + 1;
    const double base_w = lo < 0 ? 0.0 : sf * (double)(pos[lo] - lo) + sep[lo + 1];
    if (base_w >= T) return base_j;
    const double jd = (double)base_j + ceil((T - base_w) / sf);
    const int64_t lim = lo + 1 < ne ? pos[lo + 1] : n;
    if (!(jd < (double)lim)) return lim;
    return (int64_t)jd;
  };
  SelBound* bd = bnd + g.c_off;
  int C = 0;
  // Records a cluster start; false when the segment's capacity is used up.
  auto record = [&](int64_t b) -> bool {
    if (C >= g.cap_c) return false;
    if (lane == 0) bd[C].b = b;
    C++;
    return true;
  };
  // Exact search for the cluster started at s (k value ks): a window of 64
  // candidates at the estimate, then 64 spread evenly over the bracket left.
  // Returns the first closing candidate (n when none) and its k value.
  auto exact_next = [&](int64_t s, double ks, double* ks_next) -> int64_t {
    // Bracket: pred(lo) false (or lo == s), pred(hi) true (or hi == n);
    // klo = k(P(lo + 1) / N) when lo > s.
    int64_t lo = s, hi = n;
    double klo = 0.0;
    bool klo_set = false;
    const double qt = chain_kinv(scale, ks + 1.0, nrm);
    int64_t gg;
    if (!(qt < 1.0)) gg = n - 1;
    else if (!(qt > 0.0)) gg = s + 1;
    else gg = index_est(qt * N) - 1;
    int64_t base = gg - 31;
    bool first = true;
    while (hi - lo > 1) {
      const int64_t span = hi - lo - 1;
      int64_t cand;
      if (first) {
        if (base > hi - 64) base = hi - 64;
        if (base < lo + 1) base = lo + 1;
        cand = base + lane;
      } else {
        cand = lo + 1 + (span >= 64 ? (int64_t)lane * span / 64 : (int64_t)lane);
      }
      first = false;
      const bool live = cand < hi;
      double kr = 0.0;
      bool p = true;  // lanes past hi count as true
      if (live) {
        kr = kv(cand + 1);
        p = !(kr - ks <= 1.0);
      }
      // Candidates increase with the lane: the first true lane closes the
      // bracket from above, the live false lanes before it from below.
      const uint64_t tm = __ballot(p);
      const uint64_t fm = __ballot(live && !p);
      uint64_t below = fm;
      if (tm != 0ull) {
        const int ft = __ffsll((long long)tm) - 1;
        const int64_t cand_ft = lane_value(cand, ft);
        if (__builtin_amdgcn_readlane((int)live, ft) != 0) hi = cand_ft;
        below = fm & ((1ull << ft) - 1ull);  // ft < 64; ft == 0 leaves none
      }
      if (below != 0ull) {
        const int lf = 63 - __clzll((long long)below);
        lo = lane_value(cand, lf);
        klo = lane_value(kr, lf);
        klo_set = true;
      }
    }
    // k of the next start hi: computed for candidate lo = hi - 1 unless lo
    // is still s.
    *ks_next = hi < n ? (klo_set ? klo : kv(hi)) : 0.0;
    return hi;
  };
  int64_t s = 0;
  double ks = kv(0);
  if (!record(0)) {
    if (lane == 0) st[j].fail = 1;
    return;
  }
  while (s < n - 1) {
    // Speculative pass: lane group d tests consecutive candidates around
    // the estimated boundary of the cluster d + 1 ahead (level ks + 1 + d).
    // Each cluster's true level is at most one record's k step below the
    // previous one plus one, so the estimate of group d lands within a
    // record or two for d = 0 and drifts by up to about a record per
    // cluster after that: the groups widen downwards (4, 8, 10, 12, 14 and
    // 16 lanes, each window ending four records above its estimate).  The
    // groups are then checked in order with the exact predicate and the
    // exact k of each new start; a group whose window does not bracket its
    // boundary ends the pass.
    const int grp = (lane >= 4) + (lane >= 12) + (lane >= 22) + (lane >= 34) + (lane >= 48);
    const int gstart = kChainGroup[grp];
    const int gwidth = kChainGroup[grp + 1] - gstart;
    const double qs = chain_kinv(scale, ks + 1.0 + (double)grp, nrm);
    int64_t gs;
    if (!(qs < 1.0)) gs = n - 1;
    else if (!(qs > 0.0)) gs = s + 1;
    else gs = index_est(qs * N) - 1;
    const int64_t cand = gs + (grp == 0 ? -1 : 5 - gwidth) + (lane - gstart);
    const bool low = cand < s;    // before the current start: not evaluated
    const bool over = cand >= n;  // past the last record: a virtual true
    double kr = 0.0;
    int64_t cc = cur;  // prior records before cand + 1
    if (!low && !over) kr = kvc(cand + 1, &cc);
    int64_t sp = s;
    double ksp = ks;
    int64_t csp = cur;  // prior records before sp (the cursor for the next pass)
    int done = 0;
    bool end = false;
#pragma unroll
    for (int d = 0; d < kChainGroups; d++) {
      const uint64_t gm = ((d + 1 < kChainGroups ? 1ull << kChainGroup[d + 1] : 0ull) - 1ull) &
                          ~((1ull << kChainGroup[d]) - 1ull);
      const bool valid = cand > sp;
      const bool p = over || (!low && !(kr - ksp <= 1.0));
      const uint64_t tm = __ballot(valid && p) & gm;
      if (tm == 0ull) break;
      const int f = __ffsll((long long)tm) - 1;
      const int64_t b = lane_value(cand, f);
      // Bracketed when b is the first candidate after sp or the lane before
      // f (candidate b - 1 >= sp, evaluated) is in the group: then its
      // k(P(b) / N) is the next start's k.
      double kb;
      int64_t cb = csp;
      const double kprev = lane_value(kr, f > 0 ? f - 1 : 0);
      if (f == kChainGroup[d]) {
        if (b != sp + 1) break;
        kb = b < n ? kvc(b, &cb) : 0.0;
      } else {
        kb = kprev;
        cb = lane_value(cc, f - 1);
      }
      done++;
      if (b >= n) {  // the open cluster runs to the end
        end = true;
        break;
      }
      if (!record(b)) {
        if (lane == 0) st[j].fail = 1;
        return;
      }
      sp = b;
      ksp = kb;
      csp = cb;
      if (b == n - 1) {  // the last record closes its own cluster
        end = true;
        break;
      }
    }
    if (end) break;
    if (done == 0) {
      double kn;
      const int64_t b = exact_next(s, ks, &kn);
      if (b >= n) break;
      if (!record(b)) {
        if (lane == 0) st[j].fail = 1;
        return;
      }
      sp = b;
      ksp = kn;
      csp = -1;
    }
    s = sp;
    ks = ksp;
    if (csp >= 0) cur = csp;
    else
      while (cur < ne && pos[cur] < s) cur++;
  }
  // Unless the closed-form prefix is exact (integer weights), every boundary
  // is checked in the CPU's summation order (kd_exact.h): the running weight
  // of each cluster, the sums of the clusters before each (in order), then
  // the last absorbed candidate must not close and the first one after it
  // must.  From the first cluster in question on (rare: mostly the last
  // candidate, where q_right is about 1), the chain is walked again with the
  // exact test, each boundary found from the closed-form estimate.  The
  // cluster slices (written by k_sel_plan afterwards) hold the scratch.
  const double* ew = esw + sd.es_off;
  bool closed_exact = ne == 0 ? pow2_weight(sf) : (sf == 1.0 && N < 0x1p53);
  if (closed_exact && ne > 0) {
    bool frac = false;
    for (int64_t i = lane; i < ne; i += 64) frac = frac || ew[i] != floor(ew[i]);
    closed_exact = __ballot(frac) == 0ull;
  }
  if (!closed_exact) {
    auto seek = [&](int64_t a) { return pcur_seek(PCur{pos, sep, ne, sf, 0}, a); };
    double* wt = csl + 2 * g.c_off;
    __threadfence_block();
    // wt[2c]: the cluster's running weight (then the sum of the clusters
    // before it); wt[2c + 1]: the running weight before its last record.
    for (int c = lane; c < C; c += 64) {
      const int64_t a = bd[c].b, b = c + 1 < C ? bd[c + 1].b : n;
      double W = rec_weight(a, sf, pos, ew, ne), Wp = 0.0;
      if (b - a >= 2) {
        Wp = run_weight(a, b - 1, sf, pos, ew, ne, seek(a));
        W = Wp + rec_weight(b - 1, sf, pos, ew, ne);
      }
      wt[2 * c] = W;
      wt[2 * c + 1] = Wp;
      cwv[g.c_off + c] = W;
    }
    __threadfence_block();
    {
      // The sums of the clusters before each, in order: each lane loads one
      // cluster's weight, the wave adds them one lane at a time (the same
      // sum in every lane), each lane keeps the sum before its own.
      double acc = 0.0;
      for (int c0 = 0; c0 < C; c0 += 64) {
        const int c = c0 + lane;
        const double w = c < C ? wt[2 * c] : 0.0;
        const int cnt = C - c0 < 64 ? C - c0 : 64;
        double mine = 0.0;
        for (int l = 0; l < cnt; l++) {
          const double wl = lane_value(w, l);
          if (lane == l) mine = acc;
          acc = acc + wl;
        }
        if (c < C) wt[2 * c] = mine;
      }
    }
    __threadfence_block();
    int first_bad = C;
    for (int c0 = 0; c0 < C; c0 += 64) {
      const int c = c0 + lane;
      bool bad = false;
      if (c < C) {
        const int64_t a = bd[c].b, b = c + 1 < C ? bd[c + 1].b : n;
        const double wsf = wt[2 * c];
        const double kl = chain_k(scale, wsf / N, nrm);
        // The sweep's test with the reference's operands (kd_sweep_closes),
        // k(q_left) shared by both candidates.
        auto closes = [&](double cur, double w) {
          return !(chain_k(scale, (wsf + cur + w) / N, nrm) - kl <= 1.0);
        };
        double W = rec_weight(a, sf, pos, ew, ne);
        if (b - a >= 2) {
          const double wl = rec_weight(b - 1, sf, pos, ew, ne);
          const double Wp = wt[2 * c + 1];
          bad = closes(Wp, wl);
          W = Wp + wl;
        }
        if (b < n) bad = bad || !closes(W, rec_weight(b, sf, pos, ew, ne));
      }
      const uint64_t bm = __ballot(bad);
      if (bm != 0ull) {
        first_bad = c0 + __ffsll((long long)bm) - 1;
        break;
      }
    }
    if (first_bad < C) {
      int64_t s2 = bd[first_bad].b;
      double wsf = wt[2 * first_bad];
      C = first_bad;
      for (;;) {
        if (!record(s2)) {
          if (lane == 0) st[j].fail = 1;
          return;
        }
        const int ci = C - 1;  // the cluster starting at s2
        if (s2 >= n - 1) {
          if (lane == 0) cwv[g.c_off + ci] = run_weight(s2, n, sf, pos, ew, ne, seek(s2));
          break;
        }
        cur = seek(s2);
        double kn;
        int64_t b = exact_next(s2, kv(s2), &kn);
        const int64_t ce = seek(s2);
        auto closes = [&](int64_t i) {
          const double w0 = run_weight(s2, i, sf, pos, ew, ne, ce);
          return chain_closes(scale, nrm, N, wsf, w0, rec_weight(i, sf, pos, ew, ne)) != 0;
        };
        while (b < n && !closes(b)) b++;
        while (b - 1 > s2 && closes(b - 1)) b--;
        if (b >= n) {
          if (lane == 0) cwv[g.c_off + ci] = run_weight(s2, n, sf, pos, ew, ne, ce);
          break;
        }
        const double wc = run_weight(s2, b, sf, pos, ew, ne, ce);
        if (lane == 0) cwv[g.c_off + ci] = wc;
        wsf = wsf + wc;
        s2 = b;
      }
    }
    if (lane == 0) st[j].wdone = 1;
  }
  if (lane == 0) st[j].ncl = C;
}

// Per segment, after the chain: the boundary records, the buckets to gather,
// each other bucket's cluster, and each cluster's slice exponent.
__global__ __launch_bounds__(kSelThreads) void k_sel_plan(
    const SegDesc* segs, const SelSeg* sel, const double* spl, const unsigned int* cnt,
    const unsigned int* bst, int* binfo, const int64_t* pose, SelState* st, SelBound* bnd, int* ce,
    double* csl, SelMark* mk, int2* mlist, unsigned int* mlist_cnt) {
  __shared__ uint32_t part[kSelThreads];
  __shared__ int s_fail, s_base;
  // Bucket ranks and boundary ranks in LDS: every search below runs there.
  __shared__ unsigned int sB[kSelSplMax + 2];
  __shared__ int64_t svr[kSelCmax];
  __shared__ int sbh[kSelCmax], sbl[kSelCmax];  // buckets at and below each boundary
  // The prior centroids' merged positions, when they fit.
  constexpr int kPosLds = 2 * kSelThreads;
  __shared__ int64_t spos[kPosLds];
  // Buckets marked for the gather (one bit each).
  __shared__ uint32_t smark[(kSelSplMax + 1 + 31) / 32];
  const int j = blockIdx.x;
  if (st[j].fail) return;
  const SelSeg g = sel[j];
  const SegDesc sd = segs[g.si];
  const int tid = threadIdx.x;
  const int S = st[j].nspl;
  const int64_t nv = sd.v_cnt;
  const int ne = (int)sd.e_cnt;
  const double* sp = spl + g.b_off;
  const unsigned int* c = cnt + g.b_off;
  const unsigned int* B = bst + g.b_off;
  int* info = binfo + g.b_off;
  SelBound* bd = bnd + g.c_off;
  const int64_t* pos = pose + sd.es_off;
  const int nb = S + 1;
  const int per = (nb + kSelThreads - 1) / kSelThreads;
  const int lo0 = std::min(nb, tid * per), hi0 = std::min(nb, lo0 + per);
  if (tid == 0) s_fail = 0;
  const int C = st[j].ncl;
  // The block's first global reads all in flight together: bucket starts,
  // bucket counts (a thread's consecutive buckets, at most kPerB, held in
  // registers from here on), its boundary's merged rank, prior positions.
  constexpr int kPerB = (kSelSplMax + 1 + kSelThreads - 1) / kSelThreads;
  uint32_t cv[kPerB];
  int64_t b_own = 0;
  const bool pos_lds = ne <= kPosLds;
  {
    constexpr int kPerS = (kSelSplMax + 2 + kSelThreads - 1) / kSelThreads;
    uint32_t tb[kPerS];
    int64_t pv[2];
#pragma unroll
    for (int u = 0; u < kPerS; u++) {
      const int i = u * kSelThreads + tid;
      tb[u] = i <= S + 1 ? B[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kPerB; u++) cv[u] = lo0 + u < hi0 ? c[lo0 + u] : 0u;
    if (tid < C) b_own = bd[tid].b;
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int i = u * kSelThreads + tid;
      pv[u] = pos_lds && i < ne ? pos[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPerS; u++) {
      const int i = u * kSelThreads + tid;
      if (i <= S + 1) sB[i] = tb[u];
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int i = u * kSelThreads + tid;
      if (pos_lds && i < ne) spos[i] = pv[u];
    }
    for (int i = tid; i < (kSelSplMax + 1 + 31) / 32; i += kSelThreads) smark[i] = 0u;
  }
  const int64_t* ps = pos_lds ? spos : pos;
  __syncthreads();
  // 4. Boundary records and the buckets to gather (the buckets of the new
  // values at and just below each boundary, which the finalize reads).
  for (int i = tid; i < C; i += blockDim.x) {
    const int64_t b = i == tid ? b_own : bd[i].b;
    int l = 0, h = ne;  // prior centroids before merged rank b
    while (l < h) {
      int mid = (l + h) >> 1;
      if (ps[mid] < b) l = mid + 1;
      else h = mid;
    }
    int64_t vr = b - l;
    bd[i].vr = vr;
    svr[i] = vr;
    bd[i].v_hi = !(l < ne && ps[l] == b);
    bd[i].v_lo = b > 0 && !(l > 0 && ps[l - 1] == b - 1);
    bd[i].dup = 0;
    const int bh = vr < nv ? sel_bucket_of_rank(sB, S, vr) : -1;
    // Rank vr - 1 is mostly in vr's bucket (one LDS read tells).
    const int bl = vr < 1 ? -1
                   : (bh >= 0 && (int64_t)sB[bh] <= vr - 1) ? bh
                                                             : sel_bucket_of_rank(sB, S, vr - 1);
    bd[i].bk_hi = bh;
    bd[i].bk_lo = bl;
    sbh[i] = bh;
    sbl[i] = bl;
    if (i > 0) {
      if (bh >= 0) atomicOr(&smark[bh >> 5], 1u << (bh & 31));
      if (bl >= 0) atomicOr(&smark[bl >> 5], 1u << (bl & 31));
    }
  }
  __syncthreads();
  // 5. Gather layout of the marked buckets, in bucket order.  A thread's
  // consecutive buckets (at most kPerB) are held in registers from here on.
  int iv[kPerB];
#pragma unroll
  for (int u = 0; u < kPerB; u++) {
    const int i = lo0 + u;
    iv[u] = i < hi0 && ((smark[i >> 5] >> (i & 31)) & 1u) ? -1 : 0;
  }
  uint32_t mc = 0, mv = 0;
#pragma unroll
  for (int u = 0; u < kPerB; u++)
    if (lo0 + u < hi0 && iv[u] < 0) {
      mc++;
      mv += cv[u];
    }
  __syncthreads();
  part[tid] = mc;
  __syncthreads();
  uint32_t nmk = block_excl_scan(part, kSelThreads);
  uint32_t mbase = part[tid];
  __syncthreads();
  part[tid] = mv;
  __syncthreads();
  block_excl_scan(part, kSelThreads);
  uint32_t gbase = part[tid];
  if (tid == 0) {
    st[j].nmk = (int)nmk;
    s_base = (int)atomicAdd(mlist_cnt, nmk);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerB; u++) {
    const int i = lo0 + u;
    if (i < hi0 && iv[u] < 0) {
      SelMark m;
      m.bucket = i;
      m.goff = gbase;
      m.cnt = cv[u];
      m.vstart = sB[i];
      m.cursor = 0;
      m.first_bound = 0;
      m.sum_hi = m.sum_lo = m.vmin = m.vmax = 0.0;
      mk[g.m_off + mbase] = m;
      mlist[s_base + mbase] = make_int2(j, (int)mbase);
      iv[u] = -1 - (int)mbase;
      mbase++;
      gbase += cv[u];
    }
  }
  // 6. Cluster of every other bucket: boundaries i >= 1 with vr_i <= its
  // start (a thread's buckets are consecutive, so a cursor follows them).
  int l6 = -1;
#pragma unroll
  for (int u = 0; u < kPerB; u++) {
    const int i = lo0 + u;
    if (i < hi0 && iv[u] >= 0) {
      const int64_t r = sB[i];
      if (l6 < 0) {
        int l = 1, h = C;
        while (l < h) {
          int mid = (l + h) >> 1;
          if (svr[mid] <= r) l = mid + 1;
          else h = mid;
        }
        l6 = l;
      } else {
        while (l6 < C && svr[l6] <= r) l6++;
      }
      iv[u] = l6 - 1;
    }
  }
#pragma unroll
  for (int u = 0; u < kPerB; u++)
    if (lo0 + u < hi0) info[lo0 + u] = iv[u];
  // 7. Slice exponent per cluster from the value range of its buckets.
  const double vmin = key_value(st[j].kmin), vmax = key_value(st[j].kmax);
  for (int k = tid; k < C; k += blockDim.x) {
    int64_t va = svr[k], vb = k + 1 < C ? svr[k + 1] : nv;
    int e = 0;
    if (vb > va) {
      // Buckets of the cluster's first and last new values (va < vb <= nv).
      const int bf = sbh[k], bl = k + 1 < C ? sbl[k + 1] : sel_bucket_of_rank(sB, S, nv - 1);
      double lo = bf == 0 ? vmin : sp[bf - 1];
      double hi = bl == S ? vmax : sp[bl];
      double m = fmax(fabs(lo), fabs(hi));
      if (m == 0.0) {
        e = -900;
      } else {
        int ex = (int)((kd_d2bits(m) >> 52) & 0x7ff) - 1022;  // m < 2^ex
        if (ex < -900 || ex > 990) s_fail = 1;
        e = ex;
      }
    }
    ce[g.c_off + k] = e;
    csl[2 * (g.c_off + k)] = 0.0;
    csl[2 * (g.c_off + k) + 1] = 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    st[j].ncl = C;
    st[j].blast = nv > 0 ? sel_bucket_of_rank(sB, S, nv - 1) : -1;
    if (s_fail) st[j].fail = 1;
  }
}

// Second pass over the new values: marked buckets are gathered, the rest
// are added to their cluster's slice sums.  A block stages its tile's
// gathered values (and their marked-bucket indices) in stv / stm at the
// tile's own range, takes one place range per marked bucket once the tile is
// done, and copies them there (the order inside a bucket is free: each
// gathered bucket is sorted or selected afterwards).
template <int SMAX, int CMAX>
__global__ __launch_bounds__(kSelThreads, 8) void k_sel_gather(const SegDesc* segs, const SelSeg* sel,
                                                            int nsel, int64_t t0, int64_t t1,
                                                            const double* vin, const uint16_t* bix,
                                                            const int* binfo, const int* ce,
                                                            const SelState* st, SelMark* mk,
                                                            double* csl, double* X, double* stv,
                                                            uint16_t* stm, const int* tmap) {
  constexpr int kItems = 4;                 // values per thread per chunk
  constexpr int kMk = 2 * CMAX + 2;          // gathered buckets of a segment at most
  static_assert(kMk < 65536, "marked-bucket index width");
  // Replicas of the cluster accumulators (lane l adds into replica l % rep):
  // the central clusters of a large digest hold a third of its values each,
  // so one copy would serialize a wave's atomic adds.  The slices are exact
  // fixed-point values, so any grouping of the sums gives the same bits.
  // As many replicas as the segment's clusters leave room for (at most
  // kMaxRep), each 2C + 1 doubles long, so the replicas of one cluster sit
  // in different LDS banks as well.
  constexpr int kAcc = (CMAX <= 256 ? 8 : 1) * (2 * CMAX + 1);
  constexpr int kMaxRep = 16;
  __shared__ int inf[SMAX + 1];
  __shared__ int ex[CMAX];
  __shared__ double acc[kAcc];
  __shared__ unsigned int lcnt[kMk];
  __shared__ unsigned long long lbase[kMk];  // place of the bucket's first staged value
  __shared__ unsigned int nst;
  const int64_t t = t0 + blockIdx.x;
  if (t >= t1) return;
  const int j = tmap[t];
  if (st[j].fail) return;
  const SelSeg g = sel[j];
  const SegDesc sd = segs[g.si];
  const int S = st[j].nspl, C = st[j].ncl, nmk = st[j].nmk;
  const int64_t a = (t - g.tile0) * g.tile_len;
  const int64_t b = std::min<int64_t>(sd.v_cnt, a + g.tile_len);
  const double* v = vin + sd.v_off;
  const uint16_t* bx = bix + sd.v_off;
  // The tile's first values go out before the tables are loaded.
  double xn[kItems];
  uint16_t bn[kItems];
#pragma unroll
  for (int q = 0; q < kItems; q++) {
    const int64_t i = a + (int64_t)q * kSelThreads + threadIdx.x;
    const int64_t ic = i < b ? i : b - 1;
    xn[q] = v[ic];
    bn[q] = bx[ic];
  }
  {
    // Table loads unrolled so that they are all in flight together.
    constexpr int kPerI = (SMAX + 1 + kSelThreads - 1) / kSelThreads;
    constexpr int kPerC = (CMAX + kSelThreads - 1) / kSelThreads;
    int ti[kPerI], tc[kPerC];
#pragma unroll
    for (int u = 0; u < kPerI; u++) {
      const int i = u * kSelThreads + (int)threadIdx.x;
      ti[u] = i <= S ? binfo[g.b_off + i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPerC; u++) {
      const int i = u * kSelThreads + (int)threadIdx.x;
      tc[u] = i < C ? ce[g.c_off + i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPerI; u++) {
      const int i = u * kSelThreads + (int)threadIdx.x;
      if (i <= S) inf[i] = ti[u];
    }
#pragma unroll
    for (int u = 0; u < kPerC; u++) {
      const int i = u * kSelThreads + (int)threadIdx.x;
      if (i < C) ex[i] = tc[u];
    }
  }
  const int stride = 2 * C + 1;
  const int nrep = std::max(1, std::min(kMaxRep, kAcc / stride));
  for (int i = threadIdx.x; i < nrep * stride; i += blockDim.x) acc[i] = 0.0;
  for (int i = threadIdx.x; i < nmk; i += blockDim.x) lcnt[i] = 0u;
  if (threadIdx.x == 0) nst = 0u;
  __syncthreads();
  double* acc_r = acc + (int)(threadIdx.x % (unsigned)nrep) * stride;
  double* sv = stv + sd.v_off + a;
  uint16_t* sm = stm + sd.v_off + a;
  const int lane = (int)(threadIdx.x & 63);
  const int64_t chunk = (int64_t)kItems * kSelThreads;
  for (int64_t c0 = a; c0 < b; c0 += chunk) {
    double xr[kItems];
    uint16_t br[kItems];
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      xr[q] = xn[q];
      br[q] = bn[q];
    }
    // The next chunk's loads go out before this chunk's stores.
    const int64_t c1 = c0 + chunk;
    if (c1 < b) {
#pragma unroll
      for (int q = 0; q < kItems; q++) {
        const int64_t i = c1 + (int64_t)q * kSelThreads + threadIdx.x;
        const int64_t ic = i < b ? i : b - 1;
        xn[q] = v[ic];
        bn[q] = bx[ic];
      }
    }
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      const int64_t i = c0 + (int64_t)q * kSelThreads + threadIdx.x;
      const bool in = i < b;
      const int f = in ? inf[br[q]] : 0;
      if (in && f >= 0) {
        double s1, s2;
        sel_slices(xr[q], ex[f], s1, s2);
        unsafeAtomicAdd(&acc_r[2 * f], s1);
        unsafeAtomicAdd(&acc_r[2 * f + 1], s2);
      }
      // Staged places claimed once per wave.
      const unsigned long long mask = __ballot(in && f < 0);
      if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        unsigned int base = 0u;
        if (lane == leader) base = atomicAdd(&nst, (unsigned int)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (in && f < 0) {
          const unsigned int o = base + __builtin_amdgcn_mbcnt_hi(
                                            (uint32_t)(mask >> 32),
                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
          sv[o] = xr[q];
          sm[o] = (uint16_t)(-1 - f);
          atomicAdd(&lcnt[-1 - f], 1u);
        }
      }
    }
  }
  __syncthreads();
  // One place range per gathered bucket and tile; the counters then hand
  // out the places inside it.
  for (int i = threadIdx.x; i < nmk; i += blockDim.x) {
    const unsigned int c = lcnt[i];
    if (c) lbase[i] = (unsigned long long)mk[g.m_off + i].goff + atomicAdd(&mk[g.m_off + i].cursor, c);
    lcnt[i] = 0u;
  }
  __syncthreads();
  double* xg = X + sd.x_off;
  const int ns = (int)nst;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int m = sm[i];
    const double x = sv[i];
    xg[lbase[m] + atomicAdd(&lcnt[m], 1u)] = x;
  }
  for (int i = threadIdx.x; i < 2 * C; i += blockDim.x) {
    double sum = 0.0;
    for (int r = 0; r < nrep; r++) sum += acc[r * stride + i];
    if (sum != 0.0) unsafeAtomicAdd(&csl[2 * g.c_off + i], sum);
  }
}

// Deterministic block reduction of per-thread double-doubles (fixed tree).
__device__ inline dd block_dd_sum(dd v, double* rh, double* rl) {
  rh[threadIdx.x] = v.hi;
  rl[threadIdx.x] = v.lo;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      dd s2 = dd_add(dd{rh[threadIdx.x], rl[threadIdx.x]}, dd{rh[threadIdx.x + w], rl[threadIdx.x + w]});
      rh[threadIdx.x] = s2.hi;
      rl[threadIdx.x] = s2.lo;
    }
    __syncthreads();
  }
  dd r{rh[0], rl[0]};
  __syncthreads();
  return r;
}

// First boundary i >= 1 whose new-value rank is at least r.
__device__ inline int sel_first_bound(const SelBound* bd, int C, int64_t r) {
  int l = 1, h = C;
  while (l < h) {
    int mid = (l + h) >> 1;
    if (bd[mid].vr < r) l = mid + 1;
    else h = mid;
  }
  return l;
}

// The gathered buckets split by size into the lists of the kernels that
// finish them (at most 128 values: one wave each; 1024 and 4096: LDS sort;
// more: radix select), list c at out + c * stride with its length in
// ccnt[c]; a wave claims its places with one atomic per list.  Each
// bucket's first boundary is found here, all buckets at once, rather than
// by a chain of dependent loads in the kernel that finishes it.
constexpr int kSelClasses = 4;

__global__ __launch_bounds__(256) void k_sel_classes(const SegDesc* segs, const SelSeg* sel,
                                                     const SelState* st, SelMark* mk,
                                                     const SelBound* bnd, const int2* mlist,
                                                     const unsigned int* mlist_cnt, int2* out,
                                                     SelWave* wout, int64_t stride,
                                                     unsigned int* ccnt) {
  const unsigned int total = *mlist_cnt;
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (unsigned int i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < total;
       i0 += gridDim.x * blockDim.x) {
    const unsigned int i = i0 + (unsigned int)lane;
    int c = -1;
    int2 e = make_int2(0, 0);
    SelWave wv{};
    if (i < total) {
      e = mlist[i];
      const SelSeg g = sel[e.x];
      SelMark* mp = mk + g.m_off + e.y;
      const int64_t n = mp->cnt, vs = mp->vstart;
      const int C = st[e.x].ncl;
      const int fb = sel_first_bound(bnd + g.c_off, C, vs);
      mp->first_bound = fb;
      // Buckets of failed segments are finished by the sorting path.
      c = st[e.x].fail ? -1 : (n <= 128 ? 0 : n <= 1024 ? 1 : n <= 4096 ? 2 : 3);
      if (c == 0) {
        wv.src = segs[g.si].x_off + mp->goff;
        wv.vstart = vs;
        wv.c_off = g.c_off;
        wv.mark = g.m_off + e.y;
        wv.n = (int)n;
        wv.b0 = fb;
        wv.b1 = sel_first_bound(bnd + g.c_off, C, vs + n);
      }
    }
#pragma unroll
    for (int q = 0; q < kSelClasses; q++) {
      const uint64_t m = __ballot(c == q);
      if (m == 0ull) continue;
      unsigned int base = 0;
      const int leader = __ffsll((long long)m) - 1;
      if (lane == leader) base = atomicAdd(&ccnt[q], (unsigned int)__popcll(m));
      base = __shfl(base, leader, 64);
      const unsigned int at = base + (unsigned int)__popcll(m & below);
      if (c == q) {
        if (q == 0) wout[at] = wv;
        else out[(int64_t)q * stride + at] = e;
      }
    }
  }
}

// Per gathered bucket of at most CAP values (and more than LO): sorted in
// LDS; its sum and extremes, and for every boundary rank inside it the
// values at that rank and just below and the sum of the values below.
template <int CAP, int BLOCK, int LO>
__global__ __launch_bounds__(BLOCK) void k_sel_sorted(const SegDesc* segs, const SelSeg* sel,
                                                      const SelState* st, const int2* mlist,
                                                      const unsigned int* mlist_cnt, SelMark* mk,
                                                      SelBound* bnd, const double* X) {
  __shared__ double sa[CAP + CAP / 16], sb[CAP + CAP / 16];
  __shared__ double rh[BLOCK], rl[BLOCK];
  const unsigned int total = *mlist_cnt;
  for (unsigned int gi = blockIdx.x; gi < total; gi += gridDim.x) {
    const int2 e = mlist[gi];
    const int j = e.x;
    if (st[j].fail) continue;
    const SelSeg g = sel[j];
    SelMark* mp = mk + g.m_off + e.y;
    const int64_t n = mp->cnt;
    if (n <= LO || n > CAP) continue;
    const SegDesc sd = segs[g.si];
    const double* src = X + sd.x_off + mp->goff;
    {
      // All of the bucket's loads in flight at once (CAP / BLOCK per thread).
      constexpr int kPer = CAP / BLOCK;
      double r[kPer];
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        r[u] = i < (int)n ? src[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPer; u++) {
        const int i = u * BLOCK + (int)threadIdx.x;
        if (i < (int)n) sa[padi(i)] = r[u];
      }
    }
    __syncthreads();
    const double* s = lds_sort(sa, sb, (int)n);
    dd acc{0.0, 0.0};
    for (int i = threadIdx.x; i < (int)n; i += BLOCK) acc = dd_add(acc, dd{s[padi(i)], 0.0});
    dd tot = block_dd_sum(acc, rh, rl);
    if (threadIdx.x == 0) {
      mp->sum_hi = tot.hi;
      mp->sum_lo = tot.lo;
      mp->vmin = s[padi(0)];
      mp->vmax = s[padi((int)n - 1)];
    }
    const int C = st[j].ncl;
    SelBound* bd = bnd + g.c_off;
    const int64_t vstart = mp->vstart;
    for (int i = mp->first_bound; i < C && bd[i].vr < vstart + n; i++) {
      const int rho = (int)(bd[i].vr - vstart);
      dd sl{0.0, 0.0};
      for (int q = threadIdx.x; q < rho; q += BLOCK) sl = dd_add(sl, dd{s[padi(q)], 0.0});
      dd sls = block_dd_sum(sl, rh, rl);
      if (threadIdx.x == 0) {
        bd[i].x_hi = s[padi(rho)];
        bd[i].x_lo = rho > 0 ? s[padi(rho - 1)] : 0.0;
        bd[i].dup = rho > 0 && s[padi(rho - 1)] == s[padi(rho)];
        bd[i].sl_hi = sls.hi;
        bd[i].sl_lo = sls.lo;
      }
    }
    __syncthreads();
  }
}

// Gathered buckets of at most 64 * PER values (and more than LO), one wave
// each: the values sorted in registers by a bitonic network (lane-crossing
// stages through shuffles, no LDS and no block barrier), then the same
// outputs as k_sel_sorted: the bucket's sum and extremes, and for every
// boundary rank inside it the values at that rank and just below and the sum
// of the values below (double-double, butterfly over the wave).
__device__ inline dd wave_dd_sum(dd v) {
  for (int m = 32; m > 0; m >>= 1) {
    dd o{__shfl_xor(v.hi, m, 64), __shfl_xor(v.lo, m, 64)};
    v = dd_add(v, o);
  }
  return v;
}

// One bucket of at most KMAX <= 64 PER values, sorted by its wave (element
// lane * PER + p in v[p]).  The network stops at runs of KMAX: the first
// run (ascending) holds the bucket, the padding above it is all +inf.
template <int PER, int KMAX>
__device__ inline void sel_sorted_wave_one(const SelWave& w, int lane, SelMark* mk, SelBound* bnd,
                                           const double* X) {
  const double inf = __builtin_inf();
  const int64_t n = w.n;
  const double* src = X + w.src;
  // Element lane * PER + p sits in v[p].
  double v[PER];
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int idx = lane * PER + p;
    v[p] = idx < n ? src[idx] : inf;
  }
  SelBound* bd = bnd + w.c_off;
  // The boundaries' ranks, one per lane (64 at a time), the first 64
  // loaded beside the values.
  const int nb = w.b1 - w.b0;
  int64_t vr_l = lane < nb ? bd[w.b0 + lane].vr : 0;
  for (int k = 2; k <= KMAX; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      if (jj < PER) {
#pragma unroll
        for (int p = 0; p < PER; p++) {
          const int q = p ^ jj;
          if (q > p) {
            const int idx = lane * PER + p;
            const bool asc = (idx & k) == 0;
            const double lo = __builtin_fmin(v[p], v[q]), hi = __builtin_fmax(v[p], v[q]);
            v[p] = asc ? lo : hi;
            v[q] = asc ? hi : lo;
          }
        }
      } else {
        const int lm = jj / PER;
#pragma unroll
        for (int p = 0; p < PER; p++) {
          const int idx = lane * PER + p;
          const double o = __shfl_xor(v[p], lm, 64);
          const bool asc = (idx & k) == 0, lower = (idx & jj) == 0;
          const double lo = __builtin_fmin(v[p], o), hi = __builtin_fmax(v[p], o);
          v[p] = (asc == lower) ? lo : hi;
        }
      }
    }
  }
  dd acc{0.0, 0.0};
#pragma unroll
  for (int p = 0; p < PER; p++)
    if (lane * PER + p < n) acc = dd_add(acc, dd{v[p], 0.0});
  const dd tot = wave_dd_sum(acc);
  // Element r of the sorted bucket (every lane gets it).
  auto elem = [&](int r) {
    double x = v[0];
#pragma unroll
    for (int p = 1; p < PER; p++)
      if ((r % PER) == p) x = v[p];
    return __shfl(x, r / PER, 64);
  };
  const double vmin = elem(0), vmax = elem((int)n - 1);
  if (lane == 0) {
    SelMark* mp = mk + w.mark;
    mp->sum_hi = tot.hi;
    mp->sum_lo = tot.lo;
    mp->vmin = vmin;
    mp->vmax = vmax;
  }
  for (int q = 0; q < nb; q++) {
    if (q > 0 && (q & 63) == 0) vr_l = q + lane < nb ? bd[w.b0 + q + lane].vr : 0;
    const int rho = (int)(lane_value(vr_l, q & 63) - w.vstart);
    dd sl{0.0, 0.0};
#pragma unroll
    for (int p = 0; p < PER; p++)
      if (lane * PER + p < rho) sl = dd_add(sl, dd{v[p], 0.0});
    const dd sls = wave_dd_sum(sl);
    const double xh = elem(rho), xl = rho > 0 ? elem(rho - 1) : 0.0;
    if (lane == 0) {
      SelBound& o = bd[w.b0 + q];
      o.x_hi = xh;
      o.x_lo = xl;
      o.dup = rho > 0 && xl == xh;
      o.sl_hi = sls.hi;
      o.sl_lo = sls.lo;
    }
  }
}

__global__ __launch_bounds__(256) void k_sel_sorted_wave(const SelWave* wl,
                                                         const unsigned int* wl_cnt, SelMark* mk,
                                                         SelBound* bnd, const double* X) {
  const unsigned int total = *wl_cnt;
  const int lane = threadIdx.x & 63;
  const unsigned int waves = gridDim.x * (blockDim.x >> 6);
  for (unsigned int gi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); gi < total;
       gi += waves) {
    const SelWave w = wl[gi];
    // At most 64 values: one per lane (fewer network stages, half the
    // shuffles of two per lane); at most 32: the network's runs stop at 32.
    if (w.n <= 32) sel_sorted_wave_one<1, 32>(w, lane, mk, bnd, X);
    else if (w.n <= 64) sel_sorted_wave_one<1, 64>(w, lane, mk, bnd, X);
    else sel_sorted_wave_one<2, 128>(w, lane, mk, bnd, X);
  }
}

// Gathered buckets of more than LO values: radix select on order keys (eight
// bits per pass) of the values at every boundary rank inside, up to
// kMultiSel ranks per select (one histogram each), then one pass for the
// sums below every selected key; from LDS when the bucket fits there, else
// from memory.
// The bucket fills most of a CU's LDS, so one block runs per CU: it has
// the most threads a block may have.
constexpr int kMultiSel = 16;
constexpr int kSelectThreads = 1024;

// Sum over the block of exact fixed-point slice sums (any order of the adds
// gives the same double-double): butterfly in each wave, then the waves'.
__device__ inline dd block_exact_sum(dd v, double* wh, double* wl) {
  v = wave_dd_sum(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wh[w] = v.hi;
    wl[w] = v.lo;
  }
  __syncthreads();
  dd r{0.0, 0.0};
  for (int q = 0; q < nw; q++) r = dd_add(r, dd{wh[q], wl[q]});
  __syncthreads();
  return r;
}

template <int LO>
__global__ __launch_bounds__(kSelectThreads) void k_sel_select(const SegDesc* segs,
                                                               const SelSeg* sel,
                                                               const SelState* st,
                                                               const int2* mlist,
                                                               const unsigned int* mlist_cnt,
                                                               SelMark* mk, SelBound* bnd,
                                                               const double* X) {
  __shared__ double vals[kSelLdsVals];
  __shared__ uint32_t mh[kMultiSel][256];
  __shared__ double rh[kSelectThreads / 64], rl[kSelectThreads / 64];
  __shared__ unsigned long long smin, smax;
  __shared__ unsigned long long s_pre[kMultiSel], s_rk[kMultiSel];
  // Group sums, largest keys and counts in kGrpRep replicas (lane l adds into
  // replica l % kGrpRep): a bucket's values mostly fall in one or two groups.
  constexpr int kGrpRep = 8;
  __shared__ double g1[kMultiSel * kGrpRep], g2[kMultiSel * kGrpRep];
  __shared__ unsigned long long gmr[kMultiSel * kGrpRep];
  __shared__ unsigned int gcn[kMultiSel * kGrpRep];
  __shared__ int s_ib0, s_ib1;
  const unsigned int total = *mlist_cnt;
  for (unsigned int gi = blockIdx.x; gi < total; gi += gridDim.x) {
    const int2 e = mlist[gi];
    const int j = e.x;
    if (st[j].fail) continue;
    const SelSeg g = sel[j];
    SelMark* mp = mk + g.m_off + e.y;
    const int64_t n = mp->cnt, vstart = mp->vstart;
    if (n <= LO) continue;
    const SegDesc sd = segs[g.si];
    const double* src = X + sd.x_off + mp->goff;
    const double* a = src;
    // f(x) for every value of the bucket, each thread's in index order;
    // eight loads per thread in flight (memory latency, not bandwidth, sets
    // the pace of a lone block).
    auto each = [&](const double* p, auto&& f) {
      constexpr int kU = 8;
      for (int64_t i0 = threadIdx.x; i0 < n; i0 += kU * (int64_t)blockDim.x) {
        double r[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
          const int64_t i = i0 + u * (int64_t)blockDim.x;
          r[u] = i < n ? p[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (i0 + u * (int64_t)blockDim.x < n) f(i0 + u * (int64_t)blockDim.x, r[u]);
      }
    };
    if (n <= kSelLdsVals) {
      each(src, [&](int64_t i, double x) { vals[i] = x; });
      a = vals;
    }
    if (threadIdx.x == 0) {
      smin = ~0ull;
      smax = 0ull;
    }
    __syncthreads();
    // The gathered order is not deterministic: every sum below is of exact
    // fixed-point slices (with the bucket's magnitude bound), so it does not
    // depend on the order of the additions.
    unsigned long long kmn = ~0ull, kmx = 0ull;
    each(a, [&](int64_t, double x) {
      unsigned long long k = ord_key(x);
      kmn = k < kmn ? k : kmn;
      kmx = k > kmx ? k : kmx;
    });
    for (int d = 32; d > 0; d >>= 1) {
      const unsigned long long on = __shfl_xor(kmn, d, 64), ox = __shfl_xor(kmx, d, 64);
      kmn = on < kmn ? on : kmn;
      kmx = ox > kmx ? ox : kmx;
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&smin, kmn);
      atomicMax(&smax, kmx);
    }
    __syncthreads();
    const int ex = sel_exponent(key_value(smin), key_value(smax));
    double t1 = 0.0, t2 = 0.0;
    each(a, [&](int64_t, double x) {
      double s1, s2;
      sel_slices(x, ex, s1, s2);
      t1 += s1;
      t2 += s2;
    });
    dd tot = block_exact_sum(dd{t1, t2}, rh, rl);
    if (threadIdx.x == 0) {
      mp->sum_hi = tot.hi;
      mp->sum_lo = tot.lo;
      mp->vmin = key_value(smin);
      mp->vmax = key_value(smax);
    }
    const int C = st[j].ncl;
    SelBound* bd = bnd + g.c_off;
    if (threadIdx.x == 0) {
      int i0 = mp->first_bound, i1 = i0;
      while (i1 < C && bd[i1].vr < vstart + n) i1++;
      // A boundary at the bucket's first rank: its value is the minimum and
      // nothing of the bucket lies below it.
      if (i0 < i1 && bd[i0].vr == vstart) {
        bd[i0].x_hi = key_value(smin);
        bd[i0].sl_hi = 0.0;
        bd[i0].sl_lo = 0.0;
        i0++;
      }
      s_ib0 = i0;
      s_ib1 = i1;
    }
    __syncthreads();
    const int ib0 = s_ib0, ib1 = s_ib1;
    const int lane = threadIdx.x & 63;
    // Every key of the bucket shares the leading bits of its smallest and
    // largest: the radix passes start at the first bit where those differ.
    const unsigned long long kmin_b = smin, kmax_b = smax;
    const int common = kmin_b == kmax_b ? 64 : __clzll(kmin_b ^ kmax_b);
    const unsigned long long cmask = common == 0 ? 0ull : (~0ull << (64 - common));
    for (int gb = ib0; gb < ib1; gb += kMultiSel) {
      const int nbg = min(kMultiSel, ib1 - gb);
      if ((int)threadIdx.x < nbg) {
        s_pre[threadIdx.x] = kmin_b & cmask;
        s_rk[threadIdx.x] = (unsigned long long)(bd[gb + threadIdx.x].vr - vstart);
      }
      unsigned long long pmask = cmask;
      for (int top = 64 - common; top > 0; top -= 8) {
        const int width = top < 8 ? top : 8;
        const int shift = top - width;
        const uint32_t dmask = (1u << width) - 1u;
        for (int i = threadIdx.x; i < nbg * 256; i += blockDim.x) (&mh[0][0])[i] = 0u;
        __syncthreads();
        unsigned long long pre[kMultiSel];
#pragma unroll
        for (int q = 0; q < kMultiSel; q++) pre[q] = q < nbg ? s_pre[q] : ~0ull;
        each(a, [&](int64_t, double x) {
          const unsigned long long k = ord_key(x);
          const unsigned long long km = k & pmask;
          const uint32_t d = (uint32_t)(k >> shift) & dmask;
#pragma unroll
          for (int q = 0; q < kMultiSel; q++)
            if (q < nbg && km == pre[q]) atomicAdd(&mh[q][d], 1u);
        });
        __syncthreads();
        // Wave w finds the digit holding the rank of boundaries w, w + 4,
        // ...: lane l scans bins 4l .. 4l + 3 after the lanes before it.
        for (int q = (int)(threadIdx.x >> 6); q < nbg; q += (int)(blockDim.x >> 6)) {
          uint32_t c[4];
          uint32_t sum = 0;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            c[u] = mh[q][4 * lane + u];
            sum += c[u];
          }
          uint32_t inc = sum;
          for (int dl = 1; dl < 64; dl <<= 1) {
            const uint32_t y = __shfl_up(inc, dl, 64);
            if (lane >= dl) inc += y;
          }
          const unsigned long long r = s_rk[q];
          unsigned long long cum = inc - sum;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (cum <= r && r < cum + c[u]) {
              s_rk[q] = r - cum;
              s_pre[q] |= (unsigned long long)(4 * lane + u) << shift;
            }
            cum += c[u];
          }
        }
        pmask |= (unsigned long long)dmask << shift;
        __syncthreads();
      }
      // The selected keys ascend with the ranks: a value goes to the group
      // of the first key above it (it is below that key and every later
      // one); the slice sums are exact, so the order of the adds is free.
      unsigned long long K[kMultiSel];
#pragma unroll
      for (int q = 0; q < kMultiSel; q++) K[q] = q < nbg ? s_pre[q] : ~0ull;
      if ((int)threadIdx.x < kMultiSel * kGrpRep) {
        g1[threadIdx.x] = 0.0;
        g2[threadIdx.x] = 0.0;
        gmr[threadIdx.x] = 0ull;
        gcn[threadIdx.x] = 0u;
      }
      __syncthreads();
      const int rr = lane % kGrpRep;
      each(a, [&](int64_t, double x) {
        const unsigned long long k = ord_key(x);
        int q = 0;
#pragma unroll
        for (int u = 0; u < kMultiSel; u++) q += (u < nbg && !(k < K[u])) ? 1 : 0;
        if (q < nbg) {
          double s1, s2;
          sel_slices(x, ex, s1, s2);
          const int gi = q * kGrpRep + rr;
          unsafeAtomicAdd(&g1[gi], s1);
          unsafeAtomicAdd(&g2[gi], s2);
          atomicMax(&gmr[gi], k);
          atomicAdd(&gcn[gi], 1u);
        }
      });
      __syncthreads();
      if (threadIdx.x == 0) {
        double l1 = 0.0, l2 = 0.0;
        unsigned long long kl = 0ull;
        unsigned int cl = 0;
        for (int q = 0; q < nbg; q++) {
          for (int r = 0; r < kGrpRep; r++) {
            l1 += g1[q * kGrpRep + r];
            l2 += g2[q * kGrpRep + r];
            kl = gmr[q * kGrpRep + r] > kl ? gmr[q * kGrpRep + r] : kl;
            cl += gcn[q * kGrpRep + r];
          }
          const dd sls = two_sum(l1, l2);
          SelBound& o = bd[gb + q];
          o.x_hi = key_value(K[q]);
          o.x_lo = key_value(kl);
          o.sl_hi = sls.hi;
          o.sl_lo = sls.lo;
          o.dup = (int64_t)cl != o.vr - vstart;
        }
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

// Value of the new value at rank vr - 1 for boundary i (the previous
// bucket's largest value when vr starts its bucket).
__device__ inline double sel_value_below(const SelBound& bd, const int* info, const SelMark* m) {
  const int bh = bd.bk_hi, bl = bd.bk_lo;  // bh < 0: rank vr is past the new values
  if (bh >= 0 && bh == bl) return bd.x_lo;
  if (info[bl] >= 0) return __builtin_nan("");  // not gathered: cannot happen by construction
  return m[-1 - info[bl]].vmax;
}

// Sum of the values of a gathered bucket below its local rank r, where r is
// 0, the bucket size, or the rank of boundary `bi`.
__device__ inline dd sel_prefix(const SelMark& m, int64_t r, const SelBound* bd, int bi) {
  if (r <= 0) return dd{0.0, 0.0};
  if (r >= m.cnt) return dd{m.sum_hi, m.sum_lo};
  return dd{bd[bi].sl_hi, bd[bi].sl_lo};
}

// Tie checks at every boundary, then each cluster's weight and mean.
// cwv: the running weights the chain verified (when st[j].wdone).
__global__ __launch_bounds__(kBlock) void k_sel_finalize(SegDesc* segs, const SelSeg* sel,
                                                         const SelState* st, const double* cwv,
                                                         const SelBound* bnd,
                                                         const SelMark* mk, const int* binfo,
                                                         const double* csl,
                                                         const double* esm, const double* esw,
                                                         const int64_t* pose, double* cm,
                                                         double* cw, int64_t* ccount,
                                                         SegResult* res, unsigned int* nfallback) {
  __shared__ int s_fail;
  const int j = blockIdx.x;
  const SelSeg g = sel[j];
  const SegDesc sd = segs[g.si];
  if (threadIdx.x == 0) s_fail = st[j].fail;
  __syncthreads();
  const int C = st[j].ncl;
  const int64_t nv = sd.v_cnt, ne = sd.e_cnt, n = nv + ne;
  const SelBound* bd = bnd + g.c_off;
  const int* info = binfo + g.b_off;
  const SelMark* m = mk + g.m_off;
  const double* em = esm + sd.es_off;
  const double* ew = esw + sd.es_off;
  const int64_t* pos = pose + sd.es_off;
  if (!s_fail) {
    for (int i = 1 + threadIdx.x; i < C; i += blockDim.x) {
      const SelBound b = bd[i];
      const int64_t ce_ = b.b - b.vr;  // prior centroids before the boundary
      double xh = b.v_hi ? b.x_hi : em[ce_];
      double xl = b.v_lo ? sel_value_below(b, info, m) : em[ce_ - 1];
      bool dup = b.v_hi && b.v_lo && b.dup;
      if (dup || xh != xh || xl != xl || !(xl <= xh) || ord_key(xh) - ord_key(xl) <= 2ull) s_fail = 1;
    }
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) {
      segs[g.si].mode = 0;
      atomicAdd(nfallback, 1u);
    }
    return;
  }
  const double vmin = key_value(st[j].kmin), vmax = key_value(st[j].kmax);
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    const SelBound ba = bd[c];
    const int64_t bb = c + 1 < C ? bd[c + 1].b : n;
    const int64_t vra = ba.vr, vrb = c + 1 < C ? bd[c + 1].vr : nv;
    const int64_t ea = ba.b - vra, eb = bb - vrb;
    const int64_t nV = vrb - vra;
    dd ewt{0.0, 0.0}, esum{0.0, 0.0};
    for (int64_t i = ea; i < eb; i++) {
      ewt = dd_add(ewt, dd{ew[i], 0.0});
      esum = dd_add(esum, two_prod(em[i], ew[i]));
    }
    dd wsum = dd_add(two_prod(sd.sf, (double)nV), ewt);
    double w = wsum.hi + wsum.lo;
    // Stored weight: the running weight in record order, as the CPU sums it
    // (the chain's when it checked the boundaries with it).
    const double wrun = st[j].wdone ? cwv[g.c_off + c] : run_weight(ba.b, bb, sd.sf, pos, ew, ne, ea);
    dd sv = two_sum(csl[2 * (g.c_off + c)], csl[2 * (g.c_off + c) + 1]);
    if (nV > 0) {
      // Buckets of the first and last new values of the cluster.
      const int bf = ba.bk_hi, bl = c + 1 < C ? bd[c + 1].bk_lo : st[j].blast;
      if (info[bf] < 0) {
        const SelMark& ma = m[-1 - info[bf]];
        int64_t hi_r = std::min<int64_t>(vrb, ma.vstart + ma.cnt) - ma.vstart;
        dd hi_s = sel_prefix(ma, hi_r, bd, c + 1);
        dd lo_s = sel_prefix(ma, vra - ma.vstart, bd, c);
        sv = dd_add(sv, dd_add(hi_s, dd_neg(lo_s)));
      }
      if (bl != bf && info[bl] < 0) {
        const SelMark& mb = m[-1 - info[bl]];
        sv = dd_add(sv, sel_prefix(mb, vrb - mb.vstart, bd, c + 1));
      }
    }
    dd tot = dd_add(dd_mul_d(sv, sd.sf), esum);
    // First and last records of the cluster (merged order is sorted).
    bool first_v = c == 0 ? !(ne > 0 && pos[0] == 0) : (bool)ba.v_hi;
    double xf = first_v ? (c == 0 ? vmin : ba.x_hi) : em[ea];
    double xl;
    if (c + 1 < C) {
      const SelBound b1 = bd[c + 1];
      xl = b1.v_lo ? sel_value_below(b1, info, m) : em[eb - 1];
    } else {
      bool last_v = !(ne > 0 && pos[ne - 1] == n - 1);
      xl = last_v ? vmax : em[ne - 1];
    }
    double mean;
    if (xf == xl) mean = (xf == 0.0 && bb - ba.b >= 2) ? 0.0 : xf;
    else mean = dd_div(tot, w);
    cm[sd.x_off + c] = mean;
    cw[sd.x_off + c] = wrun;
  }
  if (threadIdx.x == 0) {
    ccount[g.si] = C;
    res[g.si].vmin = vmin;
    res[g.si].vmax = vmax;
    res[g.si].replays = 0;
  }
}

// --------------------------------------------------------- 6. finish
// Result of a segment: count, total, extremes.
__device__ inline void finish_result(const SegDesc& sd, int si, int64_t C, double total,
                                     double removed, const double* X, const double* vs,
                                     SegResult* res) {
  SegResult r;
  r.count = C;
  r.total = total;
  r.removed = removed;
  r.replays = res[si].replays;
  r.path = sd.mode;
  if (sd.mode != 0) {
    r.vmin = res[si].vmin;
    r.vmax = res[si].vmax;
  } else if (sd.v_cnt > 0) {
    const double* v = sd.e_cnt == 0 ? X + sd.x_off : vs + sd.v_off;
    r.vmin = v[0];
    r.vmax = v[sd.v_cnt - 1];
  } else {
    r.vmin = 0.0;
    r.vmax = 0.0;
  }
  res[si] = r;
}

// A batch minimum or maximum that is zero takes the sign of the first zero in
// batch order: the reference starts from values[0] and replaces it only by a
// strictly smaller (larger) value (tdigest.mojo:272-281), so among -0.0 and
// +0.0 the first one seen stays.  The paths above find the extremes from
// sorted or order-keyed values, which do not tell the two zeros apart; this
// pass rereads only the segments whose extreme is zero, each up to its first
// zero.  Blocks take 256 segments at a time and list the flagged ones.
__global__ __launch_bounds__(256) void k_zero_signs(const SegDesc* segs, int nseg,
                                                    const double* v, SegResult* res) {
  __shared__ int list[256];
  __shared__ int nl;
  __shared__ unsigned long long first;
  for (int base = blockIdx.x * 256; base < nseg; base += gridDim.x * 256) {
    if (threadIdx.x == 0) nl = 0;
    __syncthreads();
    const int si = base + (int)threadIdx.x;
    if (si < nseg && segs[si].v_cnt > 0 && (res[si].vmin == 0.0 || res[si].vmax == 0.0))
      list[atomicAdd(&nl, 1)] = si;
    __syncthreads();
    const int n = nl;
    for (int t = 0; t < n; t++) {
      const int sg = list[t];
      const double* x = v + segs[sg].v_off;
      const int64_t cnt = segs[sg].v_cnt;
      if (threadIdx.x == 0) first = (unsigned long long)cnt;
      __syncthreads();
      for (int64_t c0 = 0; c0 < cnt; c0 += 256 * 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const int64_t i = c0 + j * 256 + threadIdx.x;
          if (i < cnt && x[i] == 0.0) atomicMin(&first, (unsigned long long)i);
        }
        __syncthreads();
        const bool found = first < (unsigned long long)cnt;
        __syncthreads();
        if (found) break;
      }
      if (threadIdx.x == 0 && first < (unsigned long long)cnt) {
        const double z = x[first];
        if (res[sg].vmin == 0.0) res[sg].vmin = z;
        if (res[sg].vmax == 0.0) res[sg].vmax = z;
      }
      __syncthreads();
    }
  }
}

// Many segments: one lane each, walking its centroids.
// only: -1 every segment, 0 the sort path's (mode 0), 1 the others.
__global__ void k_finish_lane(const SegDesc* segs, int nseg, const int64_t* ccount, double* cm,
                              double* cw, const double* X, const double* vs, SegResult* res,
                              int only) {
  int si = blockIdx.x * blockDim.x + threadIdx.x;
  if (si >= nseg) return;
  SegDesc sd = segs[si];
  if (only >= 0 && (sd.mode == 0) != (only == 0)) return;
  int64_t C = ccount[si];
  double* m = cm + sd.x_off;
  double* w = cw + sd.x_off;
  // _gpu_batch_finish (tdigest.mojo:304-313): total of the new centroids.
  double total = 0.0;
  for (int64_t c = 0; c < C; c++) total += w[c];
  // _prune (tdigest.mojo:441-458) when decay is active (floor > 0).
  double removed = 0.0;
  if (sd.floor > 0.0) {
    int64_t k = 0;
    for (int64_t c = 0; c < C; c++) {
      if (w[c] >= sd.floor) {
        m[k] = m[c];
        w[k] = w[c];
        k++;
      } else {
        removed += w[c];
      }
    }
    if (removed > 0.0) {
      total -= removed;
      C = k;
    }
  }
  finish_result(sd, si, C, total, removed, X, vs, res);
}

// Few segments (a lone digest's centroids are a long dependent walk for one
// lane): one wave per segment (kFinishWaves per block); the weights are
// loaded 64 at a time and summed in centroid order through lane broadcasts
// (the same sequential sums as one lane walking them), and the prune
// compacts by ballot.
constexpr int kFinishWaves = 4;
constexpr int kFinishWaveSegs = 2048;  // at most this many segments take the wave form

// The zero-extreme rule of k_zero_signs for one segment, by one wave (the
// few-segment finish does it in the same launch).  Lane 0 wrote res[si].
__device__ inline void wave_zero_sign(const SegDesc& sd, int si, const double* v, SegResult* res,
                                      int lane) {
  int need = 0;
  if (lane == 0) need = sd.v_cnt > 0 && (res[si].vmin == 0.0 || res[si].vmax == 0.0);
  if (__shfl(need, 0, 64) == 0) return;
  const double* x = v + sd.v_off;
  const int64_t cnt = sd.v_cnt;
  for (int64_t c0 = 0; c0 < cnt; c0 += 64 * 8) {
    double r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int64_t i = c0 + j * 64 + lane;
      r[j] = i < cnt ? x[i] : 1.0;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint64_t m = __ballot(c0 + j * 64 + lane < cnt && r[j] == 0.0);
      if (m != 0ull) {
        if (lane == 0) {
          const double z = x[c0 + j * 64 + (__ffsll((long long)m) - 1)];
          if (res[si].vmin == 0.0) res[si].vmin = z;
          if (res[si].vmax == 0.0) res[si].vmax = z;
        }
        return;
      }
    }
  }
}

// v: the batch's values when the finish also applies the zero-extreme rule
// (nullptr: k_zero_signs runs after it).
__global__ __launch_bounds__(64 * kFinishWaves) void k_finish(const SegDesc* segs, int nseg,
                                                             const int64_t* ccount, double* cm,
                                                             double* cw, const double* X,
                                                             const double* vs, SegResult* res,
                                                             int only, const double* v) {
  const int lane = threadIdx.x & 63;
  const int si = blockIdx.x * kFinishWaves + (threadIdx.x >> 6);
  if (si >= nseg) return;
  SegDesc sd = segs[si];
  if (only >= 0 && (sd.mode == 0) != (only == 0)) return;
  int64_t C = ccount[si];
  double* m = cm + sd.x_off;
  double* w = cw + sd.x_off;
  // _gpu_batch_finish (tdigest.mojo:304-313): total of the new centroids.
  double total = 0.0;
  for (int64_t c0 = 0; c0 < C; c0 += 64) {
    const double wl = c0 + lane < C ? w[c0 + lane] : 0.0;
    const int cnt = (int)std::min<int64_t>(64, C - c0);
    for (int l = 0; l < cnt; l++) total += lane_value(wl, l);
  }
  // _prune (tdigest.mojo:441-458) when decay is active (floor > 0): kept
  // centroids move down in order (each chunk is read before it is written,
  // and writes land at or below the reads), pruned weights are summed in
  // order.
  double removed = 0.0;
  if (sd.floor > 0.0) {
    int64_t k = 0;
    for (int64_t c0 = 0; c0 < C; c0 += 64) {
      const bool valid = c0 + lane < C;
      const double wl = valid ? w[c0 + lane] : 0.0;
      const double ml = valid ? m[c0 + lane] : 0.0;
      const bool keep = valid && wl >= sd.floor;
      const uint64_t km = __ballot(keep);
      uint64_t rm = __ballot(valid && !keep);
      while (rm) {
        const int l = __ffsll((long long)rm) - 1;
        removed += lane_value(wl, l);
        rm &= rm - 1ull;
      }
      if (keep) {
        const int64_t dst = k + __popcll(km & ((1ull << lane) - 1ull));
        m[dst] = ml;
        w[dst] = wl;
      }
      k += __popcll(km);
    }
    if (removed > 0.0) {
      total -= removed;
      C = k;
    }
  }
  if (lane == 0) finish_result(sd, si, C, total, removed, X, vs, res);
  if (v) wave_zero_sign(sd, si, v, res, lane);
}

// ---------------------------------------------------------------- engine
void Engine::set_profiling(bool on) {
  collect();
  prof_ = on;
  stats_.clear();
}

hipEvent_t Engine::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  check_hip(hipEventCreate(&e), "hipEventCreate");
  return e;
}

int Engine::phase_begin(const char* name, double bytes, hipStream_t on) {
  if (!prof_) return -1;
  Pending p{name, get_event(), get_event(), bytes, on ? on : stream_};
  check_hip(hipEventRecord(p.a, p.on), "hipEventRecord");
  pending_.push_back(p);
  return (int)pending_.size() - 1;
}

void Engine::phase_end(int id) {
  if (id < 0) return;
  check_hip(hipEventRecord(pending_[id].b, pending_[id].on), "hipEventRecord");
}

void Engine::collect() {
  if (pending_.empty()) return;
  check_hip(hipStreamSynchronize(stream_), "profile sync");
  check_hip(hipStreamSynchronize(side_), "profile sync");
  for (auto& p : pending_) {
    float ms = 0.f;
    check_hip(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
    PhaseStat& st = stats_[p.name];
    st.ms += ms;
    st.launches += 1;
    st.bytes += p.bytes;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}

Engine::Engine(int device) : device_(device) {
  check_hip(hipSetDevice(device), "hipSetDevice");
  check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreate");
  check_hip(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
  check_hip(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreate");
  void* p = nullptr;
  check_hip(hipHostMalloc(&p, sizeof(unsigned int), hipHostMallocDefault), "hipHostMalloc");
  h_fallbacks_ = static_cast<unsigned int*>(p);
  *h_fallbacks_ = 0;
}

Engine::~Engine() {
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
  if (h_plan_) (void)hipHostFree(h_plan_);
  if (h_fallbacks_) (void)hipHostFree(h_fallbacks_);
  if (h_up_) (void)hipHostFree(h_up_);
  if (h_res_) (void)hipHostFree(h_res_);
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
  if (side_) (void)hipStreamDestroy(side_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename T>
static T* as(DevBuf& b, size_t count) {
  return static_cast<T*>(b.reserve(count * sizeof(T)));
}

void Engine::sort_values(int nseg, const double* d_v, int64_t total_v) {
  double* t1 = as<double>(t1_, std::max<int64_t>(total_v, 1));
  double* t2 = as<double>(t2_, std::max<int64_t>(total_v, 1));
  double* x = static_cast<double*>(x_.p);
  double* vs = static_cast<double*>(vs_.p);
  SegDesc* dsegs = dsegs_;
  auto* rc = as<RouteCounters>(route_, 1);
  RouteCounters hrc;

  // Sorts the final buckets listed in (fs, fl); their data sits in buffers
  // d_v / t1 / t2 that later rounds no longer write.
  auto sort_finals = [&](const FBucket* fs, const FBucket* fl, const RouteCounters& c) {
    if (c.n_small > 0) {
      int ph_ = phase_begin("bucket_sort_small", 0.0);
      k_bucket_sort<kSmallCap, 64><<<(unsigned)c.n_small, 64, 0, stream_>>>(fs, dsegs, d_v, t1, t2, x, vs);
      launched("k_bucket_sort_small");
      phase_end(ph_);
    }
    if (c.n_large > 0) {
      int ph_ = phase_begin("bucket_sort", 0.0);
      k_bucket_sort<kCap, kBlock><<<(unsigned)c.n_large, kBlock, 0, stream_>>>(fl, dsegs, d_v, t1, t2, x, vs);
      launched("k_bucket_sort");
      phase_end(ph_);
    }
  };

  int64_t cap = nseg + 1;
  FBucket* fs = as<FBucket>(fb_, cap);
  FBucket* fl = as<FBucket>(fbl_, cap);
  Part* parts = as<Part>(parts_, cap);
  check_hip(hipMemsetAsync(rc, 0, sizeof(RouteCounters), stream_), "memset counters");
  k_seed<<<(nseg + 255) / 256, 256, 0, stream_>>>(dsegs, nseg, fs, fl, parts, rc, cap);
  launched("k_seed");
  check_hip(hipMemcpyAsync(&hrc, rc, sizeof hrc, hipMemcpyDeviceToHost, stream_), "D2H counters");
  check_hip(hipStreamSynchronize(stream_), "seed sync");
  sort_finals(fs, fl, hrc);
  if (hrc.n_next > 0) {
    // The rounds below may regrow the final-bucket lists these sorts read.
    check_hip(hipStreamSynchronize(stream_), "seed sort sync");
  }

  int src_id = 0;
  int round = 0;
  // Parts of the current round alternate between two device lists.
  DevBuf* cur_list = &parts_;
  DevBuf* nxt_list = &parts2_;
  while (hrc.n_next > 0) {
    if (++round > 40) throw std::runtime_error("value sort: partition did not converge");
    int dst_id = (src_id == 1) ? 2 : 1;
    double* dst = dst_id == 1 ? t1 : t2;
    int np = (int)hrc.n_next;
    int64_t round_elems = (int64_t)hrc.next_elems;
    int64_t max_tiles = round_elems / kPTile + np;
    // Each part yields at most kMaxBkt non-empty buckets and at most one per value.
    int64_t rcap = std::min<int64_t>((int64_t)np * kMaxBkt, round_elems) + 1;
    Part* dparts = static_cast<Part*>(cur_list->p);
    int64_t* tcnt = as<int64_t>(tcnt_, 2 * (int64_t)np + 2);
    int64_t* toff = tcnt + np;
    int64_t* ntiles = toff + np;  // written by the scan as its total
    PTile* dtiles = as<PTile>(tiles_, max_tiles);
    double* dspl = as<double>(spl_, (size_t)np * kMaxSpl);
    int* dns = as<int>(nspl_, np);
    auto* dbcnt = as<unsigned long long>(bcnt_, (size_t)np * kMaxBkt);
    auto* dbcur = as<unsigned long long>(bcur_, (size_t)np * kMaxBkt);
    fs = as<FBucket>(fb_, rcap);
    fl = as<FBucket>(fbl_, rcap);
    Part* dnext = as<Part>(*nxt_list, rcap);
    check_hip(hipMemsetAsync(dbcnt, 0, (size_t)np * kMaxBkt * sizeof(unsigned long long), stream_), "memset");
    check_hip(hipMemsetAsync(rc, 0, sizeof(RouteCounters), stream_), "memset counters");
    k_part_tiles<<<(np + 255) / 256, 256, 0, stream_>>>(dparts, np, tcnt);
    launched("k_part_tiles");
    k_scan_counts<<<1, 1024, 0, stream_>>>(tcnt, np, toff, ntiles);
    launched("k_scan_counts");
    k_make_tiles<<<np, 256, 0, stream_>>>(dparts, toff, dtiles);
    launched("k_make_tiles");
    const double* s0 = d_v;
    {
      int ph_ = phase_begin("splitters", 8.0 * (double)np * kCap);
      k_splitters<<<np, kBlock, 0, stream_>>>(dparts, s0, t1, t2, dspl, dns);
      launched("k_splitters");
      phase_end(ph_);
    }
    {
      int ph_ = phase_begin("bucket_hist", 8.0 * (double)round_elems);
      k_bucket_hist<<<(unsigned)max_tiles, kBlock, 0, stream_>>>(dtiles, dparts, s0, t1, t2, dspl, dns,
                                                                   dbcnt, ntiles);
      launched("k_bucket_hist");
      phase_end(ph_);
    }
    k_bucket_scan<<<(np + 63) / 64, 64, 0, stream_>>>(dparts, np, dns, dbcnt, dbcur);
    launched("k_bucket_scan");
    {
      int ph_ = phase_begin("bucket_scatter", 16.0 * (double)round_elems);
      k_bucket_scatter<<<(unsigned)max_tiles, kBlock, 0, stream_>>>(dtiles, dparts, s0, t1, t2, dst, dspl,
                                                                      dns, dbcur, ntiles);
      launched("k_bucket_scatter");
      phase_end(ph_);
    }
    k_route<<<np, 256, 0, stream_>>>(dparts, dns, dbcnt, dbcur, dsegs, dst_id, fs, fl, dnext, rc, rcap);
    launched("k_route");
    check_hip(hipMemcpyAsync(&hrc, rc, sizeof hrc, hipMemcpyDeviceToHost, stream_), "D2H counters");
    check_hip(hipStreamSynchronize(stream_), "route sync");
    if ((int64_t)hrc.n_small > rcap || (int64_t)hrc.n_large > rcap || (int64_t)hrc.n_next > rcap)
      throw std::runtime_error("value sort: bucket list overflow");
    sort_finals(fs, fl, hrc);
    // The next round reuses fb_/fbl_: wait for this round's sorts to read them.
    check_hip(hipStreamSynchronize(stream_), "sort sync");
    std::swap(cur_list, nxt_list);
    src_id = dst_id;
  }
}

void Engine::presort_samples(const double* d_v, int64_t n, int64_t batch, double* out) {
  static_assert(kSelSampleMax == kSelSample, "presorted sample size");
  check_hip(hipSetDevice(device_), "hipSetDevice");
  const int64_t nb = (n + batch - 1) / batch;
  if (nb <= 0) return;
  k_sel_presort<<<(unsigned)nb, kSelThreads, 0, stream_>>>(d_v, n, batch, out);
  launched("k_sel_presort");
}

// Page-locked host buffer of at least `need` bytes (grown by reallocation;
// callers only grow it when no copy from it is in flight).
static void* pinned_at_least(void*& p, size_t& cap, size_t need) {
  if (need > cap) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t c = std::max<size_t>(need, 4096) * 2;
    check_hip(hipHostMalloc(&p, c, hipHostMallocDefault), "hipHostMalloc");
    cap = c;
  }
  return p;
}

void Engine::compress(int scale, SegVec& segs, const double* d_v,
                      const double* d_em, const double* d_ew, ResVec& res,
                      bool sync_results) {
  check_hip(hipSetDevice(device_), "hipSetDevice");
  int nseg = (int)segs.size();
  res.resize(nseg);
  if (nseg == 0) return;
  PlanSummary P{};
  P.nseg = nseg;
  std::vector<int> eseg;
  std::vector<SelSeg> hsel;
  for (auto& l : small_) l.clear();
  for (int s = 0; s < nseg; s++) {
    SegDesc& sd = segs[s];
    if (sd.e_cnt > kECap) throw std::runtime_error("too many explicit records in one digest");
    sd.mode = 0;
    sd.sel = -1;
    if (fused_on_ && sd.v_cnt + sd.e_cnt <= kCap) {
      sd.mode = 2;
      const int64_t nm = sd.v_cnt + sd.e_cnt;
      small_[nm <= 256 ? 0 : nm <= 512 ? 1 : nm <= 1024 ? 2 : nm <= 2048 ? 3 : 4].push_back(s);
    } else if (sel_on_ && sd.v_cnt > kCap && sd.v_cnt <= kSelMaxV && sd.e_cnt + 64 <= kSelSplMax) {
      SelSeg g{};
      g.si = s;
      g.cap_s = sel_cap_s(sd.v_cnt, sd.e_cnt);
      g.cap_c = sel_cap_c(scale, sd.nrm);
      g.b_off = P.sel_b;
      P.sel_b += g.cap_s + 2;
      g.c_off = P.sel_c;
      P.sel_c += g.cap_c + 1;
      g.m_off = P.sel_m;
      P.sel_m += 2 * (int64_t)g.cap_c + 2;
      g.tile_len = sel_tile_len(sd.v_cnt, sel_tile_min_, sel_tiles_max_);
      sd.mode = 1;
      sd.sel = (int32_t)hsel.size();
      {
        const int64_t need = std::max<int64_t>(sel_sample_size(sd.v_cnt), g.cap_s);
        P.sample_cls |= 8 | (need <= 1024 ? 1 : need <= 4096 ? 2 : 4);
        P.sample_cls |= g.cap_s <= kSelSmallS ? 16 : 32;  // count instance of the segment
      }
      hsel.push_back(g);
      P.sel_v += (double)sd.v_cnt;
    } else {
      P.n_mode0++;
    }
    P.total_v = std::max(P.total_v, sd.v_off + sd.v_cnt);
    sd.es_off = P.total_e;
    P.total_e += sd.e_cnt;
    int64_t n = sd.v_cnt + sd.e_cnt;
    sd.x_off = P.total_x;
    P.total_x += n;
    sd.t_off = P.total_t;
    P.total_t += (n + kTile - 1) / kTile;
    if (sd.e_cnt > 0) {
      eseg.push_back(s);
      P.max_e_items = std::max(P.max_e_items, n);
    }
  }
  P.nsel = (int64_t)hsel.size();
  P.n_eseg = (int64_t)eseg.size();
  for (int c = 0; c < kSmallClasses; c++) P.n_small[c] = (int64_t)small_[c].size();
  if (!hsel.empty()) {
    // LDS footprint classes of the count and gather passes, in this order:
    // few splitters and clusters, many splitters and few clusters, the rest;
    // tiles are numbered in this order.
    std::stable_sort(hsel.begin(), hsel.end(), [&](const SelSeg& x, const SelSeg& y) {
      return sel_class(x.cap_s, x.cap_c) < sel_class(y.cap_s, y.cap_c);
    });
    int64_t t = 0;
    for (size_t i = 0; i < hsel.size(); i++) {
      SelSeg& g = hsel[i];
      g.tile0 = t;
      t += (segs[g.si].v_cnt + g.tile_len - 1) / g.tile_len;
      segs[g.si].sel = (int32_t)i;
      const int cl = sel_class(g.cap_s, g.cap_c);
      if (cl == 0) P.t_small = t;
      if (cl <= 1) P.t_mid = t;
    }
    P.sel_t = t;
  }
  // Upload the plan in one copy: segment table, E-segment list, small-class
  // lists (largest class first), selection segments, each at a 256-byte
  // aligned offset of the staging buffer and of up_.
  h_small_.clear();
  for (int c = kSmallClasses - 1; c >= 0; c--) h_small_.insert(h_small_.end(), small_[c].begin(), small_[c].end());
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_e = al((size_t)nseg * sizeof(SegDesc));
  const size_t o_sm = o_e + al(std::max<size_t>(eseg.size(), 1) * sizeof(int));
  const size_t o_sel = o_sm + al(std::max<size_t>(h_small_.size(), 1) * sizeof(int));
  const size_t up = o_sel + al(std::max<size_t>(hsel.size(), 1) * sizeof(SelSeg));
  char* hu = static_cast<char*>(pinned_at_least(h_up_, h_up_cap_, up));
  memcpy(hu, segs.data(), (size_t)nseg * sizeof(SegDesc));
  if (!eseg.empty()) memcpy(hu + o_e, eseg.data(), eseg.size() * sizeof(int));
  if (!h_small_.empty()) memcpy(hu + o_sm, h_small_.data(), h_small_.size() * sizeof(int));
  if (!hsel.empty()) memcpy(hu + o_sel, hsel.data(), hsel.size() * sizeof(SelSeg));
  char* du = as<char>(up_, up);
  check_hip(hipMemcpyAsync(du, hu, up, hipMemcpyHostToDevice, stream_), "H2D plan");
  dsegs_ = reinterpret_cast<SegDesc*>(du);
  deseg_ = reinterpret_cast<int*>(du + o_e);
  dsmall_ = reinterpret_cast<int*>(du + o_sm);
  dsel_ = reinterpret_cast<SelSeg*>(du + o_sel);
  defer_pending_ = false;
  execute(scale, P, d_v, d_em, d_ew, 1);
  SegResult* dres = static_cast<SegResult*>(res_.p);
  SegResult* hres = static_cast<SegResult*>(
      pinned_at_least(h_res_, h_res_cap_, (size_t)nseg * sizeof(SegResult)));
  // The call ends synchronized: the staging read and written by the
  // asynchronous copies (plan, results) is then free to change.
  if (sync_results)
    check_hip(hipMemcpyAsync(hres, dres, nseg * sizeof(SegResult), hipMemcpyDeviceToHost, stream_), "D2H results");
  check_hip(hipStreamSynchronize(stream_), "sync results");
  // Selection fallbacks (rare): their segments go through the sort path now.
  if (defer_pending_ && *h_fallbacks_ > 0) {
    execute(scale, P, d_v, d_em, d_ew, 2);
    if (sync_results)
      check_hip(hipMemcpyAsync(hres, dres, nseg * sizeof(SegResult), hipMemcpyDeviceToHost, stream_), "D2H results");
    check_hip(hipStreamSynchronize(stream_), "sync results");
  }
  if (sync_results) memcpy(res.data(), hres, (size_t)nseg * sizeof(SegResult));
  defer_pending_ = false;
  if (sync_results)
    for (const SegResult& r : res) path_counts_[r.path == 1 ? 0 : (r.path == 2 ? 2 : 1)]++;
}

void Engine::execute(int scale, const PlanSummary& P, const double* d_v, const double* d_em,
                     const double* d_ew, int part) {
  const bool fb_only = part == 2;
  const int nseg = (int)P.nseg;
  const int64_t total_e = P.total_e, total_x = P.total_x, total_t = P.total_t, total_v = P.total_v;
  SegDesc* dsegs = dsegs_;
  int* deseg = deseg_;
  const int* dl = dsmall_;
  SegResult* dres = as<SegResult>(res_, nseg);
  double* esm = as<double>(em_s_, std::max<int64_t>(total_e, 1));
  double* esw = as<double>(ew_s_, std::max<int64_t>(total_e, 1));
  double* se = as<double>(se_, total_e + nseg);
  int64_t* pose = as<int64_t>(pose_, std::max<int64_t>(total_e, 1));
  double* vs = as<double>(vs_, std::max<int64_t>(total_v, 1));
  double* x = as<double>(x_, std::max<int64_t>(total_x, 1));
  double* cm = as<double>(cm_, std::max<int64_t>(total_x, 1));
  double* cw = as<double>(cw_, std::max<int64_t>(total_x, 1));
  int64_t* cs = as<int64_t>(cs_, std::max<int64_t>(total_x, 1));
  int64_t* ccount = as<int64_t>(cnt_, 2 * (size_t)nseg + 1);
  int64_t* cbase = ccount + nseg;
  double* tsum = as<double>(tsum_, 2 * std::max<int64_t>(total_t, 1));
  int64_t* dtotal = as<int64_t>(clist_, 2);

  // (Part 2 reuses the sorted prior records of part 1.)
  if (P.n_eseg > 0 && !fb_only) {
    int ph_ = phase_begin("sort_e", 32.0 * (double)total_e);
    k_sort_e<<<nseg, kBlock, 0, stream_>>>(dsegs, nseg, d_em, d_ew, esm, esw, se);
    launched("k_sort_e");
    phase_end(ph_);
  } else if (!fb_only) {
    check_hip(hipMemsetAsync(se, 0, (total_e + nseg) * sizeof(double), stream_), "memset se");
  }
  const int nsel = (int)P.nsel;
  const int64_t sel_b = P.sel_b, sel_c = P.sel_c, sel_m = P.sel_m;
  const int64_t t_small = P.t_small, t_mid = P.t_mid, sel_t = P.sel_t;
  SelSeg* dsel = dsel_;
  SelState* dst = nullptr;
  double *dspl = nullptr, *dcsl = nullptr, *dcwv = nullptr;
  unsigned int *dcnt = nullptr, *dbst = nullptr, *dmlc = nullptr;
  int *dinfo = nullptr, *dce = nullptr;
  SelBound* dbnd = nullptr;
  SelMark* dmk = nullptr;
  int2* dml = nullptr;
  int2* dmlx = nullptr;
  uint16_t* dbix = nullptr;
  uint16_t* dstm = nullptr;
  int* dtmap = nullptr;
  SelWave* dwv = nullptr;
  unsigned int* dlcnt = nullptr;
  int *dslist = nullptr, *dclist = nullptr;
  // Small segments: overlap mode 0 runs them on the session stream after the
  // selection samples (whose blocks then start first); mode 1 on the side
  // stream once the selection count is done (the count keeps every CU busy
  // by itself), overlapping the rest of the selection path, whose
  // single-block steps leave most CUs idle; mode 2 on the side stream from
  // the start, beside the selection path's samples and count.
  bool forked = false;
  int64_t n_small = 0;
  for (int c = 0; c < kSmallClasses; c++) n_small += P.n_small[c];
  const bool small_later = overlap_ == 1 && nsel > 0;
  const bool small_first = overlap_ == 2 && nsel > 0;
  auto run_small = [&](bool side) {
    hipStream_t on = stream_;
    if (side) {
      check_hip(hipEventRecord(fork_, stream_), "fork");
      check_hip(hipStreamWaitEvent(side_, fork_, 0), "fork wait");
      forked = true;
      on = side_;
    }
    int ph_ = phase_begin("small_compress", 0.0, on);
    int64_t at = 0;
    for (int c = kSmallClasses - 1; c >= 0; c--) {
      const int m = (int)P.n_small[c];
      if (m == 0) continue;
      const unsigned grid = (unsigned)std::min(m, 65535);
      const int* l = dl + at;
#define KD_SMALL(CAP, TH)                                                                          \
  k_small_compress<CAP, TH><<<grid, TH, 0, on>>>(scale, dsegs, l, m, d_v, esm, esw, se, pose, cm,    \
                                                     cw, ccount, dres)
      switch (c) {
        case 0: KD_SMALL(256, 64); break;
        case 1: KD_SMALL(512, 128); break;
        case 2: KD_SMALL(1024, 128); break;
        case 3: KD_SMALL(2048, kBlock); break;
        default: KD_SMALL(4096, 512); break;
      }
#undef KD_SMALL
      launched("k_small_compress");
      at += m;
    }
    phase_end(ph_);
    if (forked) check_hip(hipEventRecord(join_, side_), "join");
  };
  if (n_small > 0 && !fb_only && small_first) run_small(true);
  if (nsel > 0 && !fb_only) {
    dst = as<SelState>(sel_st_, nsel);
    dspl = as<double>(sel_spl_, sel_b);
    dcnt = as<unsigned int>(sel_cnt_, sel_b);
    dbst = as<unsigned int>(sel_bst_, sel_b);
    dinfo = as<int>(sel_info_, sel_b);
    dbnd = as<SelBound>(sel_bnd_, sel_c);
    dce = as<int>(sel_ce_, sel_c);
    dcsl = as<double>(sel_csl_, 2 * sel_c);
    dcwv = as<double>(sel_cwv_, sel_c);
    dmk = as<SelMark>(sel_mk_, sel_m);
    dml = as<int2>(sel_ml_, sel_m);
    dbix = as<uint16_t>(sel_bix_, std::max<int64_t>(total_v, 1));
    // Staged gathered values: their marked-bucket indices here, the values
    // in vs (which only the sorting path uses, for other segments).
    dstm = as<uint16_t>(sel_stm_, std::max<int64_t>(total_v, 1));
    dtmap = as<int>(sel_tmap_, std::max<int64_t>(sel_t, 1));
    dwv = as<SelWave>(sel_wv_, std::max<int64_t>(sel_m, 1));
    // Work lists of the sample and count launches and their sizes.
    dlcnt = as<unsigned int>(sel_lcnt_, 8);
    dslist = as<int>(sel_slist_, 3 * (int64_t)nsel);
    dclist = as<int>(sel_clist_, 2 * std::max<int64_t>(sel_t, 1));
    // Gathered-bucket list size, fallbacks, list sizes per size class.
    dmlc = as<unsigned int>(sel_mlc_, 2 + kSelClasses);
    dmlx = as<int2>(sel_mlx_, kSelClasses * sel_m);
    {
      const int64_t items = std::max<int64_t>(std::max<int64_t>(sel_b, nsel), 2 + kSelClasses);
      const unsigned grid = (unsigned)std::min<int64_t>((items + 255) / 256, 1024);
      check_hip(hipMemsetAsync(dlcnt, 0, 8 * sizeof(unsigned int), stream_), "memset lists");
      k_sel_init<<<grid, 256, 0, stream_>>>(dsegs, dsel, dst, nsel, dcnt, sel_b, dmlc, 2 + kSelClasses,
                                            dtmap, sel_t, dlcnt, dslist, dclist);
      launched("k_sel_init");
    }
    {
      int ph_ = phase_begin("sel_sample", 0.0);
      // The three sample sizes touch disjoint segments: with several
      // segments the two smaller run on the side stream, the largest beside
      // them on the session stream (a lone digest skips the fork's latency).
      const bool fork_s = nsel > 1 && !forked;
      hipStream_t s2 = fork_s ? side_ : stream_;
      if (fork_s) {
        check_hip(hipEventRecord(fork_, stream_), "fork samples");
        check_hip(hipStreamWaitEvent(side_, fork_, 0), "fork samples wait");
      }
      // A host plan says which classes have segments; skip the others.
      const bool cls_known = (P.sample_cls & 8) != 0;
      if (!cls_known || (P.sample_cls & 1)) {
        k_sel_sample<1024, 64, 0><<<nsel, 64, 0, s2>>>(dsegs, dsel, d_v, esm, dspl, dst, nullptr, dslist,
                                                        dlcnt);
        launched("k_sel_sample<1024>");
      }
      if (!cls_known || (P.sample_cls & 2)) {
        k_sel_sample<4096, kBlock, 1024><<<nsel, kBlock, 0, s2>>>(dsegs, dsel, d_v, esm, dspl, dst,
                                                                  nullptr, dslist + nsel, dlcnt + 1);
        launched("k_sel_sample<4096>");
      }
      // The largest class fills a CU's LDS anyway: a block of kSelThreads.
      if (!cls_known || (P.sample_cls & 4)) {
        k_sel_sample<kSelSample, kSelThreads, 4096><<<nsel, kSelThreads, 0, stream_>>>(
            dsegs, dsel, d_v, esm, dspl, dst, nsel == 1 ? presorted_ : nullptr, dslist + 2 * (int64_t)nsel,
            dlcnt + 2);
        launched("k_sel_sample<8192>");
      }
      if (fork_s) {
        check_hip(hipEventRecord(join_, side_), "join samples");
        check_hip(hipStreamWaitEvent(stream_, join_, 0), "join samples wait");
      }
      phase_end(ph_);
    }
  }
  if (n_small > 0 && !fb_only && !small_later && !small_first) run_small(false);
  if (nsel > 0 && !fb_only) {
    const double sel_bytes = 8.0 * P.sel_v;
    {
      int ph_ = phase_begin("sel_count", sel_bytes * 1.25);
      // Two LDS classes by splitter capacity over every tile (a block whose
      // segment belongs to the other class returns at once): small tables
      // run several blocks per CU, on the side stream when it is free (beside
      // the large tables' blocks, one per CU, with LDS to spare).
      // With a host plan, an instance none of whose segments is there is not
      // launched.
      const bool cnt_known = (P.sample_cls & 8) != 0;
      const bool cnt_small = !cnt_known || (P.sample_cls & 16);
      const bool cnt_large = !cnt_known || (P.sample_cls & 32);
      const bool fork_c = nsel > 1 && !forked && cnt_small && cnt_large;
      if (fork_c) {
        check_hip(hipEventRecord(fork_, stream_), "fork count");
        check_hip(hipStreamWaitEvent(side_, fork_, 0), "fork count wait");
      }
      if (cnt_small) {
        k_sel_count<kSelSmallS, 0, 256><<<(unsigned)sel_t, 256, 0, fork_c ? side_ : stream_>>>(
            dsegs, dsel, nsel, 0, sel_t, d_v, dspl, dst, dcnt, dbix, dtmap, dclist, dlcnt + 3);
        launched("k_sel_count<small>");
      }
      if (cnt_large) {
        k_sel_count<kSelSplMax, kSelSmallS, kSelThreads><<<(unsigned)sel_t, kSelThreads, 0, stream_>>>(
            dsegs, dsel, nsel, 0, sel_t, d_v, dspl, dst, dcnt, dbix, dtmap, dclist + sel_t, dlcnt + 4);
        launched("k_sel_count");
      }
      if (fork_c) {
        check_hip(hipEventRecord(join_, side_), "join count");
        check_hip(hipStreamWaitEvent(stream_, join_, 0), "join count wait");
      }
      phase_end(ph_);
    }
    if (n_small > 0 && small_later) run_small(true);
    {
      int ph_ = phase_begin("sel_plan", 0.0);
      k_sel_ranks<<<nsel, kSelThreads, 0, stream_>>>(dsegs, dsel, dspl, dcnt, dbst, dinfo, esm, pose,
                                                     dst);
      launched("k_sel_ranks");
      k_sel_chain<<<(nsel + kChainWaves - 1) / kChainWaves, 64 * kChainWaves, 0, stream_>>>(
          scale, dsegs, dsel, nsel, se, esw, pose, dst, dbnd, dcsl, dcwv);
      launched("k_sel_chain");
      k_sel_plan<<<nsel, kSelThreads, 0, stream_>>>(dsegs, dsel, dspl, dcnt, dbst, dinfo, pose, dst,
                                                    dbnd, dce, dcsl, dmk, dml, dmlc);
      launched("k_sel_plan");
      k_sel_classes<<<(unsigned)std::min<int64_t>((sel_m + 255) / 256, 1024), 256, 0, stream_>>>(
          dsegs, dsel, dst, dmk, dbnd, dml, dmlc, dmlx, dwv, sel_m, dmlc + 2);
      launched("k_sel_classes");
      phase_end(ph_);
    }
    {
      int ph_ = phase_begin("sel_gather", sel_bytes * 1.25);
      if (t_small > 0) {
        k_sel_gather<kSelSmallS, kSelSmallC><<<(unsigned)t_small, kSelThreads, 0, stream_>>>(
            dsegs, dsel, nsel, 0, t_small, d_v, dbix, dinfo, dce, dst, dmk, dcsl, x, vs, dstm, dtmap);
        launched("k_sel_gather<small>");
      }
      if (t_mid > t_small) {
        k_sel_gather<kSelSplMax, kSelSmallC><<<(unsigned)(t_mid - t_small), kSelThreads, 0, stream_>>>(
            dsegs, dsel, nsel, t_small, t_mid, d_v, dbix, dinfo, dce, dst, dmk, dcsl, x, vs, dstm, dtmap);
        launched("k_sel_gather<mid>");
      }
      if (sel_t > t_mid) {
        k_sel_gather<kSelSplMax, kSelCmax><<<(unsigned)(sel_t - t_mid), kSelThreads, 0, stream_>>>(
            dsegs, dsel, nsel, t_mid, sel_t, d_v, dbix, dinfo, dce, dst, dmk, dcsl, x, vs, dstm, dtmap);
        launched("k_sel_gather");
      }
      phase_end(ph_);
    }
    {
      int ph_ = phase_begin("sel_select", 0.0);
      // Each size class walks its own list (grids sized for the worst case;
      // blocks past a list's end return at once).
      unsigned grid = (unsigned)std::min<int64_t>(sel_m, 4096);
      const int2* l1 = dmlx + sel_m;
      const int2* l2 = dmlx + 2 * sel_m;
      const int2* l3 = dmlx + 3 * sel_m;
      // The buckets of the radix select (the largest, few) are finished on
      // the side stream beside the sorted classes (disjoint buckets and
      // boundaries) when there are several digests, also behind small
      // digests there (by now they are mostly done; the finish waits for
      // that stream in any case).
      const bool fork_b = nsel > 1;
      if (fork_b) {
        check_hip(hipEventRecord(fork_, stream_), "fork select");
        check_hip(hipStreamWaitEvent(side_, fork_, 0), "fork select wait");
      }
      k_sel_select<4096><<<std::min(grid, 1024u), kSelectThreads, 0, fork_b ? side_ : stream_>>>(
          dsegs, dsel, dst, l3, dmlc + 5, dmk, dbnd, x);
      launched("k_sel_select");
      // One wave per bucket of the smallest class (the list is long; every
      // bucket's chain of dependent loads then overlaps the others').
      const unsigned grid_w = (unsigned)std::min<int64_t>((sel_m + 3) / 4, 65535);
      k_sel_sorted_wave<<<grid_w, 256, 0, stream_>>>(dwv, dmlc + 2, dmk, dbnd, x);
      launched("k_sel_sorted_wave<128>");
      k_sel_sorted<1024, kBlock, 128><<<grid, kBlock, 0, stream_>>>(dsegs, dsel, dst, l1, dmlc + 3, dmk,
                                                                    dbnd, x);
      launched("k_sel_sorted<1024>");
      // With several digests the 1025..4096-value class follows the radix
      // select on the side stream (the two smaller classes take longer).
      k_sel_sorted<4096, kBlock, 1024><<<grid, kBlock, 0, fork_b ? side_ : stream_>>>(
          dsegs, dsel, dst, l2, dmlc + 4, dmk, dbnd, x);
      launched("k_sel_sorted<4096>");
      if (fork_b) {
        check_hip(hipEventRecord(join_, side_), "join select");
        check_hip(hipStreamWaitEvent(stream_, join_, 0), "join select wait");
      }
      phase_end(ph_);
    }
    {
      int ph_ = phase_begin("sel_finalize", 0.0);
      k_sel_finalize<<<nsel, kBlock, 0, stream_>>>(dsegs, dsel, dst, dcwv, dbnd, dmk, dinfo, dcsl,
                                                   esm, esw, pose, cm, cw, ccount, dres, dmlc + 1);
      launched("k_sel_finalize");
      phase_end(ph_);
    }
    // Whether any selected segment fell back to sorting (pinned: the copy
    // does not wait for the stream).
    check_hip(hipMemcpyAsync(h_fallbacks_, dmlc + 1, sizeof(unsigned int), hipMemcpyDeviceToHost,
                             stream_),
              "D2H fallbacks");
  }
  if (forked) check_hip(hipStreamWaitEvent(stream_, join_, 0), "join side stream");
  // The multi-kernel path runs when a segment needs it: one that neither the
  // selection nor the one-block path took, or a selection fallback.
  bool sort_path = P.n_mode0 > 0 || fb_only;
  int finish_only = fb_only ? 0 : -1;  // finish: mode-0 segments, or every one
  if (nsel > 0 && !sort_path) {
    if (part == 1) {
      // Decided after the caller's sync: finish the other segments now.
      defer_pending_ = true;
      finish_only = 1;
    } else {
      check_hip(hipStreamSynchronize(stream_), "sync fallbacks");
      sort_path = *h_fallbacks_ > 0;
    }
  }
  if (sort_path) {
    auto t0 = std::chrono::steady_clock::now();
    sort_values(nseg, d_v, total_v);
    host_phase("sort_values",
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    if (P.n_eseg > 0) {
      int64_t chunks = std::min<int64_t>((P.max_e_items + kBlock * 4 - 1) / (kBlock * 4), 65535);
      for (int64_t b0 = 0; b0 < P.n_eseg; b0 += 65535) {
        unsigned ny = (unsigned)std::min<int64_t>(65535, P.n_eseg - b0);
        dim3 grid((unsigned)std::max<int64_t>(chunks, 1), ny);
        {
          int ph_ = phase_begin("merge_e", 16.0 * (double)total_x);
          k_merge_e<<<grid, kBlock, 0, stream_>>>(dsegs, deseg + b0, (int)ny, vs, esm, x, pose);
          launched("k_merge_e");
          phase_end(ph_);
        }
      }
    }
    int64_t long_cap = total_x / kTile + nseg + 1;
    int64_t* long_list = as<int64_t>(long_, 2 * long_cap);
    auto* long_cnt = reinterpret_cast<unsigned long long*>(dtotal + 1);
    check_hip(hipMemsetAsync(long_cnt, 0, sizeof(unsigned long long), stream_), "memset long count");
    {
      int ph_ = phase_begin("chain", 0.0);
      k_chain<<<(nseg + 63) / 64, 64, 0, stream_>>>(dsegs, nseg, scale, x, pose, se, esw, cs, ccount,
                                                    dres, long_list, long_cnt);
      launched("k_chain");
      phase_end(ph_);
    }
    if (total_t > 0) {
      int ph_ = phase_begin("tile_sums", 8.0 * (double)total_x);
      k_tile_sums<<<(unsigned)total_t, kBlock, 0, stream_>>>(dsegs, nseg, x, pose, esw, tsum);
      launched("k_tile_sums");
      phase_end(ph_);
    }
    k_scan_counts<<<1, 1024, 0, stream_>>>(ccount, nseg, cbase, dtotal);
    launched("k_scan_counts");
    {
      // Upper bound of the cluster count without a host round trip.
      unsigned grid = (unsigned)std::min<int64_t>((total_x + 255) / 256, 16384);
      int ph_ = phase_begin("cluster_mean", 16.0 * (double)total_x / 8.0);
      k_cluster_mean<<<std::max(grid, 1u), 256, 0, stream_>>>(dsegs, nseg, cbase, ccount, cs, x,
                                                               pose, se, esw, cm, cw, dtotal);
      launched("k_cluster_mean");
      phase_end(ph_);
    }
    if (total_t > 2) {
      unsigned grid = (unsigned)std::min<int64_t>(long_cap, 4096);
      int ph_ = phase_begin("long_sum", 0.0);
      k_long_sum<<<grid, kBlock, 0, stream_>>>(dsegs, ccount, cs, x, pose, se, esw, tsum, long_list,
                                               long_cnt, cm, cw);
      launched("k_long_sum");
      phase_end(ph_);
    }
  }
  {
    int ph_ = phase_begin("finish", 0.0);
    if (nseg <= kFinishWaveSegs) {
      // Few segments: the zero-extreme rule in the same launch.
      k_finish<<<(nseg + kFinishWaves - 1) / kFinishWaves, 64 * kFinishWaves, 0, stream_>>>(
          dsegs, nseg, ccount, cm, cw, x, vs, dres, finish_only, d_v);
      launched("k_finish");
    } else {
      k_finish_lane<<<(nseg + 63) / 64, 64, 0, stream_>>>(dsegs, nseg, ccount, cm, cw, x, vs, dres,
                                                          finish_only);
      launched("k_finish");
      k_zero_signs<<<(unsigned)std::min((nseg + 255) / 256, 1024), 256, 0, stream_>>>(dsegs, nseg,
                                                                                     d_v, dres);
      launched("k_zero_signs");
    }
    phase_end(ph_);
  }
  check_hip(hipGetLastError(), "kernel launch");
}


// ----------------------------------------------- grouped plan on the device
// Plans a whole grouped batch on the device: every touched key's B-semantics
// prepare (tdigest.mojo:236-296 per key, as grouped.mojo:99-157 runs it),
// its segment descriptor, one-block size class or selection layout, and the
// prefix offsets of the compress buffers.  Three launches, one thread per
// key: per-block sums of every counted quantity, one block scanning those
// sums, then a rewrite pass in which each key takes its block's base plus
// its exclusive rank inside the block (segments and lists keep key order, as
// the host plan does).
constexpr int kPlanThreads = 256;
constexpr double kPlanRenorm = 400.0;  // tdigest.mojo:43
constexpr double kPlanPrune = 1e-6;    // tdigest.mojo:47

struct PlanKey {
  double sf, stored, nrm, floor, logs, clock, factor;
  bool renorm;
};

// Per-key prepare of one batch of m new values (the host rule of the grouped
// add: advance the clock by m, decay exponent, renormalize, scale factor,
// stored total, normalizer, prune floor).
__device__ inline PlanKey plan_key(const GroupedPlanArgs& a, int64_t k, int64_t m) {
  PlanKey p;
  const bool decay = a.lambda > 0.0;
  double clock = a.clock[k], logs = a.logs[k], total = a.total[k];
  const double nc = clock + (double)m;
  p.renorm = false;
  p.factor = 1.0;
  if (decay && nc > clock) {
    logs += a.lambda * (nc - clock);
    if (logs > kPlanRenorm) {
      p.factor = kd_exp(-logs);
      p.renorm = true;
      total *= p.factor;
      logs = 0.0;
    }
  }
  p.clock = nc;
  p.sf = decay ? kd_exp(logs) : 1.0;
  total += (double)m * p.sf;
  const double effn = decay ? total * kd_exp(-logs) : total;
  p.nrm = kd_normalizer(a.scale, a.compression, effn);
  p.floor = decay ? kPlanPrune * kd_exp(logs) : 0.0;
  p.logs = logs;
  p.stored = total;
  return p;
}

// Quantities counted per key, summed per block and scanned.
enum PlanQ {
  kQSeg, kQE, kQX, kQT, kQESeg, kQMode0,
  kQSmall0,                    // one-block classes 0..4
  kQSel0 = kQSmall0 + 5,       // selection segments per footprint class 0..2
  kQTile0 = kQSel0 + 3,        // selection tiles per class
  kQB0 = kQTile0 + 3,          // bucket entries per class
  kQC0 = kQB0 + 3,             // cluster entries per class
  kQM0 = kQC0 + 3,             // mark entries per class
  kQErr = kQM0 + 3,            // digests with too many prior centroids
  kQCount
};

struct PlanCfg {
  int fused, sel_on;
  int64_t tile_min, tiles_max;
};

// The quantities of key k (zero when the batch does not touch it); 0 small,
// 1 selection, 2 multi-kernel path in *path; class in *cls.
__device__ inline void plan_quant(const GroupedPlanArgs& a, const PlanCfg& c, int64_t k,
                                  int64_t (&q)[kQCount], PlanKey* pk, int* path, int* cls,
                                  double* selv) {
#pragma unroll
  for (int i = 0; i < kQCount; i++) q[i] = 0;
  *path = -1;
  *selv = 0.0;
  if (k >= a.num_keys) return;
  const int64_t m = a.kcnt[k];
  if (m == 0) return;
  const int64_t e = a.off[k + 1] - a.off[k], n = m + e;
  *pk = plan_key(a, k, m);
  q[kQErr] = e > kECap ? 1 : 0;
  q[kQSeg] = 1;
  q[kQE] = e;
  q[kQX] = n;
  q[kQT] = (n + kTile - 1) / kTile;
  q[kQESeg] = e > 0 ? 1 : 0;
  if (c.fused && n <= kCap) {
    *cls = n <= 256 ? 0 : n <= 512 ? 1 : n <= 1024 ? 2 : n <= 2048 ? 3 : 4;
    q[kQSmall0 + *cls] = 1;
    *path = 0;
  } else if (c.sel_on && m > kCap && m <= kSelMaxV && e + 64 <= kSelSplMax) {
    const int cs = sel_cap_s(m, e), cc = sel_cap_c(a.scale, pk->nrm);
    const int cl = sel_class(cs, cc);
    const int64_t tl = sel_tile_len(m, c.tile_min, c.tiles_max);
    *cls = cl;
    q[kQSel0 + cl] = 1;
    q[kQTile0 + cl] = (m + tl - 1) / tl;
    q[kQB0 + cl] = cs + 2;
    q[kQC0 + cl] = cc + 1;
    q[kQM0 + cl] = 2 * (int64_t)cc + 2;
    *selv = (double)m;
    *path = 1;
  } else {
    q[kQMode0] = 1;
    *path = 2;
  }
}

// Exclusive block scan of one int64 per thread (kPlanThreads threads).
__device__ inline int64_t plan_block_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[kPlanThreads / 64 + 1];
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
    for (int i = 0; i < kPlanThreads / 64; i++) {
      int64_t t = wsum[i];
      wsum[i] = run;
      run += t;
    }
    wsum[kPlanThreads / 64] = run;
  }
  __syncthreads();
  int64_t r = inc - v + wsum[w];
  *total = wsum[kPlanThreads / 64];
  __syncthreads();
  return r;
}

// Block partials: part[b * kQCount + q] sums; pmax[2 b], pmax[2 b + 1]: largest
// merged size of a digest with prior centroids, largest end of new values;
// pselv[b]: new values on the selection path.
__global__ __launch_bounds__(kPlanThreads) void k_plan_count(GroupedPlanArgs a, PlanCfg c,
                                                             int64_t* part, int64_t* pmax,
                                                             double* pselv) {
  const int64_t k = (int64_t)blockIdx.x * kPlanThreads + threadIdx.x;
  int64_t q[kQCount];
  PlanKey pk;
  int path, cls;
  double selv;
  plan_quant(a, c, k, q, &pk, &path, &cls, &selv);
  int64_t me = 0, mv = 0;
  if (path >= 0) {
    const int64_t m = a.kcnt[k], e = a.off[k + 1] - a.off[k];
    me = e > 0 ? m + e : 0;
    mv = a.kof[k] + m;
  }
  __shared__ double sv[kPlanThreads];
  __shared__ int64_t sm[2][kPlanThreads];
  for (int i = 0; i < kQCount; i++) {
    int64_t tot;
    plan_block_scan(q[i], &tot);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * kQCount + i] = tot;
  }
  sv[threadIdx.x] = selv;
  sm[0][threadIdx.x] = me;
  sm[1][threadIdx.x] = mv;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    int64_t a0 = 0, a1 = 0;
    for (int i = 0; i < kPlanThreads; i++) {
      s += sv[i];
      a0 = sm[0][i] > a0 ? sm[0][i] : a0;
      a1 = sm[1][i] > a1 ? sm[1][i] : a1;
    }
    pselv[blockIdx.x] = s;
    pmax[2 * blockIdx.x] = a0;
    pmax[2 * blockIdx.x + 1] = a1;
  }
}

// One block: exclusive scans of the block partials into per-block bases
// (class lists offset by the classes before them), and the summary.
__global__ __launch_bounds__(1024) void k_plan_top(int64_t nb, const int* bad, int64_t* part,
                                                   const int64_t* pmax, const double* pselv,
                                                   PlanSummary* out) {
  __shared__ int64_t tot[kQCount];
  // Column q of part: block sums -> exclusive prefixes (in place), one wave
  // per column, 64 blocks at a time.
  const int lane = threadIdx.x & 63;
  for (int q = threadIdx.x >> 6; q < kQCount; q += blockDim.x >> 6) {
    int64_t run = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += 64) {
      const int64_t b = b0 + lane;
      const int64_t v = b < nb ? part[b * kQCount + q] : 0;
      int64_t inc = v;
      for (int d = 1; d < 64; d <<= 1) {
        int64_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      if (b < nb) part[b * kQCount + q] = run + inc - v;
      run += __shfl(inc, 63, 64);
    }
    if (lane == 0) tot[q] = run;
  }
  __syncthreads();
  // Class bases: small lists largest class first; selection classes 0, 1, 2
  // with their tile / bucket / cluster / mark offsets in the same order.
  __shared__ int64_t cbase[kQCount];
  if (threadIdx.x == 0) {
    for (int q = 0; q < kQCount; q++) cbase[q] = 0;
    int64_t b = 0;
    for (int c = 4; c >= 0; c--) {
      cbase[kQSmall0 + c] = b;
      b += tot[kQSmall0 + c];
    }
    const int groups[5] = {kQSel0, kQTile0, kQB0, kQC0, kQM0};
    for (int g = 0; g < 5; g++) {
      int64_t r = 0;
      for (int c = 0; c < 3; c++) {
        cbase[groups[g] + c] = r;
        r += tot[groups[g] + c];
      }
    }
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < nb * kQCount; i += blockDim.x) part[i] += cbase[i % kQCount];
  if (threadIdx.x == 0) {
    PlanSummary P{};
    P.bad = (bad && *bad) ? 1 : (tot[kQErr] > 0 ? 2 : 0);
    P.nseg = tot[kQSeg];
    P.nsel = tot[kQSel0] + tot[kQSel0 + 1] + tot[kQSel0 + 2];
    P.n_mode0 = tot[kQMode0];
    P.n_eseg = tot[kQESeg];
    for (int c = 0; c < 5; c++) P.n_small[c] = tot[kQSmall0 + c];
    P.t_small = tot[kQTile0];
    P.t_mid = tot[kQTile0] + tot[kQTile0 + 1];
    P.sel_t = P.t_mid + tot[kQTile0 + 2];
    P.sel_b = tot[kQB0] + tot[kQB0 + 1] + tot[kQB0 + 2];
    P.sel_c = tot[kQC0] + tot[kQC0 + 1] + tot[kQC0 + 2];
    P.sel_m = tot[kQM0] + tot[kQM0 + 1] + tot[kQM0 + 2];
    P.total_e = tot[kQE];
    P.total_x = tot[kQX];
    P.total_t = tot[kQT];
    int64_t me = 0, mv = 0;
    double sv = 0.0;
    for (int64_t b = 0; b < nb; b++) {
      me = pmax[2 * b] > me ? pmax[2 * b] : me;
      mv = pmax[2 * b + 1] > mv ? pmax[2 * b + 1] : mv;
      sv += pselv[b];
    }
    P.max_e_items = me;
    P.total_v = mv;
    P.sel_v = sv;
    *out = P;
  }
}

// Rewrite: every touched key's state update, segment descriptor and list
// entries at its block base plus its rank inside the block.  Nothing happens
// when the batch failed validation.
__global__ __launch_bounds__(kPlanThreads) void k_plan_write(GroupedPlanArgs a, PlanCfg c,
                                                             const int64_t* base,
                                                             const PlanSummary* sum, SegDesc* segs,
                                                             int* small_list, SelSeg* sel,
                                                             int* eseg) {
  if (sum->bad) return;
  const int64_t k = (int64_t)blockIdx.x * kPlanThreads + threadIdx.x;
  int64_t q[kQCount];
  PlanKey pk;
  int path, cls = 0;
  double selv;
  plan_quant(a, c, k, q, &pk, &path, &cls, &selv);
  int64_t o[kQCount];
  const int64_t* bb = base + (int64_t)blockIdx.x * kQCount;
  for (int i = 0; i < kQCount; i++) {
    int64_t tot;
    o[i] = bb[i] + plan_block_scan(q[i], &tot);
  }
  if (k >= a.num_keys) return;
  if (path < 0) {
    a.seg_of[k] = -1;
    return;
  }
  const int64_t m = a.kcnt[k], e = a.off[k + 1] - a.off[k];
  if (pk.renorm)  // _renormalize (tdigest.mojo:359-371) of the prior centroids
    for (int64_t i = a.off[k]; i < a.off[k + 1]; i++) a.w[i] *= pk.factor;
  a.clock[k] = pk.clock;
  a.logs[k] = pk.logs;
  a.total[k] = pk.stored;
  const int64_t seg = o[kQSeg];
  SegDesc sd{};
  sd.v_off = a.kof[k];
  sd.v_cnt = m;
  sd.e_off = a.off[k];
  sd.e_cnt = e;
  sd.x_off = o[kQX];
  sd.t_off = o[kQT];
  sd.es_off = o[kQE];
  sd.sf = pk.sf;
  sd.stored_n = pk.stored;
  sd.nrm = pk.nrm;
  sd.floor = pk.floor;
  sd.mode = 0;
  sd.sel = -1;
  if (path == 0) {
    sd.mode = 2;
    small_list[o[kQSmall0 + cls]] = (int)seg;
  } else if (path == 1) {
    SelSeg g{};
    g.si = (int32_t)seg;
    g.cap_s = (int)(q[kQB0 + cls] - 2);
    g.cap_c = (int)(q[kQC0 + cls] - 1);
    g.tile_len = sel_tile_len(m, c.tile_min, c.tiles_max);
    g.b_off = o[kQB0 + cls];
    g.c_off = o[kQC0 + cls];
    g.m_off = o[kQM0 + cls];
    g.tile0 = o[kQTile0 + cls];
    sd.mode = 1;
    sd.sel = (int32_t)o[kQSel0 + cls];
    sel[o[kQSel0 + cls]] = g;
  }
  segs[seg] = sd;
  a.seg_key[seg] = k;
  a.seg_of[k] = seg;
  if (e > 0) eseg[o[kQESeg]] = (int)seg;
}

PlanSummary Engine::compress_grouped(const GroupedPlanArgs& a, const double* d_v,
                                     const double* d_em, const double* d_ew, bool defer) {
  check_hip(hipSetDevice(device_), "hipSetDevice");
  defer_pending_ = false;
  const int64_t max_seg = std::max<int64_t>(a.num_keys, 1);
  SegDesc* dsegs = as<SegDesc>(seg_, max_seg);
  int* dl = as<int>(small_list_, max_seg);
  int* deseg = as<int>(order_, max_seg);
  SelSeg* dsel = as<SelSeg>(sel_seg_, max_seg);
  dsegs_ = dsegs;
  dsmall_ = dl;
  deseg_ = deseg;
  dsel_ = dsel;
  const int64_t nb = (a.num_keys + kPlanThreads - 1) / kPlanThreads;
  int64_t* part = as<int64_t>(plan_part_, nb * kQCount + 3 * nb + 1);
  int64_t* pmax = part + nb * kQCount;
  double* pselv = reinterpret_cast<double*>(pmax + 2 * nb);
  PlanSummary* dsum = as<PlanSummary>(plan_sum_, 1);
  if (!h_plan_) {
    void* p = nullptr;
    check_hip(hipHostMalloc(&p, sizeof(PlanSummary), hipHostMallocDefault), "hipHostMalloc");
    h_plan_ = static_cast<PlanSummary*>(p);
  }
  const PlanCfg cfg{fused_on_ ? 1 : 0, sel_on_ ? 1 : 0, sel_tile_min_, sel_tiles_max_};
  {
    int ph_ = phase_begin("plan", 0.0);
    k_plan_count<<<(unsigned)nb, kPlanThreads, 0, stream_>>>(a, cfg, part, pmax, pselv);
    launched("k_plan_count");
    k_plan_top<<<1, 1024, 0, stream_>>>(nb, a.bad, part, pmax, pselv, dsum);
    launched("k_plan_top");
    k_plan_write<<<(unsigned)nb, kPlanThreads, 0, stream_>>>(a, cfg, part, dsum, dsegs, dl, dsel,
                                                             deseg);
    launched("k_plan_write");
    phase_end(ph_);
  }
  check_hip(hipMemcpyAsync(h_plan_, dsum, sizeof(PlanSummary), hipMemcpyDeviceToHost, stream_),
            "D2H plan");
  check_hip(hipStreamSynchronize(stream_), "plan sync");
  const PlanSummary P = *h_plan_;
  if (P.bad || P.nseg == 0) return P;
  int64_t n_small = 0;
  for (int c = 0; c < kSmallClasses; c++) n_small += P.n_small[c];
  path_counts_[2] += n_small;
  path_counts_[0] += P.nsel;
  path_counts_[1] += P.n_mode0;
  if (defer) {
    execute(a.scale, P, d_v, d_em, d_ew, 1);
    // The fallback count is still in flight: counted by finish_grouped.
    fb_count_pending_ = P.nsel > 0;
    return P;
  }
  execute(a.scale, P, d_v, d_em, d_ew);
  // execute() reads the fallback count back only when it decides the sort
  // path from it; otherwise it is still in flight.
  if (P.nsel > 0 && P.n_mode0 > 0) check_hip(hipStreamSynchronize(stream_), "fallback sync");
  if (P.nsel > 0 && *h_fallbacks_ > 0) {  // read by execute after its fallback sync
    path_counts_[0] -= *h_fallbacks_;
    path_counts_[1] += *h_fallbacks_;
  }
  return P;
}

bool Engine::finish_grouped(int scale, const PlanSummary& P, const double* d_v, const double* d_em,
                            const double* d_ew) {
  // The caller has synchronized the stream: the fallback count is here.
  if (fb_count_pending_ && *h_fallbacks_ > 0) {
    path_counts_[0] -= *h_fallbacks_;
    path_counts_[1] += *h_fallbacks_;
  }
  fb_count_pending_ = false;
  const bool run = defer_pending_ && *h_fallbacks_ > 0;
  defer_pending_ = false;
  if (run) execute(scale, P, d_v, d_em, d_ew, 2);
  return run;
}

// ------------------------------------------------------ key partition
constexpr int kHashSlots = 2048;
constexpr int kKPTile = 8192;

__device__ inline int probe_insert(unsigned int* hk, unsigned int key) {
  unsigned int h = (key * 2654435761u) >> 21;  // 11 bits
  for (int p = 0; p < 16; p++) {
    int slot = (h + p) & (kHashSlots - 1);
    unsigned int cur = hk[slot];
    if (cur == key) return slot;
    if (cur == 0xffffffffu) {
      unsigned int old = atomicCAS(&hk[slot], 0xffffffffu, key);
      if (old == 0xffffffffu || old == key) return slot;
    }
  }
  return -1;
}

__device__ inline int probe_find(const unsigned int* hk, unsigned int key) {
  unsigned int h = (key * 2654435761u) >> 21;
  for (int p = 0; p < 16; p++) {
    int slot = (h + p) & (kHashSlots - 1);
    unsigned int cur = hk[slot];
    if (cur == key) return slot;
    if (cur == 0xffffffffu) return -1;
  }
  return -1;
}

__global__ __launch_bounds__(kBlock) void k_key_hist(const uint32_t* keys, int64_t n,
                                                      unsigned long long* counts) {
  __shared__ unsigned int hk[kHashSlots];
  __shared__ unsigned int hc[kHashSlots];
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
    hk[i] = 0xffffffffu;
    hc[i] = 0;
  }
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kKPTile;
  int64_t end = std::min<int64_t>(n, base + kKPTile);
  for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x) {
    unsigned int k = keys[i];
    int slot = probe_insert(hk, k);
    if (slot >= 0) atomicAdd(&hc[slot], 1u);
    else atomicAdd(&counts[k], 1ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x)
    if (hk[i] != 0xffffffffu && hc[i]) atomicAdd(&counts[hk[i]], (unsigned long long)hc[i]);
}

__global__ __launch_bounds__(kBlock) void k_key_scatter(const uint32_t* keys, const double* vals,
                                                         int64_t n, unsigned long long* cursor,
                                                         double* out) {
  __shared__ unsigned int hk[kHashSlots];
  __shared__ unsigned int hc[kHashSlots];
  __shared__ unsigned long long hb[kHashSlots];
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
    hk[i] = 0xffffffffu;
    hc[i] = 0;
  }
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kKPTile;
  int64_t end = std::min<int64_t>(n, base + kKPTile);
  for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x) {
    int slot = probe_insert(hk, keys[i]);
    if (slot >= 0) atomicAdd(&hc[slot], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
    if (hk[i] != 0xffffffffu && hc[i]) hb[i] = atomicAdd(&cursor[hk[i]], (unsigned long long)hc[i]);
    hc[i] = 0;
  }
  __syncthreads();
  for (int64_t i = base + threadIdx.x; i < end; i += blockDim.x) {
    unsigned int k = keys[i];
    int slot = probe_find(hk, k);
    unsigned long long pos;
    if (slot >= 0) pos = hb[slot] + atomicAdd(&hc[slot], 1u);
    else pos = atomicAdd(&cursor[k], 1ull);
    out[pos] = vals[i];
  }
}

__global__ void k_excl_scan_u64(const unsigned long long* cnt, int64_t n, int64_t* off,
                                unsigned long long* cursor) {
  __shared__ unsigned long long part[1024];
  int tid = threadIdx.x;
  int64_t chunk = (n + blockDim.x - 1) / blockDim.x;
  int64_t lo = std::min<int64_t>(n, tid * chunk), hi = std::min<int64_t>(n, lo + chunk);
  unsigned long long s = 0;
  for (int64_t i = lo; i < hi; i++) s += cnt[i];
  part[tid] = s;
  __syncthreads();
  if (tid == 0) {
    unsigned long long run = 0;
    for (int i = 0; i < (int)blockDim.x; i++) {
      unsigned long long v = part[i];
      part[i] = run;
      run += v;
    }
    off[n] = (int64_t)run;
  }
  __syncthreads();
  unsigned long long run = part[tid];
  for (int64_t i = lo; i < hi; i++) {
    off[i] = (int64_t)run;
    cursor[i] = run;
    run += cnt[i];
  }
}

// ----------------------------------------------- key partition (radix)
// Most-significant-digit-first counting sort by key: pass 1 splits the batch
// by the high digit, pass 2 splits every high-digit segment by the low digit,
// so each key's values end contiguous and in key order.  A pass works on
// chunks (a few tens of thousands of items, one workgroup each): digit counts
// per chunk (k_rx_hist), one exclusive scan of the (segment, digit, chunk)
// count matrix, which yields every chunk's destination for every digit, and a
// scatter (k_rx_scatter) that walks its chunk in sub-tiles staged through LDS
// in digit order, so that consecutive lanes store to consecutive addresses,
// while per-digit cursors in LDS carry the destinations from one sub-tile to
// the next.  The ranking is stable, so every key's values keep their batch
// order.  The matrix holds one entry per (digit, chunk): with chunks of 64K
// items it is a few MB for 2^28 pairs, and each chunk reads its destinations
// once.
constexpr int kRxThreads = 256;
constexpr int kRxWaves = kRxThreads / 64;
constexpr int kRxBits = 9;
constexpr int kRxBins = 1 << kRxBits;
constexpr int kRxHistUnroll = 8;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

// The segment owning chunk t: the last s with tile_start[s] <= t.
__device__ inline int rx_seg_of_tile(const uint32_t* tile_start, int nseg, uint32_t t) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_rx_init(uint32_t* seg, uint32_t n) {
  seg[0] = 0;
  seg[1] = n;
}

// tile_start[s] = sum over earlier segments of ceil(len / chunk); one block,
// each thread owning ceil(nseg / 1024) consecutive segments.
__global__ __launch_bounds__(1024) void k_rx_tiles(const uint32_t* seg_start, int nseg,
                                                   uint32_t chunk, uint32_t* tile_start) {
  __shared__ uint32_t v[1024];
  const int per = (nseg + 1023) / 1024;
  const int lo = min(nseg, (int)threadIdx.x * per), hi = min(nseg, lo + per);
  uint32_t s = 0;
  for (int i = lo; i < hi; i++) s += (seg_start[i + 1] - seg_start[i] + chunk - 1) / chunk;
  v[threadIdx.x] = s;
  __syncthreads();
  const uint32_t total = block_excl_scan(v, 1024);
  uint32_t run = v[threadIdx.x];
  for (int i = lo; i < hi; i++) {
    tile_start[i] = run;
    run += (seg_start[i + 1] - seg_start[i] + chunk - 1) / chunk;
  }
  if (threadIdx.x == 0) tile_start[nseg] = total;
}

// Exclusive scan of cnt[0..R) into out (R <= 4 * kRxThreads); all threads call it.
__device__ inline void rx_scan_bins(const uint32_t* cnt, uint32_t* out, int R) {
  __shared__ uint32_t wsum[kRxWaves];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (R + kRxThreads - 1) / kRxThreads;
  const int lo = tid * per;
  uint32_t x[4];
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    x[j] = (j < per && lo + j < R) ? cnt[lo + j] : 0u;
    s += x[j];
  }
  uint32_t inc = s;
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (int v = 0; v < w; v++) base += wsum[v];
  uint32_t run = base + inc - s;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j < per && lo + j < R) out[lo + j] = run;
    run += x[j];
  }
}

// Heavy keys of a large batch (the few keys holding most of its pairs) get
// a first-pass digit of their own (their rank among the heavy keys, in key
// order, below kHeavyMax) and land in their final place after one pass;
// every other key takes digit kHeavyLight + its high bits and goes through
// the second pass.  They are found from a sample of the keys
// (k_heavy_sample, k_heavy_classes, k_heavy_flags, k_heavy_list) and kept as
// a bitmap over the keys with a per-word count of the heavy keys before it,
// and expanded into a per-key table of first-pass digits (k_heavy_digits),
// read through the cache by the partition kernels.
constexpr int kHeavyMax = 255;
constexpr int kHeavyLight = 256;    // first digit of the keys that take two passes
constexpr int kHeavyKeyBits = 17;   // heavy split only for keys below 2^17
constexpr int kHeavyWords = (1 << kHeavyKeyBits) / 64;

// Device layout of the heavy index (one buffer).
struct HeavyIndex {
  unsigned long long bm[kHeavyWords];  // bit k % 64 of word k / 64: key k is heavy
  uint8_t pre[kHeavyWords];            // heavy keys in the words before
  uint32_t list[kHeavyMax + 1];        // heavy keys in key order
  unsigned int cls[40];                // sample counts per power-of-two class
  int nheavy;
};

// The index staged in LDS by the histogram kernel: `words` bitmap words,
// then their prefix counts.
struct HeavyLds {
  unsigned long long* bm;
  uint8_t* pre;
};

__device__ inline HeavyLds heavy_stage(unsigned long long* lds, const HeavyIndex* h, int words) {
  HeavyLds t{lds, reinterpret_cast<uint8_t*>(lds + words)};
  for (int i = threadIdx.x; i < words; i += blockDim.x) {
    t.bm[i] = h->bm[i];
    t.pre[i] = h->pre[i];
  }
  return t;
}

// First-pass digit of key k from the staged index: its heavy rank, or
// kHeavyLight + (k >> shift) & lmask (the same digit as k_heavy_digits).
__device__ inline uint32_t heavy_digit_lds(const HeavyLds& t, int words, uint32_t k, int shift,
                                           uint32_t lmask) {
  const uint32_t w = k >> 6;
  if (w < (uint32_t)words) {
    const unsigned long long bits = t.bm[w];
    const unsigned long long below = (1ull << (k & 63)) - 1ull;
    if ((bits >> (k & 63)) & 1ull) return (uint32_t)t.pre[w] + (uint32_t)__popcll(bits & below);
  }
  return kHeavyLight + ((k >> shift) & lmask);
}

// First-pass digit of key k from the per-key digit table (k_heavy_digits);
// keys at or beyond it (out of range, flagged by validation) take the light
// digit of their high bits, so the digit stays below kHeavyLight + lmask + 1.
__device__ inline uint32_t heavy_digit(const uint16_t* dtab, uint32_t nk, uint32_t k, int shift,
                                       uint32_t lmask) {
  return k < nk ? (uint32_t)dtab[k] : kHeavyLight + ((k >> shift) & lmask);
}

// Key counts of a sample: block b takes 2048 consecutive keys starting at
// b * stride, counts them in an LDS hash and adds the counts to cnt.
__global__ __launch_bounds__(256) void k_heavy_sample(const uint32_t* keys, int64_t n, int64_t stride,
                                                      int64_t num_keys, unsigned int* cnt) {
  constexpr int kSlots = 4096;
  __shared__ uint32_t hk[kSlots];
  __shared__ uint32_t hc[kSlots];
  for (int i = threadIdx.x; i < kSlots; i += 256) {
    hk[i] = 0;
    hc[i] = 0;
  }
  __syncthreads();
  const int64_t a = (int64_t)blockIdx.x * stride;
  for (int j = threadIdx.x; j < 2048; j += 256) {
    const int64_t i = a + j;
    if (i >= n) break;
    const uint32_t k = keys[i];
    if ((int64_t)k >= num_keys) continue;
    uint32_t h = (k * 2654435761u) >> 20;
    for (;;) {
      const uint32_t old = atomicCAS(&hk[h], 0u, k + 1u);
      if (old == 0u || old == k + 1u) {
        atomicAdd(&hc[h], 1u);
        break;
      }
      h = (h + 1) & (kSlots - 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kSlots; i += 256)
    if (hk[i]) atomicAdd(&cnt[hk[i] - 1u], hc[i]);
}

// Class of a sample count c >= cmin: c in [2^(b-1), 2^b) is class b (1..32);
// 0 for counts below cmin.
__device__ inline int heavy_class(unsigned int c, unsigned int cmin) {
  return (c >= cmin && c > 0u) ? 32 - __clz(c) : 0;
}

// Keys per class (classes 1..32 in h->cls; zeroed beforehand).
__global__ __launch_bounds__(256) void k_heavy_classes(const unsigned int* cnt, int64_t num_keys,
                                                       unsigned int cmin, HeavyIndex* h) {
  __shared__ unsigned int hist[33];
  if (threadIdx.x < 33) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t k = (int64_t)blockIdx.x * 1024 + threadIdx.x;
       k < num_keys && k < (int64_t)(blockIdx.x + 1) * 1024; k += 256) {
    const int b = heavy_class(cnt[k], cmin);
    if (b) atomicAdd(&hist[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 33 && threadIdx.x > 0 && hist[threadIdx.x])
    atomicAdd(&h->cls[threadIdx.x], hist[threadIdx.x]);
}

// Lowest class t such that classes t..32 hold at most kHeavyMax keys (33:
// none).
__device__ inline int heavy_threshold(const unsigned int* cls) {
  unsigned int cum = 0;
  int thr = 33;
  for (int b = 32; b >= 1; b--) {
    if (cum + cls[b] > (unsigned int)kHeavyMax) break;
    cum += cls[b];
    thr = b;
  }
  return thr;
}

// Heavy flags as bitmap words: key k is heavy when its class is at least the
// threshold.  One wave per word (grid covers every word).
__global__ __launch_bounds__(256) void k_heavy_flags(const unsigned int* cnt, int64_t num_keys,
                                                     unsigned int cmin, HeavyIndex* h, int words) {
  __shared__ int s_thr;
  if (threadIdx.x == 0) s_thr = heavy_threshold(h->cls);
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool f = k < num_keys && heavy_class(cnt[k], cmin) >= s_thr;
  const unsigned long long m = __ballot(f);
  const int64_t w = k >> 6;
  if ((threadIdx.x & 63) == 0 && w < words) h->bm[w] = m;
}

// Per-word prefix counts and the key-ordered list (one block of 1024
// threads, two words each: words <= 2048).
__global__ __launch_bounds__(1024) void k_heavy_list(HeavyIndex* h, int words) {
  __shared__ uint32_t part[1024];
  const int w0 = 2 * threadIdx.x;
  const unsigned long long b0 = w0 < words ? h->bm[w0] : 0ull;
  const unsigned long long b1 = w0 + 1 < words ? h->bm[w0 + 1] : 0ull;
  part[threadIdx.x] = (uint32_t)(__popcll(b0) + __popcll(b1));
  __syncthreads();
  const uint32_t total = block_excl_scan(part, 1024);
  uint32_t r = part[threadIdx.x];
  if (w0 < words) h->pre[w0] = (uint8_t)r;
  if (w0 + 1 < words) h->pre[w0 + 1] = (uint8_t)(r + (uint32_t)__popcll(b0));
  for (int q = 0; q < 2; q++) {
    unsigned long long b = q ? b1 : b0;
    while (b) {
      const int i = __ffsll((long long)b) - 1;
      h->list[r++] = (uint32_t)((int64_t)(w0 + q) * 64 + i);
      b &= b - 1ull;
    }
  }
  if (threadIdx.x == 0) h->nheavy = (int)total;
}

// dtab[k]: key k's heavy rank, or kHeavyLight + (k >> shift) & lmask.
__global__ __launch_bounds__(256) void k_heavy_digits(const HeavyIndex* h, int64_t num_keys,
                                                      int shift, uint32_t lmask, uint16_t* dtab) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= num_keys) return;
  const unsigned long long bits = h->bm[k >> 6];
  const unsigned long long below = (1ull << (k & 63)) - 1ull;
  dtab[k] = ((bits >> (k & 63)) & 1ull)
                ? (uint16_t)(h->pre[k >> 6] + __popcll(bits & below))
                : (uint16_t)(kHeavyLight + (((uint32_t)k >> shift) & lmask));
}

// Digit counts of one chunk, one histogram per wave (so a hot key's digit
// contends only within its wave).
template <typename KT, bool kHeavy>
__global__ __launch_bounds__(kRxThreads) void k_rx_hist(const KT* keys,
                                                        const uint32_t* seg_start,
                                                        const uint32_t* tile_start, int nseg,
                                                        uint32_t chunk, int shift, int bits,
                                                        uint32_t* mat, const HeavyIndex* hix,
                                                        int hwords, uint32_t lmask) {
  __shared__ uint32_t lc[kRxWaves][kRxBins];
  // The heavy index in LDS (its largest size: this kernel's LDS is small),
  // where the table lookup of the scatter would stall the stream of keys.
  __shared__ unsigned long long hl[kHeavy ? (kHeavyWords * 9 + 7) / 8 : 1];
  const uint32_t t = blockIdx.x;
  if (t >= tile_start[nseg]) return;
  const int s = rx_seg_of_tile(tile_start, nseg, t);
  const uint32_t lt = t - tile_start[s], nts = tile_start[s + 1] - tile_start[s];
  const uint32_t a = seg_start[s] + lt * chunk;
  const uint32_t b = min(seg_start[s + 1], a + chunk);
  const int R = 1 << bits;
  const uint32_t mask = (uint32_t)R - 1u;
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kRxWaves * kRxBins; i += kRxThreads) (&lc[0][0])[i] = 0;
  HeavyLds ht{};
  if constexpr (kHeavy) ht = heavy_stage(hl, hix, hwords);
  __syncthreads();
  constexpr uint32_t kStep = kRxThreads * kRxHistUnroll;
  for (uint32_t i0 = a; i0 < b; i0 += kStep) {
    uint32_t k[kRxHistUnroll];
#pragma unroll
    for (int r = 0; r < kRxHistUnroll; r++) {
      uint32_t i = i0 + (uint32_t)(r * kRxThreads) + threadIdx.x;
      k[r] = i < b ? keys[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kRxHistUnroll; r++) {
      uint32_t i = i0 + (uint32_t)(r * kRxThreads) + threadIdx.x;
      const bool valid = i < b;
      uint32_t d = 0xffffffffu;
      if (valid) {
        if constexpr (kHeavy) {
          d = heavy_digit_lds(ht, hwords, (uint32_t)k[r], shift, lmask);
        } else {
          d = (k[r] >> shift) & mask;
        }
      }
      // A wave whose items all share one digit (a heavy key) adds once.
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
      const uint64_t vmask = __ballot(valid);
      if (__ballot(valid && d != d0) == 0) {
        if ((threadIdx.x & 63) == 0 && d0 != 0xffffffffu)
          atomicAdd(&lc[w][d0], (uint32_t)__popcll(vmask));
      } else if (valid) {
        atomicAdd(&lc[w][d], 1u);
      }
    }
  }
  __syncthreads();
  const size_t base = (size_t)R * tile_start[s] + lt;
  for (int d = threadIdx.x; d < R; d += kRxThreads) {
    uint32_t c = 0;
#pragma unroll
    for (int v = 0; v < kRxWaves; v++) c += lc[v][d];
    mat[base + (size_t)d * nts] = c;
  }
}

// Block sums of the matrix (its live length is R * tile_start[nseg]).
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* a,
                                                              const uint32_t* tile_start, int nseg,
                                                              int R, uint32_t* bsum) {
  __shared__ uint32_t part[kScanThreads];
  const size_t total = (size_t)R * tile_start[nseg];
  const size_t c0 = (size_t)blockIdx.x * kScanChunk;
  uint32_t acc = 0;
  for (int j = 0; j < kScanItems; j++) {
    size_t i = c0 + (size_t)j * kScanThreads + threadIdx.x;
    if (i < total) acc += a[i];
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kScanThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = part[0];
}

// Exclusive scan of the nb block sums in place (one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* bsum, int nb) {
  __shared__ uint32_t v[1024];
  const int per = (nb + 1023) / 1024;
  const int lo = min(nb, (int)threadIdx.x * per), hi = min(nb, lo + per);
  uint32_t s = 0;
  for (int i = lo; i < hi; i++) s += bsum[i];
  v[threadIdx.x] = s;
  __syncthreads();
  block_excl_scan(v, 1024);
  uint32_t run = v[threadIdx.x];
  for (int i = lo; i < hi; i++) {
    uint32_t x = bsum[i];
    bsum[i] = run;
    run += x;
  }
}

// Exclusive scan of each chunk, offset by its block's scanned sum.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* a, const uint32_t* tile_start,
                                                            int nseg, int R, const uint32_t* bsum) {
  __shared__ uint32_t v[kScanThreads];
  const size_t total = (size_t)R * tile_start[nseg];
  const size_t c0 = (size_t)blockIdx.x * kScanChunk;
  if (c0 >= total) return;
  // Each thread owns kScanItems consecutive entries.
  const size_t t0 = c0 + (size_t)threadIdx.x * kScanItems;
  uint32_t x[kScanItems];
  uint32_t s = 0;
  for (int j = 0; j < kScanItems; j++) {
    x[j] = (t0 + j < total) ? a[t0 + j] : 0u;
    s += x[j];
  }
  v[threadIdx.x] = s;
  __syncthreads();
  block_excl_scan(v, kScanThreads);
  uint32_t run = v[threadIdx.x] + bsum[blockIdx.x];
  for (int j = 0; j < kScanItems; j++) {
    if (t0 + j < total) a[t0 + j] = run;
    run += x[j];
  }
}

// One chunk, walked in sub-tiles of kRxThreads * ITEMS items.  Stable: wave w
// owns the contiguous items [w * ITEMS * 64, (w + 1) * ITEMS * 64) of a
// sub-tile, walked in ITEMS rounds of 64; lanes with equal digits are found
// with a 64-bit lane mask per (wave, digit) in LDS, ranked by lane and
// counted per wave; per digit, the waves are
// then ordered by index and the sub-tiles by position in the chunk.  The next
// sub-tile's loads are issued before the current one is written out.  With
// kValidate (first pass) a key at or above key_limit or a non-finite value
// sets *bad (grouped.mojo:91-97); positions stay in bounds either way because
// only the key's digit bits place it.  With kWriteKeys (first of two passes)
// only the key bits the later passes read (key & out_mask, at most 18 bits,
// as KO) are written beside each value.
template <int ITEMS, bool kWriteKeys, bool kValidate, typename KT, bool kHeavy, typename KO>
__global__ __launch_bounds__(kRxThreads) void k_rx_scatter(const KT* keys, const double* vals,
                                                           const uint32_t* seg_start,
                                                           const uint32_t* tile_start, int nseg,
                                                           uint32_t chunk, int shift, int bits,
                                                           const uint32_t* dst, KO* keys_out,
                                                           uint32_t out_mask, double* vals_out,
                                                           uint32_t key_limit, int* bad,
                                                           const uint16_t* dtab, uint32_t dkeys,
                                                           uint32_t lmask,
                                                           const uint32_t* pos_base,
                                                           double* vals_heavy) {
  constexpr int kSub = kRxThreads * ITEMS;
  constexpr int kPerWave = ITEMS * 64;
  constexpr int kPm = kRxWaves * kRxBins;
  __shared__ uint32_t wc[kRxWaves][kRxBins];
  __shared__ uint32_t cnt[kRxBins];
  __shared__ uint32_t lc[kRxBins];
  __shared__ uint32_t run[kRxBins];
  // The lane masks of the ranking and the staged values of the write-out
  // share this LDS (the masks are all zero again when the ranking ends and
  // are cleared again after the write-out).
  __shared__ unsigned long long stage[kSub > kPm ? kSub : kPm];
  // Staged digit (bits 18 and up; below 512) and later-pass key bits (the
  // low 18) per item.
  __shared__ uint32_t sdk[kSub];
  double* sv = reinterpret_cast<double*>(stage);
  auto pm = [&](int wv, uint32_t d) -> unsigned long long& { return stage[wv * kRxBins + (int)d]; };
  const uint32_t t = blockIdx.x;
  if (t >= tile_start[nseg]) return;
  const int s = rx_seg_of_tile(tile_start, nseg, t);
  const uint32_t lt = t - tile_start[s], nts = tile_start[s + 1] - tile_start[s];
  const uint32_t a = seg_start[s] + lt * chunk;
  const uint32_t b = min(seg_start[s + 1], a + chunk);
  const int R = 1 << bits;
  const uint32_t mask = (uint32_t)R - 1u;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t gbase = (size_t)R * tile_start[s] + lt;
  const uint32_t pbase = pos_base ? *pos_base : 0u;
  for (int d = threadIdx.x; d < R; d += kRxThreads) run[d] = pbase + dst[gbase + (size_t)d * nts];
  for (int i = threadIdx.x; i < kPm; i += kRxThreads) {
    (&wc[0][0])[i] = 0;
    stage[i] = 0ull;
  }
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // Items past the chunk's end keep a key of an earlier sub-tile (or 0):
  // any key indexes the digit table in bounds after the clamp below.
  uint32_t kr[ITEMS] = {};
  uint32_t rk[ITEMS];
  uint16_t dg[ITEMS];
  double vr[ITEMS];
  bool wrong = false;
  auto load = [&](uint32_t sub) {
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      uint32_t i = sub + (uint32_t)(w * kPerWave + r * 64 + lane);
      if (i < b) {
        kr[r] = keys[i];
        vr[r] = vals[i];
      }
    }
  };
  load(a);
  __syncthreads();
  for (uint32_t sub = a; sub < b; sub += kSub) {
    const uint32_t e = min(b, sub + (uint32_t)kSub);
    uint16_t dh[ITEMS];
    if constexpr (kHeavy) {
      // The sub-tile's digit-table lookups all issued before any is used.
#pragma unroll
      for (int r = 0; r < ITEMS; r++) dh[r] = dtab[kr[r] < dkeys ? kr[r] : 0u];
    }
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      uint32_t i = sub + (uint32_t)(w * kPerWave + r * 64 + lane);
      bool valid = i < e;
      uint32_t d = 0u;
      if (valid) {
        if constexpr (kHeavy) {
          // heavy_digit with the looked-up entry.
          const uint32_t k = kr[r];
          d = k < dkeys ? (uint32_t)dh[r] : kHeavyLight + ((k >> shift) & lmask);
        } else {
          d = (kr[r] >> shift) & mask;
        }
      }
      dg[r] = (uint16_t)d;
      if (kValidate && valid) wrong |= kr[r] >= key_limit || !(vr[r] - vr[r] == 0.0);
      // The lanes sharing this lane's digit: every valid lane ORs its bit
      // into its digit's mask, then reads it back (one wave's LDS operations
      // complete in issue order, so the read sees every lane's OR, and the
      // leader's reset lands before the next round's ORs).
      // A wave whose valid items all share one digit (a heavy key) ranks
      // them by lane directly, without the masks.
      const uint32_t dv = valid ? d : 0xffffffffu;
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(dv);
      const uint64_t vmask = __ballot(valid);
      const bool uniform = __ballot(valid && dv != d0) == 0;
      uint64_t peers;
      if (uniform) {
        peers = valid ? vmask : 0ull;
      } else {
        if (valid) atomicOr(&pm(w, d), 1ull << lane);
        peers = valid ? pm(w, d) : 0ull;
      }
      uint32_t before = (uint32_t)__popcll(peers & below);
      uint32_t base = valid ? wc[w][d] : 0u;
      if (valid) {
        rk[r] = base + before;
        if (before == 0) {
          wc[w][d] = base + (uint32_t)__popcll(peers);
          if (!uniform) pm(w, d) = 0ull;
        }
      }
    }
    __syncthreads();
    // Per digit: offsets of the waves, and the sub-tile total.
    for (int d = threadIdx.x; d < R; d += kRxThreads) {
      uint32_t acc = 0;
#pragma unroll
      for (int v = 0; v < kRxWaves; v++) {
        uint32_t c = wc[v][d];
        wc[v][d] = acc;
        acc += c;
      }
      cnt[d] = acc;
    }
    __syncthreads();
    rx_scan_bins(cnt, lc, R);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
      uint32_t i = sub + (uint32_t)(w * kPerWave + r * 64 + lane);
      if (i < e) {
        uint32_t d = dg[r];
        uint32_t pos = lc[d] + wc[w][d] + rk[r];
        sv[pos] = vr[r];
        sdk[pos] = (d << 18) | (kWriteKeys ? ((uint32_t)kr[r] & out_mask & 0x3ffffu) : 0u);
      }
    }
    __syncthreads();
    if (sub + (uint32_t)kSub < b) load(sub + (uint32_t)kSub);
    const int len = (int)(e - sub);
    for (int i = threadIdx.x; i < len; i += kRxThreads) {
      const uint32_t dk = sdk[i];
      const uint32_t d = dk >> 18, k = dk & 0x3ffffu;
      uint32_t o = run[d] + ((uint32_t)i - lc[d]);
      if (kHeavy && d < (uint32_t)kHeavyLight) {
        vals_heavy[o] = sv[i];  // final place: no second pass
      } else {
        vals_out[o] = sv[i];
        if (kWriteKeys) keys_out[o] = (KO)(k & out_mask);
      }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < R; d += kRxThreads) {
      run[d] += cnt[d];
#pragma unroll
      for (int v = 0; v < kRxWaves; v++) wc[v][d] = 0;
    }
    for (int i = threadIdx.x; i < kPm; i += kRxThreads) stage[i] = 0ull;
    __syncthreads();
  }
  if (kValidate && wrong) *bad = 1;
}

// Segments of the next pass: (segment s, digit d) starts where the scanned
// matrix puts the first tile of s for digit d.
__global__ void k_rx_segs(const uint32_t* seg_start, const uint32_t* tile_start, int nseg, int R,
                          const uint32_t* dst, uint32_t* seg_out, const uint32_t* pos_base) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t m = (int64_t)nseg * R;
  if (idx > m) return;
  if (idx == m) {
    seg_out[m] = seg_start[nseg];
    return;
  }
  int s = (int)(idx / R), d = (int)(idx % R);
  uint32_t nts = tile_start[s + 1] - tile_start[s];
  const uint32_t pbase = pos_base ? *pos_base : 0u;
  seg_out[idx] = nts ? pbase + dst[(size_t)R * tile_start[s] + (size_t)d * nts] : seg_start[s];
}

// Per key: start and count of its values.  A heavy key's values are the
// first-pass segment of its digit (seg1); the other keys' are second-pass
// segments (seg, indexed by key).
__global__ void k_rx_final(const uint32_t* seg, int64_t num_keys, int64_t* off, int64_t* cnt,
                           const uint32_t* seg1, const uint16_t* dtab, uint32_t total) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > num_keys) return;
  if (k == num_keys) {
    off[k] = total;
    return;
  }
  int r = -1;
  if (dtab && dtab[k] < kHeavyLight) r = dtab[k];
  if (r >= 0) {
    off[k] = seg1[r];
    cnt[k] = (int64_t)seg1[r + 1] - (int64_t)seg1[r];
  } else {
    off[k] = seg[k];
    cnt[k] = (int64_t)seg[k + 1] - (int64_t)seg[k];
  }
}

void Engine::key_partition_radix(const uint32_t* d_keys, const double* d_vals, int64_t n,
                                 int bits, int64_t num_keys, int64_t* d_counts,
                                 int64_t* d_offsets, double* d_out, int* d_bad) {
  // Up to 2^18 keys: one or two passes (hb high bits, then lb low bits).
  // More: three passes of bits-18, 9 and 9 bits.
  const bool three = bits > 2 * kRxBits;
  const int hb = bits > kRxBits ? bits - kRxBits : 0;  // bits above the last pass's
  const int lb = bits > kRxBits ? kRxBits : bits;      // low digit bits (last pass)
  // Heavy keys get their own first-pass digit when the batch is large
  // enough for a handful of keys to hold most of it.
  const bool heavy = heavy_min_ > 0 && hb > 0 && bits <= kHeavyKeyBits && n >= ((int64_t)1 << 22);
  const int items = rx_items_ == 4 || rx_items_ == 16 ? rx_items_ : 8;
  const int64_t sub = (int64_t)kRxThreads * items;
  // Chunk: about n / rx_chunks_ items, a whole number of sub-tiles, at most 2^20.
  int64_t chunk = (n + std::max<int64_t>(rx_chunks_, 1) - 1) / std::max<int64_t>(rx_chunks_, 1);
  chunk = std::min<int64_t>(std::max<int64_t>((chunk + sub - 1) / sub * sub, sub), (int64_t)1 << 20);
  // Tiles of a pass over nseg segments: at most one partial tile per
  // segment beyond n / chunk.
  const int64_t max_tiles1 = (n + chunk - 1) / chunk;
  const int64_t max_tiles2 = max_tiles1 + ((int64_t)1 << hb);
  const int64_t mat_max = (int64_t)kRxBins * max_tiles2;
  const int64_t nb_max = (mat_max + kScanChunk - 1) / kScanChunk;
  uint32_t* mat = as<uint32_t>(rx_mat_, mat_max);
  uint32_t* bsum = as<uint32_t>(rx_bsum_, nb_max);
  uint32_t* segA = as<uint32_t>(rx_seg1_, ((int64_t)1 << bits) + kRxBins + 2);
  uint32_t* segB = as<uint32_t>(rx_seg2_, ((int64_t)1 << bits) + kRxBins + 2);
  uint32_t* tiles = as<uint32_t>(rx_tile_, ((int64_t)1 << hb) + 2);
  const uint32_t klim = (uint32_t)std::min<int64_t>(num_keys, 0xffffffffLL);
  HeavyIndex* hix = nullptr;
  uint16_t* dtab = nullptr;
  const uint32_t dkeys = heavy ? (uint32_t)num_keys : 0u;
  const int hwords = (int)((num_keys + 63) / 64);
  if (heavy) {
    unsigned int* hc = as<unsigned int>(heavy_cnt_, num_keys);
    hix = reinterpret_cast<HeavyIndex*>(as<uint8_t>(heavy_list_, sizeof(HeavyIndex)));
    check_hip(hipMemsetAsync(hc, 0, num_keys * sizeof(unsigned int), stream_), "memset heavy counts");
    check_hip(hipMemsetAsync(hix->cls, 0, sizeof(hix->cls), stream_), "memset heavy classes");
    const int nsb = 256;
    const int64_t stride = std::max<int64_t>(2048, n / nsb);
    const int64_t sampled = std::min<int64_t>(n, (int64_t)nsb * 2048);
    // Heavy: an estimated heavy_min_ values or more (a sample count of at
    // least heavy_min_ * sampled / n).
    const unsigned int cmin = (unsigned int)std::max<double>(
        1.0, std::ceil((double)heavy_min_ * (double)sampled / (double)n));
    k_heavy_sample<<<nsb, 256, 0, stream_>>>(d_keys, n, stride, num_keys, hc);
    launched("k_heavy_sample");
    k_heavy_classes<<<(unsigned)((num_keys + 1023) / 1024), 256, 0, stream_>>>(hc, num_keys, cmin, hix);
    launched("k_heavy_classes");
    k_heavy_flags<<<(unsigned)((hwords * 64 + 255) / 256), 256, 0, stream_>>>(hc, num_keys, cmin, hix,
                                                                              hwords);
    launched("k_heavy_flags");
    k_heavy_list<<<1, 1024, 0, stream_>>>(hix, hwords);
    launched("k_heavy_list");
    dtab = as<uint16_t>(heavy_dig_, num_keys);
    k_heavy_digits<<<(unsigned)((num_keys + 255) / 256), 256, 0, stream_>>>(
        hix, num_keys, lb, (1u << hb) - 1u, dtab);
    launched("k_heavy_digits");
  }
  k_rx_init<<<1, 1, 0, stream_>>>(segA, (uint32_t)n);
  launched("k_rx_init");

  const uint32_t low_mask = (1u << lb) - 1u;
  const uint32_t high_mask = (1u << hb) - 1u;
  auto scatter = [&](bool first, bool last, unsigned grid, const auto* kin, const double* vin,
                     const uint32_t* seg, int nseg, int shift, int dbits, auto* kout,
                     uint32_t out_mask, double* vout, const uint32_t* pbase) {
    using KT = std::remove_const_t<std::remove_pointer_t<decltype(kin)>>;
    using KO = std::remove_pointer_t<decltype(kout)>;
    const uint32_t c = (uint32_t)chunk;
#define KD_RX_LAUNCH(I, WK, VA, HV)                                                               \
  k_rx_scatter<I, WK, VA, KT, HV, KO><<<grid, kRxThreads, 0, stream_>>>(                         \
      kin, vin, seg, tiles, nseg, c, shift, dbits, mat, kout, out_mask, vout, klim, d_bad, dtab,  \
      dkeys, high_mask, pbase, d_out)
#define KD_RX_ITEMS(WK, VA, HV)            \
  if (items == 4) KD_RX_LAUNCH(4, WK, VA, HV);       \
  else if (items == 16) KD_RX_LAUNCH(16, WK, VA, HV); \
  else KD_RX_LAUNCH(8, WK, VA, HV)
    const bool validate = first && d_bad != nullptr;
    if constexpr (std::is_same<KT, uint32_t>::value && std::is_same<KO, uint32_t>::value) {
      // First of three passes.
      if (validate) { KD_RX_ITEMS(true, true, false); }
      else { KD_RX_ITEMS(true, false, false); }
    } else if constexpr (std::is_same<KT, uint32_t>::value) {
      if (!last && heavy) {
        if (validate) { KD_RX_ITEMS(true, true, true); }
        else { KD_RX_ITEMS(true, false, true); }
      }
      else if (!last && validate) { KD_RX_ITEMS(true, true, false); }
      else if (!last) { KD_RX_ITEMS(true, false, false); }
      else if (validate) { KD_RX_ITEMS(false, true, false); }
      else { KD_RX_ITEMS(false, false, false); }
    } else {
      KD_RX_ITEMS(false, false, false);
    }
#undef KD_RX_ITEMS
#undef KD_RX_LAUNCH
    launched("k_rx_scatter");
  };

  // Pairs taking the second pass (all but the heavy keys'): known on the
  // device only; read back for the phase byte counts when profiling.
  int64_t n_light = n;
  auto pass = [&](const auto* kin, const double* vin, const uint32_t* seg, int nseg, int shift,
                  int dbits, int64_t tiles_max, auto* kout, uint32_t out_mask, double* vout,
                  uint32_t* seg_out, bool first, bool hv, const uint32_t* pbase) {
    const int R = 1 << dbits;
    k_rx_tiles<<<1, 1024, 0, stream_>>>(seg, nseg, (uint32_t)chunk, tiles);
    launched("k_rx_tiles");
    {
      int ph_ = phase_begin("key_hist", (double)sizeof(*kin) * (first ? (double)n : (double)n_light));
      using KT = std::remove_const_t<std::remove_pointer_t<decltype(kin)>>;
      if (hv)
        k_rx_hist<KT, true><<<(unsigned)tiles_max, kRxThreads, 0, stream_>>>(
            kin, seg, tiles, nseg, (uint32_t)chunk, shift, dbits, mat, hix, hwords, high_mask);
      else
        k_rx_hist<KT, false><<<(unsigned)tiles_max, kRxThreads, 0, stream_>>>(
            kin, seg, tiles, nseg, (uint32_t)chunk, shift, dbits, mat, nullptr, 0, 0u);
      launched("k_rx_hist");
      phase_end(ph_);
    }
    const int64_t nb = ((int64_t)R * tiles_max + kScanChunk - 1) / kScanChunk;
    k_scan_reduce<<<(unsigned)nb, kScanThreads, 0, stream_>>>(mat, tiles, nseg, R, bsum);
    launched("k_scan_reduce");
    k_scan_top<<<1, 1024, 0, stream_>>>(bsum, (int)nb);
    launched("k_scan_top");
    k_scan_down<<<(unsigned)nb, kScanThreads, 0, stream_>>>(mat, tiles, nseg, R, bsum);
    launched("k_scan_down");
    if (hv && prof_) {
      // The first light digit's offset in the scanned matrix (one segment:
      // tiles_max tiles, digit-major) is the heavy keys' pair count.
      uint32_t h = 0;
      check_hip(hipMemcpyAsync(&h, mat + (size_t)kHeavyLight * tiles_max, sizeof(h),
                               hipMemcpyDeviceToHost, stream_), "heavy count");
      check_hip(hipStreamSynchronize(stream_), "heavy count");
      n_light = n - (int64_t)h;
    }
    {
      // Bytes: the key (4 or 2 B) and value read; the value and the key bits
      // of the later passes (light keys only) written.
      const double items = first ? (double)n : (double)n_light;
      const double b = ((double)sizeof(*kin) + 16.0) * items +
                       (kout ? (double)sizeof(*kout) * (double)n_light : 0.0);
      int ph_ = phase_begin("key_scatter", b);
      scatter(first, kout == nullptr, (unsigned)tiles_max, kin, vin, seg, nseg, shift, dbits, kout,
              out_mask, vout, pbase);
      phase_end(ph_);
    }
    const int64_t m = (int64_t)nseg * R + 1;
    k_rx_segs<<<(unsigned)((m + 255) / 256), 256, 0, stream_>>>(seg, tiles, nseg, R, mat, seg_out,
                                                               pbase);
    launched("k_rx_segs");
  };

  if (three) {
    // bits-18, then 9, then 9 bits: the first pass keeps the low 18 key
    // bits (32-bit), the second the low 9 (16-bit).
    const int h1 = bits - 2 * kRxBits;
    uint32_t* k1 = as<uint32_t>(rx_keys32_, n);
    uint16_t* k2 = as<uint16_t>(rx_keys_, n);
    double* v1 = as<double>(rx_vals_, n);
    double* v2 = as<double>(rx_vals2_, n);
    const int64_t max_tiles_mid = max_tiles1 + ((int64_t)1 << h1);
    pass(d_keys, d_vals, segA, 1, 2 * kRxBits, h1, max_tiles1, k1, (1u << (2 * kRxBits)) - 1u, v1,
         segB, true, false, nullptr);
    pass((const uint32_t*)k1, v1, segB, 1 << h1, kRxBits, kRxBits, max_tiles_mid, k2, low_mask, v2,
         segA, false, false, nullptr);
    pass((const uint16_t*)k2, v2, segA, 1 << (h1 + kRxBits), 0, kRxBits, max_tiles2,
         (uint16_t*)nullptr, 0u, d_out, segB, false, false, nullptr);
    std::swap(segA, segB);
  } else if (hb > 0) {
    uint16_t* k1 = as<uint16_t>(rx_keys_, n);
    double* v1 = as<double>(rx_vals_, n);
    if (heavy) {
      // Pass 1 over digits 0..255 (heavy keys, straight to the output) and
      // 256 + high bits; pass 2 over the light segments only, placed after
      // the heavy values.
      pass(d_keys, d_vals, segA, 1, lb, kRxBits, max_tiles1, k1, low_mask, v1, segB, true, true,
           nullptr);
      pass((const uint16_t*)k1, v1, segB + kHeavyLight, 1 << hb, 0, lb, max_tiles2,
           (uint16_t*)nullptr, 0u, d_out, segA, false, false, segB + kHeavyLight);
    } else {
      pass(d_keys, d_vals, segA, 1, lb, hb, max_tiles1, k1, low_mask, v1, segB, true, false,
           nullptr);
      pass((const uint16_t*)k1, v1, segB, 1 << hb, 0, lb, max_tiles2, (uint16_t*)nullptr, 0u, d_out,
           segA, false, false, nullptr);
    }
  } else {
    pass(d_keys, d_vals, segA, 1, 0, lb, max_tiles1, (uint16_t*)nullptr, 0u, d_out, segB, true,
         false, nullptr);
    std::swap(segA, segB);
  }
  k_rx_final<<<(unsigned)((num_keys + 1 + 255) / 256), 256, 0, stream_>>>(
      segA, num_keys, d_offsets, d_counts, segB, dtab, (uint32_t)n);
  launched("k_rx_final");
}

static int key_bits(int64_t num_keys) {
  int bits = 0;
  while (bits < 31 && ((int64_t)1 << bits) < num_keys) bits++;
  return std::max(bits, 1);
}

bool Engine::key_partition_validates(int64_t n, int64_t num_keys) const {
  return n > 0 && n <= ((int64_t)1 << 31) &&
         key_bits(num_keys) <= std::min(rx_max_bits_, 3 * kRxBits);
}

void Engine::key_partition(const uint32_t* d_keys, const double* d_vals, int64_t n,
                           int64_t num_keys, int64_t* d_counts, int64_t* d_offsets,
                           double* d_out, int* d_bad) {
  if (key_partition_validates(n, num_keys)) {
    key_partition_radix(d_keys, d_vals, n, key_bits(num_keys), num_keys, d_counts, d_offsets, d_out,
                        d_bad);
    check_hip(hipGetLastError(), "key partition launch");
    return;
  }
  auto* cnt = reinterpret_cast<unsigned long long*>(d_counts);
  auto* cur = as<unsigned long long>(extra_[0], num_keys + 1);
  check_hip(hipMemsetAsync(cnt, 0, num_keys * sizeof(unsigned long long), stream_), "memset counts");
  if (n > 0) {
    unsigned g = (unsigned)((n + kKPTile - 1) / kKPTile);
    {
      int ph_ = phase_begin("key_hist", 4.0 * (double)n);
      k_key_hist<<<g, kBlock, 0, stream_>>>(d_keys, n, cnt);
      launched("k_key_hist");
      phase_end(ph_);
    }
    k_excl_scan_u64<<<1, 1024, 0, stream_>>>(cnt, num_keys, d_offsets, cur);
    launched("k_excl_scan_u64");
    {
      int ph_ = phase_begin("key_scatter", 20.0 * (double)n);
      k_key_scatter<<<g, kBlock, 0, stream_>>>(d_keys, d_vals, n, cur, d_out);
      launched("k_key_scatter");
      phase_end(ph_);
    }
  } else {
    k_excl_scan_u64<<<1, 1024, 0, stream_>>>(cnt, num_keys, d_offsets, cur);
    launched("k_excl_scan_u64");
  }
  check_hip(hipGetLastError(), "key partition launch");
}

}  // namespace kd
