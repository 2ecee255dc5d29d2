// kd_batch.hip — many independent digests fed sample by sample (add_at, A
// semantics of kdigest/tdigest.mojo:163-175, 324-350) and reduced by a
// pairwise merge_into tree (tdigest.mojo:589-623), all on the GPU.
//
// One workgroup per digest keeps the digest's centroids and ingest buffer in
// LDS for the whole stream.  Per sample, the clock is a block max-scan; the
// decay exponent and the stored total are sequential floating-point sums,
// formed by the block with the loop's exact bits (block_running_sum); the
// decay factors, the stable sort, the weight prefix, the k values, the
// next-boundary search and the cluster means run across the block.  Each
// buffer merge (tdigest.mojo:391-439) is the same greedy sweep as the CPU:
// the chain of cluster starts is walked by lane 0 over a next-boundary table
// computed in parallel, and every cluster's running mean and weight are then
// replayed record by record (bit-identical to the CPU for identical
// boundaries).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kd_batch.h"
#include "kd_exact.h"
#include "kd_math.h"

namespace kd {

namespace {

constexpr int kRB = 512;          // threads per digest (8 waves: two per SIMD)
constexpr int kRCMax = 1536;      // centroids kept per digest
constexpr int kRBMax = 1536;      // ingest buffer capacity (>= 5 * compression)
constexpr int kRNMax = kRCMax + kRBMax;
constexpr double kRenorm = 400.0;   // tdigest.mojo:43
constexpr double kPruneEps = 1e-6;  // tdigest.mojo:47

// LDS layout (bytes) of one digest.
constexpr int kOffCm = 0;
constexpr int kOffCw = kOffCm + 8 * kRCMax;
constexpr int kOffBm = kOffCw + 8 * kRCMax;
constexpr int kOffBw = kOffBm + 8 * kRBMax;
constexpr int kOffWm = kOffBw + 8 * kRBMax;
constexpr int kOffWw = kOffWm + 8 * kRNMax;
constexpr int kOffP = kOffWw + 8 * kRNMax;
constexpr int kOffKv = kOffP + 8 * (kRNMax + 8);
constexpr int kOffNx = kOffKv + 8 * (kRNMax + 8);
constexpr int kLds = kOffNx + 4 * kRNMax;
static_assert(kLds <= 160 * 1024, "LDS budget of one workgroup");

struct Lds {
  double *cm, *cw, *bm, *bw, *wm, *ww, *P, *kv;
  int* nx;
  // Sort scratch and cluster starts alias the k-value region.
  double* skey;
  int* sidx;
  int* starts;
};

__device__ inline Lds carve(char* pool) {
  Lds L;
  L.cm = reinterpret_cast<double*>(pool + kOffCm);
  L.cw = reinterpret_cast<double*>(pool + kOffCw);
  L.bm = reinterpret_cast<double*>(pool + kOffBm);
  L.bw = reinterpret_cast<double*>(pool + kOffBw);
  L.wm = reinterpret_cast<double*>(pool + kOffWm);
  L.ww = reinterpret_cast<double*>(pool + kOffWw);
  L.P = reinterpret_cast<double*>(pool + kOffP);
  L.kv = reinterpret_cast<double*>(pool + kOffKv);
  L.nx = reinterpret_cast<int*>(pool + kOffNx);
  L.skey = L.kv;
  L.sidx = reinterpret_cast<int*>(pool + kOffKv + 8 * 2048);
  L.starts = reinterpret_cast<int*>(pool + kOffKv);
  return L;
}

__device__ inline uint64_t ord_key(double x) {
  uint64_t b = (x == 0.0) ? 0ull : kd_d2bits(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Shared scalar state of the digest being processed (lane 0 writes).
struct Sh {
  BatchState st;
  int m, renorm, nclusters;
  int frac, bad;  // merge_lds: some weight is not an integer; first cluster in question
  int anytie;     // merge_lds: some next-boundary link carries the tie flag
  double renorm_ls;
};

// ---------------------------------------------------------------- scans
constexpr int kRWaves = kRB / 64;
constexpr int kRPer = (kRBMax + kRB - 1) / kRB;  // values per thread in a running-sum round

// Shared state of the block scans and running sums below.
struct RunSh {
  uint64_t ws[kRWaves];
  double wd[kRWaves];
  int code;  // round: 2 * index + 1 of the first step the units do not cover (2 m + 1: none)
  int i0, done, consumed, stopped, kind;
  double S, Sprev, stopv, Send;
  int total;                  // merge_lds: clusters kept
  uint64_t pmask[kRWaves];    // merge_lds: threads that pruned a cluster
};

// Exclusive sum of v over the block's threads in thread order.
__device__ inline uint64_t block_excl_scan_u64(uint64_t v, RunSh& rs) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) rs.ws[w] = inc;
  __syncthreads();
  uint64_t base = 0;
  for (int k = 0; k < w; k++) base += rs.ws[k];
  __syncthreads();
  return base + inc - v;
}

// max(init, v of every earlier thread): an exact exclusive max-scan.
__device__ inline double block_excl_max(double v, double init, RunSh& rs) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(inc, o, 64);
    if (lane >= o) inc = y > inc ? y : inc;
  }
  if (lane == 63) rs.wd[w] = inc;
  __syncthreads();
  double base = init;
  for (int k = 0; k < w; k++) base = rs.wd[k] > base ? rs.wd[k] : base;
  const double before = __shfl_up(inc, 1, 64);
  __syncthreads();
  return lane > 0 && before > base ? before : base;
}

// Exclusive sum of v over the block's threads (a fixed tree: wave scan,
// then the wave totals in order).
__device__ inline double block_excl_sum(double v, RunSh& rs) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(inc, o, 64);
    if (lane >= o) inc = y + inc;
  }
  if (lane == 63) rs.wd[w] = inc;
  __syncthreads();
  double base = 0.0;
  for (int k = 0; k < w; k++) base = base + rs.wd[k];
  const double before = __shfl_up(inc, 1, 64);
  __syncthreads();
  return lane > 0 ? base + before : base;
}

struct RunOut {
  int consumed, stopped;
  double S, stopv;
};

// The sequential loop S = S + x[i], i = 0..m-1 (x[i] >= 0; m <= kRB * kRPer),
// bit for bit, by the block in binade rounds (kd_sum_units): every step of a
// round is an integer number of units of the round's binade, so the sums
// are an exclusive scan of those counts; the first step they do not cover
// (an exact half, a sum leaving the binade) is added plainly by lane 0 and
// the next round starts after it -- one round per call in practice.  A start
// below 2^-960 (zero) passes zeros through and adds the first nonzero x
// plainly.  With stop_on, the first nonzero x whose sum exceeds thr is not
// consumed (its sum is stopv).  out[i] = the sum after x[i] when out is
// given; out may alias x (an index is written after its round has read it).
// kd_running_sum is the same algorithm sequentially (tested against the
// loop on the host).
__device__ RunOut block_running_sum(double S0, const double* x, double* out, int m, bool stop_on,
                                    double thr, RunSh& rs) {
  const int tid = threadIdx.x;
  if (tid == 0) {
    rs.i0 = 0;
    rs.S = S0;
    rs.done = 0;
    rs.consumed = m;
    rs.stopped = 0;
    rs.stopv = 0.0;
  }
  __syncthreads();
  const int none = 2 * m + 1;
  for (;;) {
    const int i0 = rs.i0;
    const double Sc = rs.S;
    const int done = rs.done;
    __syncthreads();  // every thread has read the round's start
    if (done || i0 >= m) break;
    if (!(Sc >= 0x1p-960)) {
      if (tid == 0) rs.code = none;
      __syncthreads();
      for (int i = i0 + tid; i < m; i += kRB)
        if (x[i] != 0.0) {
          atomicMin(&rs.code, 2 * i + 1);
          break;
        }
      __syncthreads();
      const int F = rs.code >> 1;  // first nonzero x, or m
      if (out)
        for (int i = i0 + tid; i < F; i += kRB) out[i] = Sc;
      if (tid == 0) {
        if (F >= m) {
          rs.i0 = m;
        } else {
          const double ns = Sc + x[F];
          if (stop_on && ns > thr) {
            rs.done = 1;
            rs.consumed = F;
            rs.stopped = 1;
            rs.stopv = ns;
          } else {
            if (out) out[F] = ns;
            rs.S = ns;
            rs.i0 = F + 1;
          }
        }
      }
      __syncthreads();
      continue;
    }
    const int e = kd_exponent(Sc);
    const double u = kd_pow2(e - 52), ui = kd_pow2(52 - e);
    const uint64_t A0 = (uint64_t)(Sc * ui);
    const int cnt = m - i0;
    const int per = (cnt + kRB - 1) / kRB;
    const int a = i0 + min(cnt, tid * per), b = i0 + min(cnt, tid * per + per);
    if (tid == 0) rs.code = none;
    double xv[kRPer];
    int64_t rv[kRPer];
    unsigned upm = 0;
    uint64_t part = 0;
    bool live = true;
#pragma unroll
    for (int k = 0; k < kRPer; k++) {
      const int i = a + k;
      xv[k] = 0.0;
      rv[k] = 0;
      if (i < b) {
        xv[k] = x[i];
        int64_t fl = 0;
        const int64_t r = kd_sum_units(xv[k], ui, &fl);
        rv[k] = r;
        if (r >= 0 && r != fl) upm |= 1u << k;
        if (live) {
          if (r < 0) live = false;
          else part += (uint64_t)r;
        }
      }
    }
    uint64_t A = A0 + block_excl_scan_u64(part, rs);
    int code = none, kind = 0;
    double sv[kRPer];
    double Sprev = 0.0, stopv = 0.0;
    live = true;
#pragma unroll
    for (int k = 0; k < kRPer; k++) {
      const int i = a + k;
      sv[k] = 0.0;
      if (live && i < b) {
        const int64_t r = rv[k];
        const uint64_t fl = r < 0 ? 0 : (uint64_t)(r - (int64_t)((upm >> k) & 1u));
        if (r < 0 || A + fl >= (1ull << 53)) {
          live = false;
          code = 2 * i + 1;
          kind = 2;
          Sprev = (double)A * u;
        } else {
          const uint64_t A1 = A + (uint64_t)r;
          const double s1 = (double)A1 * u;
          if (stop_on && xv[k] != 0.0 && s1 > thr) {
            live = false;
            code = 2 * i + 1;
            kind = 3;
            Sprev = (double)A * u;
            stopv = s1;
          } else {
            sv[k] = s1;
            A = A1;
          }
        }
      }
    }
    if (code < none) atomicMin(&rs.code, code);
    __syncthreads();
    const int cmin = rs.code;
    const int F = cmin >> 1;  // m when every step was covered
    if (cmin < none && code == cmin) {
      rs.kind = kind;
      rs.Sprev = Sprev;
      rs.stopv = stopv;
    }
    if (cmin == none && b == m && b > a) rs.Send = (double)A * u;
    if (out) {
#pragma unroll
      for (int k = 0; k < kRPer; k++) {
        const int i = a + k;
        if (i < b && i < F) out[i] = sv[k];
      }
    }
    __syncthreads();
    if (tid == 0) {
      if (cmin == none) {
        rs.S = rs.Send;
        rs.i0 = m;
      } else if (rs.kind == 3) {
        rs.S = rs.Sprev;
        rs.done = 1;
        rs.consumed = F;
        rs.stopped = 1;
      } else {
        const double ns = rs.Sprev + x[F];
        if (stop_on && x[F] != 0.0 && ns > thr) {
          rs.S = rs.Sprev;
          rs.done = 1;
          rs.consumed = F;
          rs.stopped = 1;
          rs.stopv = ns;
        } else {
          if (out) out[F] = ns;
          rs.S = ns;
          rs.i0 = F + 1;
        }
      }
    }
    __syncthreads();
  }
  RunOut o{rs.consumed, rs.stopped, rs.S, rs.stopv};
  __syncthreads();
  return o;
}

// Stable sort of the buffer by (mean, position): the order of the pair
// (order key of the mean, position), over kSortSlots slots padded with keys
// above every finite mean.  Thread t holds slots [kSortE t, kSortE t +
// kSortE) in registers.  Each wave first sorts its kRun slots ascending with
// a bitonic network (partners closer than kSortE in the same thread, the
// others exchanged by lane shuffles), then pairs of runs are merged through
// LDS, each thread finding its outputs on the merge path.  Leaves the sorted
// means in skey and their buffer positions in sidx.
constexpr int kSortSlots = 2048;
constexpr int kSortE = kSortSlots / kRB;
constexpr int kRun = 64 * kSortE;
static_assert(kSortSlots >= kRBMax && kSortE >= 1 && kRB * kSortE == kSortSlots, "sort slots");

__device__ inline bool sort_before(uint64_t ka, int ia, uint64_t kb, int ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// In-thread stage of partner distance J (< kSortE); kd: the direction bit.
template <int J>
__device__ inline void sort_reg_stage(uint64_t (&key)[kSortE], int (&idx)[kSortE], int base, int kd) {
#pragma unroll
  for (int r = 0; r < kSortE; r++) {
    if ((r & J) == 0 && (r | J) < kSortE) {
      const int r2 = r | J;
      const bool asc = ((base + r) & kd) == 0;
      const bool sw = asc ? sort_before(key[r2], idx[r2], key[r], idx[r])
                          : sort_before(key[r], idx[r], key[r2], idx[r2]);
      if (sw) {
        const uint64_t tk = key[r];
        key[r] = key[r2];
        key[r2] = tk;
        const int ti = idx[r];
        idx[r] = idx[r2];
        idx[r2] = ti;
      }
    }
  }
}

__device__ void sort_buffer(Lds& L, int nb) {
  const int base = (int)threadIdx.x * kSortE;
  uint64_t key[kSortE];
  int idx[kSortE];
#pragma unroll
  for (int r = 0; r < kSortE; r++) {
    const int i = base + r;
    key[r] = i < nb ? ord_key(L.bm[i]) : ~0ull;
    idx[r] = i < nb ? i : 0x7fffffff;
  }
  // 1. One ascending run of kRun slots per wave.
  for (int k = 2; k <= kRun; k <<= 1) {
    const int kd = k == kRun ? 0 : k;  // the last level sorts every run ascending
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < kSortE) {
        if constexpr (kSortE > 4) {
          if (j == 4) sort_reg_stage<4>(key, idx, base, kd);
        }
        if constexpr (kSortE > 2) {
          if (j == 2) sort_reg_stage<2>(key, idx, base, kd);
        }
        if constexpr (kSortE > 1) {
          if (j == 1) sort_reg_stage<1>(key, idx, base, kd);
        }
        continue;
      }
      // Slot p takes the first of its pair when it is the lower slot of an
      // ascending pair or the upper of a descending one, else the second.
      const int lm = j / kSortE;
#pragma unroll
      for (int r = 0; r < kSortE; r++) {
        const uint64_t pk = __shfl_xor(key[r], lm, 64);
        const int pi = __shfl_xor(idx[r], lm, 64);
        const int p = base + r;
        const bool lower = (p & j) == 0;
        const bool asc = (p & kd) == 0;
        const bool mine_first = sort_before(key[r], idx[r], pk, pi);
        const bool take = (lower == asc) ? !mine_first : mine_first;
        if (take) {
          key[r] = pk;
          idx[r] = pi;
        }
      }
    }
  }
  // 2. Runs merged pairwise: outputs base.. of the merged pair come from
  //    the first i slots of run A and the first (d - i) of run B, i the first
  //    A index that B's (d - 1 - i)-th slot precedes.
  uint64_t* sk = reinterpret_cast<uint64_t*>(L.skey);
  int* si = L.sidx;
#pragma unroll
  for (int r = 0; r < kSortE; r++) {
    sk[base + r] = key[r];
    si[base + r] = idx[r];
  }
  for (int run = kRun; run < kSortSlots; run <<= 1) {
    __syncthreads();
    const int rb = base & ~(2 * run - 1), d = base - rb;
    const uint64_t* ak = sk + rb;
    const int* ai = si + rb;
    const uint64_t* bk = sk + rb + run;
    const int* bi = si + rb + run;
    int lo = d > run ? d - run : 0, hi = d < run ? d : run;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sort_before(ak[mid], ai[mid], bk[d - 1 - mid], bi[d - 1 - mid])) lo = mid + 1;
      else hi = mid;
    }
    int ia = lo, ib = d - lo;
#pragma unroll
    for (int r = 0; r < kSortE; r++) {
      const int ca = ia < run ? ia : run - 1, cb = ib < run ? ib : run - 1;
      const uint64_t xa = ak[ca], xb = bk[cb];
      const int ya = ai[ca], yb = bi[cb];
      const bool ta = ib >= run || (ia < run && sort_before(xa, ya, xb, yb));
      key[r] = ta ? xa : xb;
      idx[r] = ta ? ya : yb;
      ia += ta ? 1 : 0;
      ib += ta ? 0 : 1;
    }
    __syncthreads();
    if (2 * run < kSortSlots) {
#pragma unroll
      for (int r = 0; r < kSortE; r++) {
        sk[base + r] = key[r];
        si[base + r] = idx[r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kSortE; r++) {
    const int i = base + r;
    if (i < nb) {
      L.skey[i] = L.bm[idx[r]];
      L.sidx[i] = idx[r];
    }
  }
  __syncthreads();
}

// merge() of tdigest.mojo:391-439 (+ prune 441-458) on the digest in LDS.
// Returns false when the result would not fit kRCMax centroids.
__device__ bool merge_lds(Lds& L, Sh& sh, RunSh& rs, int scale, double compression, double lambda) {
  const int nc = sh.st.nc, nb = sh.st.nb;
  if (nb == 0) return true;
  const int n = nc + nb;
  sort_buffer(L, nb);
  // work = centroids ++ buffer, stably sorted: centroids first on ties.
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    double v = L.skey[k];
    int lo = 0, hi = nc;  // # centroids <= v
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (!(v < L.cm[mid])) lo = mid + 1;
      else hi = mid;
    }
    L.wm[k + lo] = v;
    L.ww[k + lo] = L.bw[L.sidx[k]];
  }
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    double v = L.cm[i];
    int lo = 0, hi = nb;  // # buffer < v
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (L.skey[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    L.wm[i + lo] = v;
    L.ww[i + lo] = L.cw[i];
  }
  __syncthreads();
  // Weight prefix P[j] = sum of the first j weights (blocked scan; exact for
  // integer weights, and every boundary found from it is checked in the
  // CPU's order below otherwise).
  {
    const int chunk = (n + blockDim.x - 1) / blockDim.x;
    const int a = min(n, (int)threadIdx.x * chunk), b = min(n, a + chunk);
    double s = 0.0;
    for (int j = a; j < b; j++) s += L.ww[j];
    double run = block_excl_sum(s, rs);
    for (int j = a; j < b; j++) {
      L.P[j] = run;
      run += L.ww[j];
    }
    if (b == n && a < b) L.P[n] = run;  // the thread holding the last record
  }
  __syncthreads();
  const double N = sh.st.stored_total;
  const double eff = lambda > 0.0 ? N * kd_exp(-sh.st.log_scale) : N;
  const double nrm = kd_normalizer(scale, compression, eff);
  if (threadIdx.x == 0) sh.anytie = 0;
  for (int j = threadIdx.x; j <= n; j += blockDim.x) L.kv[j] = kd_k(scale, L.P[j] / N, nrm);
  __syncthreads();
  // nx[s]: first candidate i > s with NOT(k(P[i+1]/N) - k(P[s]/N) <= 1), or n.
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    double ks = L.kv[s];
    int lo = s + 2, hi = n + 1;  // search j = i + 1
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (!(L.kv[mid] - ks <= 1.0)) hi = mid;
      else lo = mid + 1;
    }
    // Bit 30 flags a candidate within 2 ulps of the record before it, where
    // the tie rule (tdigest.mojo:429) may extend the cluster.
    const int b = lo - 1;
    const bool tie = b < n && ord_key(L.wm[b]) - ord_key(L.wm[b - 1]) <= 2;
    L.nx[s] = b | (tie ? 0x40000000 : 0);
    if (tie) sh.anytie = 1;
  }
  __syncthreads();
  // Chain of cluster starts s -> nx[s] from 0.  Without a tie flag the links
  // are plain, and one wave walks 64 links at a time: jump tables f^2 ..
  // f^32 (16-bit, n absorbing) in the free prefix and k-value regions, lane
  // i forms f^i(s) from the bits of i.
  if (!sh.anytie) {
    constexpr int kJs = kRNMax + 2;  // table stride (entries 0..n)
    uint16_t* jt = reinterpret_cast<uint16_t*>(L.P);  // f^2, f^4, f^8, f^16
    uint16_t* j32 = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(L.kv) + 4 * kRNMax);
    static_assert(4 * kJs * 2 <= 8 * (kRNMax + 8) && kJs * 2 <= 4 * (kRNMax + 8), "jump tables");
    auto f1 = [&](int x) { return x < n ? (L.nx[x] & 0x3fffffff) : n; };
    for (int x = threadIdx.x; x <= n; x += blockDim.x) jt[x] = (uint16_t)f1(f1(x));
    __syncthreads();
    for (int t = 1; t < 5; t++) {
      const uint16_t* src = jt + (t - 1) * kJs;
      uint16_t* dst = t < 4 ? jt + t * kJs : j32;
      for (int x = threadIdx.x; x <= n; x += blockDim.x) dst[x] = src[src[x]];
      __syncthreads();
    }
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      int s = 0, c = 0;
      while (s < n) {
        int x = s;
        if (lane & 1) x = f1(x);
        if (lane & 2) x = jt[x];
        if (lane & 4) x = jt[kJs + x];
        if (lane & 8) x = jt[2 * kJs + x];
        if (lane & 16) x = jt[3 * kJs + x];
        if (lane & 32) x = j32[x];
        if (x < n) L.starts[c + lane] = x;  // f^lane(s): the chain is a prefix of the lanes
        c += __popcll(__ballot(x < n));
        const int last = __shfl(x, 63, 64);
        s = f1(last);
      }
      if (lane == 0) {
        sh.nclusters = c;
        sh.frac = 0;
        sh.bad = 0x7fffffff;
      }
    }
  } else if (threadIdx.x == 0) {
    // With the tie rule (tdigest.mojo:429), lane 0.
    int s = 0, c = 0;
    while (s < n) {
      L.starts[c++] = s;
      if (s == n - 1) break;
      const int v = L.nx[s];
      const int b = v & 0x3fffffff;
      int e = b;
      if (v & 0x40000000) {
        double xb = L.wm[b];
        {
          if (L.wm[s] == xb) {
            while (e < n && L.wm[e] == xb) e++;
          } else {
            double mean = L.wm[s], wt = L.ww[s];
            for (int j = s + 1; j < b; j++) {
              double nw = wt + L.ww[j];
              mean += (L.wm[j] - mean) * (L.ww[j] / nw);
              wt = nw;
            }
            while (e < n && L.wm[e] == mean) e++;
          }
        }
      }
      s = e;
    }
    sh.nclusters = c;
    sh.frac = 0;
    sh.bad = 0x7fffffff;
  }
  __syncthreads();
  // The prefix above is exact when every weight is an integer; otherwise
  // every boundary is checked in the CPU's summation order (kd_exact.h):
  // each cluster's running weight, the sums of the clusters before it (in
  // order), then the last absorbed candidate must not close and the next one
  // must.  From the first cluster in question on (rare: mostly the last
  // candidate, where q_right is about 1), lane 0 runs the CPU sweep itself.
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    if (L.ww[j] != floor(L.ww[j])) sh.frac = 1;
  __syncthreads();
  if ((sh.frac || !(N < 0x1p53)) && n > 1) {
    const int C0 = sh.nclusters;
    double* T0 = L.P;  // free until the means
    for (int c = threadIdx.x; c < C0; c += blockDim.x) {
      const int a = L.starts[c], b = c + 1 < C0 ? L.starts[c + 1] : n;
      double W = L.ww[a];
      for (int j = a + 1; j < b; j++) W = W + L.ww[j];
      T0[c] = W;
    }
    __syncthreads();
    // T0[c] = the sum of the cluster weights up to c, in cluster order (the
    // CPU's weight_so_far after c); wsf of cluster c is T0[c - 1].
    // (One lane: the sum grows from one cluster's weight to the total, so
    // it crosses a binade every few clusters and the block's rounds would
    // not pay.)
    if (threadIdx.x == 0) {
      double acc = 0.0;
      int c = 0;
      for (; c + 8 <= C0; c += 8) {
        double t8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t8[u] = T0[c + u];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          acc = acc + t8[u];
          T0[c + u] = acc;
        }
      }
      for (; c < C0; c++) {
        acc = acc + T0[c];
        T0[c] = acc;
      }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < C0; c += blockDim.x) {
      const int a = L.starts[c], b = c + 1 < C0 ? L.starts[c + 1] : n;
      const double wsf = c > 0 ? T0[c - 1] : 0.0;
      double W = L.ww[a];
      bool bad = false;
      if (b - a >= 2) {
        for (int j = a + 1; j < b - 1; j++) W = W + L.ww[j];
        bad = kd_sweep_closes(scale, nrm, N, wsf, W, L.ww[b - 1]) != 0;
        W = W + L.ww[b - 1];
      }
      if (b < n) bad = bad || !kd_sweep_closes(scale, nrm, N, wsf, W, L.ww[b]);
      if (bad) atomicMin(&sh.bad, c);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh.bad < C0) {
      int c = sh.bad;
      int st0 = L.starts[c];
      double wsf = c > 0 ? T0[c - 1] : 0.0;
      double cur_w = L.ww[st0], cur_m = L.wm[st0];
      L.starts[c++] = st0;
      for (int i = st0 + 1; i < n; i++) {
        const double wi = L.ww[i], xi = L.wm[i];
        if (xi == cur_m || !kd_sweep_closes(scale, nrm, N, wsf, cur_w, wi)) {
          const double nw = cur_w + wi;
          cur_m += (xi - cur_m) * (wi / nw);
          cur_w = nw;
        } else {
          wsf = wsf + cur_w;
          L.starts[c++] = i;
          cur_w = wi;
          cur_m = xi;
        }
      }
      sh.nclusters = c;
    }
    __syncthreads();
  }
  const int C = sh.nclusters;
  // Running mean and weight of every cluster, record by record
  // (centroid.mojo:34-36).  P is free now: means go there; weights go to the
  // consumed buffer (bm and bw are contiguous: kRNMax doubles).
  double* om = L.P;
  double* ow = L.bm;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    int a = L.starts[c], b = (c + 1 < C) ? L.starts[c + 1] : n;
    double mean = L.wm[a], wt = L.ww[a];
    for (int j = a + 1; j < b; j++) {
      double nw = wt + L.ww[j];
      mean += (L.wm[j] - mean) * (L.ww[j] / nw);
      wt = nw;
    }
    om[c] = mean;
    ow[c] = wt;
  }
  __syncthreads();
  // Prune (decay active, tdigest.mojo:441-458) and install: each thread
  // compacts a contiguous run of clusters at its scanned offset; the pruned
  // weight is summed in cluster order by lane 0 over the runs that have any.
  {
    const double floor = lambda > 0.0 ? kPruneEps * kd_exp(sh.st.log_scale) : 0.0;
    const int per = (C + kRB - 1) / kRB;
    const int c0 = min(C, (int)threadIdx.x * per), c1 = min(C, c0 + per);
    int kc = 0;
    for (int c = c0; c < c1; c++) kc += (lambda > 0.0 && !(ow[c] >= floor)) ? 0 : 1;
    const bool pr = (c1 - c0) - kc > 0;
    const uint64_t pmask = __ballot(pr);
    int k = (int)block_excl_scan_u64((uint64_t)kc, rs);
    if (threadIdx.x == kRB - 1) rs.total = k + kc;
    if ((threadIdx.x & 63) == 0) rs.pmask[threadIdx.x >> 6] = pmask;
    __syncthreads();
    if (threadIdx.x == 0) {
      // The pruned weight, summed in cluster order over the threads that
      // pruned any.
      double removed = 0.0;
      for (int w = 0; w < kRWaves; w++) {
        uint64_t mk = rs.pmask[w];
        while (mk) {
          const int t = w * 64 + __builtin_ctzll(mk);
          mk &= mk - 1;
          const int a0 = min(C, t * per), a1 = min(C, a0 + per);
          for (int c = a0; c < a1; c++)
            if (!(ow[c] >= floor)) removed += ow[c];
        }
      }
      if (removed > 0.0) sh.st.stored_total -= removed;
      const int run = rs.total;
      const bool fits = run <= kRCMax;
      sh.st.nc = fits ? run : kRCMax;
      sh.st.nb = 0;
      if (!fits) sh.st.err = 1;
    }
    for (int c = c0; c < c1; c++) {
      if (lambda > 0.0 && !(ow[c] >= floor)) continue;
      if (k < kRCMax) {
        L.cm[k] = om[c];
        L.cw[k] = ow[c];
      }
      k++;
    }
  }
  __syncthreads();
  return sh.st.err == 0;
}

__device__ void renormalize_lds(Lds& L, Sh& sh, double log_scale) {
  // tdigest.mojo:359-371.
  double factor = kd_exp(-log_scale);
  for (int i = threadIdx.x; i < sh.st.nc; i += blockDim.x) L.cw[i] *= factor;
  for (int i = threadIdx.x; i < sh.st.nb; i += blockDim.x) L.bw[i] *= factor;
  __syncthreads();
  if (threadIdx.x == 0) {
    sh.st.stored_total *= factor;
    sh.st.log_scale = 0.0;
  }
  __syncthreads();
}

__device__ void load_digest(Lds& L, Sh& sh, const BatchState* st, const double* gcm,
                            const double* gcw, const double* gbm, const double* gbw, int64_t d) {
  if (threadIdx.x == 0) {
    sh.st = st[d];
    sh.renorm = 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < sh.st.nc; i += blockDim.x) {
    L.cm[i] = gcm[d * kRCMax + i];
    L.cw[i] = gcw[d * kRCMax + i];
  }
  for (int i = threadIdx.x; i < sh.st.nb; i += blockDim.x) {
    L.bm[i] = gbm[d * kRBMax + i];
    L.bw[i] = gbw[d * kRBMax + i];
  }
  __syncthreads();
}

__device__ void store_digest(Lds& L, Sh& sh, BatchState* st, double* gcm, double* gcw,
                             double* gbm, double* gbw, int64_t d) {
  for (int i = threadIdx.x; i < sh.st.nc; i += blockDim.x) {
    gcm[d * kRCMax + i] = L.cm[i];
    gcw[d * kRCMax + i] = L.cw[i];
  }
  for (int i = threadIdx.x; i < sh.st.nb; i += blockDim.x) {
    gbm[d * kRBMax + i] = L.bm[i];
    gbw[d * kRBMax + i] = L.bw[i];
  }
  if (threadIdx.x == 0) st[d] = sh.st;
}

}  // namespace

// add_at for each sample of each digest, in order (tdigest.mojo:163-175).
__global__ __launch_bounds__(kRB) void k_replay(int scale, double compression, double lambda,
                                                int limit, const int64_t* order, const int64_t* off,
                                                const double* vals, const double* times,
                                                const double* wts, BatchState* st, double* gcm,
                                                double* gcw, double* gbm, double* gbw) {
  __shared__ __attribute__((aligned(16))) char pool[kLds];
  __shared__ Sh sh;
  __shared__ RunSh rsh;
  Lds L = carve(pool);
  const int64_t d = order[blockIdx.x];
  load_digest(L, sh, st, gcm, gcw, gbm, gbw, d);
  if (sh.st.err) return;
  int64_t p = off[d];
  const int64_t end = off[d + 1];
  while (p < end || sh.renorm) {
    const int nb0 = sh.st.nb;
    const int have = (int)(end - p < (int64_t)(limit - nb0) ? end - p : (int64_t)(limit - nb0));
    const int tid = threadIdx.x;
    const int per = (have + kRB - 1) / kRB;
    const int i0 = min(have, tid * per), i1 = min(have, i0 + per);
    // 0. The timestamps this walk may consume, staged in LDS (P, wm, kv and
    //    ww are free between merges).
    for (int i = tid; i < have; i += kRB) L.P[i] = times[p + i];
    __syncthreads();
    // 1. Clock and decay exponent per sample (advance_clock,
    //    tdigest.mojo:324-332).  The clock before sample i is the running
    //    maximum of the earlier timestamps (a block max-scan, exact); sample
    //    i adds lambda * (t_i - clock) to the exponent when t_i is later.
    //    Only the exponent's running sum is sequential (lane 0, one addition
    //    per sample, stopping where it crosses the renormalization
    //    threshold); the sample that crosses it is consumed by the next walk.
    {
      double lm = -__builtin_inf();
      for (int i = i0; i < i1; i++) lm = L.P[i] > lm ? L.P[i] : lm;
      double prev = block_excl_max(lm, sh.st.clock, rsh);
      for (int i = i0; i < i1; i++) {
        const double t = L.P[i];
        L.bw[nb0 + i] = (lambda > 0.0 && t > prev) ? lambda * (t - prev) : 0.0;
        prev = t > prev ? t : prev;
        L.kv[i] = prev;  // the clock after sample i
      }
    }
    __syncthreads();
    if (lambda > 0.0) {
      // The exponent's running sum, stopping where it crosses the
      // renormalization threshold (block_running_sum: the sequential
      // additions' bits).
      const RunOut o = block_running_sum(sh.st.log_scale, L.bw + nb0, L.bw + nb0, have, true,
                                         kRenorm, rsh);
      if (tid == 0) {
        // The clock after the last consumed sample, or after the one whose
        // advance triggers the renormalization.
        const int last = o.stopped ? o.consumed : o.consumed - 1;
        if (last >= 0) sh.st.clock = L.kv[last];
        sh.renorm = o.stopped;
        sh.renorm_ls = o.stopv;
        sh.m = o.consumed;
        sh.st.log_scale = o.S;
      }
    } else if (tid == 0) {
      if (have > 0) sh.st.clock = L.kv[have - 1];
      sh.renorm = 0;
      sh.m = have;
    }
    __syncthreads();
    const int m = sh.m;
    // 2. Scaled weights w * e^(log_scale) in parallel (tdigest.mojo:345).
    for (int i = tid; i < m; i += kRB) {
      L.bm[nb0 + i] = vals[p + i];
      double sf = lambda > 0.0 ? kd_exp(L.bw[nb0 + i]) : 1.0;
      L.bw[nb0 + i] = wts[p + i] * sf;
    }
    __syncthreads();
    // 3. Insert bookkeeping in order (tdigest.mojo:334-347): min and max by a
    //    block reduction in sample order that keeps the earliest of equal
    //    values (as the sequential strict comparisons do); the stored total
    //    by the sequential additions' bits (block_running_sum).
    {
      const int pm = (m + kRB - 1) / kRB;
      const int j0 = min(m, tid * pm), j1 = min(m, j0 + pm);
      double mn = 0.0, mx = 0.0;
      int has = 0;
      for (int i = j0; i < j1; i++) {
        const double v = L.bm[nb0 + i];
        if (!has) {
          mn = v;
          mx = v;
          has = 1;
        } else {
          if (v < mn) mn = v;
          if (v > mx) mx = v;
        }
      }
      // Lane l takes lane l + o's range (the later samples) at step o.
      const int lane = tid & 63;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double omn = __shfl_down(mn, o, 64), omx = __shfl_down(mx, o, 64);
        const int oh = __shfl_down(has, o, 64);
        if ((lane & (2 * o - 1)) == 0 && oh) {
          if (!has) {
            mn = omn;
            mx = omx;
            has = 1;
          } else {
            if (omn < mn) mn = omn;
            if (omx > mx) mx = omx;
          }
        }
      }
      if (lane == 0) {
        L.wm[tid >> 6] = mn;
        L.ww[tid >> 6] = mx;
        L.nx[tid >> 6] = has;
      }
    }
    const bool first = sh.st.stored_total == 0.0 && nb0 == 0;
    const RunOut tot = block_running_sum(sh.st.stored_total, L.bw + nb0, nullptr, m, false, 0.0, rsh);
    if (tid == 0 && m > 0) {
      double vmin = sh.st.vmin, vmax = sh.st.vmax;
      bool init = !first;
      for (int t = 0; t < kRWaves; t++) {
        if (!L.nx[t]) continue;
        const double a = L.wm[t], c = L.ww[t];
        if (!init) {
          vmin = a;
          vmax = c;
          init = true;
        } else {
          if (a < vmin) vmin = a;
          if (c > vmax) vmax = c;
        }
      }
      sh.st.stored_total = tot.S;
      sh.st.vmin = vmin;
      sh.st.vmax = vmax;
      sh.st.nb = nb0 + m;
    }
    __syncthreads();
    p += m;
    if (sh.renorm) {
      renormalize_lds(L, sh, sh.renorm_ls);
      if (threadIdx.x == 0) sh.renorm = 0;
      __syncthreads();
      continue;  // the sample that triggered it is inserted by the next walk
    }
    if (sh.st.nb >= limit) {
      if (!merge_lds(L, sh, rsh, scale, compression, lambda)) break;
    }
  }
  store_digest(L, sh, st, gcm, gcw, gbm, gbw, d);
}

// merge() of every digest (tdigest.mojo:391-439).
__global__ __launch_bounds__(kRB) void k_flush(int scale, double compression, double lambda,
                                               BatchState* st, double* gcm, double* gcw,
                                               double* gbm, double* gbw) {
  __shared__ __attribute__((aligned(16))) char pool[kLds];
  __shared__ Sh sh;
  __shared__ RunSh rsh;
  Lds L = carve(pool);
  const int64_t d = blockIdx.x;
  load_digest(L, sh, st, gcm, gcw, gbm, gbw, d);
  if (sh.st.err || sh.st.nb == 0) return;
  merge_lds(L, sh, rsh, scale, compression, lambda);
  store_digest(L, sh, st, gcm, gcw, gbm, gbw, d);
}

// self.merge_into(other) for digest pairs (self = base + 2 i stride, other =
// self + stride); tdigest.mojo:589-623.  Both digests are already merged.
__global__ __launch_bounds__(kRB) void k_merge_pairs(int scale, double compression, double lambda,
                                                     int64_t nd, int64_t stride, BatchState* st,
                                                     double* gcm, double* gcw, double* gbm,
                                                     double* gbw) {
  __shared__ __attribute__((aligned(16))) char pool[kLds];
  __shared__ Sh sh;
  __shared__ RunSh rsh;
  __shared__ BatchState other;
  Lds L = carve(pool);
  const int64_t self = (int64_t)blockIdx.x * 2 * stride;
  const int64_t oth = self + stride;
  if (oth >= nd) return;
  if (threadIdx.x == 0) other = st[oth];
  __syncthreads();
  if (other.err) {
    if (threadIdx.x == 0) st[self].err = 1;
    return;
  }
  // other.merge() happened in k_flush; effective count of other (:602).
  double other_eff = lambda > 0.0 ? other.stored_total * kd_exp(-other.log_scale)
                                  : other.stored_total;
  if (other_eff == 0.0) return;
  load_digest(L, sh, st, gcm, gcw, gbm, gbw, self);
  if (sh.st.err) return;
  // Renormalize both digests (decay active on both: same policy).
  double ofac = lambda > 0.0 ? kd_exp(-other.log_scale) : 1.0;
  if (lambda > 0.0) renormalize_lds(L, sh, sh.st.log_scale);
  // Append other's centroids (renormalized) to the buffer (:611-614).
  for (int i = threadIdx.x; i < other.nc; i += blockDim.x) {
    L.bm[i] = gcm[oth * kRCMax + i];
    double w = gcw[oth * kRCMax + i];
    L.bw[i] = lambda > 0.0 ? w * ofac : w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ostored = lambda > 0.0 ? other.stored_total * ofac : other.stored_total;
    bool was_empty = sh.st.stored_total == 0.0;
    sh.st.nb = other.nc;
    sh.st.stored_total += ostored;
    if (was_empty) {
      sh.st.vmin = other.vmin;
      sh.st.vmax = other.vmax;
    } else {
      sh.st.vmin = fmin(sh.st.vmin, other.vmin);
      sh.st.vmax = fmax(sh.st.vmax, other.vmax);
    }
    sh.st.clock = fmax(sh.st.clock, other.clock);
  }
  __syncthreads();
  merge_lds(L, sh, rsh, scale, compression, lambda);
  store_digest(L, sh, st, gcm, gcw, gbm, gbw, self);
  // other is renormalized in place (tdigest.mojo:608-609 mutates it).
  if (lambda > 0.0) {
    for (int i = threadIdx.x; i < other.nc; i += blockDim.x) gcw[oth * kRCMax + i] *= ofac;
    if (threadIdx.x == 0) {
      st[oth].stored_total = other.stored_total * ofac;
      st[oth].log_scale = 0.0;
    }
  }
}

int batch_capacity_centroids() { return kRCMax; }
int batch_capacity_buffer() { return kRBMax; }

void launch_replay(hipStream_t s, int scale, double compression, double lambda, int limit,
                   const int64_t* order, int64_t nd, const int64_t* off, const double* vals,
                   const double* times, const double* wts, BatchState* st, double* gcm,
                   double* gcw, double* gbm, double* gbw) {
  k_replay<<<(unsigned)nd, kRB, 0, s>>>(scale, compression, lambda, limit, order, off, vals, times,
                                       wts, st, gcm, gcw, gbm, gbw);
}

void launch_flush(hipStream_t s, int scale, double compression, double lambda, int64_t nd,
                  BatchState* st, double* gcm, double* gcw, double* gbm, double* gbw) {
  k_flush<<<(unsigned)nd, kRB, 0, s>>>(scale, compression, lambda, st, gcm, gcw, gbm, gbw);
}

void launch_merge_pairs(hipStream_t s, int scale, double compression, double lambda, int64_t nd,
                        int64_t stride, BatchState* st, double* gcm, double* gcw, double* gbm,
                        double* gbw) {
  int64_t pairs = (nd + 2 * stride - 1) / (2 * stride);
  k_merge_pairs<<<(unsigned)pairs, kRB, 0, s>>>(scale, compression, lambda, nd, stride, st, gcm,
                                               gcw, gbm, gbw);
}

}  // namespace kd
