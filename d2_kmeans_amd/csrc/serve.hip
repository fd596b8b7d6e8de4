// Assignment server: exact nearest-centroid queries for distributions that
// arrive one by one, against the resident centroids of one Euclidean phase.
// Each query is what the reference's --eval mode computes per object
// (d2_assignment -> d2_labeling, clustering.c:324-376, 528-546): the label of
// the centroid with the smallest exact transport distance (first minimum on
// ties) and that distance, sqrt of the LP value (clustering.c:126-127).
//
// Continuous batching.  Queries wait in a host queue.  Whenever a stream that
// serves a waiting query's size class (m <= 64, <= 192, larger; with three or
// more streams each class has one of its own and the rest are shared) is free
// and the HBM arena has room, that class is admitted as one launch group: up
// to max_batch waiting queries of the class, in arrival order, whose inputs,
// bounds and results fit the arena.  The group's work runs on
// its own stream, so groups admitted at different times overlap on the
// device; finished groups free their stream and memory at the next progress
// call.  Per group:
//   1. lower bounds for every (query, centroid).  W2^2 splits exactly into
//      the squared distance of the means plus W2^2 of the centred pair
//      (the cross term of any coupling vanishes), and a transport plan moves
//      each query point's mass to some centroid point and each centroid
//      point's mass from some query point, so with centred costs C~
//          LP >= |mu_x - mu_c|^2 + max( sum_j b_j min_k C~_kj ,
//                                       sum_k a_k min_j C~_kj )
//      (FP64, direct differences);
//   2. exact LPs for every centroid whose bound is within a factor tau of
//      the smallest bound (the nearest centroid is almost always among them);
//   3. exact LPs for the centroids still unsolved that the bounds cannot
//      rule out: bound at most the best first-round value, and not excluded
//      by Elkan's test W2(c1, c) > 2 W2(x, c1) (W2 is a metric; c1 = the
//      best first-round centroid, centroid-pair distances computed when the
//      server opens); usually none are left;
//   4. the smallest exact value wins, lowest centroid index on ties.
// Exclusions are strict with a relative margin, so a centroid tied with the
// winner is always solved and the reference's tie rule decides.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "d2mi_internal.h"
#include "transport.h"

namespace {

constexpr int LB_T = 256;       // threads of the bound kernel
constexpr int LB_KC = 16;       // centroid rows per register chunk
constexpr int LB_DC = 8;        // dimensions per register chunk
constexpr int LB_JPT = 8;       // query columns per thread (m <= LB_T * LB_JPT)
constexpr double EXCL_MARGIN = 1e-9;

__device__ __forceinline__ double wmin(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One workgroup per (query r, centroid c): the two nearest-point bounds.
// The centroid's s x d supports sit in LDS (uniform reads); each thread owns
// up to LB_JPT query columns and keeps LB_KC row accumulators in registers.
__global__ __launch_bounds__(LB_T) void k_srv_bounds(int K, int s, int d,
                                                     const double *__restrict__ cs,
                                                     const double *__restrict__ cw,
                                                     const int *__restrict__ p_str,
                                                     const long long *__restrict__ cum,
                                                     const double *__restrict__ w,
                                                     const double *__restrict__ supp,
                                                     double *__restrict__ lb,
                                                     double *__restrict__ vals) {
  // LDS: s*d centroid supports, 4 waves x LB_KC row minima, d mean shifts,
  // 2 x LB_T partial sums of the means
  extern __shared__ __align__(16) double shA[];
  double *rmin_w = shA + (size_t)s * d;
  double *delta = rmin_w + 4 * LB_KC;
  double *pmc = delta + d, *pmx = pmc + LB_T;
  __shared__ double red[8];
  const int r = blockIdx.x / K, c = blockIdx.x - (blockIdx.x / K) * K;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int m = p_str[r];
  const long long off = cum[r];
  const double *A = cs + (size_t)c * s * d;
  const double *aw = cw + (size_t)c * s;
  for (int q = tid; q < s * d; q += LB_T) shA[q] = A[q];
  // delta = mu_c - mu_x.  Every thread takes part: thread (part, t) sums
  // dimension t over the points part, part + PD, ... (four at a time, so
  // their loads are in flight together), then each dimension adds its PD
  // partials in a fixed order (d > LB_T: one dimension per thread, directly)
  const bool parts = d <= LB_T;
  const int PD = parts ? LB_T / d : 1;
  for (int q = tid; q < PD * d; q += LB_T) {
    const int part = q / d, t = q - part * d;
    double mc = 0.0, mx[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = part; k < s; k += PD) mc += aw[k] * A[(size_t)k * d + t];
    int j = part;
    for (; j + 3 * PD < m; j += 4 * PD) {
      double wv[4], xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        wv[u] = w[off + j + u * PD];
        xv[u] = supp[(size_t)(off + j + u * PD) * d + t];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) mx[u] += wv[u] * xv[u];
    }
    for (; j < m; j += PD) mx[0] += w[off + j] * supp[(size_t)(off + j) * d + t];
    const double sx = (mx[0] + mx[1]) + (mx[2] + mx[3]);
    if (parts) {
      pmc[q] = mc;
      pmx[q] = sx;
    } else {
      delta[t] = mc - sx;
    }
  }
  __syncthreads();
  for (int t = tid; parts && t < d; t += LB_T) {
    double mc = 0.0, mx = 0.0;
    for (int part = 0; part < PD; ++part) {
      mc += pmc[part * d + t];
      mx += pmx[part * d + t];
    }
    delta[t] = mc - mx;
  }
  __syncthreads();
  // centre the pair: (a - mu_c) - (b - mu_x) = (a - delta) - b
  for (int q = tid; q < s * d; q += LB_T) shA[q] -= delta[q % d];
  double dd = 0.0;
  for (int t = 0; t < d; ++t) dd += delta[t] * delta[t];
  __syncthreads();

  double colmin[LB_JPT];
#pragma unroll
  for (int jj = 0; jj < LB_JPT; ++jj) colmin[jj] = DBL_MAX;
  double row_part = 0.0;
  for (int kc = 0; kc < s; kc += LB_KC) {
    double rmin[LB_KC];
#pragma unroll
    for (int k = 0; k < LB_KC; ++k) rmin[k] = DBL_MAX;
#pragma unroll
    for (int jj = 0; jj < LB_JPT; ++jj) {
      const int j = tid + jj * LB_T;
      if (j < m) {
        double acc[LB_KC];
#pragma unroll
        for (int k = 0; k < LB_KC; ++k) acc[k] = 0.0;
        const double *b = supp + (size_t)(off + j) * d;
        for (int t0 = 0; t0 < d; t0 += LB_DC) {
          double bv[LB_DC];
#pragma unroll
          for (int u = 0; u < LB_DC; ++u) bv[u] = t0 + u < d ? b[t0 + u] : 0.0;
#pragma unroll
          for (int k = 0; k < LB_KC; ++k) {
            if (kc + k < s) {
              const double *a = shA + (size_t)(kc + k) * d + t0;
#pragma unroll
              for (int u = 0; u < LB_DC; ++u) {
                if (t0 + u < d) {
                  const double df = a[u] - bv[u];
                  acc[k] += df * df;
                }
              }
            }
          }
        }
#pragma unroll
        for (int k = 0; k < LB_KC; ++k) {
          if (kc + k < s) {
            rmin[k] = fmin(rmin[k], acc[k]);
            colmin[jj] = fmin(colmin[jj], acc[k]);
          }
        }
      }
    }
    // block minimum of each row over the query columns
#pragma unroll
    for (int k = 0; k < LB_KC; ++k) {
      const double v = wmin(rmin[k]);
      if (lane == 0) rmin_w[wid * LB_KC + k] = v;
    }
    __syncthreads();
    if (tid < LB_KC && kc + tid < s) {
      double v = rmin_w[tid];
      for (int q = 1; q < LB_T / 64; ++q) v = fmin(v, rmin_w[q * LB_KC + tid]);
      row_part += cw[(size_t)c * s + kc + tid] * v;
    }
    __syncthreads();
  }
  double col_part = 0.0;
#pragma unroll
  for (int jj = 0; jj < LB_JPT; ++jj) {
    const int j = tid + jj * LB_T;
    if (j < m) col_part += w[off + j] * colmin[jj];
  }
  col_part = wsum(col_part);
  row_part = wsum(row_part);
  if (lane == 0) {
    red[wid] = col_part;
    red[4 + wid] = row_part;
  }
  __syncthreads();
  if (tid == 0) {
    const double lc = red[0] + red[1] + red[2] + red[3];
    const double lr = red[4] + red[5] + red[6] + red[7];
    const size_t o = (size_t)r * K + c;
    lb[o] = dd + fmax(fmax(lc, lr), 0.0);
    vals[o] = 0.0;
  }
}

// first round: centroids whose bound is within tau of the smallest bound
// (mask = 1; the smallest-bound centroid is always in)
__global__ void k_srv_round1(int R, int K, double tau, const double *__restrict__ lb,
                             unsigned char *mask) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const double *l = lb + (size_t)r * K;
  double mn = l[0];
  for (int c = 1; c < K; ++c) mn = fmin(mn, l[c]);
  const double cut = mn * tau;
  for (int c = 0; c < K; ++c) mask[(size_t)r * K + c] = (l[c] <= cut || l[c] == mn) ? 1 : 0;
}

// best solved centroid so far (lowest index on ties) over mask != 0
__device__ __forceinline__ void best_solved(const double *v, const unsigned char *mk, int K,
                                            int &best, double &bv) {
  best = -1;
  bv = DBL_MAX;
  for (int c = 0; c < K; ++c)
    if (mk[c] && (v[c] < bv || best < 0)) { bv = v[c]; best = c; }
}

// second round: unsolved centroids the bounds cannot rule out (mask = 2)
__global__ void k_srv_round2(int R, int K, const double *__restrict__ lb,
                             const double *__restrict__ vals, const double *__restrict__ cc,
                             unsigned char *mask, unsigned char *mask2) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const size_t o = (size_t)r * K;
  int f;
  double u2;
  best_solved(vals + o, mask + o, K, f, u2);
  u2 = fmax(u2, 0.0);  // W2^2 to the best first-round centroid
  const double u = sqrt(u2);
  for (int c = 0; c < K; ++c) {
    unsigned char m2 = 0;
    if (!mask[o + c]) {
      bool out = lb[o + c] > u2 * (1.0 + EXCL_MARGIN) + 1e-300;
      if (cc) out = out || cc[(size_t)f * K + c] > 2.0 * u * (1.0 + EXCL_MARGIN);
      m2 = out ? 0 : 1;
    }
    mask2[o + c] = m2;
  }
}

__global__ void k_srv_final(int R, int K, const unsigned char *__restrict__ mask,
                            const unsigned char *__restrict__ mask2,
                            const double *__restrict__ vals, int *label, double *dist,
                            unsigned long long *nlp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const size_t o = (size_t)r * K;
  int best = -1, cnt = 0;
  double bv = DBL_MAX;
  for (int c = 0; c < K; ++c) {
    if (!mask[o + c] && !mask2[o + c]) continue;
    ++cnt;
    const double v = vals[o + c];
    if (best < 0 || v < bv) { bv = v; best = c; }
  }
  label[r] = best;
  dist[r] = bv <= 0.0 ? 0.0 : sqrt(bv);
  atomicAdd(nlp, (unsigned long long)cnt);
}

// The jobs (query, centroid) a round solves, compacted in job order into
// jobs[0 .. *count): each then gets a workgroup of its own instead of the
// launch striding over every (query, centroid) pair, where workgroups that
// drew two live jobs solved them one after the other.  One workgroup.
constexpr int CP_T = 1024;
__global__ __launch_bounds__(CP_T) void k_srv_compact(const unsigned char *__restrict__ mask,
                                                      int n, int *__restrict__ jobs,
                                                      int *__restrict__ count) {
  __shared__ int sums[CP_T];
  const int t = threadIdx.x, per = (n + CP_T - 1) / CP_T;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int c = 0;
  for (int i = lo; i < hi; ++i) c += mask[i] != 0;
  sums[t] = c;
  __syncthreads();
  for (int off = 1; off < CP_T; off <<= 1) {  // inclusive scan (Hillis-Steele)
    const int v = t >= off ? sums[t - off] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int o = sums[t] - c;
  for (int i = lo; i < hi; ++i)
    if (mask[i]) jobs[o++] = i;
  if (t == CP_T - 1) *count = sums[t];
}

__global__ void k_srv_zero_status(int *st, unsigned long long *nlp) {
  if (threadIdx.x == 0) { st[0] = 0; *nlp = 0; }
}

inline size_t a256(size_t x) { return (x + 255) & ~size_t(255); }
double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ------------------------------------------------------------------ server

struct SrvQuery {
  int m = 0;
  int cls = 0;
  std::vector<double> w, supp;
  double t_submit = 0.0, t_done = 0.0;
  int state = 0;  // 0 queued, 1 in flight, 2 done
  int label = -1;
  double dist = 0.0;
};

struct SrvGroup {
  bool busy = false;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<long long> tickets;
  size_t arena_off = 0, arena_bytes = 0;
  // pinned staging: inputs (p_str, cum, w, supp) and results
  void *h_in = nullptr;
  size_t h_in_cap = 0;
  void *h_out = nullptr;
  size_t h_out_cap = 0;
  int R = 0;
};

struct d2mi_server {
  d2mi_ctx *ctx = nullptr;
  int phase = 0, K = 0, s = 0, d = 0;
  d2mi_serve_opts opt{};
  double opt_tau = 1.1;
  void *arena = nullptr;
  std::map<size_t, size_t> free_blocks;  // offset -> length
  DevMem cc, cs_half;                    // centroid pair W2 [K*K], spare
  std::vector<SrvGroup> slots;
  std::deque<long long> queue;
  std::unordered_map<long long, SrvQuery> q;
  long long next_ticket = 1;
  d2mi_serve_stats st{};
};

namespace {

int arena_alloc(d2mi_server *S, size_t bytes, size_t *off) {
  for (auto it = S->free_blocks.begin(); it != S->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t roff = it->first + bytes;
      S->free_blocks.erase(it);
      if (rest) S->free_blocks[roff] = rest;
      return 1;
    }
  }
  return 0;
}

void arena_free(d2mi_server *S, size_t off, size_t bytes) {
  auto it = S->free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != S->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    S->free_blocks.erase(nx);
  }
  if (it != S->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      S->free_blocks.erase(it);
    }
  }
}

int size_class(int m) { return m <= 64 ? 0 : (m <= 192 ? 1 : 2); }

// With at least three streams, each size class has one of its own (stream 0
// the largest queries, 1 the middle class, 2 the small ones), so no class can
// be locked out by another; streams 3 and up are shared and go to the class of
// the oldest waiting query.  A group only holds queries of one class (its
// latency is its slowest LP), and a stream takes a new group only when its
// last one is done, so a group never queues behind another on its stream.
// Admission tries a class's own stream before the shared ones.
bool slot_serves(const d2mi_server *S, int slot, int cls) {
  const int n = (int)S->slots.size();
  if (n < 3 || slot >= 3) return true;
  return slot == 2 - cls;
}

// device bytes of a group of R queries with `col` support points in total
struct GroupLayout {
  size_t p_str, cum, w, supp, in_end;                            // inputs (one H2D copy)
  size_t lb, vals, mask, mask2, jobs, njob, label, dist, nlp, status, out_end;  // outputs from `label`
  size_t total;
  GroupLayout(int R, size_t col, int d, int K) {
    size_t o = 0;
    p_str = o; o = a256(o + sizeof(int) * R);
    cum = o; o = a256(o + sizeof(long long) * R);
    w = o; o = a256(o + sizeof(double) * col);
    supp = o; o = a256(o + sizeof(double) * col * d);
    in_end = o;
    lb = o; o = a256(o + sizeof(double) * (size_t)R * K);
    vals = o; o = a256(o + sizeof(double) * (size_t)R * K);
    mask = o; o = a256(o + (size_t)R * K);
    mask2 = o; o = a256(o + (size_t)R * K);
    jobs = o; o = a256(o + sizeof(int) * (size_t)R * K);
    njob = o; o = a256(o + sizeof(int) * 2);
    label = o; o = a256(o + sizeof(int) * R);
    dist = o; o = a256(o + sizeof(double) * R);
    nlp = o; o = a256(o + sizeof(unsigned long long));
    status = o; o = a256(o + sizeof(int) * 4);
    out_end = o;
    total = o;
  }
};

size_t query_bytes(int m, int d, int K) {
  // share of one query in a group (bounds, values, mask, per-query scalars)
  return sizeof(double) * (size_t)m * (d + 1) + (sizeof(double) * 2 + 2) * (size_t)K + 64;
}

int launch_lp_round(d2mi_server *S, SrvGroup &G, const GroupLayout &L, char *base, int round) {
  d2mi_ctx *ctx = S->ctx;
  PhaseDev &P = ctx->ph[S->phase];
  const int R = G.R;
  // the group is one size class; queries are in submission order
  int m_max = 0;
  for (long long t : G.tickets) m_max = std::max(m_max, S->q[t].m);
  LpArgs a{};
  a.cost_mode = LP_COST_EUCLID;
  a.d = S->d;
  a.sA = S->s;
  a.A_supp = P.cs64.as<double>();
  a.A_w = P.cw64.as<double>();
  a.b_is_table = 0;
  a.p_str = (const int *)(base + L.p_str);
  a.cum = (const long long *)(base + L.cum);
  a.w = (const double *)(base + L.w);
  a.supp = (const double *)(base + L.supp);
  a.n_max = S->s;
  a.m_max = m_max;
  a.status = (int *)(base + L.status);
  a.stream = G.stream;
  // threads per LP: a basis tree of at most 256 nodes is jumped by one wave,
  // so beyond ~2k cells more waves only speed up pricing (8 waves); larger
  // trees jump with every wave and a workgroup barrier per round, where 4
  // waves cost least.  Measured at K = 64 x 48 points, d = 64, 1000 q/s:
  // p50 by class 1.8 / 4.0 / 10.3 ms with 64-128 / 128-256 / 256 threads,
  // 1.5 / 3.7 / 10.3 ms with 128 / 512 / 256.
  const size_t cells = (size_t)S->s * m_max;
  a.threads_hint = S->s + m_max > 256 ? 256 : (cells >= 2048 ? 512 : 128);
  a.njobs = R * S->K;
  a.implicit_K = S->K;
  a.mask = (const unsigned char *)(base + (round == 1 ? L.mask : L.mask2));
  int *jobs = (int *)(base + L.jobs), *njob = (int *)(base + L.njob) + (round - 1);
  k_srv_compact<<<1, CP_T, 0, G.stream>>>(a.mask, a.njobs, jobs, njob);
  D2MI_HIP(hipGetLastError());
  a.job_map = jobs;
  a.njobs_dev = njob;
  a.out = (double *)(base + L.vals);
  return d2mi_launch_transport(ctx, a);
}

int launch_group(d2mi_server *S, SrvGroup &G) {
  d2mi_ctx *ctx = S->ctx;
  PhaseDev &P = ctx->ph[S->phase];
  const int R = G.R, K = S->K, d = S->d, s = S->s;
  size_t col = 0;
  int m_max = 0;
  for (long long t : G.tickets) {
    col += S->q[t].m;
    m_max = std::max(m_max, S->q[t].m);
  }
  const GroupLayout L(R, col, d, K);
  char *base = (char *)S->arena + G.arena_off;
  // pack the inputs in the device layout
  char *h = (char *)G.h_in;
  int *hp = (int *)(h + L.p_str);
  long long *hc = (long long *)(h + L.cum);
  double *hw = (double *)(h + L.w), *hs = (double *)(h + L.supp);
  long long o = 0;
  for (int r = 0; r < R; ++r) {
    const SrvQuery &Q = S->q[G.tickets[r]];
    hp[r] = Q.m;
    hc[r] = o;
    memcpy(hw + o, Q.w.data(), sizeof(double) * Q.m);
    memcpy(hs + (size_t)o * d, Q.supp.data(), sizeof(double) * Q.m * d);
    o += Q.m;
  }
  hipStream_t st = G.stream;
  D2MI_HIP(hipMemcpyAsync(base, h, L.in_end, hipMemcpyHostToDevice, st));
  k_srv_zero_status<<<1, 64, 0, st>>>((int *)(base + L.status),
                                      (unsigned long long *)(base + L.nlp));
  const size_t lds = sizeof(double) * ((size_t)s * d + 4 * LB_KC + d + 2 * LB_T);
  k_srv_bounds<<<R * K, LB_T, lds, st>>>(
      K, s, d, P.cs64.as<double>(), P.cw64.as<double>(), (const int *)(base + L.p_str),
      (const long long *)(base + L.cum), (const double *)(base + L.w),
      (const double *)(base + L.supp), (double *)(base + L.lb), (double *)(base + L.vals));
  k_srv_round1<<<(R + 63) / 64, 64, 0, st>>>(R, K, S->opt_tau, (const double *)(base + L.lb),
                                             (unsigned char *)(base + L.mask));
  D2MI_HIP(hipGetLastError());
  D2MI_TRY(launch_lp_round(S, G, L, base, 1));
  k_srv_round2<<<(R + 63) / 64, 64, 0, st>>>(
      R, K, (const double *)(base + L.lb), (const double *)(base + L.vals),
      S->opt.use_triangle ? S->cc.as<double>() : nullptr, (unsigned char *)(base + L.mask),
      (unsigned char *)(base + L.mask2));
  D2MI_HIP(hipGetLastError());
  D2MI_TRY(launch_lp_round(S, G, L, base, 2));
  k_srv_final<<<(R + 63) / 64, 64, 0, st>>>(
      R, K, (const unsigned char *)(base + L.mask), (const unsigned char *)(base + L.mask2),
      (const double *)(base + L.vals), (int *)(base + L.label), (double *)(base + L.dist),
      (unsigned long long *)(base + L.nlp));
  D2MI_HIP(hipGetLastError());
  D2MI_HIP(hipMemcpyAsync(G.h_out, base + L.label, L.out_end - L.label, hipMemcpyDeviceToHost, st));
  D2MI_HIP(hipEventRecord(G.done, st));
  S->st.groups += 1;
  return D2MI_OK;
}

int retire(d2mi_server *S, SrvGroup &G) {
  const int R = G.R;
  size_t col = 0;
  for (long long t : G.tickets) col += S->q[t].m;
  const GroupLayout L(R, col, S->d, S->K);
  const char *h = (const char *)G.h_out;
  const int *lab = (const int *)(h + (L.label - L.label));
  const double *dist = (const double *)(h + (L.dist - L.label));
  const unsigned long long nlp = *(const unsigned long long *)(h + (L.nlp - L.label));
  const int status = *(const int *)(h + (L.status - L.label));
  const double t = now_ms();
  for (int r = 0; r < R; ++r) {
    SrvQuery &Q = S->q[G.tickets[r]];
    Q.label = lab[r];
    Q.dist = dist[r];
    Q.state = 2;
    Q.t_done = t;
    Q.w.clear();
    Q.w.shrink_to_fit();
    Q.supp.clear();
    Q.supp.shrink_to_fit();
  }
  S->st.completed += R;
  S->st.lps += nlp;
  arena_free(S, G.arena_off, G.arena_bytes);
  G.busy = false;
  G.tickets.clear();
  if (status) {
    d2mi_set_error("an assignment LP hit the pivot limit");
    return D2MI_ENUMERIC;
  }
  return D2MI_OK;
}

int admit(d2mi_server *S) {
  for (;;) {
    if (S->queue.empty()) return D2MI_OK;
    // the class of the oldest waiting query that has a free stream of its own
    int slot = -1, cls = -1;
    for (long long t : S->queue) {
      const int c = S->q[t].cls;
      for (size_t i = 0; i < S->slots.size(); ++i)
        if (!S->slots[i].busy && slot_serves(S, (int)i, c)) { slot = (int)i; break; }
      if (slot >= 0) { cls = c; break; }
    }
    if (slot < 0) return D2MI_OK;
    SrvGroup &G = S->slots[slot];
    std::vector<long long> pick;
    size_t col = 0;
    for (long long t : S->queue) {
      const SrvQuery &Q = S->q[t];
      if (Q.cls != cls) continue;
      const GroupLayout L((int)pick.size() + 1, col + Q.m, S->d, S->K);
      if (L.in_end > G.h_in_cap || L.out_end - L.label > G.h_out_cap) break;
      pick.push_back(t);
      col += Q.m;
      if ((int)pick.size() >= S->opt.max_batch) break;
    }
    // shrink to what the arena holds now (admission by HBM footprint)
    size_t off = 0, bytes = 0;
    while (!pick.empty()) {
      bytes = GroupLayout((int)pick.size(), col, S->d, S->K).total;
      if (arena_alloc(S, bytes, &off)) break;
      col -= S->q[pick.back()].m;
      pick.pop_back();
    }
    if (pick.empty()) {
      S->st.memory_waits += 1;
      return D2MI_OK;  // wait for a running group to free memory
    }
    for (long long t : pick) S->q[t].state = 1;
    std::deque<long long> rest;
    for (long long t : S->queue)
      if (S->q[t].state == 0) rest.push_back(t);
    S->queue.swap(rest);
    G.busy = true;
    G.tickets = pick;
    G.R = (int)pick.size();
    G.arena_off = off;
    G.arena_bytes = bytes;
    int inflight = 0;
    for (auto &g : S->slots) inflight += g.busy;
    S->st.max_inflight_groups = std::max(S->st.max_inflight_groups, (long long)inflight);
    S->st.max_group = std::max(S->st.max_group, (long long)G.R);
    const int rc = launch_group(S, G);
    if (rc != D2MI_OK) {
      arena_free(S, off, bytes);
      G.busy = false;
      return rc;
    }
  }
}

int progress_impl(d2mi_server *S) {
  for (auto &G : S->slots) {
    if (!G.busy) continue;
    const hipError_t e = hipEventQuery(G.done);
    if (e == hipErrorNotReady) continue;
    D2MI_HIP(e);
    D2MI_TRY(retire(S, G));
  }
  return admit(S);
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" int d2mi_serve_open(d2mi_ctx *ctx, int phase, const d2mi_serve_opts *opts,
                               d2mi_server **out) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (!out) {
    d2mi_set_error("d2mi_serve_open: null output");
    return D2MI_EINVAL;
  }
  *out = nullptr;
  PhaseDev *P = d2mi_phase(ctx, phase);
  if (!P) return D2MI_EINVAL;
  if (P->metric != D2MI_EUCLIDEAN || P->K <= 0 || !P->cs64.p) {
    d2mi_set_error("the server needs resident Euclidean centroids on phase %d", phase);
    return D2MI_EINVAL;
  }
  if (P->s > 64 * 8 || (size_t)(P->s + 1) * P->dim * sizeof(double) > 48 * 1024) {
    d2mi_set_error("centroids of %d x %d do not fit the bound kernel", P->s, P->dim);
    return D2MI_EUNSUPPORTED;
  }
  d2mi_server *S = new d2mi_server();
  S->ctx = ctx;
  S->phase = phase;
  S->K = P->K;
  S->s = P->s;
  S->d = P->dim;
  d2mi_serve_opts o{};
  o.max_batch = 48;
  o.streams = 6;
  o.hbm_bytes = (size_t)1 << 30;
  o.max_m = 4096;
  o.use_triangle = 1;
  if (opts) {
    if (opts->max_batch > 0) o.max_batch = opts->max_batch;
    if (opts->streams > 0) o.streams = std::min(opts->streams, 16);
    if (opts->hbm_bytes > 0) o.hbm_bytes = opts->hbm_bytes;
    if (opts->max_m > 0) o.max_m = std::min(opts->max_m, LB_T * LB_JPT);
    o.use_triangle = opts->use_triangle;
  }
  o.max_m = std::min(o.max_m, LB_T * LB_JPT);
  S->opt = o;
  int rc = D2MI_OK;
  auto fail = [&](int code) {
    d2mi_serve_close(S);
    return code;
  };
  if (hipMalloc(&S->arena, o.hbm_bytes) != hipSuccess) {
    d2mi_set_error("cannot allocate %zu B of HBM for the server", o.hbm_bytes);
    return fail(D2MI_ENOMEM);
  }
  S->free_blocks[0] = o.hbm_bytes;
  // staging sized for a full group of the largest queries
  const size_t in_cap = std::min(o.hbm_bytes, GroupLayout(o.max_batch, (size_t)o.max_batch * 256,
                                                          S->d, S->K).in_end +
                                                  query_bytes(o.max_m, S->d, S->K));
  S->slots.resize(o.streams);
  for (auto &G : S->slots) {
    if (hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&G.done, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&G.h_in, in_cap, 0) != hipSuccess ||
        hipHostMalloc(&G.h_out, GroupLayout(o.max_batch, 0, S->d, S->K).out_end, 0) != hipSuccess) {
      d2mi_set_error("cannot create the server's streams and staging buffers");
      return fail(D2MI_EHIP);
    }
    G.h_in_cap = in_cap;
    G.h_out_cap = GroupLayout(o.max_batch, 0, S->d, S->K).out_end;
  }
  // centroid pair distances for the triangle test
  if (o.use_triangle) {
    rc = S->cc.ensure(sizeof(double) * (size_t)S->K * S->K);
    if (rc == D2MI_OK) rc = S->cs_half.ensure(sizeof(double) * S->K);
    if (rc == D2MI_OK)
      rc = d2mi_centroid_pairs(ctx, phase, S->K, S->cc.as<double>(), S->cs_half.as<double>());
    if (rc != D2MI_OK) return fail(rc);
  }
  *out = S;
  return D2MI_OK;
}

extern "C" int d2mi_serve_submit(d2mi_server *S, int m, const double *w, const double *supp,
                                 long long *ticket) {
  if (!S || !w || !supp || !ticket || m <= 0) {
    d2mi_set_error("d2mi_serve_submit: bad arguments");
    return D2MI_EINVAL;
  }
  if (m > S->opt.max_m) {
    d2mi_set_error("query of %d support points is above the server's limit of %d", m,
                   S->opt.max_m);
    return D2MI_EUNSUPPORTED;
  }
  if (GroupLayout(1, m, S->d, S->K).total > S->opt.hbm_bytes ||
      GroupLayout(1, m, S->d, S->K).in_end > S->slots[0].h_in_cap) {
    d2mi_set_error("query of %d support points does not fit the server's memory", m);
    return D2MI_ENOMEM;
  }
  // the bounds and the LP both assume unit mass on each side: weights are
  // renormalised as d2_read does for every object it loads
  // (clustering_io.c:143-151), and a weight <= 1e-9 is refused as there
  double wsum = 0.0;
  for (int j = 0; j < m; ++j) {
    if (!(w[j] > 1e-9) || !std::isfinite(w[j])) {
      d2mi_set_error("query weight %d is %g; every weight must be finite and > 1e-9", j, w[j]);
      return D2MI_EINVAL;
    }
    wsum += w[j];
  }
  const long long t = S->next_ticket++;
  SrvQuery &Q = S->q[t];
  Q.m = m;
  Q.cls = size_class(m);
  Q.w.resize(m);
  for (int j = 0; j < m; ++j) Q.w[j] = w[j] / wsum;
  Q.supp.assign(supp, supp + (size_t)m * S->d);
  Q.t_submit = now_ms();
  S->queue.push_back(t);
  S->st.submitted += 1;
  *ticket = t;
  return D2MI_OK;
}

extern "C" int d2mi_serve_progress(d2mi_server *S, int *pending) {
  if (!S) {
    d2mi_set_error("d2mi_serve_progress: null server");
    return D2MI_EINVAL;
  }
  D2MI_TRY(progress_impl(S));
  if (pending) *pending = (int)(S->st.submitted - S->st.completed);
  return D2MI_OK;
}

extern "C" int d2mi_serve_result(d2mi_server *S, long long ticket, int *label, double *dist,
                                 double *latency_ms) {
  if (!S) {
    d2mi_set_error("d2mi_serve_result: null server");
    return D2MI_EINVAL;
  }
  auto it = S->q.find(ticket);
  if (it == S->q.end()) {
    d2mi_set_error("unknown or already collected ticket %lld", ticket);
    return D2MI_EINVAL;
  }
  if (it->second.state != 2) return 0;
  if (label) *label = it->second.label;
  if (dist) *dist = it->second.dist;
  if (latency_ms) *latency_ms = it->second.t_done - it->second.t_submit;
  S->q.erase(it);
  return 1;
}

extern "C" int d2mi_serve_drain(d2mi_server *S) {
  if (!S) {
    d2mi_set_error("d2mi_serve_drain: null server");
    return D2MI_EINVAL;
  }
  while (S->st.completed < S->st.submitted) {
    D2MI_TRY(progress_impl(S));
    bool any = false;
    for (auto &G : S->slots) any |= G.busy;
    if (!any && S->st.completed < S->st.submitted && S->queue.empty()) break;
    if (any) {
      for (auto &G : S->slots)
        if (G.busy) {
          D2MI_HIP(hipEventSynchronize(G.done));
          break;
        }
    }
  }
  return D2MI_OK;
}

extern "C" int d2mi_serve_stats_get(d2mi_server *S, d2mi_serve_stats *st) {
  if (!S || !st) {
    d2mi_set_error("d2mi_serve_stats_get: bad arguments");
    return D2MI_EINVAL;
  }
  *st = S->st;
  return D2MI_OK;
}

extern "C" int d2mi_serve_close(d2mi_server *S) {
  if (!S) return D2MI_OK;
  for (auto &G : S->slots) {
    if (G.busy && G.done) (void)hipEventSynchronize(G.done);
    if (G.done) (void)hipEventDestroy(G.done);
    if (G.stream) (void)hipStreamDestroy(G.stream);
    if (G.h_in) (void)hipHostFree(G.h_in);
    if (G.h_out) (void)hipHostFree(G.h_out);
  }
  if (S->arena) (void)hipFree(S->arena);
  S->cc.release();
  S->cs_half.release();
  delete S;
  return D2MI_OK;
}
