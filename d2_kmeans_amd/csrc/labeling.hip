// Assignment on the device: d2_labeling (brute force, clustering.c:324-376),
// d2_labeling_prep / d2_labeling_post (Elkan bounds, clustering.c:141-318),
// d2_compute_distance in batches (clustering.c:33-128) and the outer loop of
// d2_clustering (clustering.c:409-525).
//
// Distances are exact: sqrt of the phase-summed transport-LP values
// (transport.hip).  Elkan's bounds l[N*K], u[N], s[K], c[K*K], r[N] stay in
// HBM across outer iterations.  The Elkan step runs in two batched rounds
// instead of the reference's per-object sequential scan: round A computes
// d(i, label) for objects whose upper bound is stale, round B computes every
// d(i, j) that the bounds cannot rule out under the tightened u.  The set of
// distances computed is a superset of what the sequential scan needs, so the
// labels are the exact nearest centroids, as in the reference; only the
// distance count can differ.
#include <cfloat>
#include <cmath>
#include <vector>

#include "d2mi_internal.h"
#include "transport.h"

namespace {

constexpr int NT = 256;

enum Pairing { OBJ_CENT = 0, CENT_CENT = 1, OLD_NEW = 2 };

__global__ void k_el_init(double *l, double *u, char *r, int N, int K) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t < (size_t)N * K) l[t] = 0.0;
  if (t < (size_t)N) {
    u[t] = DBL_MAX;  // clustering_util.c:226-227
    r[t] = 1;
  }
}

// centroid pair distances -> c (symmetric) and s[k] = min_{j != k} c/2
__global__ void k_cc_fill(const double *__restrict__ d, int K, double *c, double *s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double sk = DBL_MAX;
  for (int j = 0; j < K; ++j) {
    // job (row = max, col = min) holds d(centroid min, centroid max)
    const int hi = k > j ? k : j, lo = k > j ? j : k;
    const double v = k == j ? 0.0 : d[(size_t)lo * K + hi];
    c[(size_t)k * K + j] = v;
    if (j != k && sk > v / 2.0) sk = v / 2.0;  // clustering.c:169-170
  }
  s[k] = sk;
}

__global__ void k_cc_mask(unsigned char *mask, int K) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < K * K) mask[t] = (t % K) > (t / K);  // row centroid > column centroid
}

// Elkan round A (clustering.c:183-209): objects that need work; stale upper
// bounds are tightened only when some centroid could still be closer.
__global__ void k_el_a(int N, int K, const int *__restrict__ label, const double *__restrict__ u,
                       const double *__restrict__ s, const double *__restrict__ l,
                       const double *__restrict__ c, const char *__restrict__ r, char *lsw,
                       char *need, unsigned char *mask, double *dist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  lsw[i] = 0;  // :179-181
  for (int j = 0; j < K; ++j) {
    mask[(size_t)i * K + j] = 0;
    dist[(size_t)i * K + j] = 0.0;
  }
  const int lab = label[i];
  const bool nd = lab < 0 || u[i] > s[lab];  // :185
  need[i] = nd;
  if (!nd || r[i] != 1) return;
  const int jj = lab >= 0 ? lab : 0;
  const double U = u[i];
  for (int j = 0; j < K; ++j)
    if ((j != jj || lab < 0) && U > l[(size_t)i * K + j] && U > c[(size_t)j * K + jj] / 2.0) {
      mask[(size_t)i * K + jj] = 1;  // 3a: exact d(i, label)
      return;
    }
}

// Elkan round B (:211-219): every centroid the bounds cannot exclude.
__global__ void k_el_b(int N, int K, const int *__restrict__ label, double *u, double *l,
                       const double *__restrict__ c, char *r, const char *__restrict__ need,
                       unsigned char *mask, double *dist, unsigned long long *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !need[i]) return;
  const int lab = label[i];
  const int jj = lab >= 0 ? lab : 0;
  unsigned long long dc = 0;
  if (mask[(size_t)i * K + jj]) {
    const double d = dist[(size_t)i * K + jj];
    l[(size_t)i * K + jj] = d;
    u[i] = d;
    r[i] = 0;
    mask[(size_t)i * K + jj] = 0;
    dc = 1;
  }
  const double U = u[i];
  for (int j = 0; j < K; ++j) {
    const bool m = j != jj && U > l[(size_t)i * K + j] && U > c[(size_t)j * K + jj] / 2.0;
    mask[(size_t)i * K + j] = m;
    if (m) dist[(size_t)i * K + j] = 0.0;
  }
  if (dc) atomicAdd(&cnt[1], dc);
}

// Elkan finish (:212-227): closest computed centroid, strict < keeps the
// current label on ties.
__global__ void k_el_c(int N, int K, int *label, double *u, double *l,
                       const char *__restrict__ need, const unsigned char *__restrict__ mask,
                       const double *__restrict__ dist, char *lsw, unsigned long long *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !need[i]) return;
  const int init = label[i];
  int bj = init >= 0 ? init : 0;
  double best = u[i];
  unsigned long long dc = 0;
  for (int j = 0; j < K; ++j)
    if (mask[(size_t)i * K + j]) {
      const double d = dist[(size_t)i * K + j];
      l[(size_t)i * K + j] = d;
      ++dc;
      if (d < best) {
        best = d;
        bj = j;
      }
    }
  u[i] = best;
  if (bj != init) {
    label[i] = bj;
    lsw[i] = 1;
    atomicAdd(&cnt[0], 1ULL);
  }
  if (dc) atomicAdd(&cnt[1], dc);
}

// d2_labeling's selection (:336-361): first minimum in centroid order
__global__ void k_plain_select(int N, int K, int *label, const double *__restrict__ dist,
                               char *lsw, double *mind2, unsigned long long *cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double mind = -1.0;
  int jj = label[i] >= 0 ? label[i] : 0;
  for (int j = 0; j < K; ++j) {
    const double d = dist[(size_t)i * K + j];
    if (mind < 0 || d < mind) {
      mind = d;
      jj = j;
    }
  }
  mind2[i] = mind * mind;  // cost += min_distance^2 (:348)
  if (label[i] == jj) {
    lsw[i] = 0;
  } else {
    label[i] = jj;
    lsw[i] = 1;
    atomicAdd(&cnt[0], 1ULL);
  }
}

// fixed-order sum of n doubles with one block
__global__ void k_sum_fixed(const double *__restrict__ a, size_t n, double *out) {
  __shared__ double red[1024];
  double v = 0.0;
  for (size_t t = threadIdx.x; t < n; t += blockDim.x) v += a[t];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// d2_labeling_post (:303-314)
__global__ void k_el_post(int N, int K, const int *__restrict__ label,
                          const double *__restrict__ dch, double *l, double *u, char *r) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t < (size_t)N * K) {
    const double v = l[t] - dch[t % K];
    l[t] = v > 0 ? v : 0;
  }
  if (t < (size_t)N) {
    u[t] += dch[label[t]];
    r[t] = 1;
  }
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_fill_labels(int *label, int N, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) label[i] = v;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + NT - 1) / NT); }

int common_K(d2mi_ctx *ctx, int selected_phase, int *K) {
  int k = -1;
  for (size_t p = 0; p < ctx->ph.size(); ++p) {
    if (!(selected_phase < 0 || (int)p == selected_phase)) continue;
    const PhaseDev &P = ctx->ph[p];
    if (!P.set || !P.K) {
      d2mi_set_error("phase %zu has no data or centroids", p);
      return D2MI_EINVAL;
    }
    if (k >= 0 && P.K != k) {
      d2mi_set_error("phases disagree on K");
      return D2MI_EINVAL;
    }
    k = P.K;
  }
  if (k <= 0) {
    d2mi_set_error("no phase selected");
    return D2MI_EINVAL;
  }
  *K = k;
  return D2MI_OK;
}

// Sum over the selected phases of the LP values of `njobs` pairs into out
// (zeroed here), then sqrt (clustering.c:39-127).  Job -> pair: explicit
// arrays, or implicit (job / K, job % K) when implicit_K > 0.
// Jobs o*K + c grouped by the support count of object o (power-of-two
// buckets), built once per (phase, K) and kept on the device.
int ensure_buckets(d2mi_ctx *ctx, PhaseDev &P, int K) {
  if (P.bucket_K == K) return D2MI_OK;
  const int N = (int)P.h_pstr.size();
  std::vector<std::vector<int>> objs;
  std::vector<int> mmax;
  for (int o = 0; o < N; ++o) {
    int b = 0;
    while ((32 << b) < P.h_pstr[o] && b < 10) ++b;
    if ((int)objs.size() <= b) { objs.resize(b + 1); mmax.resize(b + 1, 0); }
    objs[b].push_back(o);
    mmax[b] = std::max(mmax[b], P.h_pstr[o]);
  }
  std::vector<int> map;
  map.reserve((size_t)N * K);
  P.bucket_off.assign(1, 0);
  P.bucket_mmax.clear();
  for (size_t b = 0; b < objs.size(); ++b) {
    if (objs[b].empty()) continue;
    for (int o : objs[b])
      for (int c = 0; c < K; ++c) map.push_back(o * K + c);
    P.bucket_off.push_back(map.size());
    P.bucket_mmax.push_back(mmax[b]);
  }
  D2MI_TRY(P.bucket_map.assign(map.data(), sizeof(int) * map.size(), ctx->stream));
  P.bucket_K = K;
  return D2MI_OK;
}

int lp_over_phases(d2mi_ctx *ctx, int selected_phase, Pairing pairing, int njobs, int implicit_K,
                   const unsigned char *mask, const int *job_a, const int *job_b, double *out) {
  D2MI_TRY(d2mi_zero_doubles(ctx, out, njobs));
  D2MI_TRY(ctx->lpstat.ensure(sizeof(int) * 4));
  D2MI_TRY(d2mi_zero_ints(ctx, ctx->lpstat.as<int>(), 1));
  for (size_t p = 0; p < ctx->ph.size(); ++p) {
    if (!(selected_phase < 0 || (int)p == selected_phase)) continue;
    PhaseDev &P = ctx->ph[p];
    LpArgs a{};
    a.njobs = njobs;
    a.job_a = job_a;
    a.job_b = job_b;
    a.implicit_K = implicit_K;
    a.mask = mask;
    a.V = P.V;
    a.dist = P.dist64.as<double>();
    a.out = out;
    a.status = ctx->lpstat.as<int>();
    switch (P.metric) {
      case D2MI_EUCLIDEAN:
        a.cost_mode = LP_COST_EUCLID;
        a.d = P.dim;
        break;
      case D2MI_HISTOGRAM:
        a.cost_mode = LP_COST_HIST;
        break;
      case D2MI_SPARSE_HISTOGRAM:
        a.cost_mode = LP_COST_SPARSE;
        break;
      default:  // word embeddings (clustering.c:63-80, _dpdist2_sym)
        a.cost_mode = LP_COST_WORD;
        a.d = P.dim;
        a.vocab = P.vocab64.as<double>();
        break;
    }
    const bool euclid = P.metric == D2MI_EUCLIDEAN || P.metric == D2MI_WORD_EMBED;
    a.sA = P.s;
    a.A_supp = euclid ? (pairing == OLD_NEW ? P.cs_old.as<double>() : P.cs64.as<double>()) : nullptr;
    a.A_w = pairing == OLD_NEW ? P.cw_old.as<double>() : P.cw64.as<double>();
    a.n_max = P.s;
    if (pairing == OBJ_CENT) {
      a.b_is_table = 0;
      a.p_str = P.p_str.as<int>();
      a.cum = P.cum.as<long long>();
      a.w = P.w64.as<double>();
      a.supp = P.metric == D2MI_EUCLIDEAN ? P.supp64.as<double>() : nullptr;
      a.sym = (P.metric == D2MI_SPARSE_HISTOGRAM || P.metric == D2MI_WORD_EMBED)
                  ? P.sym.as<int>() : nullptr;
      a.m_max = P.m_max;
      if (implicit_K > 0 && njobs == (int)P.N * implicit_K && P.metric != D2MI_HISTOGRAM) {
        // one launch per size bucket, LDS and workgroup size fitted to it
        D2MI_TRY(ensure_buckets(ctx, P, implicit_K));
        for (size_t b = 0; b + 1 < P.bucket_off.size(); ++b) {
          LpArgs ab = a;
          ab.njobs = (int)(P.bucket_off[b + 1] - P.bucket_off[b]);
          ab.job_map = P.bucket_map.as<int>() + P.bucket_off[b];
          ab.m_max = P.bucket_mmax[b];
          D2MI_TRY(d2mi_launch_transport(ctx, ab));
        }
        continue;
      }
    } else {
      a.b_is_table = 1;
      a.sB = P.s;
      a.B_supp = euclid ? P.cs64.as<double>() : nullptr;
      a.B_w = P.cw64.as<double>();
      a.m_max = P.s;
    }
    D2MI_TRY(d2mi_launch_transport(ctx, a));
  }
  D2MI_TRY(d2mi_sqrt_clamp(ctx, out, njobs));
  return D2MI_OK;
}

int lp_status(d2mi_ctx *ctx) {
  int st = 0;
  D2MI_HIP(hipMemcpyAsync(&st, ctx->lpstat.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  if (st) {
    d2mi_set_error("a transport LP hit the pivot limit");
    return D2MI_ENUMERIC;
  }
  return D2MI_OK;
}

int ensure_label_state(d2mi_ctx *ctx, int K) {
  const size_t N = ctx->N, NK = N * (size_t)K;
  if (!ctx->label.p) {
    d2mi_set_error("labels are not set");
    return D2MI_EINVAL;
  }
  D2MI_TRY(ctx->dist.ensure(sizeof(double) * std::max(NK, (size_t)K * K)));
  D2MI_TRY(ctx->mask.ensure(std::max(NK, (size_t)K * K)));
  D2MI_TRY(ctx->cnt.ensure(sizeof(unsigned long long) * 4));
  D2MI_TRY(ctx->mind2.ensure(sizeof(double) * (N + 8)));
  D2MI_TRY(ctx->red.ensure(sizeof(double) * 4 + sizeof(char) * N));
  D2MI_TRY(ctx->jobs.ensure(sizeof(int) * 2 * (size_t)K));
  return D2MI_OK;
}

int elkan_reset_impl(d2mi_ctx *ctx, int K) {
  const size_t N = ctx->N;
  D2MI_TRY(ctx->el_l.ensure(sizeof(double) * N * K));
  D2MI_TRY(ctx->el_u.ensure(sizeof(double) * N));
  D2MI_TRY(ctx->el_s.ensure(sizeof(double) * K));
  D2MI_TRY(ctx->el_c.ensure(sizeof(double) * K * K));
  D2MI_TRY(ctx->el_r.ensure(N));
  k_el_init<<<nblk(N * K), NT, 0, ctx->stream>>>(ctx->el_l.as<double>(), ctx->el_u.as<double>(),
                                                  ctx->el_r.as<char>(), (int)N, K);
  D2MI_HIP(hipGetLastError());
  ctx->elkan_K = K;
  return D2MI_OK;
}

}  // namespace

// Exact W2 between every pair of resident centroids of the selected phases:
// c_dev[a*K + b] (symmetric, zero diagonal), on the context stream.
int d2mi_centroid_pairs(d2mi_ctx *ctx, int selected_phase, int K, double *c_dev, double *s_dev) {
  D2MI_TRY(ctx->dist.ensure(sizeof(double) * (size_t)K * K));
  D2MI_TRY(ctx->mask.ensure((size_t)K * K));
  hipStream_t S = ctx->stream;
  unsigned char *mask = ctx->mask.as<unsigned char>();
  double *dist = ctx->dist.as<double>();
  k_cc_mask<<<nblk((size_t)K * K), NT, 0, S>>>(mask, K);
  D2MI_TRY(lp_over_phases(ctx, selected_phase, CENT_CENT, K * K, K, mask, nullptr, nullptr, dist));
  k_cc_fill<<<nblk(K), NT, 0, S>>>(dist, K, c_dev, s_dev);
  D2MI_HIP(hipGetLastError());
  return lp_status(ctx);
}

// ------------------------------------------------------------------ C ABI

extern "C" int d2mi_distances(d2mi_ctx *ctx, int selected_phase, size_t npairs, const int *obj,
                              const int *cent, double *out) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  int K;
  D2MI_TRY(common_K(ctx, selected_phase, &K));
  if (!obj || !cent || !out || npairs == 0 || npairs > (size_t)0x7fffffff) {
    d2mi_set_error("d2mi_distances: bad arguments");
    return D2MI_EINVAL;
  }
  for (size_t t = 0; t < npairs; ++t)
    if (obj[t] < 0 || (size_t)obj[t] >= ctx->N || cent[t] < 0 || cent[t] >= K) {
      d2mi_set_error("pair %zu = (%d, %d) out of range", t, obj[t], cent[t]);
      return D2MI_EINVAL;
    }
  DevMem ja, jb, o;
  D2MI_TRY(ja.assign(cent, sizeof(int) * npairs, ctx->stream));
  D2MI_TRY(jb.assign(obj, sizeof(int) * npairs, ctx->stream));
  D2MI_TRY(o.ensure(sizeof(double) * npairs));
  D2MI_TRY(lp_over_phases(ctx, selected_phase, OBJ_CENT, (int)npairs, 0, nullptr, ja.as<int>(),
                          jb.as<int>(), o.as<double>()));
  D2MI_HIP(hipMemcpyAsync(out, o.p, sizeof(double) * npairs, hipMemcpyDeviceToHost, ctx->stream));
  return lp_status(ctx);
}

extern "C" int d2mi_elkan_reset(d2mi_ctx *ctx, int selected_phase) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  int K;
  D2MI_TRY(common_K(ctx, selected_phase, &K));
  D2MI_TRY(elkan_reset_impl(ctx, K));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

extern "C" int d2mi_labeling(d2mi_ctx *ctx, int selected_phase, int use_elkan,
                             d2mi_label_stats *st) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  int K;
  D2MI_TRY(common_K(ctx, selected_phase, &K));
  D2MI_TRY(ensure_label_state(ctx, K));
  hipStream_t S = ctx->stream;
  const int N = (int)ctx->N;
  if ((size_t)N * K > (size_t)0x7fffffff) {
    d2mi_set_error("N*K too large");
    return D2MI_EUNSUPPORTED;
  }
  unsigned long long *cnt = ctx->cnt.as<unsigned long long>();
  double *dist = ctx->dist.as<double>();
  unsigned char *mask = ctx->mask.as<unsigned char>();
  D2MI_HIP(hipEventRecord(ctx->ev[2], S));
  k_zero_u64<<<1, 64, 0, S>>>(cnt, 4);
  double cost = NAN;
  if (!use_elkan) {
    D2MI_TRY(lp_over_phases(ctx, selected_phase, OBJ_CENT, N * K, K, nullptr, nullptr, nullptr,
                            dist));
    k_plain_select<<<nblk(N), NT, 0, S>>>(N, K, ctx->label.as<int>(), dist, ctx->lsw.as<char>(),
                                          ctx->mind2.as<double>(), cnt);
    k_sum_fixed<<<1, 1024, 0, S>>>(ctx->mind2.as<double>(), N, ctx->red.as<double>());
    D2MI_HIP(hipGetLastError());
  } else {
    if (ctx->elkan_K != K) D2MI_TRY(elkan_reset_impl(ctx, K));
    double *c = ctx->el_c.as<double>(), *s = ctx->el_s.as<double>();
    double *l = ctx->el_l.as<double>(), *u = ctx->el_u.as<double>();
    char *r = ctx->el_r.as<char>();
    char *need = ctx->red.as<char>() + 4 * sizeof(double);
    // step 1 (:154-172): centroid pair distances, d(lo, hi) at [lo*K + hi]
    k_cc_mask<<<nblk((size_t)K * K), NT, 0, S>>>(mask, K);
    D2MI_TRY(lp_over_phases(ctx, selected_phase, CENT_CENT, K * K, K, mask, nullptr, nullptr,
                            dist));
    k_cc_fill<<<nblk(K), NT, 0, S>>>(dist, K, c, s);
    D2MI_HIP(hipGetLastError());
    // rounds A, B and the selection
    k_el_a<<<nblk(N), NT, 0, S>>>(N, K, ctx->label.as<int>(), u, s, l, c, r, ctx->lsw.as<char>(),
                                  need, mask, dist);
    D2MI_TRY(lp_over_phases(ctx, selected_phase, OBJ_CENT, N * K, K, mask, nullptr, nullptr,
                            dist));
    k_el_b<<<nblk(N), NT, 0, S>>>(N, K, ctx->label.as<int>(), u, l, c, r, need, mask, dist, cnt);
    D2MI_TRY(lp_over_phases(ctx, selected_phase, OBJ_CENT, N * K, K, mask, nullptr, nullptr,
                            dist));
    k_el_c<<<nblk(N), NT, 0, S>>>(N, K, ctx->label.as<int>(), u, l, need, mask, dist,
                                  ctx->lsw.as<char>(), cnt);
    D2MI_HIP(hipGetLastError());
  }
  D2MI_HIP(hipEventRecord(ctx->ev[3], S));
  unsigned long long hc[2] = {0, 0};
  D2MI_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, S));
  if (!use_elkan) D2MI_HIP(hipMemcpyAsync(&cost, ctx->red.p, sizeof(double), hipMemcpyDeviceToHost, S));
  D2MI_TRY(lp_status(ctx));
  if (st) {
    float ms = 0.f;
    D2MI_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    st->changes = hc[0];
    st->dist_count = use_elkan ? hc[1] + (unsigned long long)K * (K - 1) / 2 : (unsigned long long)N * K;
    st->cost = cost;
    st->ms = ms;
  }
  return D2MI_OK;
}

extern "C" int d2mi_save_centroids(d2mi_ctx *ctx, int selected_phase) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  for (size_t p = 0; p < ctx->ph.size(); ++p) {
    if (!(selected_phase < 0 || (int)p == selected_phase)) continue;
    PhaseDev &P = ctx->ph[p];
    if (!P.K) continue;
    const size_t ks = (size_t)P.K * P.s;
    D2MI_TRY(P.cw_old.ensure(sizeof(double) * ks));
    D2MI_HIP(hipMemcpyAsync(P.cw_old.p, P.cw64.p, sizeof(double) * ks, hipMemcpyDeviceToDevice,
                            ctx->stream));
    if (P.cs64.p) {
      D2MI_TRY(P.cs_old.ensure(sizeof(double) * ks * P.dim));
      D2MI_HIP(hipMemcpyAsync(P.cs_old.p, P.cs64.p, sizeof(double) * ks * P.dim,
                              hipMemcpyDeviceToDevice, ctx->stream));
    }
  }
  return D2MI_OK;
}

extern "C" int d2mi_elkan_post(d2mi_ctx *ctx, int selected_phase) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  int K;
  D2MI_TRY(common_K(ctx, selected_phase, &K));
  if (ctx->elkan_K != K) {
    d2mi_set_error("Elkan bounds are not initialised");
    return D2MI_EINVAL;
  }
  for (size_t p = 0; p < ctx->ph.size(); ++p)
    if ((selected_phase < 0 || (int)p == selected_phase) && !ctx->ph[p].cw_old.p) {
      d2mi_set_error("d2mi_save_centroids was not called before the update");
      return D2MI_EINVAL;
    }
  D2MI_TRY(ensure_label_state(ctx, K));
  std::vector<int> idx(K);
  for (int k = 0; k < K; ++k) idx[k] = k;
  D2MI_TRY(ctx->jobs.assign(idx.data(), sizeof(int) * K, ctx->stream));
  double *dch = ctx->dist.as<double>();
  // a = c_new(k) (columns), b = c_old(k) (rows): clustering.c:305
  D2MI_TRY(lp_over_phases(ctx, selected_phase, OLD_NEW, K, 0, nullptr, ctx->jobs.as<int>(),
                          ctx->jobs.as<int>(), dch));
  const size_t N = ctx->N;
  k_el_post<<<nblk(N * K), NT, 0, ctx->stream>>>((int)N, K, ctx->label.as<int>(), dch,
                                                  ctx->el_l.as<double>(), ctx->el_u.as<double>(),
                                                  ctx->el_r.as<char>());
  D2MI_HIP(hipGetLastError());
  return lp_status(ctx);
}

extern "C" int d2mi_clustering(d2mi_ctx *ctx, int K, int max_iter, int selected_phase,
                               int use_triangle, int maxIters, double rhoCoeff,
                               int updatePerLoops, int precision, d2mi_cluster_stats *st) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  int k0;
  D2MI_TRY(common_K(ctx, selected_phase, &k0));
  if (k0 != K || max_iter <= 0) {
    d2mi_set_error("d2mi_clustering: centroids have K=%d, asked %d", k0, K);
    return D2MI_EINVAL;
  }
  const int N = (int)ctx->N;
  D2MI_TRY(ctx->label.ensure(sizeof(int) * ctx->N));
  D2MI_TRY(ctx->lsw.ensure(ctx->N));
  k_fill_labels<<<nblk(N), NT, 0, ctx->stream>>>(ctx->label.as<int>(), N, -1);  // :117-118
  D2MI_HIP(hipGetLastError());
  if (use_triangle) D2MI_TRY(elkan_reset_impl(ctx, K));
  if (st) *st = d2mi_cluster_stats{};
  int iter = 0;
  double ms_lab = 0.0, ms_upd = 0.0;
  for (iter = 0; iter < max_iter; ++iter) {  // :455
    d2mi_label_stats ls{};
    D2MI_TRY(d2mi_labeling(ctx, selected_phase, use_triangle, &ls));
    ms_lab += ls.ms;
    if (st && iter < 256) {
      st->changes[iter] = (long long)ls.changes;
      st->dist_count[iter] = (long long)ls.dist_count;
    }
    if (!use_triangle && st) st->final_cost = ls.cost / N;
    if ((double)ls.changes < 0.001 * N) {  // :476
      ++iter;
      break;
    }
    if (use_triangle) D2MI_TRY(d2mi_save_centroids(ctx, selected_phase));
    for (size_t p = 0; p < ctx->ph.size(); ++p) {
      if (!(selected_phase < 0 || (int)p == selected_phase)) continue;
      d2mi_badmm_stats bs{};
      D2MI_TRY(d2mi_badmm(ctx, (int)p, maxIters, rhoCoeff, updatePerLoops, precision,
                          D2MI_BADMM_ENGINE, &bs));
      ms_upd += bs.ms_total;
    }
    if (use_triangle) D2MI_TRY(d2mi_elkan_post(ctx, selected_phase));
  }
  if (use_triangle) {  // :509
    d2mi_label_stats ls{};
    D2MI_TRY(d2mi_labeling(ctx, selected_phase, 0, &ls));
    ms_lab += ls.ms;
    if (st) st->final_cost = ls.cost / N;
  }
  if (st) {
    st->iterations = iter;
    st->ms_labeling = ms_lab;
    st->ms_badmm = ms_upd;
  }
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}
