// Exact transport LP values on the device: the problem the reference hands to
// Mosek's network simplex for every distance (d2_match_by_distmat,
// solver_mosek.cc:61-143):
//     min sum_{k,j} C[k + j*n] x_kj,  x >= 0,
//     sum_j x_kj = wA[k] (all n rows),  sum_k x_kj = wB[j] (j < m-1; the last
//     column constraint is dropped, so column m-1 absorbs the remainder).
//
// One 64-lane wavefront solves one LP with the transportation simplex in
// FP64:
//   * start: north-west corner, which is a spanning tree of n+m-1 cells;
//   * pricing: every lane scans a strided share of the n*m cells for the most
//     negative reduced cost C - u_k - v_j (Dantzig), ties to the lowest cell
//     index; after many pivots the rule switches to Bland's first-negative
//     cell, which cannot cycle;
//   * pivot: lane 0 walks the basis tree (parent pointers) to find the cycle,
//     moves theta around it and re-hangs the cut subtree by reversing the
//     parent pointers on the path to the leaving edge;
//   * potentials u, v: recomputed for every node by pointer jumping, so a
//     pivot costs O(log depth) parallel rounds, not a sequential tree walk.
// Costs come from supports held in LDS (Euclidean, squared distance by
// direct differences), from a histogram ground-cost matrix, or from a
// precomputed FP64 cost array.
#include <cfloat>
#include <cmath>

#include "d2mi_internal.h"
#include "transport.h"

namespace {

constexpr int WAVE = 64;

__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~size_t(15); }

struct LpLds {
  size_t A, B, wa, wb, ids, pot, flow, acc0, acc1, par, anc0, anc1, sg0, sg1, mark, pA, pB, total;
  __host__ __device__ LpLds(int n_max, int m_max, int dsupp) {
    const size_t nn = (size_t)n_max + m_max;
    size_t o = 0;
    A = o; o = al16(o + sizeof(double) * (size_t)n_max * dsupp);
    B = o; o = al16(o + sizeof(double) * (size_t)m_max * dsupp);
    wa = o; o = al16(o + sizeof(double) * n_max);
    wb = o; o = al16(o + sizeof(double) * m_max);
    ids = o; o = al16(o + sizeof(int) * m_max);
    pot = o; o = al16(o + sizeof(double) * nn);
    flow = o; o = al16(o + sizeof(double) * nn);
    acc0 = o; o = al16(o + sizeof(double) * nn);
    acc1 = o; o = al16(o + sizeof(double) * nn);
    par = o; o = al16(o + sizeof(int) * nn);
    anc0 = o; o = al16(o + sizeof(int) * nn);
    anc1 = o; o = al16(o + sizeof(int) * nn);
    sg0 = o; o = al16(o + sizeof(int) * nn);
    sg1 = o; o = al16(o + sizeof(int) * nn);
    mark = o; o = al16(o + sizeof(int) * nn);
    pA = o; o = al16(o + sizeof(int) * (nn + 1));
    pB = o; o = al16(o + sizeof(int) * (nn + 1));
    total = o;
  }
};

__device__ __forceinline__ double wmin_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, WAVE));
  return v;
}
__device__ __forceinline__ double wsum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ __forceinline__ int wmin_i(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, WAVE));
  return v;
}

struct Cost {
  int mode, d, n, V;
  const double *A, *B;       // LDS supports (Euclidean)
  const int *ids;            // LDS bin ids (sparse histogram)
  const double *__restrict__ dist;  // ground cost [V*V]
  const double *__restrict__ g;     // precomputed [n*m], element (k,j) at k + j*n
  __device__ __forceinline__ double operator()(int k, int j) const {
    switch (mode) {
      case LP_COST_EUCLID: {
        double acc = 0.0;
        for (int t = 0; t < d; ++t) {
          const double u = A[k * d + t] - B[j * d + t];
          acc += u * u;
        }
        return acc;
      }
      case LP_COST_HIST:
        return dist[k + (size_t)j * V];
      case LP_COST_SPARSE:
        return dist[(size_t)ids[j] * V + k];
      default:
        return g[k + (size_t)j * n];
    }
  }
};

// cost of the tree edge stored at node v (edge to its parent)
__device__ __forceinline__ double edge_cost(const Cost &C, int v, int p, int n) {
  return v < n ? C(v, p - n) : C(p, v - n);
}

// Solves one LP with the wave; returns the optimal value to all lanes.
__device__ double solve_lp(const Cost &C, int n, int m, const double *wa, const double *wb,
                           unsigned char *lds, const LpLds &L, int *pivots_out, int *status) {
  const int lane = threadIdx.x;
  const int nn = n + m;
  double *pot = (double *)(lds + L.pot), *flow = (double *)(lds + L.flow);
  double *acc0 = (double *)(lds + L.acc0), *acc1 = (double *)(lds + L.acc1);
  int *par = (int *)(lds + L.par), *anc0 = (int *)(lds + L.anc0), *anc1 = (int *)(lds + L.anc1);
  int *sg0 = (int *)(lds + L.sg0), *sg1 = (int *)(lds + L.sg1), *mark = (int *)(lds + L.mark);
  int *pA = (int *)(lds + L.pA), *pB = (int *)(lds + L.pB);

  // cost scale for the reduced-cost tolerance
  double cmax = 0.0;
  for (int e = lane; e < n * m; e += WAVE) cmax = fmax(cmax, fabs(C(e % n, e / n)));
  cmax = -wmin_d(-cmax);
  const double eps = 1e-12 * (1.0 + cmax);

  for (int v = lane; v < nn; v += WAVE) mark[v] = 0;
  __syncthreads();

  // north-west corner: each new cell brings one new node, hung under the
  // endpoint already in the tree; flow[v] is the flow on v's parent edge
  if (lane == 0) {
    double sa = 0.0, sbm = 0.0;
    for (int k = 0; k < n; ++k) { acc0[k] = wa[k]; sa += wa[k]; }
    for (int j = 0; j < m - 1; ++j) { acc1[j] = wb[j]; sbm += wb[j]; }
    acc1[m - 1] = sa - sbm > 0.0 ? sa - sbm : 0.0;
    par[0] = -1;
    flow[0] = 0.0;
    int i = 0, j = 0, newnode = n;  // first cell (0,0) introduces column 0
    for (int t = 0; t < n + m - 1; ++t) {
      double q = fmin(acc0[i], acc1[j]);
      if (q < 0.0) q = 0.0;
      if (newnode >= n) par[newnode] = i;       // new column hangs under row i
      else par[newnode] = n + j;                // new row hangs under column j
      flow[newnode] = q;
      acc0[i] -= q;
      acc1[j] -= q;
      if (t == n + m - 2) break;
      if (i == n - 1) { ++j; newnode = n + j; }
      else if (j == m - 1) { ++i; newnode = i; }
      else if (acc0[i] <= acc1[j]) { ++i; newnode = i; }
      else { ++j; newnode = n + j; }
    }
  }
  __syncthreads();

  int pivots = 0;
  const int dantzig_limit = 50 * nn + 1000;
  const int hard_limit = 200000 + 100 * n * m;
  for (;;) {
    // ---- potentials by pointer jumping: pot[v] = acc[v] + sg[v] * pot[anc[v]]
    for (int v = lane; v < nn; v += WAVE) {
      const int p = par[v];
      anc0[v] = p;
      acc0[v] = p >= 0 ? edge_cost(C, v, p, n) : 0.0;
      sg0[v] = p >= 0 ? -1 : 0;
    }
    __syncthreads();
    int *ancA = anc0, *ancB = anc1, *sgA = sg0, *sgB = sg1;
    double *accA = acc0, *accB = acc1;
    for (;;) {
      int any = 0;
      for (int v = lane; v < nn; v += WAVE) {
        const int a = ancA[v];
        if (a >= 0) {
          accB[v] = accA[v] + sgA[v] * accA[a];
          sgB[v] = sgA[v] * sgA[a];
          ancB[v] = ancA[a];
          any |= ancA[a] >= 0;
        } else {
          accB[v] = accA[v];
          sgB[v] = sgA[v];
          ancB[v] = -1;
        }
      }
      __syncthreads();
      int *ti = ancA; ancA = ancB; ancB = ti;
      ti = sgA; sgA = sgB; sgB = ti;
      double *td = accA; accA = accB; accB = td;
      any = __any(any);
      if (!any) break;
    }
    for (int v = lane; v < nn; v += WAVE) pot[v] = accA[v];
    __syncthreads();

    // ---- pricing
    const bool bland = pivots >= dantzig_limit;
    double best = -eps;
    int bestc = 0x7fffffff;
    for (int e = lane; e < n * m; e += WAVE) {
      const int k = e % n, j = e / n;
      const double r = C(k, j) - pot[k] - pot[n + j];
      if (bland) {
        if (r < -eps && e < bestc) { bestc = e; best = r; }
      } else if (r < best || (r == best && e < bestc && r < -eps)) {
        best = r;
        bestc = e;
      }
    }
    if (bland) {
      const int c = wmin_i(bestc);
      if (c == 0x7fffffff) break;
      bestc = c;
    } else {
      const double bm = wmin_d(best);
      const int c = wmin_i(best == bm ? bestc : 0x7fffffff);
      if (!(bm < -eps) || c == 0x7fffffff) break;
      best = bm;
      bestc = c;
    }
    if (pivots >= hard_limit) {
      if (lane == 0) atomicOr(status, 1);
      break;
    }
    const int ei = bestc % n, ej = bestc / n;

    // ---- pivot (lane 0): cycle = tree path from column ej to row ei
    if (lane == 0) {
      const int a = n + ej, b = ei;
      int la = 0;
      for (int v = a; v >= 0; v = par[v]) { pA[la] = v; mark[v] = la + 1; ++la; }
      int lb = 0, v = b;
      while (mark[v] == 0) { pB[lb++] = v; v = par[v]; }
      const int nA = mark[v] - 1;  // edges on a's side: pA[0..nA)
      for (int t = 0; t < la; ++t) mark[pA[t]] = 0;
      // minus edges: even positions from each end
      double theta = DBL_MAX;
      int qside = -1, qt = -1, qnode = 0x7fffffff;
      for (int t = 0; t < nA; t += 2) {
        const double f = flow[pA[t]];
        if (f < theta || (bland && f == theta && pA[t] < qnode)) {
          theta = f; qside = 0; qt = t; qnode = pA[t];
        }
      }
      for (int t = 0; t < lb; t += 2) {
        const double f = flow[pB[t]];
        if (f < theta || (bland && f == theta && pB[t] < qnode)) {
          theta = f; qside = 1; qt = t; qnode = pB[t];
        }
      }
      for (int t = 0; t < nA; ++t) flow[pA[t]] += (t & 1) ? theta : -theta;
      for (int t = 0; t < lb; ++t) flow[pB[t]] += (t & 1) ? theta : -theta;
      // re-hang: the cut subtree holds the path pX[0..qt]; reverse it and
      // hang its end pX[0] (a or b) under the entering cell's other end
      int *P = qside == 0 ? pA : pB;
      const int other = qside == 0 ? b : a;
      for (int t = qt; t >= 1; --t) {
        par[P[t]] = P[t - 1];
        flow[P[t]] = flow[P[t - 1]];
      }
      par[P[0]] = other;
      flow[P[0]] = theta;
    }
    __syncthreads();
    ++pivots;
  }

  // objective: sum over tree edges
  double val = 0.0;
  for (int v = lane; v < nn; v += WAVE) {
    const int p = par[v];
    if (p >= 0) val += flow[v] * edge_cost(C, v, p, n);
  }
  val = wsum_d(val);
  if (pivots_out) *pivots_out = pivots;
  __syncthreads();
  return val;
}

__global__ __launch_bounds__(WAVE) void k_transport(LpArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = threadIdx.x;
  const int dsupp = a.cost_mode == LP_COST_EUCLID ? a.d : 0;
  const LpLds L(a.n_max, a.m_max, dsupp > 0 ? dsupp : 1);
  double *A = (double *)(lds + L.A), *B = (double *)(lds + L.B);
  double *wa = (double *)(lds + L.wa), *wb = (double *)(lds + L.wb);
  int *ids = (int *)(lds + L.ids);
  for (int job = blockIdx.x; job < a.njobs; job += gridDim.x) {
    if (a.mask && a.mask[job] == 0) continue;
    // row side: centroid `ka` of table A
    const int ka = a.implicit_K > 0 ? job % a.implicit_K : a.job_a[job];
    const int jb = a.implicit_K > 0 ? job / a.implicit_K : a.job_b[job];
    const int n = a.sA;
    const double *Asrc = a.A_supp ? a.A_supp + (size_t)ka * n * dsupp : nullptr;
    const double *wA = a.A_w + (size_t)ka * n;
    int m;
    const double *Bsrc = nullptr, *wB;
    const int *idsrc = nullptr;
    if (a.b_is_table) {
      m = a.sB;
      Bsrc = a.B_supp ? a.B_supp + (size_t)jb * m * dsupp : nullptr;
      wB = a.B_w + (size_t)jb * m;
    } else {
      m = a.p_str[jb];
      const long long off = a.cum[jb];
      Bsrc = a.supp ? a.supp + (size_t)off * dsupp : nullptr;
      wB = a.w + off;
      idsrc = a.sym ? a.sym + off : nullptr;
    }
    __syncthreads();
    for (int q = lane; q < n; q += WAVE) wa[q] = wA[q];
    for (int q = lane; q < m; q += WAVE) wb[q] = wB[q];
    if (dsupp > 0) {
      for (int q = lane; q < n * dsupp; q += WAVE) A[q] = Asrc[q];
      for (int q = lane; q < m * dsupp; q += WAVE) B[q] = Bsrc[q];
    }
    int mode = a.cost_mode;
    if (mode == LP_COST_SPARSE) {
      if (idsrc) {
        for (int q = lane; q < m; q += WAVE) ids[q] = idsrc[q];
      } else {
        mode = LP_COST_HIST;  // two centroids: both are full V-bin histograms
      }
    }
    __syncthreads();
    Cost C{mode, dsupp, n, a.V, A, B, ids, a.dist,
           a.gcost ? a.gcost + (size_t)job * a.gstride : nullptr};
    int piv = 0;
    const double v = solve_lp(C, n, m, wa, wb, lds, L, &piv, a.status);
    if (lane == 0) {
      a.out[job] += v;
      if (a.pivots) a.pivots[job] += piv;
    }
  }
}

__global__ void k_sqrt_clamp(double *d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = d[i] <= 0.0 ? 0.0 : sqrt(d[i]);  // clustering.c:126-127
}

__global__ void k_zero_d(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
__global__ void k_zero_int(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

size_t d2mi_lp_lds_bytes(int n_max, int m_max, int dsupp) {
  return LpLds(n_max, m_max, dsupp > 0 ? dsupp : 1).total;
}

int d2mi_launch_transport(d2mi_ctx *ctx, const LpArgs &a) {
  if (a.njobs <= 0) return D2MI_OK;
  const int dsupp = a.cost_mode == LP_COST_EUCLID ? a.d : 0;
  const size_t lds = d2mi_lp_lds_bytes(a.n_max, a.m_max, dsupp);
  if (lds > ctx->lds_max) {
    d2mi_set_error("transport LP of %d x %d (d=%d) needs %zu B of LDS (limit %zu)", a.n_max,
                   a.m_max, dsupp, lds, ctx->lds_max);
    return D2MI_EUNSUPPORTED;
  }
  const int grid = std::min(a.njobs, ctx->num_cu * 16);
  k_transport<<<grid, WAVE, lds, ctx->stream>>>(a);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

int d2mi_zero_doubles(d2mi_ctx *ctx, double *p, int n) {
  if (n > 0) k_zero_d<<<(n + 255) / 256, 256, 0, ctx->stream>>>(p, n);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}
int d2mi_zero_ints(d2mi_ctx *ctx, int *p, int n) {
  if (n > 0) k_zero_int<<<(n + 255) / 256, 256, 0, ctx->stream>>>(p, n);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}
int d2mi_sqrt_clamp(d2mi_ctx *ctx, double *p, int n) {
  if (n > 0) k_sqrt_clamp<<<(n + 255) / 256, 256, 0, ctx->stream>>>(p, n);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

// --------------------------------------------------------------- C ABI

extern "C" int d2mi_transport_batch(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                                    const double *wX, const double *wY, double *out,
                                    int *pivots) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (n <= 0 || m <= 0 || nbatch <= 0 || !C || !wX || !wY || !out) {
    d2mi_set_error("d2mi_transport_batch: bad arguments");
    return D2MI_EINVAL;
  }
  hipStream_t S = ctx->stream;
  DevMem dC, dA, dB, dout, dja, djb, dpiv, dst;
  const size_t nm = (size_t)n * m;
  D2MI_TRY(dC.assign(C, sizeof(double) * nm * nbatch, S));
  D2MI_TRY(dA.assign(wX, sizeof(double) * (size_t)n * nbatch, S));
  D2MI_TRY(dB.assign(wY, sizeof(double) * (size_t)m * nbatch, S));
  D2MI_TRY(dout.ensure(sizeof(double) * nbatch));
  D2MI_TRY(dpiv.ensure(sizeof(int) * nbatch));
  D2MI_TRY(dst.ensure(sizeof(int)));
  std::vector<int> idx(nbatch);
  for (int i = 0; i < nbatch; ++i) idx[i] = i;
  D2MI_TRY(dja.assign(idx.data(), sizeof(int) * nbatch, S));
  D2MI_TRY(djb.assign(idx.data(), sizeof(int) * nbatch, S));
  D2MI_TRY(d2mi_zero_doubles(ctx, dout.as<double>(), nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, dpiv.as<int>(), nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, dst.as<int>(), 1));
  LpArgs a{};
  a.njobs = nbatch;
  a.job_a = dja.as<int>();
  a.job_b = djb.as<int>();
  a.b_is_table = 1;
  a.cost_mode = LP_COST_GLOBAL;
  a.sA = n;
  a.A_w = dA.as<double>();
  a.sB = m;
  a.B_w = dB.as<double>();
  a.gcost = dC.as<double>();
  a.gstride = (long long)nm;
  a.n_max = n;
  a.m_max = m;
  a.out = dout.as<double>();
  a.pivots = dpiv.as<int>();
  a.status = dst.as<int>();
  D2MI_TRY(d2mi_launch_transport(ctx, a));
  D2MI_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * nbatch, hipMemcpyDeviceToHost, S));
  if (pivots)
    D2MI_HIP(hipMemcpyAsync(pivots, dpiv.p, sizeof(int) * nbatch, hipMemcpyDeviceToHost, S));
  int st = 0;
  D2MI_HIP(hipMemcpyAsync(&st, dst.p, sizeof(int), hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (st) {
    d2mi_set_error("a transport LP hit the pivot limit");
    return D2MI_ENUMERIC;
  }
  return D2MI_OK;
}
