// B-ADMM centroid update on the device (the reference's d2_centroid_sphBregman,
// centroid_Bregman.c:51-325).
//
// Data flow of one update (all on one stream, no host round trip between
// inner iterations):
//   1. stable counting sort of the objects by label -> perm, segment starts
//      (label counts double as the empty-cluster check of :138-144);
//   2. setup: sum of the ground costs (rho, :109), Z re-initialised for
//      switched objects (:116-124), Y = 0 (:128);
//   3. per inner iteration (:150): the plan kernel does steps 1-3 for every
//      object (X, Z, Y updates, :155-175) in one HBM pass over Z and Y, with the
//      ground cost computed on the fly from the supports, and emits per-label
//      FP64 partial sums of the normalised row sums (step 4, :179-185) and, on
//      support-update iterations, of X * supports and X row sums (step 5,
//      :196-212); the finalize kernel sums the partials in a fixed order and
//      normalises (:190, :219-221).
// Objects are processed in label-sorted order, so a workgroup accumulates
// per-label sums in LDS and writes one record per (workgroup, label) pair;
// record (g, l) lives at index g + l, which is unique because chunk label
// ranges do not interleave.  Every sum has a fixed order: results are
// bit-reproducible run to run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <type_traits>

#include "d2mi_internal.h"
#include "lds_copy.h"

namespace {

constexpr int NT = 256;        // threads per workgroup (4 waves of 64)
constexpr int MAXD_FUSED = 8;  // Euclidean cost fused on the fly for d <= 8
constexpr int FLAG_BADLABEL = 1, FLAG_EMPTY = 2, FLAG_NUMERIC = 4;

inline bool getenv_flag(const char *name) {
  const char *e = getenv(name);
  return e && atoi(e) != 0;
}

// COST_TABLE: the ground cost is read from a stored table C[s*cum[i] + k + j*s]
// (the reference's own C array, clustering_util.c:179), rebuilt after every
// support update; used for Euclidean d > MAXD_FUSED and for word embeddings.
enum CostMode { COST_EUCLID = 0, COST_HIST = 1, COST_SPARSE = 2, COST_TABLE = 3 };

template <typename T>
__device__ __forceinline__ T dexp(T x);
template <>
__device__ __forceinline__ float dexp<float>(float x) { return expf(x); }
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }

// Wave total into lane 63 with DPP row shifts and row broadcasts (GFX9
// cross-lane moves, plain VALU: no LDS round trip per step).  Lanes shifted in
// from outside a row read 0; after row_shr 1, 2, 4, 8 lane 15 of each row holds
// the row's sum, row_bcast:15 (rows 1 and 3) and row_bcast:31 (rows 2 and 3)
// carry it on to lane 63.  A fixed order: reproducible.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                               ROWS, 0xf, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  typedef int i2 __attribute__((ext_vector_type(2)));
  i2 u = __builtin_bit_cast(i2, v);
  u.x = __builtin_amdgcn_update_dpp(0, u.x, CTRL, ROWS, 0xf, true);
  u.y = __builtin_amdgcn_update_dpp(0, u.y, CTRL, ROWS, 0xf, true);
  return __builtin_bit_cast(double, u);
}
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T dpp_t(T v) {
  if constexpr (std::is_same<T, double>::value) return dpp_f64<CTRL, ROWS>(v);
  else return dpp_f32<CTRL, ROWS>(v);
}
template <typename T>
__device__ __forceinline__ T wave_sum_to63(T v) {
  v += dpp_t<0x111, 0xf>(v);  // row_shr:1
  v += dpp_t<0x112, 0xf>(v);  // row_shr:2
  v += dpp_t<0x114, 0xf>(v);  // row_shr:4
  v += dpp_t<0x118, 0xf>(v);  // row_shr:8
  v += dpp_t<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_t<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Wave total of v on every lane (all 64 lanes must be active): the DPP sum
// into lane 63, read back as a scalar.  Fixed order: reproducible.
__device__ __forceinline__ double wave_sum(double v) {
  typedef int i2 __attribute__((ext_vector_type(2)));
  i2 u = __builtin_bit_cast(i2, wave_sum_to63<double>(v));
  u.x = __builtin_amdgcn_readlane(u.x, 63);
  u.y = __builtin_amdgcn_readlane(u.y, 63);
  return __builtin_bit_cast(double, u);
}

// Block-wide sum with a fixed association order; every thread gets the
// result.  Must be called by all threads of the block.
template <int NTB = NT>
__device__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < NTB / 64; ++i) t += red[i];
  return t;
}

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// LDS layout of the plan kernel, shared by host (sizing) and device.
struct PlanLds {
  size_t A, B, X, w, xs, cs, cw, g, rr, rp, ids, acc, red, up, total;
  __host__ __device__ PlanLds(int tsize, int sp, int m_max, int s, int d, int keepx, int R,
                              bool gm = false) {
    const size_t EP = gm ? 0 : (size_t)sp * m_max;  // gm: plan scratch lives in HBM
    size_t o = 0;
    A = o; o = align16(o + EP * tsize);
    B = o; o = align16(o + EP * tsize);
    X = o; o = align16(o + (keepx ? EP : 0) * tsize);
    w = o; o = align16(o + (size_t)m_max * tsize);
    xs = o; o = align16(o + (size_t)m_max * (d > 0 ? d : 1) * tsize);
    cs = o; o = align16(o + (size_t)s * (d > 0 ? d : 1) * tsize);
    cw = o; o = align16(o + (size_t)s * tsize);
    g = o; o = align16(o + (size_t)s * tsize);
    rr = o; o = align16(o + (size_t)s * tsize);
    rp = o; o = align16(o + (size_t)4 * s * tsize);
    ids = o; o = align16(o + (size_t)m_max * sizeof(int));
    acc = o; o = align16(o + (size_t)R * sizeof(double));
    red = o; o = align16(o + 8 * sizeof(double));
    // support-update partials of the four waves (R > s only on those passes)
    up = o; o = align16(o + (R > s ? (size_t)4 * s * ((d > 0 ? d : 1) + 1) : 0) * sizeof(double));
    total = o;
  }
};

// object o = perm[p], its label, column count and first column
struct alignas(16) ObjMeta {
  int o, l, m, off;  // m carries the object's label-switch flag in bit 30
};
constexpr int META_SWITCHED = 1 << 30;

template <typename T>
struct PlanArgs {
  int N, s, d, K, sp, m_max, chunk, R, V;
  int upd, resid, keepx, first, force_init;
  int rev;  // register forms: walk each chunk's objects last to first
  const char *lsw;
  const int *__restrict__ p_str;
  const long long *__restrict__ cum;
  const int *__restrict__ perm;
  const int *__restrict__ label;
  const T *__restrict__ w;
  const T *__restrict__ supp;
  const int *__restrict__ sym;
  const T *__restrict__ dist;
  const T *__restrict__ cw;
  const T *__restrict__ csupp;
  const double *__restrict__ scal;  // [0] = rho, [1] = 1/rho
  T *__restrict__ Z;
  T *__restrict__ Y;
  const T *__restrict__ ctab;       // COST_TABLE: stored costs (plan layout)
  T *__restrict__ xout;             // COST_TABLE, support-update passes: X out (plan layout)
  T *__restrict__ xg;               // HBM-scratch kernel: X (plan layout)
  T *__restrict__ z0g;              // HBM-scratch kernel, residual passes: Z before the pass
  double *__restrict__ rec;         // [(G + K) * R]
  double *__restrict__ resid_part;  // [G * 3]
  const struct ObjMeta *__restrict__ pm;  // or null: per position of perm, its object's metadata
};

// ground cost of element (k, j) before the division by rho; DIM > 0 fixes
// the Euclidean dimension at compile time (straight-line code)
template <typename T, int CM, int DIM = 0>
__device__ __forceinline__ T cost_raw(int k, int j, int s, int d, int V, const T *cs,
                                      const T *xs, const int *ids, const T *__restrict__ dist,
                                      const T *__restrict__ tab = nullptr) {
  if constexpr (CM == COST_TABLE) {
    return tab[k + (size_t)j * s];
  } else if constexpr (CM == COST_EUCLID) {
    T acc = 0;
    if constexpr (DIM > 0) {
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) {
        const T u = cs[k * DIM + dd] - xs[j * DIM + dd];
        acc += u * u;
      }
    } else {
      for (int dd = 0; dd < d; ++dd) {
        const T u = cs[k * d + dd] - xs[j * d + dd];
        acc += u * u;
      }
    }
    return acc;
  } else if constexpr (CM == COST_HIST) {
    return dist[k + (size_t)j * s];
  } else {
    return dist[(size_t)ids[j] * V + k];
  }
}

template <typename T>
__device__ __forceinline__ T scale_cost(T c, double rho, T inv_rho) {
  if constexpr (std::is_same<T, double>::value) return c / rho;  // C[i] /= rho (:110)
  else return c * inv_rho;
}

constexpr int PRE_W = 2;  // register slots for weights / bin ids: m_max <= 512
constexpr int PRE_X = 4;  // register slots for supports: m_max * d <= 1024

#ifndef D2MI_PLAN_UNROLL
#define D2MI_PLAN_UNROLL 4
#endif
constexpr int PLAN_UNROLL = D2MI_PLAN_UNROLL;  // plan elements per thread per load batch

#ifndef D2MI_PLAN_WAVES_PER_EU
#define D2MI_PLAN_WAVES_PER_EU 1
#endif

// Plan update of every object in this workgroup's chunk of the label-sorted
// order (any s; k_badmm_plan64 is the faster form for s <= 64 with the plan
// in LDS).  Per object: weights and supports are requested at the top, before
// the workgroup barrier, so their HBM round trips overlap; then
//   pass 1 (element per thread, PLAN_UNROLL elements' loads in flight at
//          once): t = Z exp(-(C + Y)) + 1e-9 into LDS;
//   pass 2 (four threads per column): column sums by quad shuffles, then the
//          same threads form x = t / colsum * w and z' = x exp(Y) + 1e-9;
//   pass 3 (four waves): partial row sums, combined by wave 0 into the
//          normalisers and the per-label sums;
//   pass 4 (element per thread): z = z' / rowsum * c_w, Y += x - z, stored.
//
// GM (objects whose plan does not fit LDS): the same passes with the plan
// scratch in HBM -- t and z' in place in Z, y (+ x) in place in Y, x in a
// scratch array; the workgroup barriers order these accesses as they order
// the LDS ones.
//
// LEAN instances serve the passes after the first with no support update,
// residual or kept X (most passes): those branches are compiled out, which
// frees registers for more resident waves.
template <typename T, int CM, int DIM, bool GM = false, bool LEAN = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(D2MI_PLAN_WAVES_PER_EU)))
void k_badmm_plan(PlanArgs<T> a, const int *__restrict__ g_perm, const int *__restrict__ g_str,
                  const long long *__restrict__ g_cum, const int *__restrict__ g_label,
                  const char *__restrict__ g_lsw) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool F64 = std::is_same<T, double>::value;
  const int s = a.s, d = DIM > 0 ? DIM : a.d, sp = GM ? a.s : a.sp, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const bool upd = !LEAN && a.upd, resid = !LEAN && a.resid, keepx = !LEAN && a.keepx;
  const bool first = !LEAN && a.first;
  const PlanLds L(sizeof(T), a.sp, a.m_max, s, d, keepx, a.R, GM);
  T *A = (T *)(smem + L.A), *B = (T *)(smem + L.B), *Xs = (T *)(smem + L.X);
  T *wl = (T *)(smem + L.w), *xs = (T *)(smem + L.xs);
  T *cs = (T *)(smem + L.cs), *cwl = (T *)(smem + L.cw), *g = (T *)(smem + L.g);
  T *rr = (T *)(smem + L.rr), *rp = (T *)(smem + L.rp);
  int *ids = (int *)(smem + L.ids);
  double *acc = (double *)(smem + L.acc), *red = (double *)(smem + L.red);
  double *up = (double *)(smem + L.up);

  const int gi = blockIdx.x;
  const int p0 = gi * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const double rho = a.scal[0];
  const T inv_rho = (T)a.scal[1];
  const int dk = NT % s, dj = NT / s, k0 = tid % s, j0 = tid / s;
  const bool small_s = s <= 64;  // row and support sums split over the 4 waves
  double r_obj = 0.0, r_pr = 0.0, r_dr = 0.0;
  for (int q = tid; q < a.R; q += NT) acc[q] = 0.0;
  int cur = -1;

  for (int p = p0; p < p1; ++p) {
    const int o = g_perm[p];
    const int l = g_label[o];
    const int m = g_str[o];
    const long long off = g_cum[o];
    const int E = s * m;
    const size_t base = (size_t)s * off;
    if constexpr (GM) {
      A = a.Z + base;
      B = a.Y + base;
      Xs = a.xg + base;
    }
    // first pass of an update: Y = 0 (:128) and, for objects whose label
    // switched, Z = w_j * c_w[label][k] (:116-124), without touching HBM
    const bool reinit = first && (a.force_init || g_lsw[o] == 1);
    // requests first: weights and supports of this object
    T wpre[PRE_W], xpre[PRE_X];
    int ipre[PRE_W];
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NT;
      if (q < m) {
        wpre[r] = a.w[off + q];
        if constexpr (CM == COST_SPARSE) ipre[r] = a.sym[off + q];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NT;
        if (q < m * d) xpre[r] = a.supp[(size_t)off * d + q];
      }
    }
    if (l != cur) {
      __syncthreads();  // every wave is done with the previous label
      if (cur >= 0)
        for (int q = tid; q < a.R; q += NT) {
          a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
          acc[q] = 0.0;
        }
      cur = l;
      for (int q = tid; q < s; q += NT) cwl[q] = a.cw[(size_t)l * s + q];
      if constexpr (CM == COST_EUCLID)
        for (int q = tid; q < s * d; q += NT) cs[q] = a.csupp[(size_t)l * s * d + q];
    }
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NT;
      if (q < m) {
        wl[q] = wpre[r];
        if constexpr (CM == COST_SPARSE) ids[q] = ipre[r];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NT;
        if (q < m * d) xs[q] = xpre[r];
      }
    }
    __syncthreads();  // also: every wave has finished the previous object's last pass

    // pass 1 -- step 1a: t = Z exp(-(C + Y)) + 1e-9 (:155-157)
    {
      int k = k0, j = j0;
      {
        // U elements per thread at a time: all their loads are in flight
        // before the first is used, instead of one HBM round trip each
        constexpr int U = PLAN_UNROLL;
        for (int e0 = tid; e0 < E; e0 += U * NT) {
          T zv[U], yv[U], cv[U];
          {
            int kk = k, jj = j;
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int e = e0 + u * NT;
              zv[u] = (T)0;
              yv[u] = (T)0;
              cv[u] = (T)0;
              if (e < E) {
                if (!reinit) zv[u] = a.Z[base + e];
                if (!first) yv[u] = a.Y[base + e];
                cv[u] = cost_raw<T, CM, DIM>(kk, jj, s, d, a.V, cs, xs, ids, a.dist,
                                             a.ctab + base);
              }
              kk += dk;
              jj += dj;
              if (kk >= s) { kk -= s; ++jj; }
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int e = e0 + u * NT;
            if (e < E) {
              const T c = scale_cost<T>(cv[u], rho, inv_rho);
              const T z = reinit ? wl[j] * cwl[k] : zv[u];
              const int li = k + j * sp;
              if (GM && resid) a.z0g[base + e] = z;
              A[li] = z * dexp<T>(-(c + yv[u])) + (T)1e-9;
              B[li] = yv[u];
            }
            k += dk;
            j += dj;
            if (k >= s) { k -= s; ++j; }
          }
        }
      }
    }
    __syncthreads();
    // pass 2 -- step 1b/1c: column sums (_dcnorm, :158) by quads, x = t / colsum
    // * w (:159); step 2a: z' = x exp(Y) + 1e-9 (:166)
    for (int t = tid; t < 4 * m; t += NT) {
      const int j = t >> 2, q = t & 3;
      T* colA = A + j * sp;
      T part = 0;
      for (int k = q; k < s; k += 4) part += colA[k];
      part += __shfl_xor(part, 1, 64);
      part += __shfl_xor(part, 2, 64);
      const T wj = wl[j];
      const T scale = F64 ? (T)0 : wj / part;
      for (int k = q; k < s; k += 4) {
        const int li = k + j * sp;
        const T x = F64 ? (colA[k] / part) * wj : colA[k] * scale;
        const T y = B[li];
        colA[k] = x * dexp<T>(y) + (T)1e-9;
        if (keepx) Xs[li] = x;
        else B[li] = y + x;
      }
    }
    __syncthreads();
    // pass 3 -- step 2b: row sums (_drnorm, :169) split over the four waves;
    // step 5 partials on support-update iterations (:206-211)
    if (small_s) {
      if (lane < s) {
        T part = 0;
        for (int j = wv; j < m; j += 4) part += A[lane + j * sp];
        rp[wv * s + lane] = part;
      }
    } else {
      for (int k = tid; k < s; k += NT) {
        T r = 0;
        for (int j = 0; j < m; ++j) r += A[k + j * sp];
        rr[k] = r;
      }
    }
    if (upd) {
      if (small_s) {  // wave wv: columns j = wv, wv + 4, ...
        if (lane < s) {
          double P[MAXD_FUSED];
          for (int dd = 0; dd < d; ++dd) P[dd] = 0.0;
          double rx = 0.0;
          for (int j = wv; j < m; j += 4) {
            const double x = (double)Xs[lane + j * sp];
            rx += x;  // row sums of X (:211)
            for (int dd = 0; dd < d; ++dd) P[dd] += x * (double)xs[j * d + dd];  // supp X^T (:206)
          }
          double *u = up + (size_t)(wv * s + lane) * (d + 1);
          for (int dd = 0; dd < d; ++dd) u[dd] = P[dd];
          u[d] = rx;
        }
      } else {
        for (int k = tid; k < s; k += NT) {
          double P[MAXD_FUSED];
          for (int dd = 0; dd < d; ++dd) P[dd] = 0.0;
          double rx = 0.0;
          for (int j = 0; j < m; ++j) {
            const double x = (double)Xs[k + j * sp];
            rx += x;
            for (int dd = 0; dd < d; ++dd) P[dd] += x * (double)xs[j * d + dd];
          }
          for (int dd = 0; dd < d; ++dd) acc[s + k * d + dd] += P[dd];
          acc[s + s * d + k] += rx;
        }
      }
    }
    __syncthreads();
    // step 4: normalised row sums summed per label (:180-185)
    if (small_s) {
      if (wv == 0) {
        T r = 0;
        if (lane < s) {
          r = rp[lane] + rp[s + lane] + rp[2 * s + lane] + rp[3 * s + lane];
          rr[lane] = r;
          g[lane] = F64 ? r : cwl[lane] / r;
        }
        double rs = wave_sum(lane < s ? (double)r : 0.0);
        if (lane < s) acc[lane] += (double)r / rs;
        if (upd && lane < s) {
          for (int dd = 0; dd <= d; ++dd) {
            double v = 0.0;
            for (int w4 = 0; w4 < 4; ++w4) v += up[(size_t)(w4 * s + lane) * (d + 1) + dd];
            if (dd < d) acc[s + lane * d + dd] += v;
            else acc[s + s * d + lane] += v;
          }
        }
      }
    } else {
      double rsum = 0.0;
      for (int k = tid; k < s; k += NT) {
        rsum += (double)rr[k];
        g[k] = F64 ? rr[k] : cwl[k] / rr[k];
      }
      rsum = block_sum(rsum, red);
      for (int k = tid; k < s; k += NT) acc[k] += (double)rr[k] / rsum;
    }
    __syncthreads();
    // pass 4 -- step 2c + 3: z = z' / rowsum * c_w (:169-170); Y += X - Z (:175)
    {
      int k = k0, j = j0;
      for (int e = tid; e < E; e += NT) {
        const int li = k + j * sp;
        const T z = F64 ? (A[li] / g[k]) * cwl[k] : A[li] * g[k];
        if (resid) {  // step 6 (:300-305), and the exit objective
          const T c = scale_cost<T>(
              cost_raw<T, CM, DIM>(k, j, s, d, a.V, cs, xs, ids, a.dist, a.ctab + base), rho,
              inv_rho);
          const T x = Xs[li];
          const T z0 = GM ? a.z0g[base + e] : (reinit ? wl[j] * cwl[k] : a.Z[base + e]);
          r_obj += (double)c * (double)x;
          r_pr += fabs((double)x - (double)z);
          r_dr += fabs((double)z0 - (double)z);
        }
        a.Z[base + e] = z;
        a.Y[base + e] = keepx ? B[li] + (Xs[li] - z) : B[li] - z;
        if (CM == COST_TABLE && upd && !GM) a.xout[base + e] = Xs[li];  // for the support GEMM
        k += dk;
        j += dj;
        if (k >= s) { k -= s; ++j; }
      }
    }
    if (resid) __syncthreads();  // pass 4 read xs / cs: keep them until all waves are done
  }
  __syncthreads();
  if (cur >= 0)
    for (int q = tid; q < a.R; q += NT) a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
  if (resid) {
    const double t0 = block_sum(r_obj, red);
    const double t1 = block_sum(r_pr, red);
    const double t2 = block_sum(r_dr, red);
    if (tid == 0) {
      a.resid_part[gi * 3 + 0] = t0;
      a.resid_part[gi * 3 + 1] = t1;
      a.resid_part[gi * 3 + 2] = t2;
    }
  }
}

// ------------------------------------------------- plan kernel for s <= 64

// FP32 exp(x) on the hardware base-2 exponential: exp(x) = 2^p 2^r with
// p = fl(x log2 e) and r = x log2 e - p recovered exactly enough by two FMAs
// (the high part of log2 e is its float rounding, the low part the rest), then
// 2^r = 1 + r ln 2 to first order (|r| is below half an ulp of p).  Relative
// error within a few ulps over the plan's argument range.
__device__ __forceinline__ float exp_split(float x) {
  constexpr float L2E_HI = 1.44269502162933349609375f;  // log2(e) rounded to float
  constexpr float L2E_LO = 1.925963033500011e-8f;       // log2(e) - L2E_HI
  constexpr float LN2 = 0.693147180559945309f;
  const float p = x * L2E_HI;
  const float r = fmaf(x, L2E_LO, fmaf(x, L2E_HI, -p));
  return __builtin_amdgcn_exp2f(p) * fmaf(r, LN2, 1.0f);
}

template <typename T>
__device__ __forceinline__ T plan_exp(T x) {
  if constexpr (std::is_same<T, double>::value) return exp(x);
  else return exp_split(x);
}

// LDS layout of k_badmm_plan64, shared by host (sizing) and device.  Plan
// columns have stride sp (even, >= s): rows 2i, 2i + 1 of a column are one
// 8-byte (FP32) or 16-byte (FP64) word.
struct Plan64Lds {
  size_t A, B, X, w, xs, cs, cw, g, rp, ids, acc, red, up, total;
  __host__ __device__ Plan64Lds(int tsize, int sp, int m_max, int s, int d, int keepx, int R) {
    const size_t EP = (size_t)sp * m_max, dd = d > 0 ? d : 1;
    size_t o = 0;
    A = o; o = align16(o + EP * tsize);
    B = o; o = align16(o + EP * tsize);
    X = o; o = align16(o + (keepx ? EP : 0) * tsize);
    w = o; o = align16(o + (size_t)m_max * tsize);
    xs = o; o = align16(o + (size_t)m_max * dd * tsize);
    cs = o; o = align16(o + (size_t)s * dd * tsize);
    cw = o; o = align16(o + (size_t)s * tsize);
    g = o; o = align16(o + (size_t)s * tsize);
    rp = o; o = align16(o + (size_t)NT * tsize);
    ids = o; o = align16(o + (size_t)m_max * sizeof(int));
    acc = o; o = align16(o + (size_t)R * sizeof(double));
    red = o; o = align16(o + 8 * sizeof(double));
    up = o; o = align16(o + (R > s ? (size_t)4 * s * (dd + 1) : 0) * sizeof(double));
    total = o;
  }
};

template <typename T>
struct Vec2 {
  typedef float2 type;
};
template <>
struct Vec2<double> {
  typedef double2 type;
};

// The plan update of k_badmm_plan, restated for s <= 64 with fewer
// instructions per plan element:
//   * element passes (1 and 4) give each thread a fixed support row
//     k = tid % s and columns j = tid / s + c * (NT / s), so the element
//     offset k + j s runs contiguously over the workgroup (coalesced), the
//     centroid support, c_w[k] and the row normaliser sit in registers and no
//     (k, j) stepping is needed;
//   * pass 1 loads U columns' Z, Y (and stored costs) before using any;
//   * pass 2 gives each of four threads of a column a contiguous run of RQ
//     rows, read and written as 2-row LDS words and kept in registers between
//     the column sum and the update;
//   * pass 3 sums rows over all NT / s column classes in parallel.
// Every sum keeps a fixed order, so results are reproducible run to run.
// register budget: the lean passes are the bulk of the time and should keep
// five or more waves per SIMD (about the LDS limit for a typical object); the
// support-update and residual passes may use more registers
#ifndef D2MI_PLAN64_WAVES_PER_EU
#define D2MI_PLAN64_WAVES_PER_EU 5
#endif
template <typename T, int CM, int DIM, bool LEAN>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(LEAN ? D2MI_PLAN64_WAVES_PER_EU : 1)))
void k_badmm_plan64(PlanArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool F64 = std::is_same<T, double>::value;
  constexpr int U = PLAN_UNROLL;
  typedef typename Vec2<T>::type T2;
  const int s = a.s, d = DIM > 0 ? DIM : a.d, sp = a.sp, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const bool upd = !LEAN && a.upd, resid = !LEAN && a.resid, keepx = !LEAN && a.keepx;
  const bool first = !LEAN && a.first;
  const Plan64Lds L(sizeof(T), sp, a.m_max, s, d, keepx, a.R);
  T *A = (T *)(smem + L.A), *B = (T *)(smem + L.B), *Xs = (T *)(smem + L.X);
  T *wl = (T *)(smem + L.w), *xs = (T *)(smem + L.xs);
  T *cs = (T *)(smem + L.cs), *cwl = (T *)(smem + L.cw), *g = (T *)(smem + L.g);
  T *rp = (T *)(smem + L.rp);
  int *ids = (int *)(smem + L.ids);
  double *acc = (double *)(smem + L.acc), *red = (double *)(smem + L.red);
  double *up = (double *)(smem + L.up);

  const int gi = blockIdx.x;
  const int p0 = gi * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const double rho = a.scal[0];
  const T inv_rho = (T)a.scal[1];
  // element passes: fixed row ke, columns jb, jb + cpi, ...
  const int cpi = NT / s, ke = tid % s;
  const int jb = tid < cpi * s ? tid / s : 0x3fffffff;  // idle threads see no column
  // pass 2: column (tid >> 2) + 64 c, rows [q RQ, q RQ + RQ) of it
  const int q4 = tid & 3, RQ = (((s + 3) >> 2) + 1) & ~1, r0 = q4 * RQ;
  double r_obj = 0.0, r_pr = 0.0, r_dr = 0.0;
  for (int q = tid; q < a.R; q += NT) acc[q] = 0.0;
  int cur = -1;

  for (int p = p0; p < p1; ++p) {
    const int o = a.perm[p];
    const int l = a.label[o];
    const int m = a.p_str[o];
    const long long off = a.cum[o];
    const size_t base = (size_t)s * off;
    const bool reinit = first && (a.force_init || a.lsw[o] == 1);
    // requests first: weights, supports or bin ids of this object
    T wpre[PRE_W], xpre[PRE_X];
    int ipre[PRE_W];
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NT;
      if (q < m) {
        wpre[r] = a.w[off + q];
        if constexpr (CM == COST_SPARSE) ipre[r] = a.sym[off + q];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NT;
        if (q < m * d) xpre[r] = a.supp[(size_t)off * d + q];
      }
    }
    if (l != cur) {
      __syncthreads();  // every wave is done with the previous label
      if (cur >= 0)
        for (int q = tid; q < a.R; q += NT) {
          a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
          acc[q] = 0.0;
        }
      cur = l;
      for (int q = tid; q < s; q += NT) cwl[q] = a.cw[(size_t)l * s + q];
      if constexpr (CM == COST_EUCLID)
        for (int q = tid; q < s * d; q += NT) cs[q] = a.csupp[(size_t)l * s * d + q];
    }
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NT;
      if (q < m) {
        wl[q] = wpre[r];
        if constexpr (CM == COST_SPARSE) ids[q] = ipre[r];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NT;
        if (q < m * d) xs[q] = xpre[r];
      }
    }
    __syncthreads();  // also: every wave has finished the previous object's last pass
    const T cwk = cwl[ke];
    T csk[DIM > 0 ? DIM : 1];
    if constexpr (CM == COST_EUCLID && DIM > 0) {
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) csk[dd] = cs[ke * DIM + dd];
    }
    T *Zo = a.Z + base, *Yo = a.Y + base;  // this object's plan (offsets k + j s)
    // ground cost of (ke, j) before the division by rho; stored and table
    // costs come from the loads issued with the batch
    auto cost_k = [&](int j) -> T {
      if constexpr (CM == COST_EUCLID) {
        T c = 0;
        if constexpr (DIM > 0) {
#pragma unroll
          for (int dd = 0; dd < DIM; ++dd) {
            const T u = csk[dd] - xs[j * DIM + dd];
            c += u * u;
          }
        } else {
          for (int dd = 0; dd < d; ++dd) {
            const T u = cs[ke * d + dd] - xs[j * d + dd];
            c += u * u;
          }
        }
        return c;
      } else if constexpr (CM == COST_HIST) {
        return a.dist[ke + (size_t)j * s];
      } else if constexpr (CM == COST_SPARSE) {
        return a.dist[(size_t)ids[j] * a.V + ke];
      } else {
        return a.ctab[base + ke + j * s];
      }
    };

    // pass 1 -- step 1a: t = Z exp(-(C + Y)) + 1e-9 (:155-157)
    for (int j0 = jb; j0 < m; j0 += U * cpi) {
      T zv[U], yv[U], cv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int jc = min(j0 + u * cpi, m - 1);  // in bounds; results past m unused
        const int e = ke + jc * s;
        zv[u] = reinit ? (T)0 : Zo[e];
        yv[u] = first ? (T)0 : Yo[e];
        cv[u] = cost_k(jc);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * cpi;
        if (j < m) {
          const T z = reinit ? wl[j] * cwk : zv[u];
          const int li = ke + j * sp;
          A[li] = z * plan_exp<T>(-(scale_cost<T>(cv[u], rho, inv_rho) + yv[u])) + (T)1e-9;
          B[li] = yv[u];
        }
      }
    }
    __syncthreads();
    // pass 2 -- step 1b/1c: column sums (_dcnorm, :158), x = t / colsum * w
    // (:159); step 2a: z' = x exp(Y) + 1e-9 (:166)
    {
      const int nr = max(0, min(RQ, s - r0));  // rows of this thread in its column
      for (int j = tid >> 2; j < m; j += NT / 4) {
        T *colA = A + j * sp + r0;
        T part = 0;
        for (int r = 0; r < nr; r += 2) {
          const T2 v = *(const T2 *)(colA + r);
          part += v.x;
          if (r + 1 < nr) part += v.y;
        }
        part += __shfl_xor(part, 1, 64);
        part += __shfl_xor(part, 2, 64);
        const T wj = wl[j];
        const T scale = F64 ? (T)0 : wj / part;
        T *colB = B + j * sp + r0;
        for (int r = 0; r < nr; r += 2) {
          const T2 t = *(const T2 *)(colA + r);
          const T2 y = *(const T2 *)(colB + r);
          T2 x, zp;
          x.x = F64 ? (t.x / part) * wj : t.x * scale;
          x.y = F64 ? (t.y / part) * wj : t.y * scale;
          zp.x = x.x * plan_exp<T>(y.x) + (T)1e-9;
          zp.y = x.y * plan_exp<T>(y.y) + (T)1e-9;
          *(T2 *)(colA + r) = zp;  // a pad row past an odd s is written, never read
          if (keepx) {
            *(T2 *)(Xs + j * sp + r0 + r) = x;
          } else {
            T2 yx;
            yx.x = y.x + x.x;
            yx.y = y.y + x.y;
            *(T2 *)(colB + r) = yx;
          }
        }
      }
    }
    __syncthreads();
    // pass 3 -- step 2b: row sums (_drnorm, :169), by column class; step 5
    // partials on support-update iterations (:206-211), by wave
    {
      T part = 0;
      for (int j = jb; j < m; j += cpi) part += A[ke + j * sp];
      if (jb < cpi) rp[jb * s + ke] = part;
    }
    if (upd && lane < s) {  // wave wv: columns j = wv, wv + 4, ...
      double P[MAXD_FUSED];
      for (int dd = 0; dd < d; ++dd) P[dd] = 0.0;
      double rx = 0.0;
      for (int j = wv; j < m; j += 4) {
        const double x = (double)Xs[lane + j * sp];
        rx += x;  // row sums of X (:211)
        if constexpr (CM == COST_EUCLID)
          for (int dd = 0; dd < d; ++dd) P[dd] += x * (double)xs[j * d + dd];  // supp X^T (:206)
      }
      double *u = up + (size_t)(wv * s + lane) * (d + 1);
      for (int dd = 0; dd < d; ++dd) u[dd] = P[dd];
      u[d] = rx;
    }
    __syncthreads();
    // step 4: normalised row sums summed per label (:180-185)
    if (wv == 0) {
      T r = 0;
      if (lane < s) {
        for (int c = 0; c < cpi; ++c) r += rp[c * s + lane];
        g[lane] = F64 ? r : cwl[lane] / r;
      }
      const double rs = wave_sum(lane < s ? (double)r : 0.0);
      if (lane < s) acc[lane] += (double)r / rs;
      if (upd && lane < s) {
        for (int dd = 0; dd <= d; ++dd) {
          double v = 0.0;
          for (int w4 = 0; w4 < 4; ++w4) v += up[(size_t)(w4 * s + lane) * (d + 1) + dd];
          if (dd < d) acc[s + lane * d + dd] += v;
          else acc[s + s * d + lane] += v;
        }
      }
    }
    __syncthreads();
    // pass 4 -- step 2c + 3: z = z' / rowsum * c_w (:169-170); Y += X - Z (:175)
    {
      const T gk = g[ke];
      for (int j0 = jb; j0 < m; j0 += U * cpi) {
        T z0v[U], cv[U];
        if (resid) {  // step 6 (:300-305) and the exit objective: Z before this pass
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int jc = min(j0 + u * cpi, m - 1);
            z0v[u] = reinit ? wl[jc] * cwk : Zo[ke + jc * s];
            cv[u] = cost_k(jc);
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = j0 + u * cpi;
          if (j < m) {
            const int li = ke + j * sp, e = ke + j * s;
            const T z = F64 ? (A[li] / gk) * cwk : A[li] * gk;
            if (resid) {
              const T c = scale_cost<T>(cv[u], rho, inv_rho);
              const T x = Xs[li];
              r_obj += (double)c * (double)x;
              r_pr += fabs((double)x - (double)z);
              r_dr += fabs((double)z0v[u] - (double)z);
            }
            Zo[e] = z;
            Yo[e] = keepx ? B[li] + (Xs[li] - z) : B[li] - z;
            if (CM == COST_TABLE && upd) a.xout[base + e] = Xs[li];  // for the support GEMM
          }
        }
      }
    }
    if (resid) __syncthreads();  // pass 4 read wl / xs: keep them until all waves are done
  }
  __syncthreads();
  if (cur >= 0)
    for (int q = tid; q < a.R; q += NT) a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
  if (resid) {
    const double t0 = block_sum(r_obj, red);
    const double t1 = block_sum(r_pr, red);
    const double t2 = block_sum(r_dr, red);
    if (tid == 0) {
      a.resid_part[gi * 3 + 0] = t0;
      a.resid_part[gi * 3 + 1] = t1;
      a.resid_part[gi * 3 + 2] = t2;
    }
  }
}

// ------------------------------------ register-resident plan kernel (s <= 64)

// register budget of the lean passes: 6 waves per SIMD with 8 column slots;
// with 16, 4 for fused Euclidean and stored costs (about 101-128 registers,
// no spill: faster than 5 waves with 96 registers and 9-28 spilled) and 5 for
// histogram costs (their lean passes fit 96 registers without a spill)
#ifndef D2MI_PLANREG_WAVES_PER_EU
#define D2MI_PLANREG_WAVES_PER_EU(nc, cm) \
  ((nc) <= 8 ? 6 : ((cm) == COST_HIST || (cm) == COST_SPARSE) ? 5 : 4)
#endif
#ifndef D2MI_PLANREG_RESID_WAVES
// residual passes: 3 waves per SIMD for stored costs in workgroups of up to
// 256 threads (no spill), 4 otherwise (a 512-thread workgroup puts 2 waves on
// each SIMD, so 3 would hold one workgroup per CU instead of 2)
#define D2MI_PLANREG_RESID_WAVES(cm, ntb) ((cm) == COST_TABLE && (ntb) <= 256 ? 3 : 4)
#endif
#ifndef D2MI_PLANREG_SETPRIO
#define D2MI_PLANREG_SETPRIO 2  // wave priority while an object's loads are issued (0: none)
#endif
// the short phases every wave then waits for at a barrier run at raised
// priority too: the column sums, and wave 0's row sums
#ifndef D2MI_PLANREG_PRIO_P2
#define D2MI_PLANREG_PRIO_P2 2  // wave priority during the column sums (0: none)
#endif
#ifndef D2MI_PLANREG_PRIO_S4
#define D2MI_PLANREG_PRIO_S4 2  // wave priority of wave 0's row sums (0: none)
#endif
#ifndef D2MI_PLANREG_PRIO_P4
#define D2MI_PLANREG_PRIO_P4 2  // wave priority of the last pass's arithmetic (0: none)
#endif
#ifndef D2MI_PLANREG_MID_CAP
#define D2MI_PLANREG_MID_CAP 1  // FP32 support-update passes share the lean budget
#endif

// A uniform read of data the kernel never writes, through the constant
// address space, so it is a scalar load: it then waits on its own counter,
// not behind the vector stores of the previous object.
template <typename T>
__device__ __forceinline__ T const_load(const T *p) {
  return *(const __attribute__((address_space(4))) T *)p;
}

// Plan I/O of one object through a raw buffer resource on its base: every
// access is base + voff (the thread's first element, one VGPR) + soff (the
// column step, an SGPR), so no per-access address registers are live while
// a store is in flight.  Resource word 3 selects 32-bit data (gfx9 layout);
// the record count is left open: callers mask their lanes.
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes buf_res(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, 0x7fffffff, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(BufRes r, int voff, int soff) {
  if constexpr (std::is_same<T, double>::value)
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_store(T v, BufRes r, int voff, int soff) {
  if constexpr (std::is_same<T, double>::value) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
  }
}

// LDS layout of k_badmm_plan_reg: only t goes through LDS (for the column
// sums); column stride sp even, as for k_badmm_plan64.  Residual passes also
// keep the Z they started from (z0, same layout as t) for the dual residual,
// and two copies of the object's supports and bin ids (xs, ids), used by
// alternate objects.
struct PlanRegLds {
  size_t A, col, w, xs, cs, cw, g, rp, ids, acc, red, up, z0, total;
  // noA: t lives in the prefetch form's static staging array instead;
  // nt: threads per workgroup; resid: room for z0
  __host__ __device__ PlanRegLds(int tsize, int sp, int m_max, int s, int d, int R,
                                 bool noA = false, int nt = NT, bool resid = false) {
    const size_t dd = d > 0 ? d : 1;
    const int cpi = nt / s;
    size_t o = 0;
    A = o; o = align16(o + (noA ? 0 : (size_t)sp * m_max * tsize));
    col = o; o = align16(o + (size_t)m_max * tsize);
    w = o; o = align16(o + (size_t)m_max * tsize);
    xs = o; o = align16(o + (size_t)(resid ? 2 : 1) * m_max * dd * tsize);
    cs = o; o = align16(o + (size_t)s * dd * tsize);
    cw = o; o = align16(o + (size_t)s * tsize);
    g = o; o = align16(o + (size_t)s * tsize);
    rp = o; o = align16(o + (size_t)nt * tsize);
    ids = o; o = align16(o + (size_t)(resid ? 2 : 1) * m_max * sizeof(int));
    acc = o; o = align16(o + (size_t)R * sizeof(double));
    red = o; o = align16(o + (size_t)(nt / 64 > 8 ? nt / 64 : 8) * sizeof(double));
    up = o; o = align16(o + (R > s ? (size_t)cpi * s * (dd + 1) : 0) * sizeof(double));
    z0 = o; o = align16(o + (resid ? (size_t)sp * m_max * tsize : 0));
    total = o;
  }
};

// The plan update with each thread's plan elements held in registers from
// load to store, for s <= 64 and objects of at most NC * (NTB / s) supports,
// NTB threads per workgroup: 256, or 64 for small plans (the wave form: one
// wave per object, so every barrier below is a single wave's).
// Thread tid owns row k = tid % s of columns j = tid / s + c (NT / s),
// c < NC, so the element offsets k + j s are contiguous over the workgroup.
// Per object:
//   loads: all of the thread's Z, Y (and stored costs) are requested at the
//          top, with the weights and supports, so the object costs one HBM
//          round trip;
//   pass 1: t = Z exp(-(C + Y)) + 1e-9 (:155-157), kept, and written to LDS;
//   pass 2: four threads per column sum its t (_dcnorm, :158) into a column
//           scale w_j / colsum;
//   pass 3: x = t * scale (:159), z' = x exp(Y) + 1e-9 (:166), Y + X, and
//           per-thread partial row sums of z' (_drnorm, :169) and, on support
//           iterations, of X and X * supports (:206-211);
//   step 4: wave 0 adds the NT / s partials of each row in a fixed order;
//   pass 4: z = z' / rowsum * c_w (:169-170), Y += X - Z (:175), stored.
// Every sum has a fixed order: results are reproducible run to run.
//
// MODE 0 (lean): passes after the first with no support update or residual;
// MODE 1: the first pass and support-update passes; MODE 2: residual passes
// (anything).  X stays in registers until pass 4 only where pass 4 needs it:
// residuals, and FP64, whose Y update keeps the reference's Y + (X - Z).
//
// PF (lean FP32 passes): the next object's Z and Y are copied from HBM
// straight into LDS (global_load_lds, no registers) while the current object
// computes, so the workgroup's next round trip overlaps its own passes 3 and
// 4.  The copy is issued once pass 2 has read t (which lives in the Z staging
// array); the barriers after it wait on LDS only, so the copy stays in flight
// until the next object's start waits for it.
constexpr int PF_STAGE = 4608;  // staging floats per array: s * m_max and sp * m_max

// workgroup barrier that waits for this wave's LDS operations only (not for
// its outstanding vector-memory loads and stores); the empty asm statements
// keep the compiler from moving memory accesses across it
__device__ __forceinline__ void lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unconstrained (gfx9 encoding)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
// workgroup barrier after all of this wave's memory operations completed
__device__ __forceinline__ void full_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <typename T, int CM, int DIM, int NC, int MODE, bool PF = false, int NTB = NT>
__global__ __launch_bounds__(NTB)
__attribute__((amdgpu_waves_per_eu(
    PF ? 4  // the staging arrays (36 KB) allow 4 workgroups per CU
    : MODE == 0 || (MODE == 1 && D2MI_PLANREG_MID_CAP && !std::is_same<T, double>::value)
        ? D2MI_PLANREG_WAVES_PER_EU(NC, CM)
        : (MODE == 2 && !std::is_same<T, double>::value ? D2MI_PLANREG_RESID_WAVES(CM, NTB) : 1))))
void k_badmm_plan_reg(PlanArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool F64 = std::is_same<T, double>::value;
  static_assert(!PF || (MODE == 0 && !F64 && CM != COST_TABLE), "prefetch: lean FP32 passes");
  static_assert(!PF || NTB == NT, "prefetch: 256-thread workgroups");
  __shared__ __align__(16) T stZ[PF ? PF_STAGE : 1], stY[PF ? PF_STAGE : 1];
  constexpr int U = PLAN_UNROLL;
  typedef typename Vec2<T>::type T2;
  const int s = a.s, d = DIM > 0 ? DIM : a.d, sp = a.sp, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const bool upd = MODE >= 1 && a.upd, first = MODE >= 1 && a.first;
  const bool resid = MODE == 2 && a.resid;
  // x held until pass 4 in FP64 only (its Y update keeps the reference's
  // Y + (X - Z) order); FP32 residual passes take x for the residuals from
  // t (still in LDS from pass 1) times the column scale (from pass 2): the
  // product pass 3 formed, so the same value, without 16 more registers
  const bool keepx = MODE >= 1 && F64;
  const PlanRegLds L(sizeof(T), sp, a.m_max, s, d, a.R, PF, NTB, resid);
  T *A = PF ? stZ : (T *)(smem + L.A);
  T *z0l = (T *)(smem + L.z0);  // residual passes: Z before this pass
  T *csl = (T *)(smem + L.col);
  T *wl = (T *)(smem + L.w), *xs = (T *)(smem + L.xs);
  T *cs = (T *)(smem + L.cs), *cwl = (T *)(smem + L.cw), *g = (T *)(smem + L.g);
  T *rp = (T *)(smem + L.rp);
  int *ids = (int *)(smem + L.ids);
  double *acc = (double *)(smem + L.acc), *red = (double *)(smem + L.red);
  double *up = (double *)(smem + L.up);

  const int gi = blockIdx.x;
  const int p0 = gi * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const double rho = a.scal[0];
  const T inv_rho = (T)a.scal[1];
  const int cpi = NTB / s, ke_t = tid % s;
  const int jb_t = tid < cpi * s ? tid / s : 0x3fffffff;  // idle threads own no column
  const int q4 = tid & 3, RQ = (((s + 3) >> 2) + 1) & ~1, r0 = q4 * RQ;
  const int nr = max(0, min(RQ, s - r0));  // pass 2: rows of this thread in a column
  double r_obj = 0.0, r_pr = 0.0, r_dr = 0.0;
  for (int q = tid; q < a.R; q += NTB) acc[q] = 0.0;
  int cur = -1;
  // PF: object pq's Z and Y blocks (s m contiguous floats each) into stZ, stY,
  // 64 elements per wave instruction (the LDS address is wave-uniform)
  auto prefetch = [&](int pq) {
    if constexpr (PF) {
      const int oq = const_load(a.perm + pq);
      const int mq = const_load(a.p_str + oq);
      const size_t bq = (size_t)s * const_load(a.cum + oq);
      const int n = s * mq;
      for (int i = wv; i * 64 < n; i += NTB / 64) {
        const int e = i * 64 + lane;
        if (e < n) {
          __builtin_amdgcn_global_load_lds(a.Z + bq + e, stZ + i * 64, 4, 0, 0);
          __builtin_amdgcn_global_load_lds(a.Y + bq + e, stY + i * 64, 4, 0, 0);
        }
      }
    }
  };
  // odd passes walk the chunk backwards, so a pass starts on the objects the
  // previous pass wrote last, which are still in the memory-side cache
  const int np = p1 - p0, rev = a.rev;
  if (PF && np > 0) prefetch(rev ? p1 - 1 : p0);
  // with pm, each object's metadata arrives one object ahead
  // four adjacent scalar loads: one 16-byte load of the entry
  auto meta_load = [&](int q) -> ObjMeta {
    const int *e = (const int *)(a.pm + q);
    return ObjMeta{const_load(e), const_load(e + 1), const_load(e + 2), const_load(e + 3)};
  };
  ObjMeta mt_next{0, 0, 0, 0};
  if (a.pm && np > 0) mt_next = meta_load(rev ? p1 - 1 : p0);

  for (int i = 0; i < np; ++i) {
    const int p = rev ? p1 - 1 - i : p0 + i;
    // per-object copies of the thread's row and first column: without this
    // the compiler keeps every column slot's addresses live across objects
    int ke = ke_t, jb = jb_t;
    asm volatile("" : "+v"(ke), "+v"(jb));
    // residual passes: this object's supports and bin ids go to the copy the
    // previous object did not use, whose pass 4 may still read its own
    const int bo = resid ? (i & 1) : 0;
    T *xo = xs + (size_t)bo * a.m_max * (d > 0 ? d : 1);
    int *io = ids + bo * a.m_max;
    int o, l, m;
    long long off;
    bool switched;
    if (a.pm) {
      const ObjMeta mt = mt_next;
      if (i + 1 < np) mt_next = meta_load(rev ? p - 1 : p + 1);
      o = mt.o;
      l = mt.l;
      m = mt.m & (META_SWITCHED - 1);
      switched = (mt.m & META_SWITCHED) != 0;
      off = (unsigned int)mt.off;
    } else {
      o = const_load(a.perm + p);
      l = const_load(a.label + o);
      m = const_load(a.p_str + o);
      off = const_load(a.cum + o);
      switched = first && const_load(a.lsw + o) == 1;
    }
    const size_t base = (size_t)s * off;
    // the flag comes with the metadata, so no dependent load waits before
    // the object's plan loads
    const bool reinit = first && (a.force_init || switched);
    // this object's plan: element (k, j) at base + k + j s; the thread's
    // column slot c is voff + c * cstride bytes from the base
    const BufRes rz = buf_res(a.Z + base), ry = buf_res(a.Y + base);
    const int voff = (ke + jb * s) * (int)sizeof(T), cstride = cpi * s * (int)sizeof(T);
    // requests first: plan elements, weights, supports or bin ids, at raised
    // wave priority (D2MI_PLANREG_SETPRIO) so they leave before other waves' work
    if (D2MI_PLANREG_SETPRIO) __builtin_amdgcn_s_setprio(D2MI_PLANREG_SETPRIO);
    T zr[NC], yr[NC], xr[NC], cr[NC];
    if constexpr (PF) {
      full_sync();  // every wave's share of this object's copy has landed
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int j = jb + c * cpi;
        zr[c] = (T)0;
        yr[c] = (T)0;
        if (j < m) {
          zr[c] = stZ[ke + j * s];
          yr[c] = stY[ke + j * s];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int j = jb + c * cpi;
        zr[c] = (T)0;
        yr[c] = (T)0;
        if (j < m) {
          if (!reinit) zr[c] = buf_load<T>(rz, voff, c * cstride);
          if (!first) yr[c] = buf_load<T>(ry, voff, c * cstride);
          if constexpr (CM == COST_TABLE) cr[c] = a.ctab[base + ke + j * s];
        }
      }
    }
    T wpre[PRE_W], xpre[PRE_X];
    int ipre[PRE_W];
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NTB;
      if (q < m) {
        wpre[r] = a.w[off + q];
        if constexpr (CM == COST_SPARSE) ipre[r] = a.sym[off + q];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NTB;
        if (q < m * d) xpre[r] = a.supp[(size_t)off * d + q];
      }
    }
    if (D2MI_PLANREG_SETPRIO) __builtin_amdgcn_s_setprio(0);
    if (l != cur) {
      __syncthreads();  // every wave is done with the previous label
      if (cur >= 0)
        for (int q = tid; q < a.R; q += NTB) {
          a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
          acc[q] = 0.0;
        }
      cur = l;
      for (int q = tid; q < s; q += NTB) cwl[q] = a.cw[(size_t)l * s + q];
      if constexpr (CM == COST_EUCLID)
        for (int q = tid; q < s * d; q += NTB) cs[q] = a.csupp[(size_t)l * s * d + q];
    }
#pragma unroll
    for (int r = 0; r < PRE_W; ++r) {
      const int q = tid + r * NTB;
      if (q < m) {
        wl[q] = wpre[r];
        if constexpr (CM == COST_SPARSE) io[q] = ipre[r];
      }
    }
    if constexpr (CM == COST_EUCLID) {
#pragma unroll
      for (int r = 0; r < PRE_X; ++r) {
        const int q = tid + r * NTB;
        if (q < m * d) xo[q] = xpre[r];
      }
    }
    __syncthreads();  // also: every wave has finished the previous object's last pass
    const T cwk = cwl[ke];
    T csk[DIM > 0 ? DIM : 1];
    if constexpr (CM == COST_EUCLID && DIM > 0) {
#pragma unroll
      for (int dd = 0; dd < DIM; ++dd) csk[dd] = cs[ke * DIM + dd];
    }
    // ground cost of (ke, j) before the division by rho (stored costs: cr)
    auto cost_k = [&](int j) -> T {
      if constexpr (CM == COST_EUCLID) {
        T c = 0;
        if constexpr (DIM > 0) {
#pragma unroll
          for (int dd = 0; dd < DIM; ++dd) {
            const T u = csk[dd] - xo[j * DIM + dd];
            c += u * u;
          }
        } else {
          for (int dd = 0; dd < d; ++dd) {
            const T u = cs[ke * d + dd] - xo[j * d + dd];
            c += u * u;
          }
        }
        return c;
      } else if constexpr (CM == COST_HIST) {
        return a.dist[ke + (size_t)j * s];
      } else if constexpr (CM == COST_SPARSE) {
        return a.dist[(size_t)io[j] * a.V + ke];
      } else {
        return a.ctab[base + ke + j * s];
      }
    };

    // pass 1 -- step 1a: t = Z exp(-(C + Y)) + 1e-9 (:155-157)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int j = jb + c * cpi;
      if (j < m) {
        const T cost = CM == COST_TABLE ? cr[c] : cost_k(j);
        const T z = reinit ? wl[j] * cwk : zr[c];
        if (resid) z0l[ke + j * sp] = z;  // read back by this thread in pass 4
        const T t = z * plan_exp<T>(-(scale_cost<T>(cost, rho, inv_rho) + yr[c])) + (T)1e-9;
        A[ke + j * sp] = t;
        zr[c] = t;
      }
    }
    __syncthreads();
    // pass 2 -- step 1b: column sums (_dcnorm, :158) -> the column's scale
    if (D2MI_PLANREG_PRIO_P2) __builtin_amdgcn_s_setprio(D2MI_PLANREG_PRIO_P2);
    for (int j = tid >> 2; j < m; j += NTB / 4) {
      const T *colA = A + j * sp + r0;
      T part = 0;
      for (int r = 0; r < nr; r += 2) {
        const T2 v = *(const T2 *)(colA + r);
        part += v.x;
        if (r + 1 < nr) part += v.y;
      }
      part += __shfl_xor(part, 1, 64);
      part += __shfl_xor(part, 2, 64);
      if (q4 == 0) csl[j] = F64 ? part : wl[j] / part;
    }
    if (D2MI_PLANREG_PRIO_P2) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    // t (in stZ) is read: the next object's copy can start
    if (PF && i + 1 < np) prefetch(rev ? p - 1 : p + 1);
    // pass 3 -- step 1c: x = t / colsum * w (:159); step 2a: z' = x exp(Y) +
    // 1e-9 (:166); partial row sums (_drnorm, :169) and step 5 partials
    {
      T rpart = 0;
      double P[DIM > 0 ? DIM : MAXD_FUSED], rx = 0.0;
      if (upd)
        for (int dd = 0; dd < d; ++dd) P[dd] = 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int j = jb + c * cpi;
        if (j < m) {
          const T x = F64 ? (zr[c] / csl[j]) * wl[j] : zr[c] * csl[j];
          const T zp = x * plan_exp<T>(yr[c]) + (T)1e-9;
          rpart += zp;
          zr[c] = zp;
          if (keepx) xr[c] = x;
          else yr[c] += x;
          if (upd) {
            rx += (double)x;  // row sums of X (:211)
            if constexpr (CM == COST_EUCLID)
              for (int dd = 0; dd < d; ++dd) P[dd] += (double)x * (double)xo[j * d + dd];  // (:206)
            if constexpr (CM == COST_TABLE) a.xout[base + ke + j * s] = x;  // for the support GEMM
          }
        }
      }
      if (jb < cpi) {
        rp[jb * s + ke] = rpart;
        if (upd) {
          double *u = up + (size_t)(jb * s + ke) * (d + 1);
          for (int dd = 0; dd < d; ++dd) u[dd] = P[dd];
          u[d] = rx;
        }
      }
    }
    if constexpr (PF) lds_sync();
    else __syncthreads();
    // step 4: normalised row sums summed per label (:180-185)
    if (wv == 0) {
      if (D2MI_PLANREG_PRIO_S4) __builtin_amdgcn_s_setprio(D2MI_PLANREG_PRIO_S4);
      T r = 0;
      if (lane < s) {
        for (int c = 0; c < cpi; ++c) r += rp[c * s + lane];
        g[lane] = F64 ? r : cwl[lane] / r;
      }
      const double rs = wave_sum(lane < s ? (double)r : 0.0);
      if (lane < s) acc[lane] += (double)r / rs;
      if (upd && lane < s) {
        for (int dd = 0; dd <= d; ++dd) {
          double v = 0.0;
          for (int c = 0; c < cpi; ++c) v += up[(size_t)(c * s + lane) * (d + 1) + dd];
          if (dd < d) acc[s + lane * d + dd] += v;
          else acc[s + s * d + lane] += v;
        }
      }
      if (D2MI_PLANREG_PRIO_S4) __builtin_amdgcn_s_setprio(0);
    }
    if constexpr (PF) lds_sync();
    else __syncthreads();
    // pass 4 -- step 2c + 3: z = z' / rowsum * c_w (:169-170); Y += X - Z (:175).
    // Residual passes read, besides registers, LDS the next object leaves
    // alone until this pass is done everywhere: this thread's own t and Z0
    // slots, csl and g (rewritten after two more barriers), this object's copy
    // of the supports and bin ids (the next object fills the other copy) --
    // so no barrier follows this pass.  An object's terms are summed in T and
    // added to the FP64 totals once.
    {
      if (D2MI_PLANREG_PRIO_P4) __builtin_amdgcn_s_setprio(D2MI_PLANREG_PRIO_P4);
      const T gk = g[ke];
      T ob = 0, pr = 0, dr = 0;
#pragma unroll
      for (int c0 = 0; c0 < NC; c0 += U) {
        T z0v[U], cv[U];
        if (resid) {  // step 6 (:300-305) and the exit objective: Z before this pass
#pragma unroll
          for (int u = 0; u < U && c0 + u < NC; ++u) {
            const int j = jb + (c0 + u) * cpi;
            if (j < m) {
              z0v[u] = z0l[ke + j * sp];
              cv[u] = cost_k(j);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < U && c0 + u < NC; ++u) {
          const int c = c0 + u, j = jb + c * cpi;
          if (j < m) {
            const T z = F64 ? (zr[c] / gk) * cwk : zr[c] * gk;
            if (resid) {
              const T cost = scale_cost<T>(cv[u], rho, inv_rho);
              const T xv = keepx ? xr[c] : A[ke + j * sp] * csl[j];
              ob += cost * xv;
              pr += fabs(xv - z);
              dr += fabs(z0v[u] - z);
            }
            yr[c] = keepx ? yr[c] + (xr[c] - z) : yr[c] - z;
            zr[c] = z;
          }
        }
      }
      // the stores, from registers that stay untouched while they drain:
      // every value is materialised first (an instruction that rewrites a
      // register an in-flight store reads must wait for that store)
      if (resid) {
        r_obj += (double)ob;
        r_pr += (double)pr;
        r_dr += (double)dr;
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) asm volatile("" : "+v"(zr[c]), "+v"(yr[c]));
      if (D2MI_PLANREG_PRIO_P4) __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (jb + c * cpi < m) {
          buf_store<T>(zr[c], rz, voff, c * cstride);
          buf_store<T>(yr[c], ry, voff, c * cstride);
        }
      }
    }
  }
  __syncthreads();
  if (cur >= 0)
    for (int q = tid; q < a.R; q += NTB) a.rec[(size_t)(gi + cur) * a.R + q] = acc[q];
  if (resid) {
    const double t0 = block_sum<NTB>(r_obj, red);
    const double t1 = block_sum<NTB>(r_pr, red);
    const double t2 = block_sum<NTB>(r_dr, red);
    if (tid == 0) {
      a.resid_part[gi * 3 + 0] = t0;
      a.resid_part[gi * 3 + 1] = t1;
      a.resid_part[gi * 3 + 2] = t2;
    }
  }
}

// ------------------------------------------------------- tall register form
//
// Plans with many rows and few columns -- sparse-histogram phases, whose
// centroids have s = V bins (up to 1024) while objects hold 5-40 nonzeros --
// updated with one thread per row k and the object's columns j < MC in
// registers: a workgroup of ceil(s / 64) waves holds the whole s x m plan, so
// each object costs one HBM round trip (Z and Y read and written once), where
// the HBM-scratch form sweeps its plan several times.  Row sums (_drnorm,
// :169) are thread-local; column sums (_dcnorm, :158) reduce over each wave
// by shuffles and across waves through LDS, in a fixed order (results are
// reproducible).  Objects come in size classes of 8 columns (MC = 8 ... 40):
// the column loops have no per-column branches -- columns past m load the
// last column again (in bounds), are computed, masked out of the row sums and
// never stored -- so the compiler batches each group's loads.
constexpr int TALL_NT = 1024;  // rows per workgroup at most (s <= 1024)
constexpr int TALL_NW = TALL_NT / 64;
constexpr int TALL_MC32 = 40, TALL_MC64 = 16;  // most columns held in registers

// MODE 0: lean (no residual), MODE 1: first pass, MODE 2: residual pass
// (anything).  Histogram metrics have no support update.
template <typename T, int MC, int MODE>
__global__ __launch_bounds__(TALL_NT) void k_badmm_plan_tall(PlanArgs<T> a) {
  constexpr bool F64 = std::is_same<T, double>::value;
  __shared__ T colpart[TALL_NW][MC];
  __shared__ T scl[MC];
  __shared__ double red[3][TALL_NW];
  const int s = a.s, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nw = blockDim.x >> 6;
  const bool row = tid < s;
  const int k = row ? tid : s - 1;  // lanes past row s repeat the last row, unstored
  const bool first = MODE >= 1 && a.first, resid = MODE == 2 && a.resid;
  const int gi = blockIdx.x;
  const int p0 = gi * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const double rho = a.scal[0];
  const T inv_rho = (T)a.scal[1];
  const int voff = k * (int)sizeof(T), cstride = s * (int)sizeof(T);
  // ground-cost table dist[id * V + k]: the row in the VGPR, the bin in an SGPR
  const BufRes rdist = buf_res(a.dist);
  const int vstride = a.V * (int)sizeof(T);
  double acc = 0.0, r_obj = 0.0, r_pr = 0.0, r_dr = 0.0;
  int cur = -1;

  for (int p = p0; p < p1; ++p) {
    const int o = const_load(a.perm + p);
    const int l = const_load(a.label + o);
    const int m = const_load(a.p_str + o);
    const long long off = const_load(a.cum + o);
    const size_t base = (size_t)s * off;
    const bool reinit = first && (a.force_init || const_load(a.lsw + o) == 1);
    if (l != cur) {  // no load in here: a branch join would wait on the plan loads
      if (cur >= 0 && row) a.rec[(size_t)(gi + cur) * a.R + k] = acc;
      acc = 0.0;
      cur = l;
    }
    // the object's loads at raised wave priority, as in the register form
    if (D2MI_PLANREG_SETPRIO) __builtin_amdgcn_s_setprio(D2MI_PLANREG_SETPRIO);
    const T cwk = a.cw[(size_t)l * s + k];
    const BufRes rz = buf_res(a.Z + base), ry = buf_res(a.Y + base);
    T zr[MC], yr[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      const int cc = min(c, m - 1);
      zr[c] = reinit ? (T)0 : buf_load<T>(rz, voff, cc * cstride);
      yr[c] = first ? (T)0 : buf_load<T>(ry, voff, cc * cstride);
    }
    // the object's bin ids in one vector load, lane j holding column j's
    // (columns past m repeat the last); each column reads its id out with
    // v_readlane instead of a scalar load that the next one waits behind
    const int idv = a.sym[off + min(lane, m - 1)];
    if (D2MI_PLANREG_SETPRIO) __builtin_amdgcn_s_setprio(0);

    // pass 1 -- step 1a: t = Z exp(-(C + Y)) + 1e-9 (:155-157), C = dist[id_j][k],
    // 8 columns at a time (a scheduling barrier between groups keeps every
    // column's cost load from being hoisted, which would spill the plan)
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      const int cc = min(c, m - 1);
      const int id = __builtin_amdgcn_readlane(idv, c);
      const T cost = buf_load<T>(rdist, voff, id * vstride);
      const T z = reinit ? const_load(a.w + off + cc) * cwk : zr[c];
      zr[c] = z * plan_exp<T>(-(scale_cost<T>(cost, rho, inv_rho) + yr[c])) + (T)1e-9;
      if (c % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // every wave is done with the previous object's colpart, scl
    // step 1b: column sums (_dcnorm, :158): wave totals, then over the waves
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      const T v = wave_sum_to63<T>(row ? zr[c] : (T)0);
      if (lane == 63) colpart[wv][c] = v;
    }
    __syncthreads();
    if (tid < MC) {
      T sc = 0;
      if (tid < m) {
        T cs = 0;
        for (int w = 0; w < nw; ++w) cs += colpart[w][tid];
        sc = F64 ? cs : a.w[off + tid] / cs;
      }
      scl[tid] = sc;
    }
    __syncthreads();
    // pass 3 -- step 1c: x = t / colsum * w (:159); step 2a: z' = x exp(Y) +
    // 1e-9 (:166); the row sum (_drnorm, :169) is this thread's
    T rk = 0;
    T xr[MODE == 2 || F64 ? MC : 1];
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      T x;
      if constexpr (F64) x = c < m ? (zr[c] / scl[c]) * const_load(a.w + off + min(c, m - 1)) : 0.0;
      else x = zr[c] * scl[c];
      const T zp = x * plan_exp<T>(yr[c]) + (T)1e-9;
      if (c < m) rk += zp;
      zr[c] = zp;
      if constexpr (MODE == 2 || F64) xr[c] = x;
      else yr[c] += x;
      if (c % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    // step 4: this object's normalised row sums into the label's sums (:180-185)
    const double rs = wave_sum_to63<double>(row ? (double)rk : 0.0);
    if (lane == 63) red[0][wv] = rs;
    __syncthreads();
    double rtot = 0.0;
    for (int w = 0; w < nw; ++w) rtot += red[0][w];
    if (row) acc += (double)rk / rtot;
    // pass 4 -- step 2c + 3: z = z' / rowsum * c_w (:169-170); Y += X - Z (:175)
    const T gk = F64 ? rk : cwk / rk;
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      const T z = F64 ? (zr[c] / gk) * cwk : zr[c] * gk;
      if constexpr (MODE == 2 || F64) {
        if (resid && c < m) {  // step 6 (:300-305): Z before this pass, cost again
          const T z0 = reinit ? const_load(a.w + off + c) * cwk : buf_load<T>(rz, voff, c * cstride);
          const T cost = scale_cost<T>(
              buf_load<T>(rdist, voff, __builtin_amdgcn_readlane(idv, c) * vstride), rho, inv_rho);
          if (row) {
            r_obj += (double)cost * (double)xr[c];
            r_pr += fabs((double)xr[c] - (double)z);
            r_dr += fabs((double)z0 - (double)z);
          }
        }
        yr[c] = yr[c] + (xr[c] - z);
      } else {
        yr[c] = yr[c] - z;
      }
      zr[c] = z;
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) asm volatile("" : "+v"(zr[c]), "+v"(yr[c]));
    // the column offsets again from opaque copies of the stride and the
    // column count: otherwise the compiler keeps MC offsets live in scalar
    // registers from the loads to here, and spills them
    int cst = cstride, ms = m;
    asm volatile("" : "+s"(cst), "+s"(ms));
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      if (c < ms && row) {
        buf_store<T>(zr[c], rz, voff, c * cst);
        buf_store<T>(yr[c], ry, voff, c * cst);
      }
    }
  }
  if (cur >= 0 && row) a.rec[(size_t)(gi + cur) * a.R + k] = acc;
  if (resid) {
    const double v0 = wave_sum(r_obj), v1 = wave_sum(r_pr), v2 = wave_sum(r_dr);
    __syncthreads();
    if (lane == 0) {
      red[0][wv] = v0;
      red[1][wv] = v1;
      red[2][wv] = v2;
    }
    __syncthreads();
    if (tid < 3) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += red[tid][w];
      a.resid_part[gi * 3 + tid] = t;
    }
  }
}

// ------------------------------------------------------------- finalize

constexpr int MAXB = 8;  // object-size buckets of one update

template <typename T>
struct FinArgs {
  int K, s, d, R, G, upd, resid, iter;
  // records of bucket b: objects [start[b], start[b + 1]) of the order by
  // (bucket, label), chunks of chunk[b], record (g, l) at rec_off[b] + g + l;
  // label l of bucket b spans segk[b K + l] .. segk[b K + l + 1]
  int nbk;
  int chunk[MAXB], start[MAXB];
  long long rec_off[MAXB];
  const int *__restrict__ segk;
  const double *__restrict__ rec;
  const double *__restrict__ resid_part;
  double *__restrict__ cw64;
  T *__restrict__ cwT;
  double *__restrict__ cs64;
  T *__restrict__ csT;
  double *__restrict__ resid_out;  // [iters][3]
  double *__restrict__ rxout;      // COST_TABLE support passes: X row sums per label [K*s]
  int *__restrict__ flags;
};

template <typename T>
__global__ __launch_bounds__(NT) void k_badmm_finalize(FinArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  double *sums = (double *)smem;
  double *red = sums + a.R;
  double *parts = red + 8;  // [P][R]
  const int l = blockIdx.x, tid = threadIdx.x, s = a.s, d = a.d;
  // P threads per record column: each sums a contiguous share of every
  // bucket's workgroups (buckets in order), then the shares are added in
  // order -- a fixed association, and P times fewer dependent round trips to
  // records written by workgroups on other XCDs
  const int P = a.R >= NT ? 1 : NT / a.R;
  for (int t = tid; t < P * a.R; t += NT) {
    const int q = t % a.R, part = t / a.R;
    double v = 0.0;
    for (int b = 0; b < a.nbk; ++b) {
      const int lo = a.segk[b * a.K + l] - a.start[b], hi = a.segk[b * a.K + l + 1] - a.start[b];
      if (hi <= lo) continue;
      const int glo = lo / a.chunk[b], n = (hi - 1) / a.chunk[b] - glo + 1;
      const int g0 = glo + (int)((long long)n * part / P);
      const int g1 = glo + (int)((long long)n * (part + 1) / P);
      const double *r = a.rec + (size_t)(a.rec_off[b] + l) * a.R + q;  // record (g, l) at g + l
#pragma unroll 4
      for (int g = g0; g < g1; ++g) v += r[(size_t)g * a.R];
    }
    parts[t] = v;
  }
  __syncthreads();
  for (int q = tid; q < a.R; q += NT) {
    double v = 0.0;
    for (int p = 0; p < P; ++p) v += parts[p * a.R + q];
    sums[q] = v;
  }
  __syncthreads();
  double tot = 0.0;
  for (int k = tid; k < s; k += NT) tot += sums[k];
  tot = block_sum(tot, red);
  if (!(tot > 0.0)) {
    if (tid == 0) atomicOr(a.flags, FLAG_NUMERIC);
  } else {
    for (int k = tid; k < s; k += NT) {
      const double v = sums[k] / tot;  // _dcnorm over the centroid (:190)
      a.cw64[(size_t)l * s + k] = v;
      a.cwT[(size_t)l * s + k] = (T)v;
    }
  }
  if (a.upd && a.rxout)
    for (int k = tid; k < s; k += NT) a.rxout[(size_t)l * s + k] = sums[s + s * d + k];
  if (a.upd) {
    for (int q = tid; q < s * d; q += NT) {
      const int k = q / d;
      const double rx = sums[s + s * d + k];
      if (!(rx > 0.0)) {
        atomicOr(a.flags, FLAG_NUMERIC);
        continue;
      }
      const double v = sums[s + q] / rx;  // _dirms (:219-221)
      a.cs64[(size_t)l * s * d + q] = v;
      a.csT[(size_t)l * s * d + q] = (T)v;
    }
  }
  if (a.resid && l == 0) {
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int g = tid; g < a.G; g += NT) v += a.resid_part[g * 3 + c];
      v = block_sum(v, red);
      if (tid == 0) a.resid_out[a.iter * 3 + c] = v;
    }
  }
}

// ------------------------------------------------------------ stored costs

template <typename T>
struct TabArgs {
  int N, s, d, chunk;
  const int *__restrict__ perm;  // label-sorted order
  const int *__restrict__ p_str;
  const long long *__restrict__ cum;
  const int *__restrict__ label;
  const T *__restrict__ supp;    // Euclidean supports [col*d]
  const int *__restrict__ sym;   // word ids [col]
  const T *__restrict__ vocab;   // word embeddings [V*d]
  const T *__restrict__ csupp;   // centroid supports [K*s*d]
  T *__restrict__ ctab;          // out, plan layout
};

constexpr int TAB_KC = 16;  // centroid rows per thread

// Word embeddings in FP32: C of object i, element (k, j) = W[id_j * (K s) +
// label_i * s + k], where W = pdist2(vocab, all centroid supports) comes from
// the matrix cores (pdist2.hip); a thread per element, 64 consecutive k of
// one word are one contiguous read.
__global__ __launch_bounds__(NT) void k_word_gather(int N, int s, int Ks,
                                                    const int *__restrict__ p_str,
                                                    const long long *__restrict__ cum,
                                                    const int *__restrict__ label,
                                                    const int *__restrict__ sym,
                                                    const float *__restrict__ W,
                                                    float *__restrict__ ctab) {
  for (int o = blockIdx.x; o < N; o += gridDim.x) {
    const int m = p_str[o], l = label[o];
    const long long off = cum[o];
    float *C = ctab + (size_t)s * off;
    for (int e = threadIdx.x; e < s * m; e += NT) {
      const int j = e / s, k = e - (e / s) * s;
      C[e] = W[(size_t)sym[off + j] * Ks + (size_t)l * s + k];
    }
  }
}

// The same gather for s % 4 == 0, 16 bytes per thread: an object's block is
// m x s/4 float4s (four consecutive k of one word), and each thread requests
// the word ids of its WG_U float4s, then their rows, then stores them, so an
// object costs one round trip per WG_U x NT float4s rather than one per NT
// elements.
constexpr int WG_U = 4;
__global__ __launch_bounds__(NT) void k_word_gather4(int N, int s, int Ks,
                                                     const int *__restrict__ p_str,
                                                     const long long *__restrict__ cum,
                                                     const int *__restrict__ label,
                                                     const int *__restrict__ sym,
                                                     const float *__restrict__ W,
                                                     float *__restrict__ ctab) {
  const int s4 = s >> 2;
  for (int o = blockIdx.x; o < N; o += gridDim.x) {
    const int m = p_str[o], l = label[o];
    const long long off = cum[o];
    float4 *C = (float4 *)(ctab + (size_t)s * off);
    const float *Wl = W + (size_t)l * s;
    const int n4 = s4 * m;
    for (int f0 = threadIdx.x; f0 < n4; f0 += NT * WG_U) {
      int id[WG_U];
#pragma unroll
      for (int u = 0; u < WG_U; ++u) {
        const int f = f0 + u * NT;
        id[u] = f < n4 ? sym[off + f / s4] : 0;
      }
      float4 v[WG_U];
#pragma unroll
      for (int u = 0; u < WG_U; ++u) {
        const int f = f0 + u * NT;
        if (f < n4) v[u] = *(const float4 *)(Wl + (size_t)id[u] * Ks + 4 * (f % s4));
      }
#pragma unroll
      for (int u = 0; u < WG_U; ++u) {
        const int f = f0 + u * NT;
        if (f < n4) C[f] = v[u];
      }
    }
  }
}

// calculate_distmat (centroid_util.c:8-36): _dpdist2 for Euclidean phases,
// _dpdist2_sym (blas_like64.c:210-220) for word embeddings, by direct
// differences.  A workgroup walks a chunk of the label-sorted order with its
// label's centroid supports in LDS (dimension-major, so a thread's 16 rows
// are contiguous); thread <-> (column j, chunk of 16 rows).
template <typename T, bool WORD>
__global__ __launch_bounds__(NT) void k_badmm_cost_table(TabArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *cs = (T *)smem;  // [d][s] of the current label
  const int s = a.s, d = a.d, tid = threadIdx.x;
  const int nkc = (s + TAB_KC - 1) / TAB_KC;
  const int p0 = blockIdx.x * a.chunk, p1 = min(a.N, p0 + a.chunk);
  int cur = -1;
  for (int p = p0; p < p1; ++p) {
    const int o = a.perm[p], l = a.label[o], m = a.p_str[o];
    const long long off = a.cum[o];
    if (l != cur) {
      __syncthreads();
      for (int q = tid; q < s * d; q += NT) {
        const int k = q / d, t = q - k * d;
        cs[t * s + k] = a.csupp[(size_t)l * s * d + q];
      }
      __syncthreads();
      cur = l;
    }
    T *C = a.ctab + (size_t)s * off;
    for (int u = tid; u < m * nkc; u += NT) {
      const int j = u / nkc, kc = (u - j * nkc) * TAB_KC;
      const T *x = WORD ? a.vocab + (size_t)a.sym[off + j] * d : a.supp + (size_t)(off + j) * d;
      T acc[TAB_KC];
#pragma unroll
      for (int r = 0; r < TAB_KC; ++r) acc[r] = 0;
      for (int t = 0; t < d; ++t) {
        const T xv = x[t];
        const T *crow = cs + t * s + kc;
#pragma unroll
        for (int r = 0; r < TAB_KC; ++r) {
          if (kc + r < s) {
            const T df = crow[r] - xv;
            acc[r] += df * df;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < TAB_KC; ++r)
        if (kc + r < s) C[kc + r + (size_t)j * s] = acc[r];
    }
  }
}

template <typename T>
struct SuppArgs {
  int N, s, d, chunk;
  const int *__restrict__ perm;
  const int *__restrict__ p_str;
  const long long *__restrict__ cum;
  const int *__restrict__ label;
  const T *__restrict__ supp;
  const int *__restrict__ sym;
  const T *__restrict__ vocab;
  const T *__restrict__ X;  // plan layout, from the support-update pass
  double *__restrict__ rec;  // [(G + K) * s * d]
};

constexpr int SUP_KPT = 16;  // rows per thread: s <= 4 * SUP_KPT
constexpr int SUP_TT = 128;  // dimensions per tile (two per lane)

// Step 5 numerators: sum over the label's objects of supp_i X_i^T
// (centroid_Bregman.c:199-212, one dgemm per object) or of X_i[k, j] *
// vocab[id_j] (:229-247).  Records per (workgroup, label) at g + l as in the
// plan kernel.  Per dimension tile, each object's X and its supports (or
// word vectors) are staged through LDS in blocks of columns (128 in FP32,
// 64 in FP64), all loads of a block in flight at once; a thread owns
// dimensions t0 + lane and t0 + 64 + lane and rows k = wave + 4 i, with FP64
// accumulators.
template <typename T, bool WORD>
__global__ __launch_bounds__(NT) void k_badmm_supp_num(SuppArgs<T> a) {
  constexpr int SUP_JB = sizeof(T) == 8 ? 64 : 128;
  __shared__ T Xl[4 * SUP_KPT * SUP_JB];  // [k][j] of the block
  __shared__ T Vl[SUP_JB * SUP_TT];       // [j][t] of the block
  const int s = a.s, d = a.d, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int g = blockIdx.x;
  const int p0 = g * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const size_t R = (size_t)s * d;
  for (int t0 = 0; t0 < d; t0 += SUP_TT) {
    double acc[SUP_KPT][2];
    int cur = -1;
    auto flush = [&](int l) {
#pragma unroll
      for (int i = 0; i < SUP_KPT; ++i) {
        const int k = wv + 4 * i;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int t = t0 + h * 64 + lane;
          if (k < s && t < d) a.rec[(size_t)(g + l) * R + (size_t)k * d + t] = acc[i][h];
        }
      }
    };
    for (int p = p0; p < p1; ++p) {
      const int o = a.perm[p], l = a.label[o], m = a.p_str[o];
      const long long off = a.cum[o];
      if (l != cur) {
        if (cur >= 0) flush(cur);
#pragma unroll
        for (int i = 0; i < SUP_KPT; ++i) acc[i][0] = acc[i][1] = 0.0;
        cur = l;
      }
      const T *Xo = a.X + (size_t)s * off;
      for (int j0 = 0; j0 < m; j0 += SUP_JB) {
        const int jn = min(SUP_JB, m - j0);
        __syncthreads();  // the previous block is consumed
        for (int e = tid; e < s * jn; e += NT) {  // X block, column-major in HBM
          const int j = e / s, k = e - (e / s) * s;
          Xl[k * SUP_JB + j] = Xo[(size_t)(j0 + j) * s + k];
        }
        for (int e = tid; e < jn * SUP_TT; e += NT) {
          const int j = e / SUP_TT, t = e - (e / SUP_TT) * SUP_TT;
          const T *x = WORD ? a.vocab + (size_t)a.sym[off + j0 + j] * d
                            : a.supp + (size_t)(off + j0 + j) * d;
          Vl[j * SUP_TT + t] = t0 + t < d ? x[t0 + t] : (T)0;
        }
        __syncthreads();
        for (int j = 0; j < jn; ++j) {
          const double v0 = (double)Vl[j * SUP_TT + lane], v1 = (double)Vl[j * SUP_TT + 64 + lane];
#pragma unroll
          for (int i = 0; i < SUP_KPT; ++i) {
            const int k = wv + 4 * i;
            if (k < s) {
              const double xk = (double)Xl[k * SUP_JB + j];
              acc[i][0] += xk * v0;
              acc[i][1] += xk * v1;
            }
          }
        }
      }
    }
    if (cur >= 0) flush(cur);
  }
}

// Step 5 numerators on the matrix cores (FP32): per label, the split-K GEMM
// sum_i X_i (s x m_i) W_i (m_i x d), W_i the object's supports or word
// vectors (centroid_Bregman.c:199-212, :229-247).  A workgroup walks its chunk
// of the label-sorted order in blocks of SM_JB columns of one object (zero
// past m); wave w owns output column blocks w, w + 4, w + 8 of 32 (d <= 384
// per sweep) and both 32-row blocks (s <= 64), accumulating in FP32 MFMA
// tiles; a label change flushes the tiles into the FP64 record (g, l).  The
// next block's X and W are loaded into registers while the current one is
// multiplied.  MFMA 32x32x2 f32 operands: lane l holds A[l % 32][l / 32] and
// B[l / 32][l % 32]; D[i][j] sits in lane j + 32 ((i / 4) % 2), register
// (i % 4) + 4 (i / 8).
constexpr int SM_JB = 16;            // columns per block
constexpr int SM_XP = SM_JB + 1;     // LDS row stride of the X block
constexpr int SM_TD = 384;           // output columns per sweep (3 blocks of 32 per wave)
constexpr int SM_WP = SM_TD + 4;     // LDS row stride of the W block

template <bool WORD>
__global__ __launch_bounds__(NT) void k_badmm_supp_mfma(SuppArgs<float> a) {
  __shared__ float Xl[64 * SM_XP];
  __shared__ float Wl[SM_JB * SM_WP];
  typedef float v16f __attribute__((ext_vector_type(16)));
  const int s = a.s, d = a.d, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int g = blockIdx.x;
  const int p0 = g * a.chunk, p1 = min(a.N, p0 + a.chunk);
  const size_t R = (size_t)s * d;
  for (int t0 = 0; t0 < d; t0 += SM_TD) {
    v16f acc[2][3];
    auto zero = [&]() {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 3; ++cb)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[rb][cb][r] = 0.f;
    };
    // record (g, l) as a buffer: lane offset in a VGPR, tile offsets in SGPRs
    auto flush = [&](int l) {
      const BufRes rr = buf_res(a.rec + (size_t)(g + l) * R);
      const int kl = 4 * (lane >> 5), tl = t0 + wv * 32 + (lane & 31);
      const int voff = (kl * d + tl) * (int)sizeof(double);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 3; ++cb)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int kr = rb * 32 + (r & 3) + 8 * (r >> 2);  // row less the lane part
            if (kr + kl < s && tl + 128 * cb < d)
              buf_store<double>((double)acc[rb][cb][r], rr, voff,
                                (kr * d + 128 * cb) * (int)sizeof(double));
          }
    };
    // register prefetch of one block: lane k (row) of X columns
    // j = wave + 4 u, and W rows j = wave + 4 u at dimensions t0 + lane + 64 q
    constexpr int XU = SM_JB / 4, WQ = SM_TD / 64;
    float xr[XU], wr[XU][WQ];
    auto fetch = [&](int o, int j0) {
      const int m = const_load(a.p_str + o);
      const long long off = const_load(a.cum + o);
      int ln = lane;
      asm volatile("" : "+v"(ln));  // keep per-lane offsets from being hoisted
      const float *Xo = a.X + (size_t)s * off;
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int j = j0 + wv + 4 * u;
        xr[u] = (ln < s && j < m) ? Xo[(size_t)j * s + ln] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int j = j0 + wv + 4 * u;
        const bool live = j < m;
        const float *w = a.vocab;
        if (live) {
          if constexpr (WORD) w = a.vocab + (size_t)const_load(a.sym + off + j) * d;
          else w = a.supp + (size_t)(off + j) * d;
        }
#pragma unroll
        for (int q = 0; q < WQ; ++q) {
          const int t = t0 + ln + 64 * q;
          wr[u][q] = (live && t < d) ? w[t] : 0.f;
        }
      }
    };
    auto stage = [&]() {
#pragma unroll
      for (int u = 0; u < XU; ++u) Xl[lane * SM_XP + wv + 4 * u] = xr[u];
#pragma unroll
      for (int u = 0; u < XU; ++u)
#pragma unroll
        for (int q = 0; q < WQ; ++q) Wl[(wv + 4 * u) * SM_WP + lane + 64 * q] = wr[u][q];
    };
    // the chunk's blocks in order: (object p, first column j0)
    int p = p0, j0 = 0, cur = -1;
    if (p < p1) fetch(const_load(a.perm + p), 0);
    zero();
    while (p < p1) {
      const int o = const_load(a.perm + p), l = const_load(a.label + o), m = const_load(a.p_str + o);
      if (l != cur) {
        if (cur >= 0) flush(cur);
        zero();
        cur = l;
      }
      __syncthreads();  // the previous block is consumed
      stage();
      __syncthreads();
      // advance, and request the next block before multiplying this one
      int pn = p, jn = j0 + SM_JB;
      if (jn >= m) {
        ++pn;
        jn = 0;
      }
      if (pn < p1) fetch(const_load(a.perm + pn), jn);
#pragma unroll
      for (int kk = 0; kk < SM_JB; kk += 2) {
        const int kq = kk + (lane >> 5);
        float av[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = Xl[(rb * 32 + (lane & 31)) * SM_XP + kq];
#pragma unroll
        for (int cb = 0; cb < 3; ++cb) {
          const float bv = Wl[kq * SM_WP + (wv + 4 * cb) * 32 + (lane & 31)];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[rb], bv, acc[rb][cb], 0, 0, 0);
        }
      }
      p = pn;
      j0 = jn;
    }
    if (cur >= 0) flush(cur);
    __syncthreads();  // LDS reuse by the next sweep
  }
}

// ---- step 5 numerators over packed columns, copied straight into LDS
//
// The same numerators as k_badmm_supp_mfma, with the work laid out once per
// update: each label's columns (its objects in label-sorted order, each
// object's columns in order) are packed into blocks of SD_JB, the label's
// last block padded, the labels' blocks back to back.  For block b,
// desc[32 b + c] is column c as a plan column q (X at s q; -1: padding, read
// from a zero buffer) and desc[32 b + 16 + c] the row of its weight vector
// (the word id, or q for supports; 0 for padding); blab[b] is its label,
// bs[l] label l's first block and bs[K] the number of blocks.
constexpr int SD_JB = 16, SD_NT = 256, SD_DSLOTS = 5;
constexpr int SD_TPW_MAX = 5;  // 32 x 32 tiles per wave: 320 output columns per sweep

// cum[i] = sum of the first i objects' column counts in the order perm: each
// block of NT objects scanned (k_cumw_local, block totals to bt), then the
// totals of the blocks before added (k_cumw_carry)
__global__ __launch_bounds__(NT) void k_cumw_local(const int *__restrict__ perm,
                                                   const int *__restrict__ p_str, int n,
                                                   long long *__restrict__ cum,
                                                   long long *__restrict__ bt) {
  __shared__ long long buf[NT];
  const int tid = threadIdx.x, i = blockIdx.x * NT + tid;
  buf[tid] = i < n ? (long long)p_str[perm[i]] : 0;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {  // inclusive Hillis-Steele scan
    const long long v = tid >= off ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] += v;
    __syncthreads();
  }
  if (i < n) cum[i + 1] = buf[tid];
  if (i == 0) cum[0] = 0;
  if (tid == NT - 1) bt[blockIdx.x] = buf[NT - 1];
}
__global__ __launch_bounds__(NT) void k_cumw_carry(int n, const long long *__restrict__ bt,
                                                   long long *__restrict__ cum) {
  __shared__ long long red[NT];
  const int tid = threadIdx.x, i = blockIdx.x * NT + tid;
  long long t = 0;
  for (int b = tid; b < (int)blockIdx.x; b += NT) t += bt[b];
  red[tid] = t;
  __syncthreads();
  for (int h = NT / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (i < n) cum[i + 1] += red[0];
}
// bs[l]: first block of label l (labels' column counts from cum at the label
// segments seg of perm, rounded up to whole blocks)
__global__ void k_supp_blocks(int K, const int *__restrict__ seg, const long long *__restrict__ cum,
                              int *__restrict__ bs) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    long long run = 0;
    for (int l = 0; l < K; ++l) {
      bs[l] = (int)run;
      run += (cum[seg[l + 1]] - cum[seg[l]] + SD_JB - 1) / SD_JB;
    }
    bs[K] = (int)run;
  }
}
// one wave per object (perm position p): its columns' descriptors
template <bool WORD>
__global__ __launch_bounds__(64) void k_supp_desc(const int *__restrict__ perm,
                                                  const int *__restrict__ label,
                                                  const int *__restrict__ p_str,
                                                  const long long *__restrict__ ocum,
                                                  const int *__restrict__ sym,
                                                  const int *__restrict__ seg,
                                                  const long long *__restrict__ cum,
                                                  const int *__restrict__ bs,
                                                  int *__restrict__ desc) {
  const int p = blockIdx.x, o = perm[p], l = label[o], m = p_str[o];
  const long long q0 = ocum[o];
  const long long pos = (long long)SD_JB * bs[l] + (cum[p] - cum[seg[l]]);
  for (int j = threadIdx.x; j < m; j += 64) {
    const long long x = pos + j, b = x / SD_JB;
    const int c = (int)(x - b * SD_JB), q = (int)(q0 + j);
    desc[32 * b + c] = q;
    desc[32 * b + 16 + c] = WORD ? sym[q] : q;
  }
}
// one workgroup per label: its padding columns and its blocks' label
__global__ __launch_bounds__(64) void k_supp_pad(const int *__restrict__ seg,
                                                 const long long *__restrict__ cum,
                                                 const int *__restrict__ bs,
                                                 int *__restrict__ desc, int *__restrict__ blab) {
  const int l = blockIdx.x;
  const long long cols = cum[seg[l + 1]] - cum[seg[l]];
  const long long x0 = (long long)SD_JB * bs[l] + cols, x1 = (long long)SD_JB * bs[l + 1];
  for (long long x = x0 + threadIdx.x; x < x1; x += 64) {
    const long long b = x / SD_JB;
    const int c = (int)(x - b * SD_JB);
    desc[32 * b + c] = -1;
    desc[32 * b + 16 + c] = 0;
  }
  for (int b = bs[l] + threadIdx.x; b < bs[l + 1]; b += 64) blab[b] = l;
}

struct SuppDmaArgs {
  int s, d, G;
  const int *__restrict__ desc;
  const int *__restrict__ blab;
  const int *__restrict__ nblk;  // bs[K]
  const float *__restrict__ X;   // plan layout, from the support-update pass
  const float *__restrict__ W;   // word vectors or supports, rows of d
  const float *__restrict__ zeros;  // 64 zero floats
  double *__restrict__ rec;      // record (g, l) at g + l, s x d each
};

// Workgroup g takes an equal run of blocks (so equal work).  Per sweep of
// 64 TPW output columns, block i goes through three LDS stages: its
// descriptors are copied into one of five slots four blocks ahead, its X
// columns and weight rows two blocks ahead (X: wave w copies columns 4 w ..
// 4 w + 3, 16 bytes per lane; W: wave w copies pieces w + 4 u of the
// block's flattened 16 x 64 TPW rows), so blocks i + 1 and i + 2 are in
// flight while the matrix cores multiply block i.  Wave w owns 32 x 32
// output tiles f = w TPW + u (rows 32 (f % 2), columns 32 (f / 2)), FP32
// MFMA accumulators; a label change flushes them into the FP64 record
// (g, l).  Needs s % 4 == 0, s <= 64 and d % 4 == 0: rows past s and
// columns past d are read from row 0 or column 0 (finite) and never stored.
template <int TPW, int NS>
__global__ __launch_bounds__(SD_NT, NS == 2 ? 3 : 2) void k_badmm_supp_dma(SuppDmaArgs a) {
  static_assert(NS == 2 || NS == 3, "two or three LDS stages");
  constexpr int TD = 64 * TPW, XST = SD_JB * 64, STG = XST + SD_JB * TD;
  __shared__ __align__(16) float St[NS * STG];  // stage i % NS at St + (i % NS) STG
  __shared__ __align__(16) int Dsc[SD_DSLOTS][32];
  typedef float v16f __attribute__((ext_vector_type(16)));
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s = a.s, d = a.d, g = blockIdx.x;
  const size_t R = (size_t)s * d;
  const int nblk = const_load(a.nblk);
  const int nbw = (nblk + a.G - 1) / a.G;
  const int b0 = min(nblk, g * nbw), nb = min(nblk, b0 + nbw) - b0;
  if (nb <= 0) return;
  // the helpers below are forced inline: a call would keep acc in scratch
  v16f acc[TPW];
  auto zero = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < TPW; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[u][r] = 0.f;
  };
  // record (g, l) as a buffer: the lane's offset in one VGPR, made opaque so
  // the 16 TPW store addresses are not hoisted out of the block loop
  auto flush = [&](int l, int t0) __attribute__((always_inline)) {
    const BufRes rr = buf_res(a.rec + (size_t)(g + l) * R);
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int kl = 4 * (ln >> 5), tl = t0 + (ln & 31);
    const int voff = (kl * d + tl) * (int)sizeof(double);
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int f = w * TPW + u, tb = (f >> 1) * 32;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int kr = (f & 1) * 32 + (r & 3) + 8 * (r >> 2);  // row less the lane part
        if (kr + kl < s && tl + tb < d)
          buf_store<double>((double)acc[u][r], rr, voff, (kr * d + tb) * (int)sizeof(double));
      }
    }
  };
  auto stage_of = [&](int i) __attribute__((always_inline)) -> float * { return St + (i % NS) * STG; };
  auto issue_desc = [&](int j) __attribute__((always_inline)) {  // block b0 + j: wave w copies descriptor ints 8 w .. 8 w + 7
    if (lane < 8) lds_copy4(a.desc + (size_t)(b0 + j) * 32 + 8 * w + lane, &Dsc[j % SD_DSLOTS][8 * w]);
  };
  auto issue_data = [&](int j, int t0) __attribute__((always_inline)) {
    float *S = stage_of(j);
    const int *D = Dsc[j % SD_DSLOTS];
    {
      const int c = 4 * w + (lane >> 4), r4 = 4 * (lane & 15);
      const int q = D[c];
      const float *src = q >= 0 ? a.X + (size_t)s * q + (r4 < s ? r4 : 0) : a.zeros + r4;
      lds_copy16(src, S + 4 * w * 64);
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int f = (w + 4 * u) * 64 + lane;  // 16-byte piece of the flattened block
      const int r = f / (TD / 4), t = t0 + 4 * (f - r * (TD / 4));
      const float *src = a.W + (size_t)D[16 + r] * d + (t < d ? t : 0);
      lds_copy16(src, S + XST + (w + 4 * u) * 256);
    }
  };
  auto products = [&](int i) __attribute__((always_inline)) {
    const float *Xs = stage_of(i), *Ws = Xs + XST;
#pragma unroll
    for (int kk = 0; kk < SD_JB; kk += 2) {
      const int kq = kk + (lane >> 5);
      const float a0 = Xs[kq * 64 + (lane & 31)], a1 = Xs[kq * 64 + 32 + (lane & 31)];
#pragma unroll
      for (int u = 0; u < TPW; ++u) {
        const int f = w * TPW + u;
        const float bv = Ws[kq * TD + (f >> 1) * 32 + (lane & 31)];
        acc[u] = __builtin_amdgcn_mfma_f32_32x32x2f32((f & 1) ? a1 : a0, bv, acc[u], 0, 0, 0);
      }
    }
  };
  // Step i issues block i + NS + 1's descriptors, then block i + NS - 1's
  // data.  Block i's copies have landed once at most the copies issued after
  // them are outstanding -- with three stages, block i + 1's data and block
  // i + 3's descriptors; with two, none (one counter orders them; stores
  // issued in between only make the wait stricter) -- and then a barrier,
  // after which every wave is also done with block i - 1's stage and
  // descriptor slot, refilled below.
  constexpr int PER_BLOCK = 1 + TPW;  // data copies per wave per block
  auto wait_block = [&](int i) __attribute__((always_inline)) {
    asm volatile("" ::: "memory");
    if (NS == 3 && i + 3 < nb) __builtin_amdgcn_s_waitcnt(0x0070 | (PER_BLOCK + 1));  // vmcnt(n) lgkmcnt(0)
    else if (NS == 3 && i + 1 < nb) __builtin_amdgcn_s_waitcnt(0x0070 | PER_BLOCK);
    else __builtin_amdgcn_s_waitcnt(0x0070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
  };
  static_assert(PER_BLOCK + 1 <= 15, "vmcnt in the low field");
  for (int t0 = 0; t0 < d; t0 += TD) {
    zero();
    int cur = -1;
    issue_desc(0);
    if (nb > 1) issue_desc(1);
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0x0070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    for (int j = 1 - NS; j < 0; ++j) {  // the steps before block 0's
      if (j + NS + 1 < nb) issue_desc(j + NS + 1);
      if (j + NS - 1 < nb) issue_data(j + NS - 1, t0);
    }
    for (int i = 0; i < nb; ++i) {
      wait_block(i);
      if (i + NS + 1 < nb) issue_desc(i + NS + 1);
      if (i + NS - 1 < nb) issue_data(i + NS - 1, t0);
      const int l = const_load(a.blab + b0 + i);
      if (l != cur) {
        if (cur >= 0) flush(cur, t0);
        zero();
        cur = l;
      }
      products(i);
    }
    flush(cur, t0);
    __syncthreads();  // stages and descriptor slots are reused by the next sweep
  }
}

// new supports = numerators / X row sums (_dirms, :219-221), records summed
// in workgroup order
template <typename T>
__global__ __launch_bounds__(NT) void k_badmm_supp_fin(int s, int d, int chunk,
                                                       const int *__restrict__ seg,
                                                       const int *__restrict__ bs, int G,
                                                       const double *__restrict__ rec,
                                                       const double *__restrict__ rx,
                                                       double *__restrict__ cs64,
                                                       T *__restrict__ csT, int *flags) {
  // grid (ceil(s d / NT), K): one thread per (label, support coordinate); each
  // sums its label's workgroup records in workgroup order.  Records of
  // k_badmm_supp_dma (bs: its label block starts; G workgroups of equal runs
  // of blocks), else of chunks of chunk objects over the label segments seg.
  const int l = blockIdx.y;
  int glo, ghi;
  if (bs) {
    const int nbw = (bs[gridDim.y] + G - 1) / G;
    glo = bs[l] / nbw;
    ghi = (bs[l + 1] - 1) / nbw;
  } else {
    glo = seg[l] / chunk;
    ghi = (seg[l + 1] - 1) / chunk;
  }
  const size_t R = (size_t)s * d;
  const size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  if (q < R) {
    double v = 0.0;
#pragma unroll 4
    for (int g = glo; g <= ghi; ++g) v += rec[(size_t)(g + l) * R + q];
    const double r = rx[(size_t)l * s + q / d];
    if (!(r > 0.0)) {
      atomicOr(flags, FLAG_NUMERIC);
      return;
    }
    const double c = v / r;
    cs64[(size_t)l * R + q] = c;
    csT[(size_t)l * R + q] = (T)c;
  }
}

// ---------------------------------------------------------------- setup

template <typename T>
struct SetupArgs {
  int N, s, d, V, m_max, per_block;
  const int *__restrict__ p_str;
  const long long *__restrict__ cum;
  const int *__restrict__ label;
  const char *__restrict__ lsw;
  int force_init, write_state;
  const T *__restrict__ w;
  const T *__restrict__ supp;
  const int *__restrict__ sym;
  const T *__restrict__ dist;
  const T *__restrict__ cw;
  const T *__restrict__ csupp;
  T *__restrict__ Z;
  T *__restrict__ Y;
  const T *__restrict__ ctab;
  double *__restrict__ part;  // [gridDim.x]
};

// Ground-cost sum (rho, :109), Z re-initialised as w_i[j] * c_w[label][k]
// for switched objects (:116-124), Y = 0 (:128).
template <typename T, int CM>
__global__ __launch_bounds__(NT) void k_badmm_setup(SetupArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int s = a.s, d = a.d, tid = threadIdx.x;
  T *xs = (T *)smem;
  T *cs = xs + (size_t)a.m_max * (d > 0 ? d : 1);
  int *ids = (int *)(cs + (size_t)s * (d > 0 ? d : 1));
  double *red = (double *)(smem + align16((size_t)((char *)(ids + a.m_max) - (char *)smem)));
  double csum = 0.0;
  const int o0 = blockIdx.x * a.per_block, o1 = min(a.N, o0 + a.per_block);
  for (int o = o0; o < o1; ++o) {
    const int l = a.label[o], m = a.p_str[o];
    const long long off = a.cum[o];
    __syncthreads();
    if constexpr (CM == COST_EUCLID) {
      for (int q = tid; q < m * d; q += NT) xs[q] = a.supp[(size_t)off * d + q];
      for (int q = tid; q < s * d; q += NT) cs[q] = a.csupp[(size_t)l * s * d + q];
    }
    if constexpr (CM == COST_SPARSE)
      for (int q = tid; q < m; q += NT) ids[q] = a.sym[off + q];
    __syncthreads();
    const size_t base = (size_t)s * off;
    const bool init = a.force_init || a.lsw[o] == 1;
    for (int e = tid; e < s * m; e += NT) {
      const int j = e / s, k = e - j * s;
      const T c = cost_raw<T, CM>(k, j, s, d, a.V, cs, xs, ids, a.dist, a.ctab + base);
      csum += fabs((double)c);
      if (a.write_state) {  // the engine does this inside its first pass instead
        if (init) a.Z[base + e] = a.w[off + j] * a.cw[(size_t)l * s + k];
        a.Y[base + e] = (T)0;
      }
    }
  }
  csum = block_sum(csum, red);
  if (tid == 0) a.part[blockIdx.x] = csum;
}

// Ground-cost sum of a stored cost table (rho, :109): the table holds every
// object's s x m block back to back, so the sum is a flat reduction, 16 bytes
// (FP32) or 32 bytes (FP64) per thread per step, one FP64 partial per block.
template <typename T>
__global__ __launch_bounds__(NT) void k_abs_sum(const T *__restrict__ c, size_t n,
                                                double *__restrict__ part) {
  __shared__ double red[8];
  double v = 0.0;
  const size_t nq = n / 4;
  for (size_t q = blockIdx.x * (size_t)NT + threadIdx.x; q < nq; q += (size_t)gridDim.x * NT) {
    if constexpr (std::is_same<T, float>::value) {
      const float4 x = ((const float4 *)c)[q];
      v += (fabs((double)x.x) + fabs((double)x.y)) + (fabs((double)x.z) + fabs((double)x.w));
    } else {
      const double2 x = ((const double2 *)c)[2 * q], y = ((const double2 *)c)[2 * q + 1];
      v += (fabs(x.x) + fabs(x.y)) + (fabs(y.x) + fabs(y.y));
    }
  }
  if (blockIdx.x == 0)
    for (size_t i = nq * 4 + threadIdx.x; i < n; i += NT) v += fabs((double)c[i]);
  v = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Ground-cost sum of a Euclidean phase (rho, :109) from each object's mean
// support point mu_i and spread sum_j |x_j - mu_i|^2 (computed at upload):
// sum_k sum_j |c_k - x_j|^2 = m_i sum_k |c_k - mu_i|^2 + s spread_i, every
// term non-negative.  One thread per object, one FP64 partial per block.
__global__ __launch_bounds__(NT) void k_badmm_rho_euclid(int N, int s, int d,
                                                         const int *__restrict__ p_str,
                                                         const int *__restrict__ label,
                                                         const double *__restrict__ mean,
                                                         const double *__restrict__ spread,
                                                         const double *__restrict__ cs64,
                                                         double *__restrict__ part) {
  __shared__ double red[8];
  double v = 0.0;
  for (int o = blockIdx.x * NT + threadIdx.x; o < N; o += gridDim.x * NT) {
    const double *mu = mean + (size_t)o * d;
    const double *c = cs64 + (size_t)label[o] * s * d;
    double a = 0.0;
    for (int k = 0; k < s; ++k)
      for (int dd = 0; dd < d; ++dd) {
        const double u = c[k * d + dd] - mu[dd];
        a += u * u;
      }
    v += (double)p_str[o] * a + (double)s * spread[o];
  }
  v = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Ground-cost sums of a histogram phase (rho, :109) without its s x m costs.
// Sparse histograms: object i's cost rows are dist[id_j * V + k], k < s, so
// its sum is sum_j L[id_j] with L[v] = sum_{k<s} |dist[v V + k]|.  Dense
// histograms: cost (k, j) is dist[k + j s], so the sum is sum_{j<m} L[j] with
// L[j] = sum_{k<s} |dist[k + j s]|.  FP64, from the FP64 table.
__global__ __launch_bounds__(NT) void k_hist_line_sums(const double *__restrict__ dist, int V,
                                                       int s, int nlines, int sparse,
                                                       double *__restrict__ L) {
  const int v = blockIdx.x * NT + threadIdx.x;
  if (v >= nlines) return;
  double a = 0.0;
  for (int k = 0; k < s; ++k) a += fabs(sparse ? dist[(size_t)v * V + k] : dist[k + (size_t)v * s]);
  L[v] = a;
}

__global__ __launch_bounds__(NT) void k_badmm_rho_hist(int N, const int *__restrict__ p_str,
                                                       const long long *__restrict__ cum,
                                                       const int *__restrict__ sym, int sparse,
                                                       const double *__restrict__ L,
                                                       double *__restrict__ part) {
  __shared__ double red[8];
  double v = 0.0;
  for (int o = blockIdx.x * NT + threadIdx.x; o < N; o += gridDim.x * NT) {
    const int m = p_str[o];
    const long long off = cum[o];
    double a = 0.0;
    for (int j = 0; j < m; ++j) a += L[sparse ? sym[off + j] : j];
    v += a;
  }
  v = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_badmm_rho(const double *__restrict__ part, int nparts, double rhoCoeff,
                            double denom, double *__restrict__ scal, int *flags) {
  __shared__ double red[8];
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += NT) v += part[i];
  v = block_sum(v, red);
  if (threadIdx.x == 0) {
    const double rho = rhoCoeff * v / denom;
    scal[0] = rho;
    scal[1] = 1.0 / rho;
    if (!(rho > 0.0)) atomicOr(flags, FLAG_NUMERIC);
  }
}

// --------------------------------------------- stable counting sort by label

__global__ __launch_bounds__(NT) void k_label_hist(const int *__restrict__ label, int N, int K,
                                                   int *__restrict__ blkcnt, int *flags) {
  extern __shared__ int hist[];
  for (int l = threadIdx.x; l < K; l += NT) hist[l] = 0;
  __syncthreads();
  const int o = blockIdx.x * NT + threadIdx.x;
  if (o < N) {
    const int l = label[o];
    if (l < 0 || l >= K) atomicOr(flags, FLAG_BADLABEL);
    else atomicAdd(&hist[l], 1);
  }
  __syncthreads();
  for (int l = threadIdx.x; l < K; l += NT) blkcnt[(size_t)blockIdx.x * K + l] = hist[l];
}

// one block per label: exclusive scan over blocks; totals into seg[l]
__global__ __launch_bounds__(NT) void k_label_scan(int nb, int K, int *__restrict__ blkcnt,
                                                   int *__restrict__ total) {
  __shared__ int buf[NT];
  __shared__ int carry;
  const int l = blockIdx.x;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += NT) {
    const int b = b0 + threadIdx.x;
    const int v = b < nb ? blkcnt[(size_t)b * K + l] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {  // inclusive Hillis-Steele scan
      const int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (b < nb) blkcnt[(size_t)b * K + l] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == NT - 1) carry += buf[NT - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) total[l] = carry;
}

__global__ void k_label_seg(int K, const int *__restrict__ total, int *__restrict__ seg,
                            int *flags) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int run = 0;
    for (int l = 0; l < K; ++l) {
      seg[l] = run;
      if (total[l] == 0) atomicOr(flags, FLAG_EMPTY);
      run += total[l];
    }
    seg[K] = run;
  }
}

__global__ __launch_bounds__(NT) void k_label_scatter(const int *__restrict__ label, int N, int K,
                                                      const int *__restrict__ blkoff,
                                                      const int *__restrict__ seg,
                                                      int *__restrict__ perm) {
  __shared__ int lab[NT];
  const int o = blockIdx.x * NT + threadIdx.x;
  lab[threadIdx.x] = o < N ? label[o] : -1;
  __syncthreads();
  if (o < N) {
    const int l = lab[threadIdx.x];
    if (l < 0 || l >= K) return;
    int rank = 0;
    for (int t = 0; t < (int)threadIdx.x; ++t) rank += lab[t] == l;
    perm[seg[l] + blkoff[(size_t)blockIdx.x * K + l] + rank] = o;
  }
}

// sort key of the per-bucket order: bucket(m) * K + label, bucket b the
// first with m <= lim[b]
struct BucketLims {
  int n, lim[MAXB];
};
__global__ __launch_bounds__(NT) void k_bucket_key(const int *__restrict__ label,
                                                   const int *__restrict__ p_str, int N, int K,
                                                   BucketLims bl, int *__restrict__ key) {
  const int o = blockIdx.x * NT + threadIdx.x;
  if (o >= N) return;
  const int m = p_str[o];
  int b = 0;
  while (b + 1 < bl.n && m > bl.lim[b]) ++b;
  key[o] = b * K + label[o];
}

// ------------------------------------------ plain one-thread-per-object form

template <typename T>
struct PlainArgs {
  int N, s, d, V, upd, keepx;
  const int *__restrict__ p_str;
  const long long *__restrict__ cum;
  const int *__restrict__ label;
  const T *__restrict__ w;
  const T *__restrict__ supp;
  const int *__restrict__ sym;
  const T *__restrict__ dist;
  const T *__restrict__ cw;
  const T *__restrict__ csupp;
  const double *__restrict__ scal;
  T *__restrict__ Z;
  T *__restrict__ Y;
  T *__restrict__ X;       // [s*col] scratch
  double *__restrict__ acc;  // [K * s * (d + 2)]: weights, supports, X row sums
  const T *__restrict__ ctab;   // COST_TABLE costs
  const T *__restrict__ vocab;  // word embeddings: support j is vocab[sym[j]]
};

// The CPU loops of centroid_Bregman.c:155-212 with one thread owning one
// object, state in global memory and per-label sums by FP64 atomics.
template <typename T, int CM>
__global__ void k_badmm_plain(PlainArgs<T> a) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= a.N) return;
  const int s = a.s, d = a.d, l = a.label[o], m = a.p_str[o];
  const long long off = a.cum[o];
  const size_t base = (size_t)s * off;
  const double rho = a.scal[0];
  const T inv_rho = (T)a.scal[1];
  const T *cs = a.csupp + (size_t)l * s * d;
  const T *xs = a.supp + (size_t)off * d;
  const int *ids = a.sym + off;
  for (int j = 0; j < m; ++j) {
    T csum = 0;
    for (int k = 0; k < s; ++k) {
      const size_t e = base + k + (size_t)j * s;
      const T c = scale_cost<T>(cost_raw<T, CM>(k, j, s, d, a.V, cs, xs, ids, a.dist,
                                                a.ctab + base), rho, inv_rho);
      const T t = a.Z[e] * dexp<T>(-(c + a.Y[e])) + (T)1e-9;
      a.X[e] = t;
      csum += t;
    }
    const T wj = a.w[off + j];
    for (int k = 0; k < s; ++k) {
      const size_t e = base + k + (size_t)j * s;
      const T x = (a.X[e] / csum) * wj;
      a.X[e] = x;
      a.Z[e] = x * dexp<T>(a.Y[e]) + (T)1e-9;
    }
  }
  const size_t stride = (size_t)s * (d + 2);
  double *acc = a.acc + (size_t)l * stride;
  double rsum = 0.0;
  for (int k = 0; k < s; ++k) {
    T r = 0;
    for (int j = 0; j < m; ++j) r += a.Z[base + k + (size_t)j * s];
    rsum += (double)r;
  }
  for (int k = 0; k < s; ++k) {
    T r = 0;
    for (int j = 0; j < m; ++j) r += a.Z[base + k + (size_t)j * s];
    const T ck = a.cw[(size_t)l * s + k];
    for (int j = 0; j < m; ++j) {
      const size_t e = base + k + (size_t)j * s;
      const T z = (a.Z[e] / r) * ck;
      a.Z[e] = z;
      a.Y[e] += a.X[e] - z;
    }
    atomicAdd(&acc[k], (double)r / rsum);
    if (a.upd) {
      double rx = 0.0;
      for (int j = 0; j < m; ++j) {
        const double x = (double)a.X[base + k + (size_t)j * s];
        rx += x;
        const T *xj = a.vocab ? a.vocab + (size_t)ids[j] * d : xs + (size_t)j * d;
        for (int dd = 0; dd < d; ++dd) atomicAdd(&acc[s + k * d + dd], x * (double)xj[dd]);
      }
      atomicAdd(&acc[s + s * d + k], rx);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_badmm_plain_finalize(int K, int s, int d, int upd,
                                                             double *__restrict__ acc,
                                                             double *cw64, T *cwT, double *cs64,
                                                             T *csT, int *flags) {
  __shared__ double red[8];
  const int l = blockIdx.x, tid = threadIdx.x;
  double *A = acc + (size_t)l * s * (d + 2);
  double tot = 0.0;
  for (int k = tid; k < s; k += NT) tot += A[k];
  tot = block_sum(tot, red);
  if (!(tot > 0.0) && tid == 0) atomicOr(flags, FLAG_NUMERIC);
  for (int k = tid; k < s; k += NT) {
    const double v = A[k] / tot;
    cw64[(size_t)l * s + k] = v;
    cwT[(size_t)l * s + k] = (T)v;
  }
  if (upd)
    for (int q = tid; q < s * d; q += NT) {
      const double rx = A[s + s * d + q / d];
      const double v = A[s + q] / rx;
      cs64[(size_t)l * s * d + q] = v;
      csT[(size_t)l * s * d + q] = (T)v;
    }
  __syncthreads();
  for (int q = tid; q < s * (d + 2); q += NT) A[q] = 0.0;
}

__global__ void k_zero_f64(double *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Object metadata in the plan kernels' order: pm[p] = {object, label, column
// count, first column} of perm[p], one 16-byte scalar load per object instead
// of a load of perm[p] and three loads that wait for it
__global__ __launch_bounds__(NT) void k_plan_meta(const int *__restrict__ perm, int n,
                                                  const int *__restrict__ label,
                                                  const int *__restrict__ p_str,
                                                  const long long *__restrict__ cum,
                                                  const char *__restrict__ lsw,
                                                  ObjMeta *__restrict__ pm) {
  const int p = blockIdx.x * NT + threadIdx.x;
  if (p < n) {
    const int o = perm[p];
    pm[p] = ObjMeta{o, label[o], p_str[o] | (lsw[o] == 1 ? META_SWITCHED : 0), (int)cum[o]};
  }
}

__global__ void k_zero_i32(int *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

}  // namespace

// ------------------------------------------------------------------ host side

int d2mi_ensure_plans(d2mi_ctx *ctx, PhaseDev *p, int precision) {
  const size_t n = (size_t)p->s * p->col;
  if (precision == 64) {
    if (p->plan_s64 != p->s) {
      D2MI_TRY(p->Z64.ensure(n * sizeof(double)));
      D2MI_TRY(p->Y64.ensure(n * sizeof(double)));
      p->plan_s64 = p->s;
      return 1;  // freshly allocated
    }
  } else {
    if (p->plan_s32 != p->s) {
      D2MI_TRY(p->Z32.ensure(n * sizeof(float)));
      D2MI_TRY(p->Y32.ensure(n * sizeof(float)));
      p->plan_s32 = p->s;
      return 1;
    }
  }
  return 0;
}

namespace {

// resident workgroups per CU of a plan kernel at this LDS size (HIP's
// occupancy calculator), so one launch is exactly one round of workgroups
template <typename F>
int occupancy_of(F kernel, size_t lds, int threads = NT) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, lds) != hipSuccess || n <= 0)
    n = 1;
  return n;
}

// the plan kernel for an update: k_badmm_plan_reg (plan elements in
// registers; s <= 64, objects of up to NC * (NT / s) supports), k_badmm_plan64
// (plan in LDS, s <= 64), k_badmm_plan (any s) or its HBM-scratch form (GM);
// LEAN instances serve passes after the first with no support update,
// residual or kept X
enum PlanKind { PK_LEGACY, PK_GM, PK_64, PK_REG8, PK_REG16, PK_TALL, PK_REG8P, PK_REG16P,
                PK_WAVE8, PK_WAVE16, PK_REG16W, PK_REG16X };

const char *plan_kind_name(PlanKind k) {
  switch (k) {
    case PK_GM: return "hbm";
    case PK_64: return "lds";
    case PK_REG8: return "reg8";
    case PK_REG16: return "reg16";
    case PK_TALL: return "tall";
    case PK_REG8P: return "reg8p";
    case PK_REG16P: return "reg16p";
    case PK_WAVE8: return "wave8";
    case PK_WAVE16: return "wave16";
    case PK_REG16W: return "reg16w";
    case PK_REG16X: return "reg16x";
    default: return "general";
  }
}

inline bool is_tall(PlanKind k) { return k == PK_TALL; }
// register forms; the P kinds run their lean passes with the next object's
// plan copied into LDS ahead of time (k_badmm_plan_reg<..., PF = true>)
inline bool is_reg(PlanKind k) {
  return k == PK_REG8 || k == PK_REG16 || k == PK_REG8P || k == PK_REG16P || k == PK_WAVE8 ||
         k == PK_WAVE16 || k == PK_REG16W || k == PK_REG16X;
}
inline bool is_pf(PlanKind k) { return k == PK_REG8P || k == PK_REG16P; }
// the register form with 64-thread workgroups: one wave per object
inline bool is_wave(PlanKind k) { return k == PK_WAVE8 || k == PK_WAVE16; }
// the register form with 512- and 1024-thread workgroups (FP32): two and
// four times the columns of the 256-thread form at the same registers per
// thread, for objects past its limit (word-embedding documents of 65-128 and
// 129-256 words against s = 64)
constexpr int WIDE_NT = 512, XWIDE_NT = 1024;
// workgroup size of a plan kernel: one wave per 64 rows for the tall form,
// one wave for the wave forms
inline int plan_threads(PlanKind k, int s) {
  return is_tall(k)          ? (s + 63) / 64 * 64
         : is_wave(k)        ? 64
         : k == PK_REG16W    ? WIDE_NT
         : k == PK_REG16X    ? XWIDE_NT
                             : NT;
}

// d2mi_badmm_tune's request: passes timed per candidate, and the report
struct TuneReq {
  int reps;
  std::string report;
};

template <typename T, int CM, int DIM>
struct PlanKernels {
  // lean instances exist for FP32 only (FP64 passes always keep X), and FP64
  // register forms for up to 8 columns per thread
  static constexpr bool F32 = !std::is_same<T, double>::value;
  static constexpr bool PFOK = F32 && CM != COST_TABLE;  // prefetch form instances
  // register-form mode: 0 lean, 1 first / support-update passes, 2 residual
  template <bool LEAN>
  static int reg_mode(const PlanArgs<T> &pa) {
    return LEAN ? 0 : pa.resid ? 2 : 1;
  }
  // the tall form (sparse histograms, s > 64): MC = the size class's column
  // count (a multiple of 8), passed as pa.m_max
  static constexpr bool TALL = CM == COST_SPARSE && DIM == 0;
  static constexpr int TALL_MC = F32 ? TALL_MC32 : TALL_MC64;
  template <int MC>
  static int tall_occ1(int mode, int threads) {
    if constexpr (TALL && MC <= TALL_MC) {
      if (mode == 0) return occupancy_of(k_badmm_plan_tall<T, MC, 0>, 0, threads);
      if (mode == 2) return occupancy_of(k_badmm_plan_tall<T, MC, 2>, 0, threads);
      return occupancy_of(k_badmm_plan_tall<T, MC, 1>, 0, threads);
    }
    return 1;
  }
  template <int MC>
  static void tall_run1(int G, hipStream_t S, const PlanArgs<T> &pa, int mode) {
    if constexpr (TALL && MC <= TALL_MC) {
      const int th = plan_threads(PK_TALL, pa.s);
      if (mode == 0) k_badmm_plan_tall<T, MC, 0><<<G, th, 0, S>>>(pa);
      else if (mode == 2) k_badmm_plan_tall<T, MC, 2><<<G, th, 0, S>>>(pa);
      else k_badmm_plan_tall<T, MC, 1><<<G, th, 0, S>>>(pa);
    }
  }
  static int tall_occ(int mc, int mode, int threads) {
    switch (mc) {
      case 8: return tall_occ1<8>(mode, threads);
      case 16: return tall_occ1<16>(mode, threads);
      case 24: return tall_occ1<24>(mode, threads);
      case 32: return tall_occ1<32>(mode, threads);
      default: return tall_occ1<40>(mode, threads);
    }
  }
  static void tall_run(int G, hipStream_t S, const PlanArgs<T> &pa, int mode) {
    switch (pa.m_max) {
      case 8: tall_run1<8>(G, S, pa, mode); break;
      case 16: tall_run1<16>(G, S, pa, mode); break;
      case 24: tall_run1<24>(G, S, pa, mode); break;
      case 32: tall_run1<32>(G, S, pa, mode); break;
      default: tall_run1<40>(G, S, pa, mode);
    }
  }
  template <bool LEAN>
  static int occ(size_t lds, PlanKind k, int mode, int s, int mc) {
    switch (k) {
      case PK_TALL: return tall_occ(mc, LEAN ? 0 : mode, plan_threads(k, s));
      case PK_WAVE16:
        if constexpr (F32) {
          if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 0, false, 64>, lds, 64);
          return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 2, false, 64>, lds, 64)
                           : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 1, false, 64>, lds, 64);
        }
        return 0;
      case PK_WAVE8:
        if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 0, false, 64>, lds, 64);
        return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 2, false, 64>, lds, 64)
                         : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 1, false, 64>, lds, 64);
      case PK_REG16P:
        if constexpr (PFOK && LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 0, true>, lds);
        [[fallthrough]];
      case PK_REG16:
        if constexpr (F32) {
          if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 0>, lds);
          return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 2>, lds)
                           : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 1>, lds);
        }
        return 0;
      case PK_REG16X:
        if constexpr (F32) {
          if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 0, false, XWIDE_NT>, lds, XWIDE_NT);
          return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 2, false, XWIDE_NT>, lds, XWIDE_NT)
                           : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 1, false, XWIDE_NT>, lds, XWIDE_NT);
        }
        return 0;
      case PK_REG16W:
        if constexpr (F32) {
          if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 0, false, WIDE_NT>, lds, WIDE_NT);
          return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 2, false, WIDE_NT>, lds, WIDE_NT)
                           : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 16, 1, false, WIDE_NT>, lds, WIDE_NT);
        }
        return 0;
      case PK_REG8P:
        if constexpr (PFOK && LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 0, true>, lds);
        [[fallthrough]];
      case PK_REG8:
        if constexpr (LEAN) return occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 0>, lds);
        return mode == 2 ? occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 2>, lds)
                         : occupancy_of(k_badmm_plan_reg<T, CM, DIM, 8, 1>, lds);
      case PK_64: return occupancy_of(k_badmm_plan64<T, CM, DIM, LEAN>, lds);
      case PK_GM: return occupancy_of(k_badmm_plan<T, CM, 0, true>, lds);
      default: return occupancy_of(k_badmm_plan<T, CM, DIM, false, LEAN>, lds);
    }
  }
  template <bool LEAN>
  static void run(int G, size_t lds, hipStream_t S, const PlanArgs<T> &pa, PlanKind k) {
    const int mode = reg_mode<LEAN>(pa);
    switch (k) {
      case PK_TALL: tall_run(G, S, pa, mode); break;
      case PK_WAVE16:
        if constexpr (F32) {
          if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 16, 0, false, 64><<<G, 64, lds, S>>>(pa);
          else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 16, 2, false, 64><<<G, 64, lds, S>>>(pa);
          else k_badmm_plan_reg<T, CM, DIM, 16, 1, false, 64><<<G, 64, lds, S>>>(pa);
        }
        break;
      case PK_WAVE8:
        if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 8, 0, false, 64><<<G, 64, lds, S>>>(pa);
        else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 8, 2, false, 64><<<G, 64, lds, S>>>(pa);
        else k_badmm_plan_reg<T, CM, DIM, 8, 1, false, 64><<<G, 64, lds, S>>>(pa);
        break;
      case PK_REG16P:
        if constexpr (PFOK && LEAN) {
          k_badmm_plan_reg<T, CM, DIM, 16, 0, true><<<G, NT, lds, S>>>(pa);
          break;
        }
        [[fallthrough]];
      case PK_REG16:
        if constexpr (F32) {
          if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 16, 0><<<G, NT, lds, S>>>(pa);
          else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 16, 2><<<G, NT, lds, S>>>(pa);
          else k_badmm_plan_reg<T, CM, DIM, 16, 1><<<G, NT, lds, S>>>(pa);
        }
        break;
      case PK_REG16X:
        if constexpr (F32) {
          if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 16, 0, false, XWIDE_NT><<<G, XWIDE_NT, lds, S>>>(pa);
          else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 16, 2, false, XWIDE_NT><<<G, XWIDE_NT, lds, S>>>(pa);
          else k_badmm_plan_reg<T, CM, DIM, 16, 1, false, XWIDE_NT><<<G, XWIDE_NT, lds, S>>>(pa);
        }
        break;
      case PK_REG16W:
        if constexpr (F32) {
          if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 16, 0, false, WIDE_NT><<<G, WIDE_NT, lds, S>>>(pa);
          else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 16, 2, false, WIDE_NT><<<G, WIDE_NT, lds, S>>>(pa);
          else k_badmm_plan_reg<T, CM, DIM, 16, 1, false, WIDE_NT><<<G, WIDE_NT, lds, S>>>(pa);
        }
        break;
      case PK_REG8P:
        if constexpr (PFOK && LEAN) {
          k_badmm_plan_reg<T, CM, DIM, 8, 0, true><<<G, NT, lds, S>>>(pa);
          break;
        }
        [[fallthrough]];
      case PK_REG8:
        if constexpr (LEAN) k_badmm_plan_reg<T, CM, DIM, 8, 0><<<G, NT, lds, S>>>(pa);
        else if (mode == 2) k_badmm_plan_reg<T, CM, DIM, 8, 2><<<G, NT, lds, S>>>(pa);
        else k_badmm_plan_reg<T, CM, DIM, 8, 1><<<G, NT, lds, S>>>(pa);
        break;
      case PK_64: k_badmm_plan64<T, CM, DIM, LEAN><<<G, NT, lds, S>>>(pa); break;
      case PK_GM:
        k_badmm_plan<T, CM, 0, true><<<G, NT, lds, S>>>(pa, pa.perm, pa.p_str, pa.cum, pa.label,
                                                        pa.lsw);
        break;
      default:
        k_badmm_plan<T, CM, DIM, false, LEAN><<<G, NT, lds, S>>>(pa, pa.perm, pa.p_str, pa.cum,
                                                                 pa.label, pa.lsw);
    }
  }
  static int occupancy(size_t lds, PlanKind k, bool lean, bool resid, int s, int mc) {
    if constexpr (F32)
      if (lean && k != PK_GM) return occ<true>(lds, k, 0, s, mc);
    return occ<false>(lds, k, resid ? 2 : 1, s, mc);
  }
  static void launch(int G, size_t lds, hipStream_t S, const PlanArgs<T> &pa, PlanKind k,
                     bool lean) {
    if constexpr (F32)
      if (lean && k != PK_GM) return run<true>(G, lds, S, pa, k);
    run<false>(G, lds, S, pa, k);
  }
};

template <typename T, int CM>
int run_badmm(d2mi_ctx *ctx, PhaseDev *p, int maxIters, double rhoCoeff, int updatePerLoops,
              int variant, d2mi_badmm_stats *st, bool fresh, TuneReq *tune) {
  constexpr bool F64 = std::is_same<T, double>::value;
  hipStream_t S = ctx->stream;
  const int N = (int)p->N, K = p->K, s = p->s, d = CM == COST_EUCLID ? p->dim : 0;
  const int dsup = (CM == COST_EUCLID || CM == COST_TABLE) ? p->dim : 0;  // support dimension
  const bool word = p->metric == D2MI_WORD_EMBED;
  const int m_max = p->m_max;
  T *Z = F64 ? (T *)p->Z64.p : (T *)p->Z32.p;
  T *Y = F64 ? (T *)p->Y64.p : (T *)p->Y32.p;
  const T *w = F64 ? (const T *)p->w64.p : (const T *)p->w32.p;
  const T *supp = F64 ? (const T *)p->supp64.p : (const T *)p->supp32.p;
  const T *dist = F64 ? (const T *)p->dist64.p : (const T *)p->dist32.p;
  T *cwT = F64 ? (T *)p->cw64.p : (T *)p->cw32.p;
  T *csT = F64 ? (T *)p->cs64.p : (T *)p->cs32.p;
  const int *sym = p->sym.as<int>();
  const T *vocab = F64 ? (const T *)p->vocab64.p : (const T *)p->vocab32.p;
  const bool euclid_upd = CM == COST_EUCLID || CM == COST_TABLE;
  T *ctab = nullptr, *xout = nullptr;
  if constexpr (CM == COST_TABLE) {
    D2MI_TRY(p->Ctab.ensure(sizeof(T) * (size_t)s * p->col));
    D2MI_TRY(p->Xout.ensure(sizeof(T) * (size_t)s * p->col));
    D2MI_TRY(ctx->rxbuf.ensure(sizeof(double) * (size_t)K * s));
    ctab = p->Ctab.as<T>();
    xout = p->Xout.as<T>();
  }
  // COST_TABLE: rebuild C from the current centroid supports (calculate_distmat)
  const int tab_G = std::max(1, std::min(N, ctx->num_cu * 4));
  const int tab_chunk = (N + tab_G - 1) / tab_G;
  // support-numerator records, allocated here: nothing allocates inside the
  // (possibly captured) inner loop
  if (CM == COST_TABLE) D2MI_TRY(ctx->rec2.ensure(sizeof(double) * (size_t)(tab_G + K) * s * dsup));
  const bool mfma_words = word && !F64 && !getenv_flag("D2MI_WORD_DIRECT");
  const bool bf16_words = mfma_words && getenv_flag("D2MI_PDIST2_BF16");
  if (CM == COST_TABLE && mfma_words) {
    D2MI_TRY(ctx->wtab.ensure(sizeof(float) * (size_t)p->V * K * s));
    D2MI_TRY(ctx->wnorm.ensure(sizeof(float) * ((size_t)p->V + (size_t)K * s)));
    if (bf16_words) {
      // the vocabulary in BF16 once per upload; centroid supports per rebuild
      const size_t dp = d2mi_pdist2_pad(dsup);
      D2MI_TRY(ctx->cs16.ensure(2 * (size_t)K * s * dp));
      if (!p->vocab16.p) {
        D2MI_TRY(p->vocab16.ensure(2 * (size_t)p->V * dp));
        D2MI_TRY(p->vnorm16.ensure(sizeof(float) * (size_t)p->V));
        D2MI_TRY(d2mi_rows_to_bf16_dev(ctx->stream, p->V, dsup, (const float *)p->vocab32.p,
                                       p->vocab16.as<__bf16>(), p->vnorm16.as<float>()));
      }
    }
  }
  auto build_table = [&](hipStream_t S) -> int {
    if (mfma_words) {  // one GEMM for every (word, centroid point) pair, then a gather
      float *nv = ctx->wnorm.as<float>(), *nc = nv + p->V;
      if (bf16_words) {
        D2MI_TRY(d2mi_rows_to_bf16_dev(S, K * s, dsup, (const float *)csT,
                                       ctx->cs16.as<__bf16>(), nc));
        D2MI_TRY(d2mi_pdist2_bf16_dev(S, p->V, K * s, dsup, p->vocab16.as<__bf16>(),
                                      p->vnorm16.as<float>(), ctx->cs16.as<__bf16>(), nc,
                                      (const float *)p->vocab32.p, (const float *)csT,
                                      ctx->wtab.as<float>()));
      } else {
        D2MI_TRY(d2mi_pdist2_mfma_dev(S, p->V, K * s, dsup, (const float *)p->vocab32.p,
                                      (const float *)csT, nv, nc, ctx->wtab.as<float>()));
      }
      if (s % 4 == 0 && !getenv_flag("D2MI_GATHER_SCALAR"))
        k_word_gather4<<<std::min(N, ctx->num_cu * 16), NT, 0, S>>>(
            N, s, K * s, p->p_str.as<int>(), p->cum.as<long long>(), ctx->label.as<int>(), sym,
            ctx->wtab.as<float>(), (float *)ctab);
      else
        k_word_gather<<<std::min(N, ctx->num_cu * 16), NT, 0, S>>>(
            N, s, K * s, p->p_str.as<int>(), p->cum.as<long long>(), ctx->label.as<int>(), sym,
            ctx->wtab.as<float>(), (float *)ctab);
      D2MI_HIP(hipGetLastError());
      return D2MI_OK;
    }
    TabArgs<T> ta{};
    ta.N = N; ta.s = s; ta.d = dsup; ta.chunk = tab_chunk; ta.perm = ctx->perm.as<int>();
    ta.p_str = p->p_str.as<int>(); ta.cum = p->cum.as<long long>(); ta.label = ctx->label.as<int>();
    ta.supp = supp; ta.sym = sym; ta.vocab = vocab; ta.csupp = csT; ta.ctab = ctab;
    const size_t lds = sizeof(T) * (size_t)s * dsup;
    const int G = (N + tab_chunk - 1) / tab_chunk;
    if (word) k_badmm_cost_table<T, true><<<G, NT, lds, S>>>(ta);
    else k_badmm_cost_table<T, false><<<G, NT, lds, S>>>(ta);
    D2MI_HIP(hipGetLastError());
    return D2MI_OK;
  };
  // COST_TABLE support update: numerators from X, then supports
  // FP32 stored-cost phases with s, d multiples of 4: step 5 numerators over
  // the label-sorted columns packed into blocks (k_badmm_supp_dma), from a
  // descriptor table built after the label sort; D2MI_SUPP_PER_OBJECT=1 keeps
  // the per-object blocks of k_badmm_supp_mfma
  bool supp_dma = false;
  int supp_tpw = 0;
  // LDS stages: 3 (two workgroups per CU) or 2 (three; D2MI_SUPP_STAGES=2)
  const int supp_ns = [] {
    const char *e = getenv("D2MI_SUPP_STAGES");
    return e && atoi(e) == 2 ? 2 : 3;
  }();
  const int supp_G = std::min(tab_G, ctx->num_cu * (supp_ns == 2 ? 3 : 2));
  auto update_supports = [&](hipStream_t S, int *seg, int *flags) -> int {
    D2MI_TRY(ctx->rec2.ensure(sizeof(double) * (size_t)(tab_G + K) * s * dsup));
    SuppArgs<T> sa2{};
    sa2.N = N; sa2.s = s; sa2.d = dsup; sa2.chunk = tab_chunk; sa2.perm = ctx->perm.as<int>();
    sa2.p_str = p->p_str.as<int>(); sa2.cum = p->cum.as<long long>(); sa2.label = ctx->label.as<int>();
    sa2.supp = supp; sa2.sym = sym; sa2.vocab = vocab; sa2.X = xout; sa2.rec = ctx->rec2.as<double>();
    const int G = (N + tab_chunk - 1) / tab_chunk;
    // FP32: matrix cores; FP64 (and s > 64): FP64 FMAs
    if constexpr (!F64) {
      if (supp_dma) {
        SuppDmaArgs da{};
        da.s = s; da.d = dsup; da.G = supp_G;
        da.desc = ctx->sdesc.as<int>(); da.blab = ctx->sblab.as<int>();
        da.nblk = ctx->sbs.as<int>() + K; da.X = xout; da.W = word ? vocab : supp;
        da.zeros = ctx->szero.as<float>(); da.rec = ctx->rec2.as<double>();
        auto go = [&](auto ns_c) {
          constexpr int NSc = decltype(ns_c)::value;
          switch (supp_tpw) {
            case 1: k_badmm_supp_dma<1, NSc><<<supp_G, SD_NT, 0, S>>>(da); break;
            case 2: k_badmm_supp_dma<2, NSc><<<supp_G, SD_NT, 0, S>>>(da); break;
            case 3: k_badmm_supp_dma<3, NSc><<<supp_G, SD_NT, 0, S>>>(da); break;
            case 4: k_badmm_supp_dma<4, NSc><<<supp_G, SD_NT, 0, S>>>(da); break;
            default: k_badmm_supp_dma<5, NSc><<<supp_G, SD_NT, 0, S>>>(da); break;
          }
        };
        if (supp_ns == 2) go(std::integral_constant<int, 2>());
        else go(std::integral_constant<int, 3>());
      } else if (s <= 64 && !getenv_flag("D2MI_SUPP_VALU")) {
        if (word) k_badmm_supp_mfma<true><<<G, NT, 0, S>>>(sa2);
        else k_badmm_supp_mfma<false><<<G, NT, 0, S>>>(sa2);
      } else if (word) {
        k_badmm_supp_num<T, true><<<G, NT, 0, S>>>(sa2);
      } else {
        k_badmm_supp_num<T, false><<<G, NT, 0, S>>>(sa2);
      }
    } else if (word) {
      k_badmm_supp_num<T, true><<<G, NT, 0, S>>>(sa2);
    } else {
      k_badmm_supp_num<T, false><<<G, NT, 0, S>>>(sa2);
    }
    k_badmm_supp_fin<T><<<dim3((unsigned)((s * dsup + NT - 1) / NT), K), NT, 0, S>>>(
        s, dsup, tab_chunk, seg, supp_dma ? ctx->sbs.as<int>() : nullptr, supp_G,
        ctx->rec2.as<double>(),
                                         ctx->rxbuf.as<double>(), p->cs64.as<double>(), csT, flags);
    D2MI_HIP(hipGetLastError());
    return D2MI_OK;
  };

  // ---- scratch
  const int nb = (N + NT - 1) / NT;
  D2MI_TRY(ctx->perm.ensure(sizeof(int) * (size_t)N));
  D2MI_TRY(ctx->blkcnt.ensure(sizeof(int) * (size_t)nb * K));
  D2MI_TRY(ctx->segstart.ensure(sizeof(int) * (size_t)(2 * K + 2)));
  D2MI_TRY(ctx->scal.ensure(sizeof(double) * 8));
  D2MI_TRY(ctx->flags.ensure(sizeof(int) * 4));
  D2MI_TRY(ctx->resid.ensure(sizeof(double) * 3 * (size_t)(maxIters + 1)));
  int *seg = ctx->segstart.as<int>();
  int *total = seg + K + 1;
  int *flags = ctx->flags.as<int>();
  double *scal = ctx->scal.as<double>();
  k_zero_i32<<<1, 64, 0, S>>>(flags, 4);

  // ---- 1. label sort
  k_label_hist<<<nb, NT, sizeof(int) * K, S>>>(ctx->label.as<int>(), N, K, ctx->blkcnt.as<int>(),
                                               flags);
  k_label_scan<<<K, NT, 0, S>>>(nb, K, ctx->blkcnt.as<int>(), total);
  k_label_seg<<<1, 64, 0, S>>>(K, total, seg, flags);
  k_label_scatter<<<nb, NT, 0, S>>>(ctx->label.as<int>(), N, K, ctx->blkcnt.as<int>(), seg,
                                    ctx->perm.as<int>());
  D2MI_HIP(hipGetLastError());
  int hflags = 0;
  D2MI_HIP(hipMemcpyAsync(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (hflags & FLAG_BADLABEL) {
    d2mi_set_error("a label is outside [0, %d)", K);
    return D2MI_EINVAL;
  }
  if (hflags & FLAG_EMPTY) {
    d2mi_set_error("a cluster has no object (the reference asserts, centroid_Bregman.c:144)");
    return D2MI_EEMPTY;
  }
  if constexpr (CM == COST_TABLE && !F64) {
    supp_dma = s <= 64 && s % 4 == 0 && dsup % 4 == 0 && p->col < (size_t)INT32_MAX &&
               !getenv_flag("D2MI_SUPP_VALU") && !getenv_flag("D2MI_SUPP_PER_OBJECT");
    if (supp_dma) {
      const int sweeps = (dsup + 64 * SD_TPW_MAX - 1) / (64 * SD_TPW_MAX);
      supp_tpw = ((dsup + sweeps - 1) / sweeps + 63) / 64;
      const size_t nblk_max = p->col / SD_JB + (size_t)K;  // sum over labels of ceil(cols / SD_JB)
      D2MI_TRY(ctx->sdesc.ensure(sizeof(int) * 32 * nblk_max));
      D2MI_TRY(ctx->sblab.ensure(sizeof(int) * nblk_max));
      D2MI_TRY(ctx->sbs.ensure(sizeof(int) * ((size_t)K + 1)));
      D2MI_TRY(ctx->scum.ensure(sizeof(long long) * ((size_t)N + 1)));
      D2MI_TRY(ctx->scbt.ensure(sizeof(long long) * (size_t)nb));
      D2MI_TRY(ctx->szero.ensure(sizeof(float) * 64));
      const int *perm = ctx->perm.as<int>();
      long long *scum = ctx->scum.as<long long>();
      int *sbs = ctx->sbs.as<int>();
      k_cumw_local<<<nb, NT, 0, S>>>(perm, p->p_str.as<int>(), N, scum, ctx->scbt.as<long long>());
      k_cumw_carry<<<nb, NT, 0, S>>>(N, ctx->scbt.as<long long>(), scum);
      k_supp_blocks<<<1, 64, 0, S>>>(K, seg, scum, sbs);
      if (word)
        k_supp_desc<true><<<N, 64, 0, S>>>(perm, ctx->label.as<int>(), p->p_str.as<int>(),
                                           p->cum.as<long long>(), sym, seg, scum, sbs,
                                           ctx->sdesc.as<int>());
      else
        k_supp_desc<false><<<N, 64, 0, S>>>(perm, ctx->label.as<int>(), p->p_str.as<int>(),
                                            p->cum.as<long long>(), sym, seg, scum, sbs,
                                            ctx->sdesc.as<int>());
      k_supp_pad<<<K, 64, 0, S>>>(seg, scum, sbs, ctx->sdesc.as<int>(), ctx->sblab.as<int>());
      k_zero_i32<<<1, 64, 0, S>>>(ctx->szero.as<int>(), 64);
      D2MI_HIP(hipGetLastError());
    }
  }

  // ---- 2. setup
  D2MI_HIP(hipEventRecord(ctx->ev[0], S));
  if constexpr (CM == COST_TABLE) D2MI_TRY(build_table(S));
  const int setup_per_block = std::max(1, N / (ctx->num_cu * 8));
  const int setup_blocks = (N + setup_per_block - 1) / setup_per_block;
  D2MI_TRY(ctx->part.ensure(sizeof(double) * (size_t)std::max(setup_blocks, 3 * 65536)));
  {
    SetupArgs<T> sa{};
    sa.N = N; sa.s = s; sa.d = d; sa.V = p->V; sa.m_max = m_max; sa.per_block = setup_per_block;
    sa.p_str = p->p_str.as<int>(); sa.cum = p->cum.as<long long>(); sa.label = ctx->label.as<int>();
    sa.lsw = ctx->lsw.as<char>(); sa.force_init = fresh ? 1 : 0;
    sa.write_state = variant == D2MI_BADMM_PLAIN ? 1 : 0;
    sa.w = w; sa.supp = supp; sa.sym = sym; sa.dist = dist; sa.cw = cwT; sa.csupp = csT;
    sa.Z = Z; sa.Y = Y; sa.ctab = ctab; sa.part = ctx->part.as<double>();
    const size_t lds = align16(sizeof(T) * ((size_t)m_max * std::max(d, 1) + (size_t)s * std::max(d, 1)) +
                               sizeof(int) * m_max) + 8 * sizeof(double);
    int nparts = setup_blocks;
    if (variant == D2MI_BADMM_ENGINE && p->metric == D2MI_EUCLIDEAN && p->xmean.p &&
        !getenv_flag("D2MI_RHO_DIRECT")) {
      // the engine initialises Z and Y in its first pass: only rho is needed
      nparts = std::max(1, std::min((N + NT - 1) / NT, ctx->num_cu * 4));
      k_badmm_rho_euclid<<<nparts, NT, 0, S>>>(N, s, p->dim, p->p_str.as<int>(),
                                              ctx->label.as<int>(), p->xmean.as<double>(),
                                              p->xspread.as<double>(), p->cs64.as<double>(),
                                              ctx->part.as<double>());
    } else if (variant == D2MI_BADMM_ENGINE && (CM == COST_SPARSE || CM == COST_HIST) &&
               !getenv_flag("D2MI_RHO_DIRECT")) {
      const bool sparse = CM == COST_SPARSE;
      const int nlines = sparse ? p->V : m_max;
      D2MI_TRY(ctx->rowsum.ensure(sizeof(double) * (size_t)nlines));
      k_hist_line_sums<<<(nlines + NT - 1) / NT, NT, 0, S>>>(p->dist64.as<double>(), p->V, s,
                                                             nlines, sparse ? 1 : 0,
                                                             ctx->rowsum.as<double>());
      nparts = std::max(1, std::min((N + NT - 1) / NT, ctx->num_cu * 4));
      k_badmm_rho_hist<<<nparts, NT, 0, S>>>(N, p->p_str.as<int>(), p->cum.as<long long>(), sym,
                                             sparse ? 1 : 0, ctx->rowsum.as<double>(),
                                             ctx->part.as<double>());
    } else if (variant == D2MI_BADMM_ENGINE && CM == COST_TABLE && !getenv_flag("D2MI_RHO_DIRECT")) {
      nparts = ctx->num_cu * 4;
      k_abs_sum<T><<<nparts, NT, 0, S>>>(ctab, (size_t)s * p->col, ctx->part.as<double>());
    } else {
      k_badmm_setup<T, CM><<<setup_blocks, NT, lds, S>>>(sa);
    }
    k_badmm_rho<<<1, NT, 0, S>>>(ctx->part.as<double>(), nparts, rhoCoeff,
                                 (double)s * (double)p->col, scal, flags);
    D2MI_HIP(hipGetLastError());
  }

  auto is_resid = [&](int iter) {
    return (iter % 100 == 99) || (iter < 100 && iter % 20 == 19) || iter == maxIters;
  };

  int n_launch = 0;
  double ms_plan = 0.0;
  bool graphed = false, timed = true;
  std::string forms;

  // per-pass timing events of the eager loop, kept by the context.  A
  // captured graph records none: HIP cannot report elapsed times between
  // events recorded by graph nodes (hipEventElapsedTime answers
  // hipErrorInvalidHandle, or the times of an earlier eager record of the
  // same event), so a graphed update reports no per-pass time.
  while (ctx->pass_ev.size() < 2 * (size_t)(maxIters + 1)) {
    hipEvent_t e;
    D2MI_HIP(hipEventCreate(&e));
    ctx->pass_ev.push_back(e);
  }
  hipEvent_t *evs = ctx->pass_ev.data();

  // the inner iterations (:150-320), launched by `loop`; captured into one HIP
  // graph and replayed where the context allows it (results are the same)
  auto launch_loop = [&](auto &&loop) -> int {
    const char *genv = getenv("D2MI_BADMM_GRAPH");
    const bool want = ctx->badmm_graph >= 0 ? ctx->badmm_graph == 1 : !(genv && atoi(genv) == 0);
    if (want && !tune) {
      hipGraph_t g = nullptr;
      const auto key = std::make_tuple((int)(p - ctx->ph.data()), (int)sizeof(T) * 8 + 100 * CM,
                                       maxIters, updatePerLoops);
      auto &entry = ctx->badmm_exec[key];
      const bool dbg = getenv_flag("D2MI_DEBUG_GRAPH");
      const hipError_t eb = hipStreamBeginCapture(S, hipStreamCaptureModeThreadLocal);
      if (dbg) fprintf(stderr, "[d2mi graph] begin capture: %s\n", hipGetErrorName(eb));
      if (eb == hipSuccess) {
        timed = false;  // no event nodes in the graph
        const int rc = loop();
        const hipError_t ec = hipStreamEndCapture(S, &g);
        if (dbg)
          fprintf(stderr, "[d2mi graph] loop rc %d (%s), end capture: %s\n", rc,
                  rc ? d2mi_last_error() : "", hipGetErrorName(ec));
        if (rc == D2MI_OK && ec == hipSuccess && g) {
          hipGraphExec_t &ex = entry.exec;
          bool ok = false;
          if (ex) {
            hipGraphNode_t en = nullptr;
            hipGraphExecUpdateResult ur;
            const hipError_t eu = hipGraphExecUpdate(ex, g, &en, &ur);
            ok = eu == hipSuccess && ur == hipGraphExecUpdateSuccess;
            if (dbg) fprintf(stderr, "[d2mi graph] exec update: %s result %d\n", hipGetErrorName(eu), (int)ur);
            if (!ok) {
              (void)hipGetLastError();
              (void)hipGraphExecDestroy(ex);
              ex = nullptr;
            }
          }
          if (!ok) {
            const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            ok = ei == hipSuccess;
            if (dbg) fprintf(stderr, "[d2mi graph] instantiate: %s\n", hipGetErrorName(ei));
            if (!ok) {
              (void)hipGetLastError();
              ex = nullptr;
            }
          }
          (void)hipGraphDestroy(g);
          if (ok) {
            D2MI_HIP(hipGraphLaunch(ex, S));
            graphed = true;
            return D2MI_OK;
          }
        } else if (g) {
          (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();  // capture refused: launch one by one
        timed = true;
      } else {
        (void)hipGetLastError();
      }
      n_launch = 0;
    }
    return loop();
  };

  if (variant == D2MI_BADMM_ENGINE) {
    const int Rmax = euclid_upd ? s * (d + 2) : s;
    if (m_max > NT * PRE_W || (size_t)m_max * std::max(d, 1) > (size_t)NT * PRE_X) {
      d2mi_set_error("objects of up to %d supports (d=%d) exceed this build's plan kernel "
                     "(m <= %d, m*d <= %d)", m_max, d, NT * PRE_W, NT * PRE_X);
      return D2MI_EUNSUPPORTED;
    }
    static const int pad64 = [] {
      const char *e = getenv("D2MI_PLAN64_PAD");  // extra column padding (tuning)
      return e ? 2 * (atoi(e) / 2) : 0;
    }();
    // launch geometry: contiguous chunks of a bucket's label-sorted order,
    // one chunk per resident workgroup (D2MI_PLAN_WG_PER_CU overrides)
    static const int wg_env = [] {
      const char *e = getenv("D2MI_PLAN_WG_PER_CU");
      return e ? atoi(e) : 0;
    }();
    const bool force_gm = getenv_flag("D2MI_BADMM_HBM_SCRATCH");
    const bool no_64 = getenv_flag("D2MI_PLAN_LEGACY"), no_reg = getenv_flag("D2MI_PLAN_NOREG");
    const bool no_lean = getenv_flag("D2MI_PLAN_NO_LEAN");
    // odd passes of the register forms walk their chunks backwards
    // (D2MI_PLAN_NO_REVERSE=1: every pass forwards)
    const bool reverse_odd = !getenv_flag("D2MI_PLAN_NO_REVERSE");
    const bool forced = force_gm || no_64 || no_reg;  // tuned choices give way to these
    // fixed Euclidean dimensions get straight-line cost code
    const int dim_sel = CM == COST_EUCLID && (d == 2 || d == 3) ? d : 0;

    // ---- object-size buckets: when the largest object is too big for the
    // register form, objects are launched by size class, each class with the
    // kernel form and LDS its own largest object allows. For s <= 64 the
    // classes end at the register form's limit (16 columns per thread in
    // FP32, 8 in FP64), then twice, four and eight times that, then m_max.
    // D2MI_PLAN_NO_BUCKETS=1 launches all objects together.
    // For sparse histograms with 64 < s <= 1024 (tall plans) the classes end
    // at the tall register form's column counts (16, and 40 in FP32).
    std::vector<int> lims;
    if (s <= 64 && !getenv_flag("D2MI_PLAN_NO_BUCKETS")) {
      const int cpi = NT / s;
      for (int c = F64 ? 8 : 16; c * cpi < m_max && (int)lims.size() < 4; c *= 2)
        lims.push_back(c * cpi);
    }
    if (CM == COST_SPARSE && s > 64 && s <= TALL_NT && !getenv_flag("D2MI_PLAN_NO_BUCKETS")) {
      const int mc_max = F64 ? TALL_MC64 : TALL_MC32;
      for (int c = 8; c <= mc_max && c < m_max; c += 8) lims.push_back(c);
    }
    lims.push_back(m_max);
    const int nbk = (int)lims.size();
    struct Bucket {
      int m_max, start, count, sp, Gmax, wg_mode;
      bool k64, gm;
      PlanKind kind;
      long long rec_off;
    };
    std::vector<Bucket> bk(nbk);
    const int *permk = ctx->perm.as<int>(), *segk = seg;
    std::vector<int> hsegk;
    if (nbk > 1) {
      const int Kb = nbk * K;
      D2MI_TRY(ctx->bkey.ensure(sizeof(int) * (size_t)N));
      D2MI_TRY(ctx->permk.ensure(sizeof(int) * (size_t)N));
      D2MI_TRY(ctx->blkcnt2.ensure(sizeof(int) * (size_t)nb * Kb));
      D2MI_TRY(ctx->segk.ensure(sizeof(int) * (size_t)(2 * Kb + 2)));
      BucketLims bl{};
      bl.n = nbk;
      for (int b = 0; b < nbk; ++b) bl.lim[b] = lims[b];
      int *sk = ctx->segk.as<int>(), *tot2 = sk + Kb + 1;
      k_bucket_key<<<nb, NT, 0, S>>>(ctx->label.as<int>(), p->p_str.as<int>(), N, K, bl,
                                     ctx->bkey.as<int>());
      k_label_hist<<<nb, NT, sizeof(int) * Kb, S>>>(ctx->bkey.as<int>(), N, Kb,
                                                    ctx->blkcnt2.as<int>(), flags + 2);
      k_label_scan<<<Kb, NT, 0, S>>>(nb, Kb, ctx->blkcnt2.as<int>(), tot2);
      k_label_seg<<<1, 64, 0, S>>>(Kb, tot2, sk, flags + 2);  // empty classes are expected
      k_label_scatter<<<nb, NT, 0, S>>>(ctx->bkey.as<int>(), N, Kb, ctx->blkcnt2.as<int>(), sk,
                                        ctx->permk.as<int>());
      D2MI_HIP(hipGetLastError());
      hsegk.resize(Kb + 1);
      D2MI_HIP(hipMemcpyAsync(hsegk.data(), sk, sizeof(int) * (Kb + 1), hipMemcpyDeviceToHost, S));
      D2MI_HIP(hipStreamSynchronize(S));
      permk = ctx->permk.as<int>();
      segk = sk;
      for (int b = 0; b < nbk; ++b) {
        bk[b].start = hsegk[b * K];
        bk[b].count = hsegk[(b + 1) * K] - hsegk[b * K];
      }
    } else {
      bk[0].start = 0;
      bk[0].count = N;
    }
    // per-object metadata in the plan order (D2MI_PLAN_NO_PMETA=1: the kernels
    // look it up through perm)
    const ObjMeta *pmeta = nullptr;
    if (p->col < (size_t)INT32_MAX && !getenv_flag("D2MI_PLAN_NO_PMETA")) {
      D2MI_TRY(ctx->pmeta.ensure(sizeof(ObjMeta) * (size_t)N));
      k_plan_meta<<<(N + NT - 1) / NT, NT, 0, S>>>(permk, N, ctx->label.as<int>(),
                                                  p->p_str.as<int>(), p->cum.as<long long>(),
                                                  ctx->lsw.as<char>(), ctx->pmeta.as<ObjMeta>());
      D2MI_HIP(hipGetLastError());
      pmeta = ctx->pmeta.as<ObjMeta>();
    }

    // ---- kernel forms: which fit a class (LDS, columns per thread), the
    // rule-based choice, and the tuner's choice where it has one
    const int sp64 = s + (s & 1) + pad64;
    const int spl = (s % 2 == 0) ? s + 1 : s;  // odd column stride: conflict-free sums
    auto set_kind = [&](Bucket &B, PlanKind k) {
      B.kind = k;
      B.k64 = k == PK_64 || is_reg(k);
      B.gm = k == PK_GM;
      B.sp = B.k64 ? sp64 : spl;
    };
    std::function<bool(int, PlanKind)> fits = [&](int m_lim, PlanKind k) -> bool {
      const int ncols = s <= 64 ? (m_lim + NT / s - 1) / (NT / s) : 1 << 30;
      switch (k) {
        case PK_TALL:  // class limits are multiples of 8 (the kernel's column slots)
          return CM == COST_SPARSE && s > 64 && s <= TALL_NT && m_lim % 8 == 0 &&
                 m_lim <= (F64 ? TALL_MC64 : TALL_MC32);
        case PK_REG16P:
        case PK_REG8P:  // also: the staging arrays hold t (sp x m) and the block (s x m)
          if (F64 || CM == COST_TABLE || s < 7 || (size_t)sp64 * m_lim > PF_STAGE ||
              (size_t)s * m_lim > PF_STAGE)
            return false;
          return fits(m_lim, k == PK_REG16P ? PK_REG16 : PK_REG8);
        case PK_REG16:
          if (F64) return false;
          [[fallthrough]];
        case PK_REG8:
          return s <= 64 && ncols <= (k == PK_REG8 ? 8 : 16) &&
                 Plan64Lds(sizeof(T), sp64, m_lim, s, d, 1, Rmax).total <= ctx->lds_max &&
                 PlanRegLds(sizeof(T), sp64, m_lim, s, d, Rmax, false, NT, true).total <= ctx->lds_max;
        case PK_REG16W:
        case PK_REG16X: {  // 512 or 1024 threads: nt / s columns per slot
          if (F64 || s > 64) return false;
          const int nt = plan_threads(k, s);
          const int wcols = (m_lim + nt / s - 1) / (nt / s);
          return wcols <= 16 && m_lim <= nt * PRE_W &&
                 (size_t)m_lim * std::max(d, 1) <= (size_t)nt * PRE_X &&
                 PlanRegLds(sizeof(T), sp64, m_lim, s, d, Rmax, false, nt, true).total <= ctx->lds_max;
        }
        case PK_WAVE16:
          if (F64) return false;
          [[fallthrough]];
        case PK_WAVE8: {  // one wave: 64 / s columns per slot, weights and supports 64 wide
          if (s > 64) return false;
          const int wcols = (m_lim + 64 / s - 1) / (64 / s);
          return wcols <= (k == PK_WAVE8 ? 8 : 16) && m_lim <= 64 * PRE_W &&
                 (size_t)m_lim * std::max(d, 1) <= (size_t)64 * PRE_X &&
                 PlanRegLds(sizeof(T), sp64, m_lim, s, d, Rmax, false, 64, true).total <= ctx->lds_max;
        }
        case PK_64:
          return s <= 64 && Plan64Lds(sizeof(T), sp64, m_lim, s, d, 1, Rmax).total <= ctx->lds_max;
        case PK_LEGACY:
          return PlanLds(sizeof(T), spl, m_lim, s, d, 1, Rmax).total <= ctx->lds_max;
        default:
          return PlanLds(sizeof(T), spl, m_lim, s, d, 1, Rmax, true).total <= ctx->lds_max;
      }
    };
    auto tune_key = [&](int m_lim) {
      return PlanTuneKey{(int)sizeof(T) * 8, CM, dim_sel, s, d, m_lim};
    };
    bool any_gm = false;
    for (int b = 0; b < nbk; ++b) {
      Bucket &B = bk[b];
      B.m_max = lims[b];
      B.wg_mode = 0;
      PlanKind k = PK_GM;
      const bool no_tall = getenv_flag("D2MI_PLAN_NO_TALL");
      if (!force_gm && !no_tall && fits(B.m_max, PK_TALL)) {
        k = PK_TALL;
      } else if (!force_gm) {
        if (!no_64 && fits(B.m_max, PK_64)) {
          k = PK_64;
          // small plans (at most 8 elements per lane of one wave): the wave
          // form, one object per wave; otherwise the 256-thread register form
          const bool small = (size_t)s * B.m_max <= 512 && !getenv_flag("D2MI_PLAN_NO_WAVE");
          if (!no_reg && small && fits(B.m_max, PK_WAVE8)) k = PK_WAVE8;
          else if (!no_reg && small && fits(B.m_max, PK_WAVE16)) k = PK_WAVE16;
          else if (!no_reg && fits(B.m_max, PK_REG8)) k = PK_REG8;
          else if (!no_reg && fits(B.m_max, PK_REG16)) k = PK_REG16;
          else if (!no_reg && !getenv_flag("D2MI_PLAN_NO_WIDE") && fits(B.m_max, PK_REG16W))
            k = PK_REG16W;
          else if (!no_reg && !getenv_flag("D2MI_PLAN_NO_WIDE") && fits(B.m_max, PK_REG16X))
            k = PK_REG16X;
          // D2MI_PLAN_PREFETCH=1: the register forms' prefetching variants
          if (getenv_flag("D2MI_PLAN_PREFETCH")) {
            if (k == PK_REG8 && fits(B.m_max, PK_REG8P)) k = PK_REG8P;
            if (k == PK_REG16 && fits(B.m_max, PK_REG16P)) k = PK_REG16P;
          }
        } else if (fits(B.m_max, PK_LEGACY)) {
          k = PK_LEGACY;
        }
      }
      if (!forced && !tune) {
        auto it = ctx->plan_tune.find(tune_key(B.m_max));
        if (it != ctx->plan_tune.end() && fits(B.m_max, (PlanKind)it->second.kind)) {
          k = (PlanKind)it->second.kind;
          B.wg_mode = it->second.wg_mode;
        }
      }
      set_kind(B, k);
      any_gm = any_gm || B.gm;
    }
    // dynamic LDS of bucket B's kernel for a pass; lean: the prefetch form's
    // lean passes keep t in their static staging array
    auto lds_of = [&](const Bucket &B, int keepx, int R, bool lean = false,
                      bool resid = false) -> size_t {
      if (is_tall(B.kind)) return 0;  // static LDS only
      if (is_reg(B.kind))
        return PlanRegLds(sizeof(T), B.sp, B.m_max, s, d, R, lean && is_pf(B.kind),
                          plan_threads(B.kind, s), resid)
            .total;
      return B.k64 ? Plan64Lds(sizeof(T), B.sp, B.m_max, s, d, keepx, R).total
                   : PlanLds(sizeof(T), B.sp, B.m_max, s, d, keepx, R, B.gm).total;
    };
    auto occupancy = [&](const Bucket &B, size_t lds, bool lean, bool resid) {
      if constexpr (CM == COST_EUCLID) {
        if (dim_sel == 2) return PlanKernels<T, CM, 2>::occupancy(lds, B.kind, lean, resid, s, B.m_max);
        if (dim_sel == 3) return PlanKernels<T, CM, 3>::occupancy(lds, B.kind, lean, resid, s, B.m_max);
      }
      return PlanKernels<T, CM, 0>::occupancy(lds, B.kind, lean, resid, s, B.m_max);
    };
    auto per_cu_of = [&](const Bucket &B, size_t lds, bool lean, bool resid) {
      if (wg_env > 0) return wg_env;
      const int occ = occupancy(B, lds, lean, resid);
      return B.wg_mode == 1 ? std::max(1, occ - 1) : B.wg_mode == 2 ? 2 * occ : occ;
    };
    auto geometry = [&](const Bucket &B, size_t lds, int *G_out, int *chunk_out, bool lean,
                        bool resid) {
      const int Gt = std::max(1, ctx->num_cu * per_cu_of(B, lds, lean, resid));
      *chunk_out = std::max(1, (B.count + Gt - 1) / Gt);
      *G_out = (B.count + *chunk_out - 1) / *chunk_out;
    };
    auto launch_plan = [&](int G, size_t lds, const PlanArgs<T> &pa, PlanKind kind, bool lean) {
      if constexpr (CM == COST_EUCLID) {
        if (dim_sel == 2) return PlanKernels<T, CM, 2>::launch(G, lds, S, pa, kind, lean);
        if (dim_sel == 3) return PlanKernels<T, CM, 3>::launch(G, lds, S, pa, kind, lean);
      }
      PlanKernels<T, CM, 0>::launch(G, lds, S, pa, kind, lean);
    };
    // the plan-kernel arguments of bucket B for one pass
    auto plan_args = [&](const Bucket &B, int chunk, int R, int upd, int resid, int keepx,
                         int first) {
      PlanArgs<T> pa{};
      pa.N = B.count; pa.s = s; pa.d = d; pa.K = K; pa.sp = B.sp; pa.m_max = B.m_max;
      pa.chunk = chunk; pa.R = R; pa.V = p->V; pa.upd = upd; pa.resid = resid; pa.keepx = keepx;
      pa.p_str = p->p_str.as<int>(); pa.cum = p->cum.as<long long>(); pa.perm = permk + B.start;
      pa.label = ctx->label.as<int>(); pa.w = w; pa.supp = supp; pa.sym = sym; pa.dist = dist;
      pa.cw = cwT; pa.csupp = csT; pa.scal = scal; pa.Z = Z; pa.Y = Y;
      pa.ctab = ctab; pa.xout = xout;
      pa.xg = B.gm ? p->Xout.as<T>() : nullptr; pa.z0g = B.gm ? p->Z0.as<T>() : nullptr;
      pa.rec = ctx->rec.as<double>() + (size_t)B.rec_off * R;
      pa.first = first; pa.force_init = fresh ? 1 : 0; pa.lsw = ctx->lsw.as<char>();
      pa.pm = pmeta ? pmeta + B.start : nullptr;
      return pa;
    };
    if (any_gm) {
      D2MI_TRY(p->Xout.ensure(sizeof(T) * (size_t)s * p->col));
      D2MI_TRY(p->Z0.ensure(sizeof(T) * (size_t)s * p->col));
      xout = p->Xout.as<T>();
    }

    // ---- tuner: time every feasible (form, workgroups per CU) of each class
    // on lean passes over the resident plans, then put the plans back
    if (tune) {
      if (fresh) {
        d2mi_set_error("d2mi_badmm_tune: run one update of this phase at this precision first");
        return D2MI_EINVAL;
      }
      const size_t bytes = sizeof(T) * (size_t)s * p->col;
      D2MI_TRY(ctx->tuneZ.ensure(bytes));
      D2MI_TRY(ctx->tuneY.ensure(bytes));
      D2MI_HIP(hipMemcpyAsync(ctx->tuneZ.p, Z, bytes, hipMemcpyDeviceToDevice, S));
      D2MI_HIP(hipMemcpyAsync(ctx->tuneY.p, Y, bytes, hipMemcpyDeviceToDevice, S));
      const int reps = std::max(1, tune->reps);
      char line[256];
      for (int b = 0; b < nbk; ++b) {
        Bucket B = bk[b];
        if (B.count == 0) continue;
        snprintf(line, sizeof line, "m<=%d (%d objects):", B.m_max, B.count);
        tune->report += line;
        // candidates: the rule-based form first (the incumbent), then every
        // other feasible (form, workgroups per CU); a challenger must be more
        // than 3% faster on the plan passes, since more workgroups also cost
        // the per-pass reduction (more records) that these timings leave out
        std::vector<PlanTuneVal> cands{PlanTuneVal{(int)B.kind, 0}};
        const PlanKind kinds[] = {PK_TALL,  PK_WAVE8,  PK_WAVE16, PK_REG8,   PK_REG16, PK_REG16W,
                                  PK_REG16X, PK_REG8P, PK_REG16P, PK_64, PK_LEGACY, PK_GM};
        bool any_fast = false;
        for (PlanKind k : kinds)
          if (k != PK_GM && fits(B.m_max, k)) any_fast = true;
        for (PlanKind k : kinds) {
          if (!fits(B.m_max, k) || (k == PK_GM && any_fast)) continue;
          for (int wm = 0; wm < 3; ++wm)
            if (!(k == B.kind && wm == 0)) cands.push_back(PlanTuneVal{(int)k, wm});
        }
        float best = 0.f;
        PlanTuneVal bestv = cands[0];
        for (size_t ci = 0; ci < cands.size(); ++ci) {
          const PlanKind k = (PlanKind)cands[ci].kind;
          const int wm = cands[ci].wg_mode;
          if (k == PK_GM) {  // the HBM-scratch form keeps x and Z0 in plan-sized arrays
            D2MI_TRY(p->Xout.ensure(sizeof(T) * (size_t)s * p->col));
            D2MI_TRY(p->Z0.ensure(sizeof(T) * (size_t)s * p->col));
            xout = p->Xout.as<T>();
          }
          set_kind(B, k);
          B.wg_mode = wm;
          B.rec_off = 0;
          const bool lean = !F64 && !B.gm && !no_lean;
          const size_t lds = lds_of(B, F64 ? 1 : 0, s, lean);
          if (wm == 1 && occupancy(B, lds, lean, false) <= 1) continue;
          int G, chunk;
          geometry(B, lds, &G, &chunk, lean, false);
          D2MI_TRY(ctx->rec.ensure(sizeof(double) * (size_t)(G + K) * s));
          const PlanArgs<T> pa = plan_args(B, chunk, s, 0, 0, F64 ? 1 : 0, 0);
          launch_plan(G, lds, pa, k, lean);  // warm
          D2MI_HIP(hipEventRecord(ctx->ev[2], S));
          for (int r = 0; r < reps; ++r) launch_plan(G, lds, pa, k, lean);
          D2MI_HIP(hipEventRecord(ctx->ev[3], S));
          D2MI_HIP(hipGetLastError());
          D2MI_HIP(hipEventSynchronize(ctx->ev[3]));
          float ms = 0.f;
          D2MI_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
          ms /= reps;
          snprintf(line, sizeof line, " %s/%s %.4f", plan_kind_name(k),
                   wm == 0 ? "occ" : wm == 1 ? "occ-1" : "2occ", ms);
          tune->report += line;
          if (ci == 0 || ms < 0.97f * best) {
            best = ms;
            bestv = cands[ci];
          }
        }
        ctx->plan_tune[tune_key(B.m_max)] = bestv;
        snprintf(line, sizeof line, " -> %s/%s\n", plan_kind_name((PlanKind)bestv.kind),
                 bestv.wg_mode == 0 ? "occ" : bestv.wg_mode == 1 ? "occ-1" : "2occ");
        tune->report += line;
      }
      D2MI_HIP(hipMemcpyAsync(Z, ctx->tuneZ.p, bytes, hipMemcpyDeviceToDevice, S));
      D2MI_HIP(hipMemcpyAsync(Y, ctx->tuneY.p, bytes, hipMemcpyDeviceToDevice, S));
      D2MI_HIP(hipStreamSynchronize(S));
      return D2MI_OK;
    }

    // lean passes: FP32 passes after the first with no support update,
    // residual or kept X (not the HBM-scratch form)
    long long rec_total = 0;
    int part_total = 0;
    for (int b = 0; b < nbk; ++b) {
      Bucket &B = bk[b];
      const size_t lds_full = lds_of(B, 1, Rmax, false, true);
      if (lds_full > ctx->lds_max) {
        d2mi_set_error("B-ADMM with s = %d, d = %d needs %zu B of LDS (limit %zu)", s, d,
                       lds_full, ctx->lds_max);
        return D2MI_EUNSUPPORTED;
      }
      int chunk_tmp = 0;
      B.Gmax = 0;
      if (B.count > 0) {
        const bool lean0 = !F64 && !B.gm && !no_lean;
        geometry(B, lds_of(B, 0, s, lean0), &B.Gmax, &chunk_tmp, lean0, false);
      }
      B.rec_off = rec_total;
      rec_total += B.Gmax + K;
      part_total += B.Gmax;
      if (B.count > 0) {
        char f[64];
        const bool lean0 = !F64 && !B.gm && !no_lean;
        snprintf(f, sizeof f, "%sm<=%d:%s/%d(%d)", forms.empty() ? "" : " ", B.m_max,
                 plan_kind_name(B.kind), per_cu_of(B, lds_of(B, 0, s, lean0), lean0, false),
                 B.count);
        forms += f;
      }
    }
    D2MI_TRY(ctx->rec.ensure(sizeof(double) * (size_t)rec_total * Rmax));
    D2MI_TRY(ctx->part.ensure(sizeof(double) * (size_t)std::max(setup_blocks, 3 * part_total)));
    auto loop = [&]() -> int {
      n_launch = 0;
      for (int iter = 0; iter <= maxIters; ++iter) {
        const int upd = euclid_upd && (iter % updatePerLoops == 0);
        const int resid = is_resid(iter);
        const int keepx = F64 || upd || resid;
        const int R = upd ? s * (d + 2) : s;
        FinArgs<T> fa{};
        fa.nbk = nbk;
        fa.segk = segk;
        int part_off = 0;
        if (timed) D2MI_HIP(hipEventRecord(evs[2 * iter], S));
        for (int b = 0; b < nbk; ++b) {
          const Bucket &B = bk[b];
          fa.start[b] = B.start;
          fa.rec_off[b] = B.rec_off;
          fa.chunk[b] = 1;
          if (B.count == 0) continue;
          const bool lean = !F64 && !B.gm && !no_lean && iter > 0 && !upd && !resid && !keepx;
          const size_t lds = lds_of(B, keepx, R, lean, resid);
          int G, chunk;
          geometry(B, lds, &G, &chunk, lean, resid);
          if (G > B.Gmax) {  // records are sized for Gmax workgroups
            chunk = std::max(1, (B.count + B.Gmax - 1) / B.Gmax);
            G = (B.count + chunk - 1) / chunk;
          }
          fa.chunk[b] = chunk;
          PlanArgs<T> pa = plan_args(B, chunk, R, upd, resid, keepx, iter == 0);
          pa.resid_part = ctx->part.as<double>() + 3 * part_off;
          pa.rev = reverse_odd && (iter & 1);
          launch_plan(G, lds, pa, B.kind, lean);
          part_off += G;
        }
        if (timed) D2MI_HIP(hipEventRecord(evs[2 * iter + 1], S));
        fa.K = K; fa.s = s; fa.d = d; fa.R = R; fa.G = part_off; fa.upd = upd;
        fa.resid = resid; fa.iter = iter; fa.rec = ctx->rec.as<double>();
        fa.resid_part = ctx->part.as<double>(); fa.cw64 = p->cw64.as<double>(); fa.cwT = cwT;
        fa.cs64 = p->cs64.as<double>(); fa.csT = csT; fa.resid_out = ctx->resid.as<double>();
        fa.rxout = CM == COST_TABLE ? ctx->rxbuf.as<double>() : nullptr;
        fa.flags = flags;
        k_badmm_finalize<T><<<K, NT, sizeof(double) * (R + 8 + std::max(NT, R)), S>>>(fa);
        D2MI_HIP(hipGetLastError());
        if (CM == COST_TABLE && upd) {  // step 5 (:196-227): supports, then C again
          D2MI_TRY(update_supports(S, seg, flags));
          D2MI_TRY(build_table(S));
        }
        ++n_launch;
      }
      return D2MI_OK;
    };
    D2MI_TRY(launch_loop(loop));
  } else {
    // ---- plain form
    const size_t n = (size_t)s * p->col;
    DevMem &Xm = F64 ? ctx->X64 : ctx->X32;
    D2MI_TRY(Xm.ensure(n * sizeof(T)));
    const size_t accn = (size_t)K * s * (dsup + 2);
    D2MI_TRY(ctx->rec.ensure(sizeof(double) * accn));
    k_zero_f64<<<256, NT, 0, S>>>(ctx->rec.as<double>(), accn);
    forms = "plain";
    for (int iter = 0; iter <= maxIters; ++iter) {
      const int upd = euclid_upd && (iter % updatePerLoops == 0);
      PlainArgs<T> pa{};
      pa.N = N; pa.s = s; pa.d = dsup; pa.V = p->V; pa.upd = upd;
      pa.ctab = ctab; pa.vocab = word ? vocab : nullptr;
      pa.p_str = p->p_str.as<int>(); pa.cum = p->cum.as<long long>(); pa.label = ctx->label.as<int>();
      pa.w = w; pa.supp = supp; pa.sym = sym; pa.dist = dist; pa.cw = cwT; pa.csupp = csT;
      pa.scal = scal; pa.Z = Z; pa.Y = Y; pa.X = (T *)Xm.p; pa.acc = ctx->rec.as<double>();
      D2MI_HIP(hipEventRecord(evs[2 * iter], S));
      k_badmm_plain<T, CM><<<(N + 127) / 128, 128, 0, S>>>(pa);
      D2MI_HIP(hipEventRecord(evs[2 * iter + 1], S));
      k_badmm_plain_finalize<T><<<K, NT, 0, S>>>(K, s, dsup, upd, ctx->rec.as<double>(),
                                                 p->cw64.as<double>(), cwT,
                                                 p->cs64.as<double>(), csT, flags);
      D2MI_HIP(hipGetLastError());
      if (CM == COST_TABLE && upd) D2MI_TRY(build_table(S));
      ++n_launch;
    }
  }
  D2MI_HIP(hipEventRecord(ctx->ev[1], S));
  D2MI_HIP(hipStreamSynchronize(S));
  for (int i = 0; i < n_launch && timed; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, evs[2 * i], evs[2 * i + 1]) != hipSuccess) {
      // a graph that cannot report its recorded events' times: no per-pass time
      (void)hipGetLastError();
      timed = false;
      break;
    }
    ms_plan += ms;
  }
  if (!timed) ms_plan = -1.0;
  D2MI_HIP(hipMemcpy(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost));
  if (hflags & FLAG_NUMERIC) {
    d2mi_set_error("a normaliser was not positive (blas_like64.c:131/148 assert)");
    return D2MI_ENUMERIC;
  }
  double hscal[2];
  D2MI_HIP(hipMemcpy(hscal, scal, sizeof(hscal), hipMemcpyDeviceToHost));
  if (st) {
    float ms = 0.f;
    D2MI_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    st->rho = hscal[0];
    st->ms_total = ms;
    st->ms_plan = ms_plan;
    st->n_plan_launches = n_launch;
    st->graph = graphed ? 1 : 0;
    snprintf(st->plan_forms, sizeof st->plan_forms, "%s", forms.c_str());
    st->plan_elems = (double)s * (double)p->col;
    st->n_trace = 0;
    st->obj = 0.0;
    if (variant == D2MI_BADMM_ENGINE) {
      std::vector<double> res(3 * (size_t)(maxIters + 1));
      D2MI_HIP(hipMemcpy(res.data(), ctx->resid.p, sizeof(double) * res.size(),
                         hipMemcpyDeviceToHost));
      const double Nn = (double)N;
      for (int iter = 0; iter <= maxIters; ++iter) {
        const bool ref_row = (iter % 100 == 99) || (iter < 100 && iter % 20 == 19);
        if (ref_row && st->n_trace < 64) {
          st->trace[st->n_trace][0] = iter + 1;
          st->trace[st->n_trace][1] = res[3 * iter] * hscal[0] / Nn;
          st->trace[st->n_trace][2] = res[3 * iter + 1] / Nn;
          st->trace[st->n_trace][3] = res[3 * iter + 2] / Nn;
          st->n_trace++;
        }
      }
      st->obj = res[3 * maxIters] * hscal[0] / Nn;
    }
  }
  return D2MI_OK;
}

}  // namespace

namespace {
int badmm_dispatch(d2mi_ctx *ctx, int phase, int maxIters, double rhoCoeff, int updatePerLoops,
                   int precision, int variant, d2mi_badmm_stats *stats, TuneReq *tune) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  PhaseDev *p = d2mi_phase(ctx, phase);
  if (!p) return D2MI_EINVAL;
  if (!p->K || !ctx->label.p) {
    d2mi_set_error("set centroids and labels before d2mi_badmm");
    return D2MI_EINVAL;
  }
  if (maxIters < 0 || updatePerLoops <= 0 || !(rhoCoeff > 0) ||
      (precision != 32 && precision != 64) ||
      (variant != D2MI_BADMM_ENGINE && variant != D2MI_BADMM_PLAIN)) {
    d2mi_set_error("d2mi_badmm: bad options");
    return D2MI_EINVAL;
  }
  // stored-cost path: word embeddings, Euclidean d > MAXD_FUSED, or on request
  const char *force_env = getenv("D2MI_BADMM_TABLE");
  const bool force_table = force_env && atoi(force_env) != 0;
  const bool table = p->metric == D2MI_WORD_EMBED ||
                     (p->metric == D2MI_EUCLIDEAN && (p->dim > MAXD_FUSED || force_table));
  if (table) {
    if (p->s > 4 * SUP_KPT) {
      d2mi_set_error("stored-cost B-ADMM handles s <= %d", 4 * SUP_KPT);
      return D2MI_EUNSUPPORTED;
    }
    if ((size_t)p->s * p->dim * (precision == 64 ? 8 : 4) > ctx->lds_max) {
      d2mi_set_error("centroid supports of %d x %d do not fit LDS", p->s, p->dim);
      return D2MI_EUNSUPPORTED;
    }
  }
  if (p->K > 16384) {
    d2mi_set_error("K = %d too large", p->K);
    return D2MI_EUNSUPPORTED;
  }
  if (tune && (precision == 64 ? p->plan_s64 : p->plan_s32) != p->s) {
    d2mi_set_error("d2mi_badmm_tune: run one update of this phase at this precision first");
    return D2MI_EINVAL;
  }
  const int fresh = d2mi_ensure_plans(ctx, p, precision);
  if (fresh < 0) return fresh;
  const bool f = fresh == 1;
  if (table) {
    if (precision == 64)
      return run_badmm<double, COST_TABLE>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
    return run_badmm<float, COST_TABLE>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
  }
  if (precision == 64) {
    switch (p->metric) {
      case D2MI_EUCLIDEAN:
        return run_badmm<double, COST_EUCLID>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
      case D2MI_HISTOGRAM:
        return run_badmm<double, COST_HIST>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
      default:
        return run_badmm<double, COST_SPARSE>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
    }
  }
  switch (p->metric) {
    case D2MI_EUCLIDEAN:
      return run_badmm<float, COST_EUCLID>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
    case D2MI_HISTOGRAM:
      return run_badmm<float, COST_HIST>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
    default:
      return run_badmm<float, COST_SPARSE>(ctx, p, maxIters, rhoCoeff, updatePerLoops, variant, stats, f, tune);
  }
}
}  // namespace

extern "C" int d2mi_badmm(d2mi_ctx *ctx, int phase, int maxIters, double rhoCoeff,
                          int updatePerLoops, int precision, int variant,
                          d2mi_badmm_stats *stats) {
  if (stats) memset(stats, 0, sizeof *stats);
  return badmm_dispatch(ctx, phase, maxIters, rhoCoeff, updatePerLoops, precision, variant,
                        stats, nullptr);
}

extern "C" int d2mi_badmm_tune(d2mi_ctx *ctx, int phase, int precision, int reps, char *report,
                               size_t report_len) {
  TuneReq tq{reps, std::string()};
  // the setup of an update (label order, rho, stored costs) with the
  // reference's default options, then the timing passes; no iteration runs
  const int rc = badmm_dispatch(ctx, phase, 10, 2.0, 10, precision, D2MI_BADMM_ENGINE, nullptr,
                                &tq);
  if (report && report_len) snprintf(report, report_len, "%s", tq.report.c_str());
  return rc;
}

extern "C" int d2mi_badmm_tune_clear(d2mi_ctx *ctx) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  ctx->plan_tune.clear();
  return D2MI_OK;
}

extern "C" int d2mi_badmm_set_graph(d2mi_ctx *ctx, int on) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  ctx->badmm_graph = on ? 1 : 0;
  return D2MI_OK;
}
