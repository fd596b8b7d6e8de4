// Internal declarations shared by the libd2mi.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "d2mi.h"

void d2mi_set_error(const char *fmt, ...);

#define D2MI_HIP(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      d2mi_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return D2MI_EHIP;                                                               \
    }                                                                                 \
  } while (0)

#define D2MI_TRY(x)        \
  do {                     \
    int r_ = (x);          \
    if (r_ != D2MI_OK) return r_; \
  } while (0)

// Device allocation that remembers its size and can grow.
struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want);  // grows (contents not kept) to at least `want` bytes
  int assign(const void *host, size_t nbytes, hipStream_t s);  // ensure + H2D copy
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

// One phase of the data set resident in HBM (the reference's `sph`,
// include/d2/clustering.h:14-95) plus its centroids and B-ADMM state.
struct PhaseDev {
  bool set = false;
  int metric = 0, dim = 0, V = 0, m_max = 0;
  size_t N = 0, col = 0;
  DevMem p_str;                 // int[N]
  DevMem cum;                   // int64[N]
  DevMem w64, w32;              // weights  double/float [col]
  DevMem supp64, supp32;        // supports double/float [col*dim] (Euclidean)
  DevMem xmean, xspread;        // Euclidean: per object, FP64 mean support point [N*dim]
                                // and sum of squared distances to it [N]
  DevMem sym;                   // int[col] word ids / histogram bins
  DevMem vocab64, vocab32;      // [V*dim]
  DevMem vocab16, vnorm16;      // word embeddings: BF16 rows padded to d2mi_pdist2_pad(dim),
                                // FP32 norms of the unrounded FP32 rows (made on first use)
  DevMem dist64, dist32;        // [V*V]
  // centroids (K x s)
  int K = 0, s = 0;
  DevMem cw64, cw32;            // [K*s]
  DevMem cs64, cs32;            // [K*s*dim]
  DevMem cw_old, cs_old;        // FP64 copy taken before a centroid update (Elkan)
  // resident transport plans, element (k, j) of object i at s*cum[i] + k + j*s
  DevMem Z32, Y32, Z64, Y64;
  int plan_s32 = 0, plan_s64 = 0;  // s the plans were sized for (0 = none)
  // stored ground costs and the X of support-update passes (plan layout),
  // for phases whose cost is not recomputed inside the plan kernel
  DevMem Ctab, Xout, Z0;
  // size buckets for batched LPs: objects grouped by support count, so each
  // launch sizes LDS for its own largest object; jobs o*K + c per bucket
  std::vector<int> h_pstr;
  int bucket_K = 0;
  std::vector<int> bucket_mmax;
  std::vector<size_t> bucket_off;  // into bucket_map, in jobs
  DevMem bucket_map;
};

// Plan-kernel choice of one object-size class found by d2mi_badmm_tune.
struct PlanTuneKey {
  int prec, cm, dim, s, d, m_max;
  bool operator<(const PlanTuneKey &o) const {
    return std::tie(prec, cm, dim, s, d, m_max) < std::tie(o.prec, o.cm, o.dim, o.s, o.d, o.m_max);
  }
};
struct PlanTuneVal {
  int kind;     // kernel form (PlanKind in badmm.hip)
  int wg_mode;  // workgroups per CU: 0 occupancy limit, 1 one less, 2 twice
};

struct d2mi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t N = 0;
  DevMem label, lsw;            // int[N], char[N]
  std::vector<PhaseDev> ph;
  // B-ADMM scratch
  DevMem perm, blkcnt, segstart, rec, part, scal, flags, resid, X32, X64;
  DevMem rec2, rxbuf;           // support-numerator records, X row sums per label
  DevMem bkey, permk, blkcnt2, segk;  // B-ADMM order by (object-size bucket, label)
  DevMem sdesc, sblab, sbs, scum, scbt, szero;  // step 5 over packed columns: block descriptors
  DevMem pmeta;                       // B-ADMM: {object, label, m, first column} in plan order
  DevMem wtab, wnorm;           // word embeddings: pdist2(vocab, centroid supports), norms
  DevMem cs16;                  // BF16 centroid supports for the BF16 pdist2
  DevMem rowsum;                // histogram phases: ground-cost line sums for rho
  // pdist2 scratch
  DevMem pA, pB, pC;
  // labeling / Elkan state (clustering.h:106-112), resident in HBM
  int elkan_K = 0;
  DevMem el_l, el_u, el_s, el_c, el_r;  // double[N*K], double[N], double[K], double[K*K], char[N]
  DevMem dist, mask, jobs, lpstat, cnt, mind2, red;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // B-ADMM inner loop as HIP graphs, per (phase, precision and cost kind,
  // maxIters, updatePerLoops); pass_ev: the eager loop's per-pass events
  int badmm_graph = -1;  // -1: environment decides, 0 off, 1 on
  struct BadmmGraph {
    hipGraphExec_t exec = nullptr;
  };
  std::map<std::tuple<int, int, int, int>, BadmmGraph> badmm_exec;
  std::vector<hipEvent_t> pass_ev;
  // plan-kernel tuner: choices per size class, plan copies taken while timing
  std::map<PlanTuneKey, PlanTuneVal> plan_tune;
  DevMem tuneZ, tuneY;
  size_t lds_max = 0;
  int num_cu = 0;
};

int d2mi_check_ctx(d2mi_ctx *ctx);

// pdist2 tables on the matrix cores (pdist2.hip): T (P x Q, FP32) from FP32
// rows x (P x d) and y (Q x d) with FP32 products (nx, ny: norm scratch), or
// with BF16 products from rows padded to d2mi_pdist2_pad(d) and the FP32
// norms (d2mi_rows_to_bf16_dev makes both; x and y stay the FP32 rows for
// the cancellation recompute)
int d2mi_pdist2_mfma_dev(hipStream_t S, int P, int Q, int d, const float *x, const float *y,
                         float *nx, float *ny, float *T);
int d2mi_pdist2_pad(int d);
int d2mi_row_norms_dev(hipStream_t S, int R, int d, const float *x, float *nx);
int d2mi_rows_to_bf16_dev(hipStream_t S, int R, int d, const float *x, __bf16 *xb, float *nx);
int d2mi_pdist2_bf16_dev(hipStream_t S, int P, int Q, int d, const __bf16 *xb, const float *nx,
                         const __bf16 *yb, const float *ny, const float *x, const float *y,
                         float *T);
// takes the device address of the no-allocation scratch array of `device`
// (which must be current); once per device, from d2mi_create (transport.hip)
int d2mi_noalloc_setup(int device);
PhaseDev *d2mi_phase(d2mi_ctx *ctx, int phase);
