// Batched exact transport LPs on the device (transport.hip).
#pragma once

#include "d2mi_internal.h"

enum LpCostMode {
  LP_COST_EUCLID = 0,  // squared Euclidean distance of supports held in LDS
  LP_COST_HIST = 1,    // dense histogram ground cost dist[k + j*V]
  LP_COST_SPARSE = 2,  // sparse histogram: dist[id_j * V + k]
  LP_COST_GLOBAL = 3,  // precomputed costs gcost[job*gstride + k + j*n]
  LP_COST_WORD = 4,    // word embeddings: centroid rows against vocab[sym[j]]
};

// One LP per job: rows = the sA supports of entry job_a[job] of centroid
// table A, columns = object job_b[job] of a phase (b_is_table == 0) or the sB
// supports of entry job_b[job] of centroid table B.  The LP value is ADDED
// to out[job] (phases accumulate in call order).
struct LpArgs {
  int njobs;
  const int *job_a;
  const int *job_b;
  int implicit_K;              // > 0: job -> (job_b = job / K, job_a = job % K), no job arrays
  const int *job_map;          // with implicit_K: launch job t is global job job_map[t]
                               // (mask, out and the pair are those of the global job)
  const unsigned char *mask;   // optional: jobs with mask[job] == 0 are skipped
  const int *njobs_dev;        // optional: the launch-job count, read on the device
                               // (at most njobs; with job_map, a compacted job list)
  int b_is_table;
  int cost_mode;
  int d;
  int sA;
  const double *A_supp;
  const double *A_w;
  const int *p_str;
  const long long *cum;
  const double *w;
  const double *supp;
  const int *sym;
  int sB;
  const double *B_supp;
  const double *B_w;
  int V;
  const double *dist;
  const double *vocab;  // LP_COST_WORD: [V*d]
  const double *gcost;
  long long gstride;
  int n_max, m_max;
  // launch shape, filled by d2mi_plan_transport
  int threads, cpt, table, stage_a, stage_b, vogel;
  // Vogel's reduced-cost start: each LP takes lp::dual_schedule(n * m) of its
  // own size (so its result does not depend on the other LPs of the launch)
  // when dual_ok and n * m >= dual_min; the fields below override the
  // schedule when >= 0 (dual_iters 0: plain Vogel)
  int dual_ok;
  long long dual_min;
  int dual_iters;
  float dual_eps0, dual_eps1;
  int dual_reduce;
  size_t lds;
  int threads_hint;     // > 0: workgroup size to use
  hipStream_t stream;   // null: the context stream
  double *out;
  int *pivots;  // may be null
  int *status;  // bit 0: an LP hit the pivot limit
  long long *prof;  // optional: lp::PROF_N phase times per job (wall-clock ticks)
};

int d2mi_plan_transport(d2mi_ctx *ctx, LpArgs &a);
// Exact W2 between all pairs of resident centroids (labeling.hip).
int d2mi_centroid_pairs(d2mi_ctx *ctx, int selected_phase, int K, double *c_dev, double *s_dev);
int d2mi_launch_transport(d2mi_ctx *ctx, const LpArgs &a);
int d2mi_zero_doubles(d2mi_ctx *ctx, double *p, int n);
int d2mi_zero_ints(d2mi_ctx *ctx, int *p, int n);
int d2mi_sqrt_clamp(d2mi_ctx *ctx, double *p, int n);
