/*
 * d2mi.h -- C ABI of libd2mi.so, the MI355X (gfx950) engine for the
 * AD2-clustering hot path of bobye/d2_kmeans.
 *
 * Plain pointers and sizes only.  Host arrays use the reference's layouts
 * (include/d2/clustering.h:14-124 with SCALAR = double): per phase, CSR-style
 * ragged objects (p_str[i] supports starting at column p_str_cum[i]), Euclidean
 * supports row-major per support (p_supp[col*dim]), word ids / histogram bins
 * 0-based in p_supp_sym[col]; centroids are K objects of exactly s supports
 * (centroid_rand.c:177-183): weights p_w[K*s], supports p_supp[K*s*dim].
 *
 * Every function returns 0 on success and a negative D2MI_E* code on error;
 * d2mi_last_error() then returns a message.  A context owns one HIP stream
 * on one device; calls on one context must not overlap.
 */
#ifndef D2MI_H
#define D2MI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define D2MI_OK 0
#define D2MI_EINVAL -1      /* bad argument */
#define D2MI_EHIP -2        /* HIP runtime error */
#define D2MI_ENOMEM -3      /* device allocation failed */
#define D2MI_EEMPTY -4      /* a cluster has no object (centroid_Bregman.c:144) */
#define D2MI_ENUMERIC -5    /* a normaliser was not positive (blas_like64.c:131,148) */
#define D2MI_EUNSUPPORTED -6 /* shape or metric this build does not handle */

/* metric ids, include/d2/param.h:23-28 */
#define D2MI_EUCLIDEAN 0
#define D2MI_HISTOGRAM 5
#define D2MI_WORD_EMBED 7
#define D2MI_SPARSE_HISTOGRAM 12

/* B-ADMM kernel variants */
#define D2MI_BADMM_ENGINE 0 /* LDS-staged workgroup-per-object kernels */
#define D2MI_BADMM_PLAIN 1  /* one thread per object (comparison baseline) */

typedef struct d2mi_ctx d2mi_ctx;

const char *d2mi_last_error(void);
int d2mi_device_count(int *count);

/* Context on `device`.  The reference keeps this state in d2_clustering's
 * locals and var_mph (clustering.c:421-449, clustering_util.c:152-230). */
int d2mi_create(d2mi_ctx **out, int device);
int d2mi_destroy(d2mi_ctx *ctx);
int d2mi_synchronize(d2mi_ctx *ctx);

/* Uploads one phase of the data set (the `sph` filled by d2_read,
 * clustering_io.c:14-202).  Arrays are copied; the caller keeps ownership.
 * Unused arrays may be NULL (p_supp for symbolic metrics, etc.).  Replaces
 * any earlier upload of that phase and discards its B-ADMM state.  Uploading
 * phase 0 starts a new data set: other phases, labels and Elkan bounds are
 * dropped; phases > 0 must then hold the same number of objects. */
int d2mi_set_phase(d2mi_ctx *ctx, int phase, int metric, int dim, size_t size, size_t col,
                   const int *p_str, const size_t *p_str_cum, const double *p_w,
                   const double *p_supp, const int *p_supp_sym, int vocab_size,
                   const double *vocab_vec, const double *dist_mat);

/* Labels (int[N], -1 = none) and label-switch flags (char[N]) shared by all
 * phases (mph.label, var_mph.label_switch; clustering.h:121, :198). */
int d2mi_set_labels(d2mi_ctx *ctx, const int *label, const char *label_switch);
int d2mi_get_labels(d2mi_ctx *ctx, int *label, char *label_switch);

/* Centroids of one phase: K x s supports (the `sph` centroids of
 * d2_init_centroid, clustering.c:379-404). p_supp may be NULL for
 * histogram metrics. */
int d2mi_set_centroids(d2mi_ctx *ctx, int phase, int K, int s, const double *p_w,
                       const double *p_supp);
int d2mi_get_centroids(d2mi_ctx *ctx, int phase, double *p_w, double *p_supp);

typedef struct {
  double rho;          /* rho chosen at setup (centroid_Bregman.c:109) */
  double obj;          /* rho * <C, X> / N at loop exit */
  int n_trace;         /* residual rows recorded (centroid_Bregman.c:300-316) */
  double trace[64][4]; /* (iter+1, obj, primres, dualres) */
  double ms_total;     /* device time of the whole update (HIP events) */
  double ms_plan;      /* summed device time of the plan-update kernel; -1 when the
                          update ran as a graph (HIP events recorded by graph nodes
                          cannot be timed): d2mi_badmm_set_graph(ctx, 0) times it */
  int n_plan_launches; /* launches of the plan-update kernel */
  double plan_elems;   /* s * col, plan elements per inner iteration */
  int graph;           /* 1: the inner iterations ran as one replayed HIP graph */
  char plan_forms[200]; /* per object-size class: "m<=M:form/wgpcu(objects)" */
} d2mi_badmm_stats;

/* One B-ADMM centroid update of `phase` (d2_centroid_sphBregman,
 * centroid_Bregman.c:51-325, warm start c0 == c): centroids are updated in
 * place on the device, the transport plan Z stays resident across calls
 * (warm start) and is re-initialised for objects whose label switched.
 * precision: 32 (FP32 plans, FP64 reductions) or 64 (all FP64).
 * variant: D2MI_BADMM_ENGINE or D2MI_BADMM_PLAIN.  stats may be NULL. */
int d2mi_badmm(d2mi_ctx *ctx, int phase, int maxIters, double rhoCoeff, int updatePerLoops,
               int precision, int variant, d2mi_badmm_stats *stats);

/* The inner iterations of an update (every pass's plan kernels, the
 * per-pass reductions and, for stored costs, the support update and cost
 * rebuild) are captured into one HIP graph and replayed; the executable graph
 * is kept per (phase, precision, maxIters, updatePerLoops) and refreshed in
 * place with hipGraphExecUpdate.  on = 0 launches the same kernels one by one
 * (results are identical) and times every pass (stats ms_plan).  Default: on
 * (environment D2MI_BADMM_GRAPH=0 turns it off).  No counterpart in the
 * reference, whose loop is host code (centroid_Bregman.c:150-320). */
int d2mi_badmm_set_graph(d2mi_ctx *ctx, int on);

/* Per object-size class autotuner of the plan kernel.  For the current
 * labels and resident plans of `phase` at `precision` (one d2mi_badmm call
 * must have run since the last upload), every feasible kernel form
 * (register-resident columns, plan in LDS, general, HBM scratch) and
 * workgroups-per-CU setting (the occupancy limit, one less, twice) is timed
 * over `reps` inner passes per class; the fastest is kept in the context,
 * keyed by (precision, cost kind, s, d, class size limit), and used by later
 * d2mi_badmm calls.  The plans are saved before and restored after, so the
 * call changes no result state; a later update may differ from an untuned
 * one in the last bits (another summation order).  report (may be NULL)
 * receives one line per class.  No counterpart in the reference. */
int d2mi_badmm_tune(d2mi_ctx *ctx, int phase, int precision, int reps, char *report,
                    size_t report_len);
/* Drops every tuned choice of the context (rule-based forms again). */
int d2mi_badmm_tune_clear(d2mi_ctx *ctx);

/* Copies the resident plan (which: 0 = Z, 1 = Y) of `phase` at `precision`
 * to host memory as double[s*col]. */
int d2mi_get_plan(d2mi_ctx *ctx, int phase, int precision, int which, double *out);
/* Sets the resident plan Z (double[s*col]) at `precision`. */
int d2mi_set_plan(d2mi_ctx *ctx, int phase, int precision, const double *Z);

/* Device characterisation, no counterpart in the reference: the best rate at
 * which this GPU streams two FP32 arrays of n elements in place (each element
 * read and written once in both arrays, 16 B per element), over dword and
 * float4 grid-stride kernels at 8, 16 and 32 workgroups per CU, best of `reps`
 * launches each after one warm-up.  That is the access pattern of a B-ADMM
 * plan pass without its arithmetic, so it is the practical ceiling for the
 * plan kernels.  Allocates 8n bytes for the call.  report (may be NULL)
 * receives one line per variant. */
int d2mi_stream_probe(d2mi_ctx *ctx, size_t n, int reps, double *gbps, char *report,
                      int report_len);

/* Squared Euclidean distance matrix, the reference's _dpdist2
 * (blas_like64.c:199-208, include/utils/blas_like.h:48):
 * C[i*n + j] = sum_k (A[j*d + k] - B[i*d + k])^2, A: n x d, B: m x d,
 * host arrays, computed on the device in FP64. */
int d2mi_pdist2(d2mi_ctx *ctx, int d, size_t n, size_t m, const double *A, const double *B,
                double *C);

/* Exact transport LP values of `nbatch` independent problems of one shape,
 * the LP the reference's d2_match_by_distmat poses to Mosek
 * (solver_mosek.cc:61-143, include/d2/solver.h:14-15):
 *   min sum C[b][i + j*n] x_ij, x >= 0, row sums = wX[b][i] (n rows),
 *   column sums = wY[b][j] for j < m-1 (the last column constraint dropped).
 * C: nbatch*n*m, wX: nbatch*n, wY: nbatch*m host arrays; out[b] = optimum
 * (FP64, solved on the device); pivots (may be NULL) = simplex pivots. */
int d2mi_transport_batch(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                         const double *wX, const double *wY, double *out, int *pivots);

/* Phases of one transport LP, in the order d2mi_transport_profile reports them. */
#define D2MI_LP_PHASES 12
/* d2mi_transport_batch that also reports where each LP spent its time; no
 * counterpart in the reference (Mosek is closed).  phase_us[b*12 + p] (host,
 * nbatch*12 doubles) receives LP b's wall time in microseconds in phase p:
 * 0 cost table, 1 entropic duals, 2 starting basis (Vogel or north-west),
 * 3 start rooted into a tree, 4 pointer jumping (summed over pivots),
 * 5 pricing (summed), 6 pivots (apex, ratio test, flows, re-hang; summed),
 * 7 objective; and within phase 2, Vogel's steps split into 8 finding the
 * line of largest penalty, 9 taking its cell, 10 finding the lines whose
 * two cheapest cells were crossed out, 11 walking them again.  Values and
 * pivots equal d2mi_transport_batch's. */
int d2mi_transport_profile(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                           const double *wX, const double *wY, double *out, int *pivots,
                           double *phase_us);

/* Size of the fixed per-GPU device array behind
 * d2mi_transport_batch_noalloc (32 MiB). */
#define D2MI_NOALLOC_SCRATCH_BYTES (32u << 20)

/* d2mi_transport_batch without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as d2mi_transport_batch; the device staging of
 * the costs, weights, values, pivot counts, job lists and status lives in
 * one fixed array of D2MI_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends.  A batch needing more (about 8*nbatch*(n*m + n + m + 1) +
 * 12*nbatch bytes plus alignment) is refused with D2MI_EINVAL.  Every call
 * on a GPU shares that GPU's array, so calls that overlap on one GPU (from
 * two threads or two contexts) race and must be serialised by the caller. */
int d2mi_transport_batch_noalloc(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                                 const double *wX, const double *wY, double *out, int *pivots);

/* pdist2 on the matrix cores (_dpdist2, blas_like64.c:199-208, as a GEMM):
 * C[i*n + j] = |A_j - B_i|^2 for A: n x d, B: m x d (row-major, host), from
 * FP32 inputs with FP32 norms and accumulation; bf16 = 1 runs the products
 * in BF16 (operands rounded to BF16 once, norms of the FP32 rows).  Entries
 * small next to |A_j|^2 + |B_i|^2 are recomputed by direct differences.
 * ms (may be NULL) = time of the GEMM kernel alone (norms and the BF16
 * conversion run before it). */
int d2mi_pdist2_mfma(d2mi_ctx *ctx, int d, size_t n, size_t m, const double *A, const double *B,
                     double *C, int bf16, float *ms);

/* ---------------------------------------------------------- assignment */

typedef struct {
  unsigned long long changes;    /* objects whose label changed */
  unsigned long long dist_count; /* exact distances computed (clustering.c:237) */
  double cost;                   /* sum of min d^2 (brute force only; NaN with Elkan) */
  double ms;                     /* device time */
} d2mi_label_stats;

/* Batched d2_compute_distance (clustering.c:33-128) between resident objects
 * and resident centroids: out[t] = sqrt(sum over the selected phases of the
 * exact transport-LP value between object obj[t] and centroid cent[t]).
 * selected_phase < 0 means all phases. */
int d2mi_distances(d2mi_ctx *ctx, int selected_phase, size_t npairs, const int *obj,
                   const int *cent, double *out);

/* Resets the resident Elkan bounds: l = 0, u = +max, r = 1
 * (d2_allocate_work with use_triangle, clustering_util.c:218-228). */
int d2mi_elkan_reset(d2mi_ctx *ctx, int selected_phase);

/* Relabels every object against the resident centroids: use_elkan = 0 is
 * d2_labeling (clustering.c:324-376, brute force, first minimum wins, cost
 * = sum of min d^2); use_elkan = 1 is d2_labeling_prep (clustering.c:141-242)
 * with the bounds resident on the device.  Labels and label-switch flags are
 * updated on the device; stats may be NULL. */
int d2mi_labeling(d2mi_ctx *ctx, int selected_phase, int use_elkan, d2mi_label_stats *stats);

/* Keeps a copy of the current centroids (d2_copy before the update,
 * clustering.c:484) for d2mi_elkan_post. */
int d2mi_save_centroids(d2mi_ctx *ctx, int selected_phase);

/* d2_labeling_post (clustering.c:293-318): centroid drifts between the saved
 * and the current centroids relax the resident bounds. */
int d2mi_elkan_post(d2mi_ctx *ctx, int selected_phase);

typedef struct {
  int iterations;                /* outer iterations run */
  long long changes[256];        /* label changes per outer iteration */
  long long dist_count[256];     /* distances computed per outer iteration */
  double final_cost;             /* mean of min d^2 of the closing labeling */
  double ms_labeling, ms_badmm;  /* device time per stage */
} d2mi_cluster_stats;

/* The d2_clustering outer loop (clustering.c:409-525) on resident data and
 * centroids: labels start at -1; each iteration relabels (Elkan when
 * use_triangle), stops when fewer than 0.001*N labels change, and runs one
 * B-ADMM update per selected phase; with use_triangle a closing brute-force
 * labeling gives the final cost.  B-ADMM options as in d2mi_badmm. */
int d2mi_clustering(d2mi_ctx *ctx, int K, int max_iter, int selected_phase, int use_triangle,
                    int maxIters, double rhoCoeff, int updatePerLoops, int precision,
                    d2mi_cluster_stats *stats);

/* ------------------------------------------------------------- serving */

/* Exact nearest-centroid queries with continuous batching against the
 * resident centroids of one Euclidean phase: per query, the label and the
 * distance that the reference's --eval mode computes for one object
 * (d2_assignment -> d2_labeling, clustering.c:324-376, 528-546; main.cc:253-275):
 * smallest sqrt(transport-LP value), first minimum on ties.  Queries are
 * copied at submission; progress() admits waiting queries as streams and HBM
 * free up (one size class per launch group), launches their work and
 * collects finished groups without blocking. */
typedef struct d2mi_server d2mi_server;

typedef struct {
  int max_batch;     /* queries per launch group (0: 48) */
  int streams;       /* launch groups in flight at once (0: 6, at most 16); with 3 or
                        more, streams 0-2 are kept for the large, middle and small size
                        class and the rest go to the class of the oldest waiting query */
  size_t hbm_bytes;  /* device memory for admitted queries (0: 1 GiB) */
  int max_m;         /* largest query accepted (0 or above 2048: 2048 support points) */
  int use_triangle;  /* 1: also prune with W2(c1, c) > 2 W2(x, c1) (Elkan's test) */
} d2mi_serve_opts;

typedef struct {
  long long submitted, completed;
  long long groups;               /* launch groups run */
  long long lps;                  /* exact transport LPs solved */
  long long memory_waits;         /* admissions deferred for lack of HBM */
  long long max_inflight_groups;  /* most groups on the device at once */
  long long max_group;            /* largest group admitted */
} d2mi_serve_stats;

/* Opens a server on the centroids resident for `phase` (d2mi_set_centroids);
 * opts may be NULL.  max_batch = 1 with streams = 1 serves one query per
 * launch, in arrival order (the plain form). */
int d2mi_serve_open(d2mi_ctx *ctx, int phase, const d2mi_serve_opts *opts, d2mi_server **out);
/* Queues one query: m support points (supp: m x d row-major, w: m weights).
 * The weights are renormalised to sum 1, as d2_read does for every object
 * (clustering_io.c:143-151); a weight that is not finite or is <= 1e-9 is
 * refused with D2MI_EINVAL (the reference asserts there, :145). */
int d2mi_serve_submit(d2mi_server *srv, int m, const double *w, const double *supp,
                      long long *ticket);
/* Collects finished groups and admits waiting queries; *pending (may be
 * NULL) = submitted minus completed. */
int d2mi_serve_progress(d2mi_server *srv, int *pending);
/* 1 and the result when the query is done (the ticket is then released),
 * 0 while it is pending; latency_ms = completion minus submission. */
int d2mi_serve_result(d2mi_server *srv, long long ticket, int *label, double *dist,
                      double *latency_ms);
/* Blocks until every submitted query is done. */
int d2mi_serve_drain(d2mi_server *srv);
int d2mi_serve_stats_get(d2mi_server *srv, d2mi_serve_stats *stats);
int d2mi_serve_close(d2mi_server *srv);

#ifdef __cplusplus
}
#endif
#endif
