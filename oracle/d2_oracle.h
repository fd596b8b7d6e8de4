/*
 * d2_oracle.h -- FP64 CPU restatement of the AD2-clustering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (d2_kmeans_amd/)
 * links, loads or calls this code: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * Each function names the reference lines (bobye/d2_kmeans) it restates.
 * Data layout follows the reference's D2 model: per phase, CSR-style ragged
 * objects (p_str[i] supports starting at column p_str_cum[i]); per-object
 * s x m_i matrices (cost C, plans X/Z/Y) are column-major, element (k, j) at
 * k + j*s, object i's block starting at s*p_str_cum[i]
 * (centroid_Bregman.c:118-122, blas_like64.c:207).
 */
#ifndef D2_ORACLE_H
#define D2_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* metric ids, param.h:23-28 */
#define OR_EUCLIDEAN 0
#define OR_HISTOGRAM 5
#define OR_WORD_EMBED 7
#define OR_SPARSE_HISTOGRAM 12

typedef struct {
  int metric;               /* OR_* */
  int dim;                  /* support dimension (0 for histograms) */
  size_t size;              /* number of objects */
  size_t col;               /* total supports = sum p_str */
  const int *p_str;         /* [size] */
  const size_t *p_str_cum;  /* [size] */
  const double *p_w;        /* [col] */
  const double *p_supp;     /* [col*dim] Euclidean */
  const int *p_supp_sym;    /* [col] word ids / histogram bins (0-based) */
  int vocab_size;           /* V */
  const double *vocab_vec;  /* [V*dim] word embedding */
  const double *dist_mat;   /* [V*V] histogram ground cost */
} or_phase;

typedef struct {
  int K, s, dim;
  double *p_w;     /* [K*s] */
  double *p_supp;  /* [K*s*dim] (Euclidean and word-embed) */
} or_centroids;

typedef struct {
  int maxIters;
  double rhoCoeff;
  int updatePerLoops;
} or_badmm_opts;

/* pdist2 family, blas_like64.c:199-230 */
void or_pdist2(int d, size_t n, size_t m, const double *A, const double *B, double *C);
void or_pdist2_sym(int d, size_t n, size_t m, const double *A, const int *Bi, double *C,
                   const double *vocab);
void or_pdist2_submat(size_t m, const int *Bi, double *C, int V, const double *dist_mat);

/* calculate_distmat, centroid_util.c:8-36: C for every object against its label */
void or_calculate_distmat(const or_phase *ph, const int *label, const or_centroids *c,
                          double *C);

/* B-ADMM centroid update, centroid_Bregman.c:51-325 (warm start c0 == c).
 * Z (in/out, warm-started), X, Y: s*col work arrays; C: s*col cost scratch.
 * rho_out: rho chosen at setup (:109).  trace: up to max_trace rows of
 * (iter+1, obj, primres, dualres) as printed at :315.  obj_out: rho*<C,X>/N
 * for the last plan X and the cost C it was computed with (before the last
 * support update rebuilds C).  Returns 0, or -1 when a cluster is empty
 * (:144). */
int or_badmm(const or_phase *ph, const int *label, const char *label_switch,
             or_centroids *c, double *C, double *X, double *Z, double *Y,
             const or_badmm_opts *opt, double *rho_out, double *trace, int max_trace,
             int *n_trace, double *obj_out);

/* Exact transport LP value, the problem d2_match_by_distmat poses
 * (solver_mosek.cc:61-143): min sum C[i + j*n] x_ij, x >= 0,
 * row sums = wX (n rows), column sums = wY for j < m-1 (last one dropped). */
double or_transport(int n, int m, const double *C, const double *wX, const double *wY);
/* same, also returning the simplex pivot count */
double or_transport_ex(int n, int m, const double *C, const double *wX, const double *wY,
                       int *pivots);

/* d2_compute_distance, clustering.c:33-128: sqrt of the phase-summed LP
 * values between object i of (a) and object j of (b).  b_is_centroid tells
 * which of the two holds centroid supports. */
double or_distance_obj_cent(int s_ph, const or_phase *ph, int selected_phase, size_t i,
                            const or_centroids *c, int j);
double or_distance_cent_cent(int s_ph, const or_phase *ph, int selected_phase,
                             const or_centroids *ca, int i, const or_centroids *cb, int j);

/* d2_labeling, clustering.c:324-376 (brute force, first minimum wins).
 * Returns the number of label changes; *cost = sum of min d^2. */
size_t or_labeling(int s_ph, const or_phase *ph, int selected_phase, const or_centroids *c,
                   int K, int *label, char *label_switch, double *cost);

/* Elkan state, clustering.h:106-112 */
typedef struct {
  double *l; /* [N*K] */
  double *u; /* [N] */
  double *s; /* [K] */
  double *c; /* [K*K] */
  char *r;   /* [N] */
} or_elkan;

/* d2_labeling_prep, clustering.c:141-242; returns label changes,
 * *dist_count = number of distance evaluations. */
size_t or_labeling_prep(int s_ph, const or_phase *ph, int selected_phase,
                        const or_centroids *c, int K, int *label, char *label_switch,
                        or_elkan *tr, size_t *dist_count);
/* d2_labeling_post, clustering.c:293-318 */
void or_labeling_post(int s_ph, const or_phase *ph, int selected_phase,
                      const or_centroids *c_old, const or_centroids *c_new, int K,
                      const int *label, or_elkan *tr);

/* shuffle, math.c:39-56 (glibc rand, srand(time*rank) == srand(0) serially) */
void or_shuffle(size_t *array, size_t n, unsigned seed);
/* merge, centroid_rand.c:104-158 */
void or_merge(int dim, const double *m_supp, const double *m_w, int m, double *c_supp,
              double *c_w, int n);
/* d2_centroid_rands, centroid_rand.c:161-296 (Euclidean, word-embed,
 * histogram).  Returns 0, or 1 when too few objects have >= s supports. */
int or_centroid_rands(const or_phase *ph, int K, int s, or_centroids *c, unsigned seed);

/* d2_clustering outer loop, clustering.c:409-525.  centroids: one per phase
 * (entries of unselected phases are ignored); they must hold the initial
 * centroids.  label: [N] out.  Per-outer-iteration trace: label changes and
 * distance counts; final_cost = mean of min d^2 from the closing labeling.
 * Returns the number of outer iterations run. */
int or_clustering(int s_ph, const or_phase *ph, int selected_phase, int K, int max_iter,
                  int use_triangle, const or_badmm_opts *opt, or_centroids *cents,
                  int *label, size_t *trace_changes, int max_trace, double *final_cost);

#ifdef __cplusplus
}
#endif
#endif
