/*
 * d2mi_clustering.h -- the reference-compatible C ABI of libd2mi.so.
 *
 * A program written against bobye/d2_kmeans's libad2c (include/d2/clustering.h,
 * include/d2/solver.h, include/utils/blas_like.h, built with SCALAR = double)
 * links against libd2mi.so instead and runs the hot path on the MI355X: the
 * types below have the reference's member order and types, so its structs
 * pass through unchanged, and every function keeps its name, arguments and
 * return convention.
 *
 * Device state: one engine context per process (device 0, or $D2MI_DEVICE),
 * created on first use.  Data, labels, centroids, the B-ADMM plans and the
 * Elkan bounds live in HBM between calls; the host arrays of var_mph (C, X,
 * Z, Y, the trieq bounds) are not read or written by the device path.
 * $D2MI_PRECISION = 64 runs B-ADMM in FP64 (default 32: FP32 plans, FP64
 * reductions).  Only the B-ADMM centroid method (d2_alg_type == 0) exists on
 * the device; the deprecated ADMM / gradient methods (clustering.h:233-246)
 * return -1.
 */
#ifndef D2MI_CLUSTERING_H
#define D2MI_CLUSTERING_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* metric ids (param.h:23-28) and centroid methods (param.h:17-19) */
#define D2_EUCLIDEAN_L2 0
#define D2_CITYBLOCK_L1 1
#define D2_HISTOGRAM 5
#define D2_N_GRAM 6
#define D2_WORD_EMBED 7
#define D2_SPARSE_HISTOGRAM 12
#define D2_CENTROID_BADMM 0
#define D2_CENTROID_ADMM 1
#define D2_CENTROID_GRADDEC 2
#define D2_GROUP_SIZE 7

/* BADMM_options, param.h:4-8 */
typedef struct {
  int maxIters;
  double rhoCoeff;
  int updatePerLoops;
} BADMM_options;

/* sph, clustering.h:14-95 */
typedef struct {
  int dim, str, max_str;
  size_t col, max_col;
  int *p_str;
  size_t *p_str_cum;
  double *p_w;
  int metric_type;
  double *p_supp;
  int *p_supp_sym;
  double *dist_mat;
  int vocab_size;
  double *vocab_vec;
  char is_meta_allocated;
} sph;

/* trieq, clustering.h:106-112 */
typedef struct {
  double *l, *u, *s, *c;
  char *r;
} trieq;

/* mph, clustering.h:117-124 */
typedef struct {
  int s_ph;
  size_t size;
  size_t global_size;
  int *label;
  int num_of_labels;
  sph *ph;
} mph;

/* var_sph, var_sphBregman, var_mph: clustering.h:176-200 */
typedef struct {
  double *C, *X, *L;
} var_sph;
typedef struct {
  double *X, *Z, *Y, *Xc, *Zr;
} var_sphBregman;
typedef struct {
  int s_ph;
  var_sph *g_var;
  var_sphBregman *l_var_sphBregman;
  char *label_switch;
  trieq tr;
} var_mph;

/* globals the reference's callers set (main.cc:24, 118, 234-236;
 * centroid_Bregman.c:17-22; clustering.c:24-26) */
extern int d2_alg_type;
extern BADMM_options badmm_clu_options; /* {100, 2.0, 10} */
extern BADMM_options badmm_cen_options; /* {2000, 1.0, 10} */
extern BADMM_options *p_badmm_options;
extern int world_rank;
extern int nprocs;

/* allocation and text I/O (clustering.h:130-150; clustering_util.c:18-144,
 * clustering_io.c:14-324) */
int d2_allocate_sph(sph *p_data_sph, const int d, const int stride, const size_t num,
                    const double semicol, const int type);
int d2_allocate(mph *p_data, const int size_of_phases, const size_t size_of_samples,
                const int *avg_strides, const int *dimension_of_phases,
                const int *type_of_phases);
int d2_read(const char *filename, const char *meta_filename, mph *p_data);
int d2_write(const char *filename, mph *p_data);
int d2_write_labels(const char *filename, mph *p_data);
int d2_write_labels_serial(const char *filename_ind, const char *filename, mph *p_data);
/* shards for the reference's MPI runs (clustering.h:151, clustering_io.c:329-480):
 * <filename>.ind and <filename>.0 ... <filename>.<splits-1> */
int d2_write_split(const char *filename, mph *p_data, int splits, char is_pre_processed);
int d2_free(mph *p_data);

/* centroids (clustering.h:152-155, 216-218; clustering.c:379-404,
 * centroid_rand.c:104-296) */
int d2_init_centroid(mph *p_data, mph *centroids, int selected_phase, int allocate_only);
int d2_centroid_rands(mph *p_data, int idx_ph, sph *c);

/* the clustering driver and assignment (clustering.h:158-171) */
int d2_clustering(int k, int max_iter, mph *p_data, mph *centroids, int selected_phase,
                  char use_triangle, const char *log_file);
int d2_assignment(int k, mph *p_data, mph *centroids, int selected_phase,
                  const char *centroid_filename, const char *meta_filename);

/* work arrays (clustering.h:202-203, clustering_util.c:152-259): host-side
 * bookkeeping only -- the device keeps its own copies */
int d2_allocate_work(mph *p_data, var_mph *var_work, char use_triangle, int selected_phase);
int d2_free_work(var_mph *var_work, int selected_phase);

/* the B-ADMM plug point (clustering.h:226-230, centroid_Bregman.c:51-325) */
int d2_centroid_sphBregman(mph *p_data, var_mph *var_work, int idx_ph, sph *c0, sph *c);

/* labeling with external linkage but no header in the reference
 * (clustering.c:33, 141, 293, 324) */
double d2_compute_distance(mph *a, size_t i, mph *b, size_t j, int selected_phase,
                           var_mph *var_work, size_t index_task);
size_t d2_labeling_prep(mph *p_data, mph *centroids, var_mph *var_work, int selected_phase);
size_t d2_labeling_post(mph *p_data, mph *c_old, mph *c_new, var_mph *var_work,
                        int selected_phase);
size_t d2_labeling(mph *p_data, mph *centroids, var_mph *var_work, int selected_phase);

/* the solver boundary (include/d2/solver.h:10-15; solver_mosek.cc:25-226):
 * exact transport LP value, solved on the device; x and lambda are not
 * produced (the hot path passes NULL, clustering.c:58-59) */
void d2_solver_setup(void);
void d2_solver_release(void);
double d2_match_by_distmat(int n, int m, double *C, double *wX, double *wY, double *x,
                           double *lambda, size_t index);

/* pdist2 (include/utils/blas_like.h:48, blas_like64.c:199-208) */
void _dpdist2(int d, size_t n, size_t m, double *A, double *B, double *C);

#ifdef __cplusplus
}
#endif
#endif
