/*
 * d2_oracle.c -- FP64 CPU restatement of the AD2-clustering hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  Written from the reference's equations and loop order; every
 * function cites the bobye/d2_kmeans lines it follows.
 *
 * Deliberate differences from the reference, each one documented:
 *   - no wall-clock budget break in B-ADMM (centroid_Bregman.c:318): it is
 *     nondeterministic; the oracle always runs all maxIters+1 passes;
 *   - the exact transport LP is this file's own transportation simplex, not
 *     Mosek's network simplex (solver_mosek.cc:121): the optimum value is
 *     unique, so only rounding (~1e-15 relative) can differ;
 *   - dense histograms use the phase's own ground-cost matrix where the
 *     reference indexes the phase array by object (clustering_util.c:187).
 */
#include "d2_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ROUNDOFF 1e-9 /* centroid_Bregman.c:20 */

/* ---------------------------------------------------------------- pdist2 */

/* blas_like64.c:199-208: C[i*n + j] = sum_k (A[j*d+k] - B[i*d+k])^2 */
void or_pdist2(int d, size_t n, size_t m, const double *A, const double *B, double *C) {
  for (size_t i = 0; i < m; ++i)
    for (size_t j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int k = 0; k < d; ++k) {
        double t = A[j * d + k] - B[i * d + k];
        acc += t * t;
      }
      C[i * n + j] = acc;
    }
}

/* blas_like64.c:210-220 (negative id: squared norm of the A column) */
void or_pdist2_sym(int d, size_t n, size_t m, const double *A, const int *Bi, double *C,
                   const double *vocab) {
  for (size_t i = 0; i < m; ++i)
    for (size_t j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int k = 0; k < d; ++k) {
        if (Bi[i] < 0) {
          acc += A[j * d + k] * A[j * d + k];
        } else {
          double t = A[j * d + k] - vocab[(size_t)Bi[i] * d + k];
          acc += t * t;
        }
      }
      C[i * n + j] = acc;
    }
}

/* blas_like64.c:222-230 */
void or_pdist2_submat(size_t m, const int *Bi, double *C, int V, const double *dist_mat) {
  for (size_t i = 0; i < m; ++i)
    for (int j = 0; j < V; ++j) C[i * V + j] = dist_mat[(size_t)Bi[i] * V + j];
}

/* centroid_util.c:8-36 (histograms: C is the fixed ground cost, built once
 * as clustering_util.c:183-199 does) */
void or_calculate_distmat(const or_phase *ph, const int *label, const or_centroids *c,
                          double *C) {
  const int s = c->s, d = ph->dim;
  for (size_t i = 0; i < ph->size; ++i) {
    const size_t off = ph->p_str_cum[i];
    const int m = ph->p_str[i];
    double *Ci = C + (size_t)s * off;
    switch (ph->metric) {
      case OR_EUCLIDEAN:
        or_pdist2(d, s, m, c->p_supp + (size_t)label[i] * s * d, ph->p_supp + off * d, Ci);
        break;
      case OR_WORD_EMBED:
        or_pdist2_sym(d, s, m, c->p_supp + (size_t)label[i] * s * d, ph->p_supp_sym + off, Ci,
                      ph->vocab_vec);
        break;
      case OR_HISTOGRAM:
        /* object has m == V bins, centroid s == V: C = dist_mat */
        memcpy(Ci, ph->dist_mat, sizeof(double) * (size_t)s * m);
        break;
      case OR_SPARSE_HISTOGRAM:
        or_pdist2_submat(m, ph->p_supp_sym + off, Ci, ph->vocab_size, ph->dist_mat);
        break;
    }
  }
}

/* ---------------------------------------------------------------- B-ADMM */

/* _dcnorm over an (rows x ncols) column-major block, blas_like64.c:122-136 */
static int cnorm(size_t rows, size_t ncols, double *a) {
  for (size_t j = 0; j < ncols; ++j) {
    double s = 0.0;
    for (size_t k = 0; k < rows; ++k) s += a[j * rows + k];
    if (!(s > 0)) return -1;
    for (size_t k = 0; k < rows; ++k) a[j * rows + k] /= s;
  }
  return 0;
}

int or_badmm(const or_phase *ph, const int *label, const char *label_switch,
             or_centroids *c, double *C, double *X, double *Z, double *Y,
             const or_badmm_opts *opt, double *rho_out, double *trace, int max_trace,
             int *n_trace, double *obj_out) {
  const size_t N = ph->size, col = ph->col;
  const int K = c->K, s = c->s, d = ph->dim;
  const size_t sc = (size_t)s * col;
  const int *p_str = ph->p_str;
  const size_t *cum = ph->p_str_cum;
  double *Z0 = (double *)malloc(sizeof(double) * sc);
  double *Zr = (double *)malloc(sizeof(double) * (size_t)s * (N > (size_t)K ? N : (size_t)K));
  size_t *count = (size_t *)calloc(K, sizeof(size_t));
  int ntr = 0, rc = 0;

  /* setup: :106-110 */
  or_calculate_distmat(ph, label, c, C);
  double asum = 0.0;
  for (size_t e = 0; e < sc; ++e) asum += fabs(C[e]);
  const double rho = opt->rhoCoeff * asum / (double)sc;
  for (size_t e = 0; e < sc; ++e) C[e] /= rho;

  /* Z re-initialised for objects whose label switched: :116-124 */
  for (size_t i = 0; i < N; ++i)
    if (label_switch[i] == 1) {
      double *Zi = Z + (size_t)s * cum[i];
      const double *wi = ph->p_w + cum[i];
      const double *cw = c->p_w + (size_t)label[i] * s;
      for (int j = 0; j < p_str[i]; ++j)
        for (int k = 0; k < s; ++k) Zi[k + (size_t)j * s] = wi[j] * cw[k];
    }
  for (size_t e = 0; e < sc; ++e) Y[e] = 0.0; /* :128 */

  for (size_t i = 0; i < N; ++i) ++count[label[i]]; /* :138-144 */
  for (int k = 0; k < K; ++k)
    if (count[k] == 0) rc = -1;
  if (rc) goto done;

  for (int iter = 0; iter <= opt->maxIters; ++iter) { /* :150 */
    /* step 1, :155-159 */
    for (size_t e = 0; e < sc; ++e) X[e] = Z[e] * exp(-(C[e] + Y[e])) + ROUNDOFF;
    if (cnorm(s, col, X)) { rc = -2; goto done; }
    for (size_t j = 0; j < col; ++j)
      for (int k = 0; k < s; ++k) X[j * s + k] *= ph->p_w[j];

    /* step 2, :164-171 */
    for (size_t e = 0; e < sc; ++e) {
      Z0[e] = Z[e];
      Z[e] = X[e] * exp(Y[e]) + ROUNDOFF;
    }
    for (size_t i = 0; i < N; ++i) {
      double *Zi = Z + (size_t)s * cum[i];
      double *zr = Zr + (size_t)s * i;
      const double *cw = c->p_w + (size_t)label[i] * s;
      for (int k = 0; k < s; ++k) zr[k] = 0.0;
      for (int j = 0; j < p_str[i]; ++j)
        for (int k = 0; k < s; ++k) zr[k] += Zi[k + (size_t)j * s];
      for (int k = 0; k < s; ++k)
        if (!(zr[k] > 0)) { rc = -2; goto done; }
      for (int j = 0; j < p_str[i]; ++j)
        for (int k = 0; k < s; ++k) Zi[k + (size_t)j * s] /= zr[k];
      for (int j = 0; j < p_str[i]; ++j)
        for (int k = 0; k < s; ++k) Zi[k + (size_t)j * s] *= cw[k];
    }

    /* step 3, :175 */
    for (size_t e = 0; e < sc; ++e) Y[e] += X[e] - Z[e];

    /* step 4, :179-190 */
    for (size_t e = 0; e < (size_t)K * s; ++e) c->p_w[e] = 0.0;
    if (cnorm(s, N, Zr)) { rc = -2; goto done; }
    for (size_t i = 0; i < N; ++i) {
      double *cw = c->p_w + (size_t)label[i] * s;
      const double *zr = Zr + (size_t)s * i;
      for (int k = 0; k < s; ++k) cw[k] += zr[k];
    }
    if (cnorm(s, K, c->p_w)) { rc = -2; goto done; }

    /* objective of the last plan under the cost it was computed with
     * (before step 5 rebuilds C): rho * <C, X> / N */
    if (iter == opt->maxIters && obj_out) {
      double obj = 0.0;
      for (size_t e = 0; e < sc; ++e) obj += C[e] * X[e];
      *obj_out = obj * rho / (double)N;
    }

    /* step 5, :194-267 */
    if (iter % opt->updatePerLoops == 0 &&
        (ph->metric == OR_EUCLIDEAN || ph->metric == OR_WORD_EMBED)) {
      const size_t sd = (size_t)s * d;
      for (size_t e = 0; e < sd * K; ++e) c->p_supp[e] = 0.0;
      for (size_t e = 0; e < (size_t)K * s; ++e) Zr[e] = 0.0;
      for (size_t i = 0; i < N; ++i) {
        double *cs = c->p_supp + (size_t)label[i] * sd;
        double *zr = Zr + (size_t)label[i] * s;
        const double *Xi = X + (size_t)s * cum[i];
        for (int j = 0; j < p_str[i]; ++j) {
          const double *xj = ph->metric == OR_EUCLIDEAN
                                 ? ph->p_supp + (cum[i] + j) * d
                                 : ph->vocab_vec + (size_t)ph->p_supp_sym[cum[i] + j] * d;
          for (int k = 0; k < s; ++k) {
            const double x = Xi[k + (size_t)j * s];
            /* supp_i * X_i^T (:206-208) / the vocab gather (:238-245) */
            for (int dd = 0; dd < d; ++dd) cs[(size_t)k * d + dd] += x * xj[dd];
            zr[k] += x; /* :211 / :244 */
          }
        }
      }
      for (int l = 0; l < K; ++l) /* _dirms, :219-221 */
        for (int k = 0; k < s; ++k) {
          const double z = Zr[(size_t)l * s + k];
          if (!(z > 0)) { rc = -2; goto done; }
          for (int dd = 0; dd < d; ++dd) c->p_supp[(size_t)l * sd + (size_t)k * d + dd] /= z;
        }
      or_calculate_distmat(ph, label, c, C); /* :225-227 */
      for (size_t e = 0; e < sc; ++e) C[e] /= rho;
    }

    /* step 6, :300-316 */
    if ((iter % 100 == 99) || (iter < 100 && iter % 20 == 19)) {
      double obj = 0.0, pr = 0.0, dr = 0.0;
      for (size_t e = 0; e < sc; ++e) {
        obj += C[e] * X[e];
        X[e] -= Z[e];
        Z0[e] -= Z[e];
        pr += fabs(X[e]);
        dr += fabs(Z0[e]);
      }
      if (trace && ntr < max_trace) {
        trace[4 * ntr + 0] = iter + 1;
        trace[4 * ntr + 1] = obj * rho / (double)N;
        trace[4 * ntr + 2] = pr / (double)N;
        trace[4 * ntr + 3] = dr / (double)N;
      }
      ++ntr;
    }
  }
  if (rho_out) *rho_out = rho;
done:
  if (n_trace) *n_trace = ntr;
  free(Z0);
  free(Zr);
  free(count);
  return rc;
}

/* ------------------------------------------------------ exact transport LP */

/*
 * Transportation simplex (MODI) on the LP of solver_mosek.cc:61-143.
 * Dropping the last column constraint makes column m-1 absorb
 * sum(wX) - sum_{j<m-1} wY[j]; the problem is then a balanced
 * transportation problem.  Start: north-west corner (a spanning tree of
 * n+m-1 basic cells).  Entering cell: most negative reduced cost (Dantzig),
 * switching to Bland's smallest-index rule after many pivots to exclude
 * cycling.  Leaving cell: smallest flow on the cycle's minus cells.
 */
double or_transport_ex(int n, int m, const double *C, const double *wX, const double *wY,
                       int *pivots_out) {
  const int nb = n + m - 1, nn = n + m;
  double *ra = (double *)malloc(sizeof(double) * n);
  double *rb = (double *)malloc(sizeof(double) * m);
  int *bi = (int *)malloc(sizeof(int) * nb), *bj = (int *)malloc(sizeof(int) * nb);
  double *bx = (double *)malloc(sizeof(double) * nb);
  double *pot = (double *)malloc(sizeof(double) * nn);
  int *deg = (int *)malloc(sizeof(int) * nn), *adj = (int *)malloc(sizeof(int) * 2 * nb);
  int *start = (int *)malloc(sizeof(int) * (nn + 1));
  int *par = (int *)malloc(sizeof(int) * nn), *parcell = (int *)malloc(sizeof(int) * nn);
  int *queue = (int *)malloc(sizeof(int) * nn);
  char *seen = (char *)malloc(nn);
  int *cyc = (int *)malloc(sizeof(int) * (nb + 1));
  int *cyc_a = (int *)malloc(sizeof(int) * (nn + 1)), *cyc_b = (int *)malloc(sizeof(int) * (nn + 1));
  double cmax = 0.0;
  for (int e = 0; e < n * m; ++e) cmax = fmax(cmax, fabs(C[e]));
  const double eps = 1e-12 * (1.0 + cmax);

  double sa = 0.0, sbm = 0.0;
  for (int i = 0; i < n; ++i) { ra[i] = wX[i]; sa += wX[i]; }
  for (int j = 0; j < m - 1; ++j) { rb[j] = wY[j]; sbm += wY[j]; }
  rb[m - 1] = sa - sbm > 0 ? sa - sbm : 0.0;

  /* north-west corner */
  {
    int i = 0, j = 0;
    for (int t = 0; t < nb; ++t) {
      double q = ra[i] < rb[j] ? ra[i] : rb[j];
      if (q < 0) q = 0;
      bi[t] = i; bj[t] = j; bx[t] = q;
      ra[i] -= q; rb[j] -= q;
      if (i == n - 1) ++j;
      else if (j == m - 1) ++i;
      else if (ra[i] <= rb[j]) ++i;
      else ++j;
    }
  }

  int pivots = 0;
  const int dantzig_limit = 50 * (n + m) + 1000;
  for (;;) {
    /* adjacency of the basis tree; node r < n is row r, node n + j column j */
    for (int v = 0; v < nn; ++v) deg[v] = 0;
    for (int t = 0; t < nb; ++t) { deg[bi[t]]++; deg[n + bj[t]]++; }
    start[0] = 0;
    for (int v = 0; v < nn; ++v) start[v + 1] = start[v] + deg[v];
    for (int v = 0; v < nn; ++v) deg[v] = start[v];
    for (int t = 0; t < nb; ++t) { adj[deg[bi[t]]++] = t; adj[deg[n + bj[t]]++] = t; }

    /* potentials u_i + v_j = c_ij on basic cells, BFS from row 0 */
    for (int v = 0; v < nn; ++v) seen[v] = 0;
    int qh = 0, qt = 0;
    queue[qt++] = 0; seen[0] = 1; pot[0] = 0.0; par[0] = -1; parcell[0] = -1;
    while (qh < qt) {
      int v = queue[qh++];
      for (int a = start[v]; a < start[v + 1]; ++a) {
        int t = adj[a];
        int w = v < n ? n + bj[t] : bi[t];
        if (seen[w]) continue;
        seen[w] = 1; par[w] = v; parcell[w] = t;
        pot[w] = C[bi[t] + (size_t)bj[t] * n] - pot[v];
        queue[qt++] = w;
      }
    }

    /* pricing */
    int ei = -1, ej = -1;
    double best = -eps;
    const int bland = pivots >= dantzig_limit;
    for (int j = 0; j < m && !(bland && ei >= 0); ++j)
      for (int i = 0; i < n; ++i) {
        double r = C[i + (size_t)j * n] - pot[i] - pot[n + j];
        if (r < best) {
          best = r; ei = i; ej = j;
          if (bland) break;
        }
      }
    if (ei < 0) break;

    /* cycle: tree path from column node n+ej up/down to row node ei */
    int a = n + ej, b = ei, la = 0, lb = 0;
    /* depth via walking to root */
    int da = 0, db = 0;
    for (int v = a; par[v] >= 0; v = par[v]) ++da;
    for (int v = b; par[v] >= 0; v = par[v]) ++db;
    int *pa = cyc_a, *pb = cyc_b; /* cells on each side of the meeting node */
    while (da > db) { pa[la++] = parcell[a]; a = par[a]; --da; }
    while (db > da) { pb[lb++] = parcell[b]; b = par[b]; --db; }
    while (a != b) {
      pa[la++] = parcell[a]; a = par[a];
      pb[lb++] = parcell[b]; b = par[b];
    }
    /* cycle order: entering (+), then from column ej towards the meeting
     * node (pa), then back down to row ei (pb reversed); signs alternate */
    int nc = 0;
    for (int t = 0; t < la; ++t) cyc[nc++] = pa[t];
    for (int t = lb - 1; t >= 0; --t) cyc[nc++] = pb[t];
    double theta = DBL_MAX;
    int leave = -1;
    for (int t = 0; t < nc; t += 2) /* minus cells: positions 0, 2, ... */
      if (bx[cyc[t]] < theta || (bland && bx[cyc[t]] == theta && cyc[t] < leave)) {
        theta = bx[cyc[t]]; leave = cyc[t];
      }
    for (int t = 0; t < nc; ++t) bx[cyc[t]] += (t % 2 == 0) ? -theta : theta;
    bi[leave] = ei; bj[leave] = ej; bx[leave] = theta;
    ++pivots;
    if (pivots > 100000 + 100 * n * m) break; /* safety net, never expected */
  }

  double val = 0.0;
  for (int t = 0; t < nb; ++t) val += C[bi[t] + (size_t)bj[t] * n] * bx[t];
  if (pivots_out) *pivots_out = pivots;
  free(ra); free(rb); free(bi); free(bj); free(bx); free(pot); free(deg); free(adj);
  free(start); free(par); free(parcell); free(queue); free(seen); free(cyc);
  free(cyc_a); free(cyc_b);
  return val;
}

double or_transport(int n, int m, const double *C, const double *wX, const double *wY) {
  return or_transport_ex(n, m, C, wX, wY, NULL);
}

/* ----------------------------------------------------------- distances */

/* one phase's LP value between a centroid (supports A, weights wA, s) and
 * the object/centroid described by (B or ids, wB, m): clustering.c:46-101 */
static double phase_lp(const or_phase *ph, int s, const double *A, const double *wA, int m,
                       const double *B, const int *Bsym, const double *wB) {
  double *Cb = (double *)malloc(sizeof(double) * (size_t)s * m);
  double v;
  switch (ph->metric) {
    case OR_EUCLIDEAN:
      or_pdist2(ph->dim, s, m, A, B, Cb);
      v = or_transport(s, m, Cb, wA, wB);
      break;
    case OR_WORD_EMBED:
      if (Bsym) or_pdist2_sym(ph->dim, s, m, A, Bsym, Cb, ph->vocab_vec);
      else or_pdist2(ph->dim, s, m, A, B, Cb);
      v = or_transport(s, m, Cb, wA, wB);
      break;
    case OR_HISTOGRAM:
      v = or_transport(s, m, ph->dist_mat, wA, wB);
      break;
    case OR_SPARSE_HISTOGRAM:
      if (Bsym) or_pdist2_submat(m, Bsym, Cb, ph->vocab_size, ph->dist_mat);
      else memcpy(Cb, ph->dist_mat, sizeof(double) * (size_t)s * m);
      v = or_transport(s, m, Cb, wA, wB);
      break;
    default:
      v = 0.0;
  }
  free(Cb);
  return v;
}

double or_distance_obj_cent(int s_ph, const or_phase *ph, int selected_phase, size_t i,
                            const or_centroids *c, int j) {
  double d = 0.0;
  for (int n = 0; n < s_ph; ++n) {
    if (!(selected_phase < 0 || n == selected_phase)) continue;
    const or_phase *p = ph + n;
    const or_centroids *cn = c + n;
    const size_t off = p->p_str_cum[i];
    const int s = cn->s, m = p->p_str[i];
    const double *A = cn->p_supp ? cn->p_supp + (size_t)j * s * p->dim : NULL;
    d += phase_lp(p, s, A, cn->p_w + (size_t)j * s, m,
                  p->metric == OR_EUCLIDEAN ? p->p_supp + off * p->dim : NULL,
                  (p->metric == OR_WORD_EMBED || p->metric == OR_SPARSE_HISTOGRAM)
                      ? p->p_supp_sym + off : NULL,
                  p->p_w + off);
  }
  if (d <= 0) return 0.0; /* :126-127 */
  return sqrt(d);
}

/* clustering.c:33-128 called with a = centroids(i), b = centroids(j):
 * the LP's rows are b's supports, its columns a's. */
double or_distance_cent_cent(int s_ph, const or_phase *ph, int selected_phase,
                             const or_centroids *ca, int i, const or_centroids *cb, int j) {
  double d = 0.0;
  for (int n = 0; n < s_ph; ++n) {
    if (!(selected_phase < 0 || n == selected_phase)) continue;
    const or_phase *p = ph + n;
    const int s = cb[n].s, m = ca[n].s;
    const double *A = cb[n].p_supp ? cb[n].p_supp + (size_t)j * s * p->dim : NULL;
    const double *B = ca[n].p_supp ? ca[n].p_supp + (size_t)i * m * p->dim : NULL;
    d += phase_lp(p, s, A, cb[n].p_w + (size_t)j * s, m, B, NULL, ca[n].p_w + (size_t)i * m);
  }
  if (d <= 0) return 0.0;
  return sqrt(d);
}

/* ------------------------------------------------------------ labeling */

size_t or_labeling(int s_ph, const or_phase *ph, int selected_phase, const or_centroids *c,
                   int K, int *label, char *label_switch, double *cost) {
  size_t count = 0;
  double acc = 0.0;
  const size_t N = ph[selected_phase < 0 ? 0 : selected_phase].size;
  for (size_t i = 0; i < N; ++i) {
    double mind = -1.0;
    int jj = label[i] >= 0 ? label[i] : 0;
    for (int j = 0; j < K; ++j) {
      double d = or_distance_obj_cent(s_ph, ph, selected_phase, i, c, j);
      if (mind < 0 || d < mind) { mind = d; jj = j; } /* :344 */
    }
    acc += mind * mind;
    if (label[i] == jj) {
      if (label_switch) label_switch[i] = 0;
    } else {
      label[i] = jj;
      if (label_switch) label_switch[i] = 1;
      ++count;
    }
  }
  if (cost) *cost = acc;
  return count;
}

size_t or_labeling_prep(int s_ph, const or_phase *ph, int selected_phase,
                        const or_centroids *c, int K, int *label, char *label_switch,
                        or_elkan *tr, size_t *dist_count) {
  const size_t N = ph[selected_phase < 0 ? 0 : selected_phase].size;
  size_t count = 0, dc = 0;
  for (int k = 0; k < K; ++k) tr->s[k] = DBL_MAX;
  for (int k = 0; k < K * K; ++k) tr->c[k] = 0.0;
  for (int i = 0; i < K; ++i) /* :158-172 */
    for (int j = i + 1; j < K; ++j) {
      double d = or_distance_cent_cent(s_ph, ph, selected_phase, c, i, c, j);
      ++dc;
      tr->c[i * K + j] = d;
      tr->c[i + j * K] = d;
      if (tr->s[i] > d / 2.f) tr->s[i] = d / 2.f;
      if (tr->s[j] > d / 2.f) tr->s[j] = d / 2.f;
    }
  for (size_t i = 0; i < N; ++i) label_switch[i] = 0;
  for (size_t i = 0; i < N; ++i) {
    if (!(label[i] < 0 || tr->u[i] > tr->s[label[i]])) continue; /* :185 */
    const int init_label = label[i];
    int jj = init_label >= 0 ? init_label : 0;
    double mind = 0.0;
    double *U = tr->u + i;
    double *L = tr->l + i * (size_t)K;
    for (int j = 0; j < K; ++j) {
      if (!((j != jj || init_label < 0) && *U > L[j] && *U > tr->c[j * K + jj] / 2.)) continue;
      if (tr->r[i] == 1) { /* 3a */
        double d = or_distance_obj_cent(s_ph, ph, selected_phase, i, c, jj);
        ++dc;
        L[jj] = d; *U = d; mind = d; tr->r[i] = 0;
      } else {
        mind = *U;
      }
      if ((mind > L[j] || mind > tr->c[j * K + jj] / 2.) && j != jj) { /* 3b */
        double d = or_distance_obj_cent(s_ph, ph, selected_phase, i, c, j);
        ++dc;
        L[j] = d;
        if (d < mind) { jj = j; mind = d; *U = d; }
      }
    }
    if (jj != init_label) {
      label[i] = jj;
      label_switch[i] = 1;
      ++count;
    }
  }
  if (dist_count) *dist_count = dc;
  return count;
}

void or_labeling_post(int s_ph, const or_phase *ph, int selected_phase,
                      const or_centroids *c_old, const or_centroids *c_new, int K,
                      const int *label, or_elkan *tr) {
  const size_t N = ph[selected_phase < 0 ? 0 : selected_phase].size;
  double *dch = (double *)malloc(sizeof(double) * K);
  for (int k = 0; k < K; ++k) /* a = c_new(k), b = c_old(k): :303-307 */
    dch[k] = or_distance_cent_cent(s_ph, ph, selected_phase, c_new, k, c_old, k);
  for (size_t j = 0; j < N; ++j) {
    double *L = tr->l + j * (size_t)K;
    for (int k = 0; k < K; ++k) L[k] = (L[k] - dch[k]) > 0 ? (L[k] - dch[k]) : 0;
    tr->u[j] += dch[label[j]];
    tr->r[j] = 1;
  }
  free(dch);
}

/* -------------------------------------------------------- initialisation */

/* math.c:39-56 */
void or_shuffle(size_t *array, size_t n, unsigned seed) {
  srand(seed);
  if (n > 1)
    for (int k = 0; k < 10; ++k)
      for (size_t i = 0; i < n - 1; i++) {
        size_t j = i + rand() / (RAND_MAX / (n - i) + 1);
        size_t t = array[j];
        array[j] = array[i];
        array[i] = t;
      }
}

/* centroid_rand.c:104-158: greedy Ward-cost merging down to n supports */
void or_merge(int dim, const double *m_supp, const double *m_w, int m, double *c_supp,
              double *c_w, int n) {
  double *D = (double *)malloc(sizeof(double) * (size_t)m * m);
  double *supp = (double *)malloc(sizeof(double) * (size_t)dim * m);
  double *w = (double *)malloc(sizeof(double) * m);
  int mm = m;
  memcpy(supp, m_supp, sizeof(double) * (size_t)dim * m);
  memcpy(w, m_w, sizeof(double) * m);
  or_pdist2(dim, m, m, supp, supp, D);
  while (mm > n) {
    double best = DBL_MAX;
    int bidx = -1;
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j) {
        double cost = w[i] * w[j] * D[i * m + j] / (w[i] + w[j]);
        if (w[i] > 0 && w[j] > 0 && best > cost) { best = cost; bidx = i * m + j; }
      }
    int i = bidx / m, j = bidx % m;
    for (int k = 0; k < dim; ++k)
      supp[i * dim + k] = (supp[i * dim + k] * w[i] + supp[j * dim + k] * w[j]) / (w[i] + w[j]);
    or_pdist2(dim, 1, m, supp + i * dim, supp, D + (size_t)i * m);
    for (int k = 0; k < m; ++k) D[(size_t)k * m + i] = D[(size_t)i * m + k];
    w[i] = w[i] + w[j];
    w[j] = 0;
    --mm;
  }
  for (int i = 0, j = 0; i < m; ++i)
    if (w[i] > 0) {
      memcpy(c_supp + (size_t)j * dim, supp + (size_t)i * dim, sizeof(double) * dim);
      c_w[j] = w[i];
      ++j;
    }
  free(D); free(supp); free(w);
}

/* centroid_rand.c:161-296 for the serial case (rank 0, one process) */
int or_centroid_rands(const or_phase *ph, int K, int s, or_centroids *c, unsigned seed) {
  const size_t N = ph->size;
  const int d = ph->dim;
  size_t *arr = (size_t *)malloc(sizeof(size_t) * N);
  for (size_t i = 0; i < N; ++i) arr[i] = i;
  or_shuffle(arr, N, seed);
  for (size_t e = 0; e < (size_t)K * s; ++e) c->p_w[e] = 0;
  if (c->p_supp)
    for (size_t e = 0; e < (size_t)K * s * d; ++e) c->p_supp[e] = 0;
  size_t i = 0;
  int j = 0;
  while (i < N && j < K) {
    while (i < N && ph->p_str[arr[i]] < s && ph->metric != OR_SPARSE_HISTOGRAM) ++i;
    if (i == N) break;
    const int ms = ph->p_str[arr[i]];
    const size_t off = ph->p_str_cum[arr[i]];
    const double *mw = ph->p_w + off;
    switch (ph->metric) {
      case OR_HISTOGRAM:
        memcpy(c->p_w + (size_t)j * s, mw, sizeof(double) * s);
        break;
      case OR_EUCLIDEAN:
        if (ms == s) {
          memcpy(c->p_supp + (size_t)j * s * d, ph->p_supp + off * d, sizeof(double) * s * d);
          memcpy(c->p_w + (size_t)j * s, mw, sizeof(double) * s);
        } else {
          or_merge(d, ph->p_supp + off * d, mw, ms, c->p_supp + (size_t)j * s * d,
                   c->p_w + (size_t)j * s, s);
        }
        break;
      case OR_WORD_EMBED: {
        double *tmp = (double *)malloc(sizeof(double) * (size_t)ms * d);
        for (int k = 0; k < ms; ++k)
          memcpy(tmp + (size_t)k * d, ph->vocab_vec + (size_t)ph->p_supp_sym[off + k] * d,
                 sizeof(double) * d);
        if (ms == s) {
          memcpy(c->p_supp + (size_t)j * s * d, tmp, sizeof(double) * s * d);
          memcpy(c->p_w + (size_t)j * s, mw, sizeof(double) * s);
        } else {
          or_merge(d, tmp, mw, ms, c->p_supp + (size_t)j * s * d, c->p_w + (size_t)j * s, s);
        }
        free(tmp);
        break;
      }
      case OR_SPARSE_HISTOGRAM:
        /* the reference writes centroid 0 only (centroid_rand.c:265-266);
         * here each centroid j gets its own sample's bins */
        for (int k = 0; k < s; ++k) c->p_w[(size_t)j * s + k] = 0;
        for (int k = 0; k < ms; ++k) c->p_w[(size_t)j * s + ph->p_supp_sym[off + k]] = mw[k];
        break;
    }
    ++i;
    ++j;
  }
  free(arr);
  return j < K ? 1 : 0;
}

/* ------------------------------------------------------------ outer loop */

static void copy_cents(int s_ph, const or_phase *ph, int selected_phase,
                       const or_centroids *src, or_centroids *dst) {
  for (int n = 0; n < s_ph; ++n) {
    if (!(selected_phase < 0 || n == selected_phase)) continue;
    const size_t ks = (size_t)src[n].K * src[n].s;
    memcpy(dst[n].p_w, src[n].p_w, sizeof(double) * ks);
    if (src[n].p_supp && dst[n].p_supp)
      memcpy(dst[n].p_supp, src[n].p_supp, sizeof(double) * ks * ph[n].dim);
  }
}

int or_clustering(int s_ph, const or_phase *ph, int selected_phase, int K, int max_iter,
                  int use_triangle, const or_badmm_opts *opt, or_centroids *cents,
                  int *label, size_t *trace_changes, int max_trace, double *final_cost) {
  const size_t N = ph[0].size;
  double **C = (double **)calloc(s_ph, sizeof(double *));
  double **X = (double **)calloc(s_ph, sizeof(double *));
  double **Z = (double **)calloc(s_ph, sizeof(double *));
  double **Y = (double **)calloc(s_ph, sizeof(double *));
  or_centroids *old = (or_centroids *)calloc(s_ph, sizeof(or_centroids));
  char *lsw = (char *)calloc(N, 1);
  or_elkan tr = {0};
  for (int n = 0; n < s_ph; ++n) {
    if (!(selected_phase < 0 || n == selected_phase)) continue;
    const size_t sc = (size_t)cents[n].s * ph[n].col;
    C[n] = (double *)malloc(sizeof(double) * sc);
    X[n] = (double *)malloc(sizeof(double) * sc);
    Z[n] = (double *)calloc(sc, sizeof(double));
    Y[n] = (double *)malloc(sizeof(double) * sc);
    old[n] = cents[n];
    old[n].p_w = (double *)malloc(sizeof(double) * (size_t)K * cents[n].s);
    old[n].p_supp = cents[n].p_supp
                        ? (double *)malloc(sizeof(double) * (size_t)K * cents[n].s * ph[n].dim)
                        : NULL;
  }
  for (size_t i = 0; i < N; ++i) label[i] = -1; /* clustering_util.c:117-118 */
  if (use_triangle) { /* clustering_util.c:218-228 */
    tr.l = (double *)calloc(N * (size_t)K, sizeof(double));
    tr.u = (double *)malloc(sizeof(double) * N);
    tr.s = (double *)malloc(sizeof(double) * K);
    tr.c = (double *)malloc(sizeof(double) * K * K);
    tr.r = (char *)malloc(N);
    for (size_t i = 0; i < N; ++i) { tr.u[i] = DBL_MAX; tr.r[i] = 1; }
  }
  int iter;
  double cost = 0.0;
  for (iter = 0; iter < max_iter; ++iter) { /* :455 */
    size_t changes, dc = 0;
    if (use_triangle)
      changes = or_labeling_prep(s_ph, ph, selected_phase, cents, K, label, lsw, &tr, &dc);
    else
      changes = or_labeling(s_ph, ph, selected_phase, cents, K, label, lsw, &cost);
    if (trace_changes && iter < max_trace) trace_changes[iter] = changes;
    if ((double)changes < 0.001 * (double)N) { ++iter; break; } /* :476 */
    if (use_triangle) copy_cents(s_ph, ph, selected_phase, cents, old);
    for (int n = 0; n < s_ph; ++n) {
      if (!(selected_phase < 0 || n == selected_phase)) continue;
      or_badmm(ph + n, label, lsw, cents + n, C[n], X[n], Z[n], Y[n], opt, NULL, NULL, 0,
               NULL, NULL);
    }
    if (use_triangle) or_labeling_post(s_ph, ph, selected_phase, old, cents, K, label, &tr);
  }
  if (use_triangle) /* :509 */
    or_labeling(s_ph, ph, selected_phase, cents, K, label, lsw, &cost);
  if (final_cost) *final_cost = cost / (double)N;
  for (int n = 0; n < s_ph; ++n) {
    free(C[n]); free(X[n]); free(Z[n]); free(Y[n]);
    if (old[n].p_w) free(old[n].p_w);
    if (old[n].p_supp) free(old[n].p_supp);
  }
  free(C); free(X); free(Z); free(Y); free(old); free(lsw);
  free(tr.l); free(tr.u); free(tr.s); free(tr.c); free(tr.r);
  return iter;
}
