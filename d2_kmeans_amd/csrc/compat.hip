// The reference-compatible C ABI (include/d2mi_clustering.h) on top of the
// device engine (include/d2mi.h): libad2c's data model, allocation, text I/O,
// initialisation, B-ADMM plug point, labeling, clustering driver, solver and
// pdist2 entry points, with the compute on the MI355X.
//
// Host code here only moves data and runs the reference's control flow; the
// arithmetic of the hot path (costs, transport plans, LP values, labels) runs
// in the kernels.  One engine context per process, created on first use.
#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "d2mi.h"
#include "d2mi_clustering.h"

// ------------------------------------------------------------------ globals

extern "C" {
int d2_alg_type = D2_CENTROID_BADMM;  // clustering.c:24
int world_rank = 0;                   // clustering.c:25
int nprocs = 1;                       // clustering.c:26
BADMM_options badmm_clu_options = {100, 2.0, 10};   // centroid_Bregman.c:17
BADMM_options badmm_cen_options = {2000, 1.0, 10};  // centroid_Bregman.c:18
BADMM_options *p_badmm_options = &badmm_clu_options;  // centroid_Bregman.c:22
}

namespace {

d2mi_ctx *g_ctx = nullptr;
bool g_elkan_fresh = false;

struct DataKey {
  const mph *p = nullptr;
  size_t size = 0;
  std::vector<const void *> pw;
  std::vector<size_t> cols;
};
DataKey g_key;

int precision() {
  const char *e = getenv("D2MI_PRECISION");
  return (e && atoi(e) == 64) ? 64 : 32;
}

d2mi_ctx *ctx() {
  if (!g_ctx) {
    const char *e = getenv("D2MI_DEVICE");
    const int dev = e ? atoi(e) : 0;
    if (d2mi_create(&g_ctx, dev) != D2MI_OK) {
      fprintf(stderr, "d2mi: cannot create a device context: %s\n", d2mi_last_error());
      exit(-1);
    }
  }
  return g_ctx;
}

bool ok(int rc, const char *what) {
  if (rc == D2MI_OK) return true;
  fprintf(stderr, "d2mi: %s failed (%d): %s\n", what, rc, d2mi_last_error());
  return false;
}

// uploads every phase of p_data unless the device already holds it
bool sync_data(mph *p) {
  DataKey k;
  k.p = p;
  k.size = p->size;
  for (int n = 0; n < p->s_ph; ++n) {
    k.pw.push_back(p->ph[n].p_w);
    k.cols.push_back(p->ph[n].col);
  }
  if (g_key.p == k.p && g_key.size == k.size && g_key.pw == k.pw && g_key.cols == k.cols)
    return true;
  for (int n = 0; n < p->s_ph; ++n) {
    sph &s = p->ph[n];
    if (s.col == 0) continue;
    const bool sym = s.metric_type == D2_WORD_EMBED || s.metric_type == D2_SPARSE_HISTOGRAM;
    if (!ok(d2mi_set_phase(ctx(), n, s.metric_type, s.dim, p->size, s.col, s.p_str, s.p_str_cum,
                           s.p_w, s.metric_type == D2_EUCLIDEAN_L2 ? s.p_supp : nullptr,
                           sym ? s.p_supp_sym : nullptr, s.vocab_size,
                           s.metric_type == D2_WORD_EMBED ? s.vocab_vec : nullptr,
                           (s.metric_type == D2_HISTOGRAM || s.metric_type == D2_SPARSE_HISTOGRAM)
                               ? s.dist_mat
                               : nullptr),
            "uploading data"))
      return false;
  }
  g_key = k;
  g_elkan_fresh = true;
  return true;
}

bool put_centroid_phase(int n, int K, sph *c) {
  const bool supp = c->metric_type == D2_EUCLIDEAN_L2 || c->metric_type == D2_WORD_EMBED;
  return ok(d2mi_set_centroids(ctx(), n, K, c->str, c->p_w, supp ? c->p_supp : nullptr),
            "uploading centroids");
}

bool put_centroids(mph *c, int selected_phase) {
  for (int n = 0; n < c->s_ph; ++n)
    if ((selected_phase < 0 || n == selected_phase) && c->ph[n].col > 0)
      if (!put_centroid_phase(n, (int)c->size, c->ph + n)) return false;
  return true;
}

bool get_centroids(mph *c, int selected_phase) {
  for (int n = 0; n < c->s_ph; ++n)
    if ((selected_phase < 0 || n == selected_phase) && c->ph[n].col > 0) {
      const bool supp = c->ph[n].metric_type == D2_EUCLIDEAN_L2;
      if (!ok(d2mi_get_centroids(ctx(), n, c->ph[n].p_w, supp ? c->ph[n].p_supp : nullptr),
              "reading centroids"))
        return false;
    }
  return true;
}

double now() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

void shuffle(size_t *array, size_t n) {
  // math.c:39-56: glibc rand, seeded with time * rank (0 for one process)
  srand((unsigned)(time(nullptr) * world_rank));
  if (n > 1)
    for (int k = 0; k < 10; ++k)
      for (size_t i = 0; i < n - 1; i++) {
        const size_t j = i + rand() / (RAND_MAX / (n - i) + 1);
        const size_t t = array[j];
        array[j] = array[i];
        array[i] = t;
      }
}

// merge of centroid_rand.c:104-158: greedily fuse the pair of supports with
// the least Ward cost w_i w_j |x_i - x_j|^2 / (w_i + w_j) until n remain
void merge(int dim, const double *m_supp, const double *m_w, int m, double *c_supp,
           double *c_w, int n) {
  std::vector<double> D((size_t)m * m), supp(m_supp, m_supp + (size_t)dim * m), w(m_w, m_w + m);
  auto d2 = [&](int i, int j) {
    double a = 0.0;
    for (int k = 0; k < dim; ++k) {
      const double u = supp[(size_t)i * dim + k] - supp[(size_t)j * dim + k];
      a += u * u;
    }
    return a;
  };
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) D[(size_t)i * m + j] = d2(i, j);
  for (int mm = m; mm > n; --mm) {
    double best = DBL_MAX;
    int bi = -1, bj = -1;
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j) {
        const double cost = w[i] * w[j] * D[(size_t)i * m + j] / (w[i] + w[j]);
        if (w[i] > 0 && w[j] > 0 && best > cost) {
          best = cost;
          bi = i;
          bj = j;
        }
      }
    for (int k = 0; k < dim; ++k)
      supp[(size_t)bi * dim + k] =
          (supp[(size_t)bi * dim + k] * w[bi] + supp[(size_t)bj * dim + k] * w[bj]) / (w[bi] + w[bj]);
    for (int k = 0; k < m; ++k) D[(size_t)bi * m + k] = D[(size_t)k * m + bi] = d2(bi, k);
    w[bi] += w[bj];
    w[bj] = 0;
  }
  for (int i = 0, j = 0; i < m; ++i)
    if (w[i] > 0) {
      memcpy(c_supp + (size_t)j * dim, supp.data() + (size_t)i * dim, sizeof(double) * dim);
      c_w[j] = w[i];
      ++j;
    }
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ allocation

int d2_allocate_sph(sph *ph, const int d, const int stride, const size_t num,
                    const double semicol, const int type) {
  assert(stride > 0 && num > 0 && semicol >= 0);
  const size_t m = (size_t)(num * (stride + semicol));
  ph->max_col = m;
  ph->dim = d;
  ph->str = stride;
  ph->max_str = stride;
  ph->col = 0;
  ph->p_str = (int *)calloc(num, sizeof(int));
  ph->p_str_cum = (size_t *)calloc(num, sizeof(size_t));
  ph->p_w = (double *)malloc(sizeof(double) * m);
  ph->p_supp = nullptr;
  ph->p_supp_sym = nullptr;
  ph->dist_mat = nullptr;
  ph->vocab_vec = nullptr;
  ph->vocab_size = 0;
  ph->metric_type = type;
  if (type == D2_EUCLIDEAN_L2) ph->p_supp = (double *)malloc(sizeof(double) * m * d);
  else if (type == D2_WORD_EMBED || type == D2_SPARSE_HISTOGRAM)
    ph->p_supp_sym = (int *)malloc(sizeof(int) * m);
  ph->is_meta_allocated = 0;
  return 0;
}

int d2_allocate(mph *p, const int s_ph, const size_t n, const int *avg_strides,
                const int *dims, const int *types) {
  p->s_ph = s_ph;
  p->size = n;
  p->global_size = n;
  p->ph = (sph *)malloc(s_ph * sizeof(sph));
  p->num_of_labels = 0;
  p->label = (int *)malloc(sizeof(int) * n);
  for (size_t i = 0; i < n; ++i) p->label[i] = -1;  // clustering_util.c:117-118
  for (int k = 0; k < s_ph; ++k)
    if (d2_allocate_sph(p->ph + k, dims[k], avg_strides[k], n, 0.6, types[k])) return -1;
  return 0;
}

static void free_sph(sph *s) {
  free(s->p_w);
  free(s->p_str);
  free(s->p_str_cum);
  free(s->p_supp);
  free(s->p_supp_sym);
  if (s->is_meta_allocated) {
    free(s->dist_mat);
    free(s->vocab_vec);
  }
  memset(s, 0, sizeof(*s));
}

int d2_free(mph *p) {
  if (g_key.p == p) g_key = DataKey();
  if (p->ph) {
    for (int i = 0; i < p->s_ph; ++i)
      if (p->ph[i].p_str) free_sph(p->ph + i);
    free(p->ph);
    p->ph = nullptr;
  }
  free(p->label);
  p->label = nullptr;
  return 0;
}

// ------------------------------------------------------------------ text I/O

int d2_read(const char *filename, const char *meta_filename, mph *p) {
  const double t0 = now();
  std::string fn = filename;
  FILE *fp = nullptr;
  if (nprocs > 1) fp = fopen((fn + "." + std::to_string(world_rank)).c_str(), "r");
  if (!fp) fp = fopen(filename, "r");
  if (!fp) {
    fprintf(stderr, "d2_read: cannot open %s\n", filename);
    return -1;
  }
  const int s_ph = p->s_ph;
  for (int n = 0; n < s_ph; ++n) {  // side files (clustering_io.c:54-90)
    sph &s = p->ph[n];
    s.col = 0;
    if (s.metric_type == D2_HISTOGRAM || s.metric_type == D2_SPARSE_HISTOGRAM ||
        s.metric_type == D2_WORD_EMBED) {
      const bool hist = s.metric_type != D2_WORD_EMBED;
      std::string side = (meta_filename && s_ph == 1)
                             ? std::string(meta_filename)
                             : fn + (hist ? ".hist" : ".vocab") + std::to_string(n);
      FILE *f2 = fopen(side.c_str(), "r");
      if (!f2) {
        fprintf(stderr, "d2_read: cannot open %s\n", side.c_str());
        fclose(fp);
        return -1;
      }
      if (hist) {
        int V = 0;
        if (fscanf(f2, "%d", &V) != 1) V = 0;
        s.dist_mat = (double *)malloc(sizeof(double) * V * V);
        for (int i = 0; i < V * V; ++i)
          if (fscanf(f2, "%lf", &s.dist_mat[i]) != 1) s.dist_mat[i] = 0;
        s.vocab_size = V;
      } else {
        int d = 0;
        if (fscanf(f2, "%d", &d) != 1 || d != s.dim) {
          fprintf(stderr, "d2_read: %s has dimension %d, expected %d\n", side.c_str(), d, s.dim);
          fclose(f2);
          fclose(fp);
          return -1;
        }
        if (fscanf(f2, "%d", &s.vocab_size) != 1) s.vocab_size = 0;
        s.vocab_vec = (double *)malloc(sizeof(double) * (size_t)d * s.vocab_size);
        for (size_t i = 0; i < (size_t)d * s.vocab_size; ++i)
          if (fscanf(f2, "%lf", &s.vocab_vec[i]) != 1) s.vocab_vec[i] = 0;
      }
      s.is_meta_allocated = 1;
      fclose(f2);
    }
  }
  size_t size = p->size;
  std::vector<int> cursor(s_ph, 0);
  for (size_t i = 0; i < size; ++i) {  // objects (clustering_io.c:93-172)
    for (int n = 0; n < s_ph; ++n) {
      sph &s = p->ph[n];
      int dim, str;
      if (fscanf(fp, "%d", &dim) != 1) {
        fprintf(stderr, "rank %d warning: only read %zu d2!\n", world_rank, i);
        p->size = size = i;
        goto done;
      }
      assert(dim == s.dim);
      if (fscanf(fp, "%d", &str) != 1 || str < 0) str = 0;
      if (str == 0) continue;
      s.p_str[i] = str;
      if (str > s.max_str) s.max_str = str;
      if (s.col + str > s.max_col) {  // grow (:113-135)
        s.max_col = 2 * (s.col + str);
        s.p_w = (double *)realloc(s.p_w, sizeof(double) * s.max_col);
        if (s.p_supp) s.p_supp = (double *)realloc(s.p_supp, sizeof(double) * s.max_col * dim);
        if (s.p_supp_sym) s.p_supp_sym = (int *)realloc(s.p_supp_sym, sizeof(int) * s.max_col);
      }
      double *w = s.p_w + s.col, wsum = 0.0;
      for (int j = 0; j < str; ++j) {
        if (fscanf(fp, "%lf", &w[j]) != 1) w[j] = 0;
        if (s.metric_type != D2_HISTOGRAM && !(w[j] > 1e-9)) {  // :145
          fprintf(stderr, "d2_read: object %zu phase %d has a weight <= 1e-9\n", i, n);
          fclose(fp);
          return -1;
        }
        wsum += w[j];
      }
      for (int j = 0; j < str; ++j) w[j] /= wsum;  // :149-151
      if (s.metric_type == D2_EUCLIDEAN_L2) {
        double *x = s.p_supp + s.col * dim;
        for (int j = 0; j < str * dim; ++j)
          if (fscanf(fp, "%lf", &x[j]) != 1) x[j] = 0;
      } else if (s.metric_type == D2_WORD_EMBED || s.metric_type == D2_SPARSE_HISTOGRAM) {
        int *ids = s.p_supp_sym + s.col;
        for (int j = 0; j < str; ++j) {
          if (fscanf(fp, "%d", &ids[j]) != 1) ids[j] = 0;
          ids[j]--;                                    // 1-based on disk (:164)
          if (ids[j] < 0) ids[j] = s.vocab_size - 1;   // :165-166
        }
      }
      s.col += str;
    }
  }
done:
  for (int n = 0; n < s_ph; ++n) {
    sph &s = p->ph[n];
    if (s.col == 0) continue;
    s.p_str_cum[0] = 0;
    for (size_t i = 1; i < size; ++i) s.p_str_cum[i] = s.p_str_cum[i - 1] + s.p_str[i - 1];
    if (s.metric_type == D2_SPARSE_HISTOGRAM) s.str = s.vocab_size;  // :182-185
  }
  fclose(fp);
  p->global_size = p->size;
  printf("IO time: %lf\n", now() - t0);
  return 0;
}

int d2_write(const char *filename, mph *p) {
  if (world_rank != 0) return 0;
  FILE *fp = filename ? fopen(filename, "w+") : stdout;
  if (!fp) return -1;
  for (size_t i = 0; i < p->size; ++i)  // clustering_io.c:213-244
    for (int n = 0; n < p->s_ph; ++n) {
      sph &s = p->ph[n];
      if (s.col == 0) {
        fprintf(fp, "%d\n0\n\n", s.dim);
        continue;
      }
      const int str = s.p_str[i];
      const size_t pos = s.p_str_cum[i];
      fprintf(fp, "%d\n%d\n", s.dim, str);
      for (int k = 0; k < str; ++k) fprintf(fp, "%lf ", s.p_w[pos + k]);
      fprintf(fp, "\n");
      if (s.metric_type == D2_EUCLIDEAN_L2) {
        for (int k = 0; k < str; ++k) {
          for (int d = 0; d < s.dim; ++d) fprintf(fp, "%lf ", s.p_supp[(pos + k) * s.dim + d]);
          fprintf(fp, "\n");
        }
      } else if (s.metric_type == D2_WORD_EMBED || s.metric_type == D2_SPARSE_HISTOGRAM) {
        for (int k = 0; k < str; ++k) fprintf(fp, "%d ", s.p_supp_sym[pos + k] + 1);
        fprintf(fp, "\n");
      }
    }
  if (filename) fclose(fp);
  printf("Write centroids to %s\n", filename ? filename : "stdout");
  return 0;
}

int d2_write_labels(const char *filename, mph *p) {
  FILE *fp = fopen(filename, "w");
  if (!fp) return -1;
  for (size_t i = 0; i < p->size; ++i) fprintf(fp, "%d\n", p->label[i]);
  fclose(fp);
  printf("Write data partitioned labels to %s\n", filename);
  return 0;
}

int d2_write_labels_serial(const char *filename_ind, const char *filename, mph *p) {
  // clustering_io.c:277-324: the labels of <filename>.label (as written by
  // d2_write_labels, in shard order) are put back in the original object
  // order given by the shard index file and written as <filename>.label_o;
  // without an index file nothing is written
  if (world_rank != 0) return 0;
  const size_t n = p->global_size;
  const std::string lf = std::string(filename) + ".label";
  FILE *fl = fopen(lf.c_str(), "r");
  if (!fl) {
    fprintf(stderr, "d2_write_labels_serial: cannot open %s\n", lf.c_str());
    return -1;
  }
  std::vector<int> lab(n, -1);
  for (size_t i = 0; i < n; ++i)
    if (fscanf(fl, "%d", &lab[i]) != 1) lab[i] = -1;
  fclose(fl);
  FILE *fi = fopen(filename_ind, "r");
  if (!fi) return 0;
  std::vector<size_t> ind(n);
  for (size_t i = 0; i < n; ++i)
    if (fscanf(fi, "%zu", &ind[i]) != 1) ind[i] = i;
  fclose(fi);
  std::vector<int> out_lab(n, -1);
  for (size_t i = 0; i < n; ++i)
    if (ind[i] < n) out_lab[ind[i]] = lab[i];
  std::string out = std::string(filename) + ".label_o";
  FILE *fo = fopen(out.c_str(), "w");
  if (!fo) return -1;
  for (size_t i = 0; i < n; ++i) fprintf(fo, "%d\n", out_lab[i]);
  fclose(fo);
  printf("Write serial data labels to %s\n", out.c_str());
  return 0;
}

int d2_write_split(const char *filename, mph *p, int splits, char is_pre_processed) {
  // clustering_io.c:329-480: shuffle the objects (the serial glibc sequence
  // of math.c:39-56), write the permutation to <filename>.ind and the
  // objects in that order to <filename>.0 ... <filename>.<splits-1>, as
  // they are or pre-processed (word embeddings become Euclidean supports,
  // supports beyond the phase's stride are merged down to it, dense
  // histograms become sparse ones)
  if (!filename || splits <= 0) return -1;
  const size_t size = p->size;
  std::vector<size_t> idx(size);
  for (size_t n = 0; n < size; ++n) idx[n] = n;
  shuffle(idx.data(), size);
  const std::string base(filename);
  FILE *fp = fopen((base + ".ind").c_str(), "w+");
  if (!fp) return -1;
  for (size_t n = 0; n < size; ++n) fprintf(fp, "%zd\n", idx[n]);
  fclose(fp);
  fprintf(stderr, "\twrite %zd indices to %s.ind\n", size, filename);
  const size_t batch = size ? 1 + (size - 1) / splits : 0;
  printf("batch_size: %zd\n", batch);
  std::vector<double> supp, w, csupp, cw;
  std::vector<int> sym;
  for (int k = 0; k < splits; ++k) {
    const std::string fn = base + "." + std::to_string(k);
    FILE *fo = fopen(fn.c_str(), "w+");
    if (!fo) return -1;
    size_t q = batch * k;
    for (; q < size && q < batch * (k + 1); ++q) {
      const size_t i = idx[q];
      for (int n = 0; n < p->s_ph; ++n) {
        const sph &s = p->ph[n];
        if (s.col == 0) continue;
        const int dim = s.dim;
        int str = s.p_str[i];
        const size_t pos = s.p_str_cum[i];
        const double *pw = s.p_w + pos;
        if (!is_pre_processed) {
          if (s.metric_type == D2_HISTOGRAM) fprintf(fo, "%d %d ", dim, str);
          else fprintf(fo, "%d\n%d\n", dim, str);
          for (int j = 0; j < str; ++j) fprintf(fo, "%lf ", pw[j]);
          fprintf(fo, "\n");
          if (s.metric_type == D2_EUCLIDEAN_L2) {
            for (int j = 0; j < str; ++j) {
              for (int d = 0; d < dim; ++d) fprintf(fo, "%lf ", s.p_supp[(pos + j) * dim + d]);
              fprintf(fo, "\n");
            }
          } else if (s.metric_type == D2_WORD_EMBED || s.metric_type == D2_SPARSE_HISTOGRAM) {
            for (int j = 0; j < str; ++j) fprintf(fo, "%d ", s.p_supp_sym[pos + j] + 1);
            fprintf(fo, "\n");
          }
          continue;
        }
        // pre-processed (:386-470)
        const bool points = s.metric_type == D2_EUCLIDEAN_L2 || s.metric_type == D2_WORD_EMBED;
        if (points) {
          supp.resize((size_t)str * dim);
          w.assign(pw, pw + str);
          for (int j = 0; j < str; ++j)
            memcpy(supp.data() + (size_t)j * dim,
                   s.metric_type == D2_EUCLIDEAN_L2
                       ? s.p_supp + (pos + j) * dim
                       : s.vocab_vec + (size_t)s.p_supp_sym[pos + j] * dim,
                   sizeof(double) * dim);
          if (str > s.str) {
            csupp.assign((size_t)s.str * dim, 0.0);
            cw.assign(s.str, 0.0);
            merge(dim, supp.data(), w.data(), str, csupp.data(), cw.data(), s.str);
            supp.swap(csupp);
            w.swap(cw);
            str = s.str;
          }
        } else if (s.metric_type == D2_HISTOGRAM) {
          sym.clear();
          w.clear();
          for (int j = 0; j < str; ++j)
            if (pw[j] > 0) {
              sym.push_back(j);
              w.push_back(pw[j]);
            }
          str = (int)sym.size();
        } else {
          // a sparse histogram is already in the form pre-processing makes
          // of a dense one: its weights and bin ids go out as they are (the
          // reference leaves its weight pointer NULL in this case and
          // crashes writing it)
          w.assign(pw, pw + str);
          sym.assign(s.p_supp_sym + pos, s.p_supp_sym + pos + str);
        }
        fprintf(fo, "%d\n%d\n", dim, str);
        for (int j = 0; j < str; ++j) fprintf(fo, "%lf ", w[j]);
        fprintf(fo, "\n");
        if (points) {
          for (int j = 0; j < str; ++j) {
            for (int d = 0; d < dim; ++d) fprintf(fo, "%lf ", supp[(size_t)j * dim + d]);
            fprintf(fo, "\n");
          }
        } else {
          for (int j = 0; j < str; ++j) fprintf(fo, "%d ", sym[j] + 1);
          fprintf(fo, "\n");
        }
      }
    }
    fclose(fo);
    fprintf(stderr, "\twrite %zd objects to %s\n", q - batch * k, fn.c_str());
  }
  return 0;
}

// ---------------------------------------------------------- initialisation

int d2_centroid_rands(mph *p_data, int idx_ph, sph *c) {
  // centroid_rand.c:161-296, one process
  sph *ph = p_data->ph + idx_ph;
  const size_t K = p_data->num_of_labels, N = p_data->size;
  const int dim = ph->dim, str = ph->str;
  for (size_t i = 0; i < K; ++i) {
    c->p_str[i] = str;
    c->p_str_cum[i] = i * str;
  }
  c->col = (size_t)str * K;
  if (ph->metric_type == D2_HISTOGRAM || ph->metric_type == D2_SPARSE_HISTOGRAM) {
    c->dist_mat = ph->dist_mat;
    c->vocab_size = ph->vocab_size;
  }
  std::vector<size_t> arr(N);
  for (size_t i = 0; i < N; ++i) arr[i] = i;
  shuffle(arr.data(), N);
  for (size_t e = 0; e < c->col; ++e) c->p_w[e] = 0;
  if (c->p_supp)
    for (size_t e = 0; e < c->col * dim; ++e) c->p_supp[e] = 0;
  size_t i = 0, j = world_rank;
  while (i < N && j < K) {
    while (i < N && ph->p_str[arr[i]] < str && ph->metric_type != D2_SPARSE_HISTOGRAM) ++i;
    if (i == N) break;
    const int ms = ph->p_str[arr[i]];
    const size_t off = ph->p_str_cum[arr[i]];
    const double *mw = ph->p_w + off;
    switch (ph->metric_type) {
      case D2_HISTOGRAM:
        memcpy(c->p_w + j * str, mw, sizeof(double) * str);
        break;
      case D2_EUCLIDEAN_L2:
        if (ms == str) {
          memcpy(c->p_supp + j * str * dim, ph->p_supp + off * dim, sizeof(double) * str * dim);
          memcpy(c->p_w + j * str, mw, sizeof(double) * str);
        } else {
          merge(dim, ph->p_supp + off * dim, mw, ms, c->p_supp + j * str * dim, c->p_w + j * str, str);
        }
        break;
      case D2_WORD_EMBED: {
        std::vector<double> tmp((size_t)ms * dim);
        for (int k = 0; k < ms; ++k)
          memcpy(tmp.data() + (size_t)k * dim, ph->vocab_vec + (size_t)ph->p_supp_sym[off + k] * dim,
                 sizeof(double) * dim);
        if (ms == str) {
          memcpy(c->p_supp + j * str * dim, tmp.data(), sizeof(double) * str * dim);
          memcpy(c->p_w + j * str, mw, sizeof(double) * str);
        } else {
          merge(dim, tmp.data(), mw, ms, c->p_supp + j * str * dim, c->p_w + j * str, str);
        }
        break;
      }
      case D2_SPARSE_HISTOGRAM:
        // each centroid gets its own sample's bins (the reference fills only
        // centroid 0, centroid_rand.c:265-266)
        for (int k = 0; k < str; ++k) c->p_w[j * str + k] = 0;
        for (int k = 0; k < ms; ++k) c->p_w[j * str + ph->p_supp_sym[off + k]] = mw[k];
        break;
      default:
        fprintf(stderr, "Unknown type of data!");
        return -1;
    }
    ++i;
    j += nprocs;
  }
  if (j < K) {
    fprintf(stderr,
            "rank %d error: couldn't find enough samples with stride >= %d for initializating %zu "
            "centroids\n",
            world_rank, str, K);
    exit(1);  // centroid_rand.c:290-293
  }
  return 0;
}

int d2_init_centroid(mph *p_data, mph *c, int selected_phase, int allocate_only) {
  if (!allocate_only) {
    printf("Initializing centroids ... ");
    fflush(stdout);
  }
  c->s_ph = p_data->s_ph;
  c->size = p_data->num_of_labels;
  c->global_size = c->size;
  c->label = nullptr;
  c->num_of_labels = 0;
  c->ph = (sph *)calloc(p_data->s_ph, sizeof(sph));
  for (int i = 0; i < p_data->s_ph; ++i) {
    const int type = p_data->ph[i].metric_type % D2_GROUP_SIZE;  // clustering.c:390
    if (selected_phase < 0 || i == selected_phase) {
      d2_allocate_sph(c->ph + i, p_data->ph[i].dim, p_data->ph[i].str, p_data->num_of_labels, 0.,
                      type);
      if (!allocate_only && d2_centroid_rands(p_data, i, c->ph + i)) return -1;
    } else {
      c->ph[i].dim = p_data->ph[i].dim;
      c->ph[i].col = 0;
      c->ph[i].metric_type = type;
    }
  }
  if (!allocate_only) printf("[done]\n");
  return 0;
}

int d2_allocate_work(mph *p_data, var_mph *w, char use_triangle, int selected_phase) {
  (void)selected_phase;
  memset(w, 0, sizeof(*w));
  w->s_ph = p_data->s_ph;
  w->label_switch = (char *)calloc(p_data->size, 1);
  if (use_triangle) g_elkan_fresh = true;  // bounds start at l = 0, u = max, r = 1
  return 0;
}

int d2_free_work(var_mph *w, int selected_phase) {
  (void)selected_phase;
  free(w->label_switch);
  free(w->g_var);
  free(w->l_var_sphBregman);
  memset(w, 0, sizeof(*w));
  return 0;
}

// ------------------------------------------------------------ the hot path

int d2_centroid_sphBregman(mph *p_data, var_mph *var_work, int idx_ph, sph *c0, sph *c) {
  if (!c0) {
    if (d2_centroid_rands(p_data, idx_ph, c)) return -1;
  } else if (c != c0) {
    *c = *c0;  // warm start (centroid_Bregman.c:100)
  }
  if (!sync_data(p_data)) return -1;
  if (!ok(d2mi_set_labels(ctx(), p_data->label, var_work ? var_work->label_switch : nullptr),
          "uploading labels"))
    return -1;
  if (!put_centroid_phase(idx_ph, p_data->num_of_labels, c)) return -1;
  const BADMM_options *o = p_badmm_options;
  d2mi_badmm_stats st;
  if (!ok(d2mi_badmm(ctx(), idx_ph, o->maxIters, o->rhoCoeff, o->updatePerLoops, precision(),
                     D2MI_BADMM_ENGINE, &st),
          "B-ADMM update"))
    return -1;
  const bool supp = c->metric_type == D2_EUCLIDEAN_L2 || c->metric_type == D2_WORD_EMBED;
  if (!ok(d2mi_get_centroids(ctx(), idx_ph, c->p_w, supp ? c->p_supp : nullptr),
          "reading centroids"))
    return -1;
  for (int r = 0; r < st.n_trace; ++r)
    printf("\t%d\t%f\t%f\t%f\n", (int)st.trace[r][0], st.trace[r][1], st.trace[r][2],
           st.trace[r][3]);
  return 0;
}

static size_t labeling_impl(mph *p_data, mph *centroids, var_mph *var_work, int sel,
                            int use_elkan, double *cost) {
  if (!sync_data(p_data)) return 0;
  if (!put_centroids(centroids, sel)) return 0;
  if (!ok(d2mi_set_labels(ctx(), p_data->label, nullptr), "uploading labels")) return 0;
  if (use_elkan && g_elkan_fresh) {
    if (!ok(d2mi_elkan_reset(ctx(), sel), "resetting bounds")) return 0;
    g_elkan_fresh = false;
  }
  d2mi_label_stats st;
  if (!ok(d2mi_labeling(ctx(), sel, use_elkan, &st), "labeling")) return 0;
  if (!ok(d2mi_get_labels(ctx(), p_data->label, var_work ? var_work->label_switch : nullptr),
          "reading labels"))
    return 0;
  if (cost) *cost = st.cost;
  if (use_elkan)
    printf("\n\t\t\t\t %llu objects change their labels\n\t\t\t\t %llu distance pairs computed\n",
           st.changes, st.dist_count);
  return (size_t)st.changes;
}

size_t d2_labeling(mph *p_data, mph *centroids, var_mph *var_work, int selected_phase) {
  double cost = 0.0;
  const size_t ch = labeling_impl(p_data, centroids, var_work, selected_phase, 0, &cost);
  printf("\t %zu labels change.\tmean cost %lf [done]\n", ch, cost / p_data->global_size);
  return ch;
}

size_t d2_labeling_prep(mph *p_data, mph *centroids, var_mph *var_work, int selected_phase) {
  return labeling_impl(p_data, centroids, var_work, selected_phase, 1, nullptr);
}

size_t d2_labeling_post(mph *p_data, mph *c_old, mph *c_new, var_mph *var_work,
                        int selected_phase) {
  (void)var_work;
  if (!sync_data(p_data)) return 0;
  if (!put_centroids(c_old, selected_phase)) return 0;
  if (!ok(d2mi_save_centroids(ctx(), selected_phase), "saving centroids")) return 0;
  if (!put_centroids(c_new, selected_phase)) return 0;
  ok(d2mi_elkan_post(ctx(), selected_phase), "relaxing bounds");
  return 0;
}

void d2_solver_setup(void) { ctx(); }
void d2_solver_release(void) {}

double d2_match_by_distmat(int n, int m, double *C, double *wX, double *wY, double *x,
                           double *lambda, size_t index) {
  (void)x;
  (void)lambda;
  (void)index;
  double v = 0.0;
  if (!ok(d2mi_transport_batch(ctx(), n, m, 1, C, wX, wY, &v, nullptr), "transport LP"))
    return 0.0;  // the reference returns 0 when the LP fails (solver_mosek.cc:194-219)
  return v;
}

void _dpdist2(int d, size_t n, size_t m, double *A, double *B, double *C) {
  ok(d2mi_pdist2(ctx(), d, n, m, A, B, C), "pdist2");
}

double d2_compute_distance(mph *a, size_t i, mph *b, size_t j, int selected_phase,
                           var_mph *var_work, size_t index_task) {
  // clustering.c:33-128: sum over phases of the LP value between object i of
  // a and object j of b, then sqrt; costs and LP on the device
  (void)var_work;
  (void)index_task;
  double d = 0.0;
  for (int n = 0; n < a->s_ph; ++n) {
    if (!(selected_phase < 0 || n == selected_phase)) continue;
    sph *A = a->ph + n, *Bp = b->ph + n;
    const int nb = Bp->p_str[j], ma = A->p_str[i];
    std::vector<double> C((size_t)nb * ma);
    switch (A->metric_type) {
      case D2_EUCLIDEAN_L2:
        _dpdist2(A->dim, nb, ma, Bp->p_supp + Bp->p_str_cum[j] * A->dim,
                 A->p_supp + A->p_str_cum[i] * A->dim, C.data());
        break;
      case D2_WORD_EMBED: {
        // _dpdist2_sym (blas_like64.c:210-220): the object's words are
        // gathered from the vocabulary, a negative id standing for the zero
        // vector, and the costs come from the device pdist2
        const int d = A->dim;
        std::vector<double> X((size_t)ma * d, 0.0);
        for (int q = 0; q < ma; ++q) {
          const int id = A->p_supp_sym[A->p_str_cum[i] + q];
          if (id >= 0)
            memcpy(X.data() + (size_t)q * d, A->vocab_vec + (size_t)id * d, sizeof(double) * d);
        }
        _dpdist2(d, nb, ma, Bp->p_supp + Bp->p_str_cum[j] * d, X.data(), C.data());
        break;
      }
      case D2_HISTOGRAM:
        memcpy(C.data(), A->dist_mat, sizeof(double) * nb * ma);
        break;
      case D2_SPARSE_HISTOGRAM: {
        // _dpdist2_submat (blas_like64.c:222-230, precomputed at
        // clustering_util.c:189-199): column q is the ground-cost row of the
        // object's bin, row k the centroid's bin k (a centroid holds all V
        // bins).  A side with fewer than V bins is a sparse object and
        // names its bins by id, so two objects can be compared as well.
        const int V = A->vocab_size;
        auto bins = [V](const sph *s, size_t o, int cnt, int q) {
          return cnt < V ? s->p_supp_sym[s->p_str_cum[o] + q] : q;
        };
        for (int q = 0; q < ma; ++q) {
          const size_t row = (size_t)bins(A, i, ma, q) * V;
          for (int k = 0; k < nb; ++k) C[(size_t)q * nb + k] = A->dist_mat[row + bins(Bp, j, nb, k)];
        }
        break;
      }
      default:
        fprintf(stderr, "d2_compute_distance: metric %d is not handled here\n", A->metric_type);
        return NAN;
    }
    d += d2_match_by_distmat(nb, ma, C.data(), Bp->p_w + Bp->p_str_cum[j],
                             A->p_w + A->p_str_cum[i], nullptr, nullptr, 0);
  }
  return d <= 0 ? 0.0 : sqrt(d);
}

int d2_clustering(int num_of_clusters, int max_iter, mph *p_data, mph *centroids,
                  int selected_phase, char use_triangle, const char *log_file) {
  // clustering.c:409-525 with every stage on the device
  printf("Accelerated Discrete Distribution Clustering under Wasserstein Distance "
         "(MI355X engine)\n\n");
  if (d2_alg_type != D2_CENTROID_BADMM) {
    fprintf(stderr, "d2mi: only the B-ADMM centroid method (-M 0) runs on the device\n");
    return -1;
  }
  assert(num_of_clusters > 0 && max_iter > 0 && selected_phase < p_data->s_ph);
  p_data->num_of_labels = num_of_clusters;
  if (!centroids->ph) {
    if (d2_init_centroid(p_data, centroids, selected_phase, false)) return -1;
  } else {
    printf("Centroid initialization provided\n");
  }
  var_mph work;
  d2_allocate_work(p_data, &work, use_triangle, selected_phase);
  if (!sync_data(p_data) || !put_centroids(centroids, selected_phase)) return -1;
  if (!ok(d2mi_set_labels(ctx(), p_data->label, nullptr), "uploading labels")) return -1;
  if (use_triangle && !ok(d2mi_elkan_reset(ctx(), selected_phase), "resetting bounds")) return -1;
  g_elkan_fresh = false;
  const double t0 = now();
  const BADMM_options *o = p_badmm_options;
  int iter;
  for (iter = 0; iter < max_iter; ++iter) {
    printf("Round %d ... \n\tRe-labeling all instances ... ", iter);
    fflush(stdout);
    d2mi_label_stats ls;
    if (!ok(d2mi_labeling(ctx(), selected_phase, use_triangle, &ls), "labeling")) return -1;
    if (use_triangle)
      printf("\n\t\t\t\t %llu objects change their labels\n\t\t\t\t %llu distance pairs "
             "computed\n\t\t\t\t seconds: %f\n",
             ls.changes, ls.dist_count, now() - t0);
    else
      printf("\t %llu labels change.\tmean cost %lf\ttime %f s [done]\n", ls.changes,
             ls.cost / p_data->global_size, now() - t0);
    if (log_file && (iter + 1) % 10 == 0) {  // snapshots (:464-470)
      if (!get_centroids(centroids, selected_phase)) return -1;
      ok(d2mi_get_labels(ctx(), p_data->label, nullptr), "reading labels");
      std::string lf(log_file);
      d2_write((lf + "_c.d2").c_str(), centroids);
      d2_write_labels((lf + ".label").c_str(), p_data);
    }
    if ((double)ls.changes < 0.001 * p_data->global_size) {  // :476
      printf("Terminate!\n");
      break;
    }
    if (use_triangle && !ok(d2mi_save_centroids(ctx(), selected_phase), "saving centroids"))
      return -1;
    printf("\tUpdate centroids ... \n");
    for (int n = 0; n < p_data->s_ph; ++n) {
      if (!(selected_phase < 0 || n == selected_phase)) continue;
      printf("\t phase %d: \n", n);
      d2mi_badmm_stats bs;
      if (!ok(d2mi_badmm(ctx(), n, o->maxIters, o->rhoCoeff, o->updatePerLoops, precision(),
                         D2MI_BADMM_ENGINE, &bs),
              "B-ADMM update"))
        return -1;
      for (int r = 0; r < bs.n_trace; ++r)
        printf("\t%d\t%f\t%f\t%f\n", (int)bs.trace[r][0], bs.trace[r][1], bs.trace[r][2],
               bs.trace[r][3]);
    }
    if (use_triangle && !ok(d2mi_elkan_post(ctx(), selected_phase), "relaxing bounds")) return -1;
  }
  printf("Iteration time: %lf\n", now() - t0);
  if (use_triangle) {  // :509
    d2mi_label_stats ls;
    if (!ok(d2mi_labeling(ctx(), selected_phase, 0, &ls), "labeling")) return -1;
    printf("\t %llu labels change.\tmean cost %lf\ttime %f s [done]\n", ls.changes,
           ls.cost / p_data->global_size, now() - t0);
  }
  if (!get_centroids(centroids, selected_phase)) return -1;
  if (!ok(d2mi_get_labels(ctx(), p_data->label, nullptr), "reading labels")) return -1;
  d2_free_work(&work, selected_phase);
  std::vector<size_t> cnt(num_of_clusters, 0);  // :516-522
  for (size_t i = 0; i < p_data->size; ++i)
    if (p_data->label[i] >= 0) ++cnt[p_data->label[i]];
  for (int i = 0; i < num_of_clusters; ++i) printf("%zu ", cnt[i]);
  printf("\n");
  return 0;
}

int d2_assignment(int num_of_clusters, mph *p_data, mph *centroids, int selected_phase,
                  const char *centroid_filename, const char *meta_filename) {
  // clustering.c:528-546
  p_data->num_of_labels = num_of_clusters;
  if (d2_init_centroid(p_data, centroids, selected_phase, true)) return -1;
  if (d2_read(centroid_filename, meta_filename, centroids)) return -1;
  var_mph work;
  d2_allocate_work(p_data, &work, 0, selected_phase);
  d2_labeling(p_data, centroids, &work, selected_phase);
  d2_free_work(&work, selected_phase);
  return 0;
}

}  // extern "C"
