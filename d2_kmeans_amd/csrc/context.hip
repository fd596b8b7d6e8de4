// Device context, uploads and the small C-ABI entry points of libd2mi.so.
#include <cstdio>
#include <cstring>
#include <string>

#include "d2mi_internal.h"

static thread_local std::string g_err;

void d2mi_set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

extern "C" const char *d2mi_last_error(void) { return g_err.c_str(); }

int DevMem::ensure(size_t want) {
  if (want <= bytes && p) return D2MI_OK;
  release();
  if (want == 0) want = 16;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    d2mi_set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
    return D2MI_ENOMEM;
  }
  bytes = want;
  return D2MI_OK;
}

int DevMem::assign(const void *host, size_t nbytes, hipStream_t s) {
  D2MI_TRY(ensure(nbytes));
  if (nbytes) D2MI_HIP(hipMemcpyAsync(p, host, nbytes, hipMemcpyHostToDevice, s));
  return D2MI_OK;
}

void DevMem::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

int d2mi_check_ctx(d2mi_ctx *ctx) {
  if (!ctx) {
    d2mi_set_error("null context");
    return D2MI_EINVAL;
  }
  D2MI_HIP(hipSetDevice(ctx->device));
  return D2MI_OK;
}

PhaseDev *d2mi_phase(d2mi_ctx *ctx, int phase) {
  if (phase < 0 || phase >= (int)ctx->ph.size() || !ctx->ph[phase].set) {
    d2mi_set_error("phase %d has not been uploaded", phase);
    return nullptr;
  }
  return &ctx->ph[phase];
}

extern "C" int d2mi_device_count(int *count) {
  int n = 0;
  D2MI_HIP(hipGetDeviceCount(&n));
  if (count) *count = n;
  return D2MI_OK;
}

extern "C" int d2mi_create(d2mi_ctx **out, int device) {
  if (!out) {
    d2mi_set_error("null output pointer");
    return D2MI_EINVAL;
  }
  int n = 0;
  D2MI_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    d2mi_set_error("device %d out of range (%d devices)", device, n);
    return D2MI_EINVAL;
  }
  D2MI_HIP(hipSetDevice(device));
  D2MI_TRY(d2mi_noalloc_setup(device));
  d2mi_ctx *c = new d2mi_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  D2MI_HIP(hipGetDeviceProperties(&prop, device));
  c->lds_max = prop.sharedMemPerBlock;
  c->num_cu = prop.multiProcessorCount;
  D2MI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto &e : c->ev) D2MI_HIP(hipEventCreate(&e));
  *out = c;
  return D2MI_OK;
}

static void release_phase(PhaseDev &p) {
  DevMem *all[] = {&p.p_str, &p.cum, &p.w64, &p.w32, &p.supp64, &p.supp32, &p.sym,
                   &p.vocab64, &p.vocab32, &p.vocab16, &p.vnorm16, &p.dist64, &p.dist32, &p.cw64, &p.cw32,
                   &p.cs64, &p.cs32, &p.Z32, &p.Y32, &p.Z64, &p.Y64, &p.cw_old, &p.cs_old,
                   &p.Ctab, &p.Xout, &p.Z0, &p.bucket_map, &p.xmean, &p.xspread};
  for (DevMem *m : all) m->release();
  p = PhaseDev();
}

extern "C" int d2mi_destroy(d2mi_ctx *ctx) {
  if (!ctx) return D2MI_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->ph) release_phase(p);
  DevMem *all[] = {&ctx->label, &ctx->lsw, &ctx->perm, &ctx->blkcnt, &ctx->segstart,
                   &ctx->rec, &ctx->part, &ctx->scal, &ctx->flags, &ctx->resid,
                   &ctx->X32, &ctx->X64, &ctx->pA, &ctx->pB, &ctx->pC, &ctx->el_l,
                   &ctx->el_u, &ctx->el_s, &ctx->el_c, &ctx->el_r, &ctx->dist, &ctx->mask,
                   &ctx->jobs, &ctx->lpstat, &ctx->cnt, &ctx->mind2, &ctx->red,
                   &ctx->rec2, &ctx->rxbuf, &ctx->bkey, &ctx->permk, &ctx->blkcnt2,
                   &ctx->segk, &ctx->wtab, &ctx->wnorm, &ctx->cs16, &ctx->tuneZ, &ctx->tuneY,
                   &ctx->rowsum, &ctx->sdesc, &ctx->sblab, &ctx->sbs, &ctx->scum,
                   &ctx->scbt, &ctx->szero, &ctx->pmeta};
  for (DevMem *m : all) m->release();
  for (auto &e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &kv : ctx->badmm_exec)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto &e : ctx->pass_ev) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return D2MI_OK;
}

extern "C" int d2mi_synchronize(d2mi_ctx *ctx) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

// FP64 -> FP32 element-wise conversion
__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// Per object of a Euclidean phase: the mean support point mu (unweighted)
// and sum_j |x_j - mu|^2, in FP64.  They give the ground-cost sum of an
// object against any s centroid points without forming its s x m costs:
// sum_k sum_j |c_k - x_j|^2 = m sum_k |c_k - mu|^2 + s sum_j |x_j - mu|^2
// (the cross term vanishes because sum_j (x_j - mu) = 0), a sum of
// non-negative terms with no cancellation.
__global__ void k_obj_moments(size_t N, int d, const int *__restrict__ p_str,
                              const long long *__restrict__ cum, const double *__restrict__ supp,
                              double *__restrict__ mean, double *__restrict__ spread) {
  for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < N;
       o += (size_t)gridDim.x * blockDim.x) {
    const int m = p_str[o];
    const double *x = supp + (size_t)cum[o] * d;
    double *mu = mean + o * d;
    for (int dd = 0; dd < d; ++dd) {
      double a = 0.0;
      for (int j = 0; j < m; ++j) a += x[(size_t)j * d + dd];
      mu[dd] = a / m;
    }
    double v = 0.0;
    for (int j = 0; j < m; ++j)
      for (int dd = 0; dd < d; ++dd) {
        const double u = x[(size_t)j * d + dd] - mu[dd];
        v += u * u;
      }
    spread[o] = v;
  }
}

static int upload_pair(DevMem &d64, DevMem &d32, const double *h, size_t n, hipStream_t s) {
  D2MI_TRY(d64.assign(h, n * sizeof(double), s));
  D2MI_TRY(d32.ensure(n * sizeof(float)));
  if (n) {
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    k_to_f32<<<blocks, 256, 0, s>>>(d64.as<double>(), d32.as<float>(), n);
    D2MI_HIP(hipGetLastError());
  }
  return D2MI_OK;
}

extern "C" int d2mi_set_phase(d2mi_ctx *ctx, int phase, int metric, int dim, size_t size,
                              size_t col, const int *p_str, const size_t *p_str_cum,
                              const double *p_w, const double *p_supp, const int *p_supp_sym,
                              int vocab_size, const double *vocab_vec,
                              const double *dist_mat) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (phase < 0 || phase > 64 || size == 0 || !p_str || !p_str_cum || !p_w) {
    d2mi_set_error("d2mi_set_phase: bad arguments");
    return D2MI_EINVAL;
  }
  if (metric != D2MI_EUCLIDEAN && metric != D2MI_HISTOGRAM && metric != D2MI_WORD_EMBED &&
      metric != D2MI_SPARSE_HISTOGRAM) {
    d2mi_set_error("d2mi_set_phase: metric %d not supported", metric);
    return D2MI_EUNSUPPORTED;
  }
  if (phase == 0) {
    // phase 0 starts a new data set: drop the other phases and per-object state
    for (size_t k = 1; k < ctx->ph.size(); ++k) release_phase(ctx->ph[k]);
    if (!ctx->ph.empty()) ctx->ph.resize(1);
    ctx->label.release();
    ctx->lsw.release();
    ctx->elkan_K = 0;
    ctx->N = 0;
  } else if (ctx->N && size != ctx->N) {
    d2mi_set_error("phase %d has %zu objects, phase 0 has %zu", phase, size, ctx->N);
    return D2MI_EINVAL;
  }
  // validate the CSR layout on the host: the kernels trust it
  size_t acc = 0;
  int m_max = 0;
  for (size_t i = 0; i < size; ++i) {
    if (p_str[i] <= 0 || p_str_cum[i] != acc) {
      d2mi_set_error("object %zu: p_str=%d p_str_cum=%zu (expected %zu)", i, p_str[i],
                     p_str_cum[i], acc);
      return D2MI_EINVAL;
    }
    acc += (size_t)p_str[i];
    m_max = std::max(m_max, p_str[i]);
  }
  if (acc != col) {
    d2mi_set_error("sum of p_str = %zu but col = %zu", acc, col);
    return D2MI_EINVAL;
  }
  if (metric == D2MI_EUCLIDEAN && (!p_supp || dim <= 0)) {
    d2mi_set_error("Euclidean phase needs p_supp and dim > 0");
    return D2MI_EINVAL;
  }
  if ((metric == D2MI_WORD_EMBED || metric == D2MI_SPARSE_HISTOGRAM) && !p_supp_sym) {
    d2mi_set_error("symbolic phase needs p_supp_sym");
    return D2MI_EINVAL;
  }
  if (metric == D2MI_WORD_EMBED && (!vocab_vec || vocab_size <= 0 || dim <= 0)) {
    d2mi_set_error("word-embedding phase needs vocab_vec");
    return D2MI_EINVAL;
  }
  if ((metric == D2MI_HISTOGRAM || metric == D2MI_SPARSE_HISTOGRAM) &&
      (!dist_mat || vocab_size <= 0)) {
    d2mi_set_error("histogram phase needs dist_mat and vocab_size");
    return D2MI_EINVAL;
  }
  if (p_supp_sym)
    for (size_t e = 0; e < col; ++e)
      if (p_supp_sym[e] < 0 || p_supp_sym[e] >= vocab_size) {
        d2mi_set_error("p_supp_sym[%zu] = %d outside [0, %d)", e, p_supp_sym[e], vocab_size);
        return D2MI_EINVAL;
      }
  if (metric == D2MI_HISTOGRAM && m_max != vocab_size) {
    d2mi_set_error("dense histogram objects must have V = %d bins", vocab_size);
    return D2MI_EINVAL;
  }
  if ((int)ctx->ph.size() <= phase) ctx->ph.resize(phase + 1);
  PhaseDev &p = ctx->ph[phase];
  release_phase(p);
  hipStream_t s = ctx->stream;
  p.metric = metric;
  p.dim = dim;
  p.N = size;
  p.col = col;
  p.V = vocab_size;
  p.m_max = m_max;
  p.h_pstr.assign(p_str, p_str + size);
  p.bucket_K = 0;
  std::vector<long long> cum64(size);
  for (size_t i = 0; i < size; ++i) cum64[i] = (long long)p_str_cum[i];
  D2MI_TRY(p.p_str.assign(p_str, size * sizeof(int), s));
  D2MI_TRY(p.cum.assign(cum64.data(), size * sizeof(long long), s));
  D2MI_TRY(upload_pair(p.w64, p.w32, p_w, col, s));
  if (metric == D2MI_EUCLIDEAN) {
    D2MI_TRY(upload_pair(p.supp64, p.supp32, p_supp, col * dim, s));
    D2MI_TRY(p.xmean.ensure(sizeof(double) * size * dim));
    D2MI_TRY(p.xspread.ensure(sizeof(double) * size));
    const unsigned blocks = (unsigned)std::min<size_t>((size + 255) / 256, 4096);
    k_obj_moments<<<blocks, 256, 0, s>>>(size, dim, p.p_str.as<int>(), p.cum.as<long long>(),
                                         p.supp64.as<double>(), p.xmean.as<double>(),
                                         p.xspread.as<double>());
    D2MI_HIP(hipGetLastError());
  }
  if (p_supp_sym && metric != D2MI_EUCLIDEAN)
    D2MI_TRY(p.sym.assign(p_supp_sym, col * sizeof(int), s));
  if (metric == D2MI_WORD_EMBED)
    D2MI_TRY(upload_pair(p.vocab64, p.vocab32, vocab_vec, (size_t)vocab_size * dim, s));
  if (metric == D2MI_HISTOGRAM || metric == D2MI_SPARSE_HISTOGRAM)
    D2MI_TRY(upload_pair(p.dist64, p.dist32, dist_mat, (size_t)vocab_size * vocab_size, s));
  D2MI_HIP(hipStreamSynchronize(s));  // host arrays may be freed after return
  p.set = true;
  ctx->N = size;
  return D2MI_OK;
}

extern "C" int d2mi_set_labels(d2mi_ctx *ctx, const int *label, const char *label_switch) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (!ctx->N || !label) {
    d2mi_set_error("d2mi_set_labels: upload a phase first / null labels");
    return D2MI_EINVAL;
  }
  D2MI_TRY(ctx->label.assign(label, ctx->N * sizeof(int), ctx->stream));
  if (label_switch) {
    D2MI_TRY(ctx->lsw.assign(label_switch, ctx->N, ctx->stream));
  } else {
    D2MI_TRY(ctx->lsw.ensure(ctx->N));
    D2MI_HIP(hipMemsetAsync(ctx->lsw.p, 0, ctx->N, ctx->stream));
  }
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

extern "C" int d2mi_get_labels(d2mi_ctx *ctx, int *label, char *label_switch) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (!ctx->label.p) {
    d2mi_set_error("no labels on the device");
    return D2MI_EINVAL;
  }
  if (label)
    D2MI_HIP(hipMemcpyAsync(label, ctx->label.p, ctx->N * sizeof(int), hipMemcpyDeviceToHost,
                            ctx->stream));
  if (label_switch && ctx->lsw.p)
    D2MI_HIP(hipMemcpyAsync(label_switch, ctx->lsw.p, ctx->N, hipMemcpyDeviceToHost,
                            ctx->stream));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

extern "C" int d2mi_set_centroids(d2mi_ctx *ctx, int phase, int K, int s, const double *p_w,
                                  const double *p_supp) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  PhaseDev *p = d2mi_phase(ctx, phase);
  if (!p) return D2MI_EINVAL;
  if (K <= 0 || s <= 0 || !p_w) {
    d2mi_set_error("d2mi_set_centroids: bad K=%d s=%d", K, s);
    return D2MI_EINVAL;
  }
  const bool needs_supp = p->metric == D2MI_EUCLIDEAN || p->metric == D2MI_WORD_EMBED;
  if (needs_supp && !p_supp) {
    d2mi_set_error("centroids of this metric need supports");
    return D2MI_EINVAL;
  }
  if ((p->metric == D2MI_HISTOGRAM || p->metric == D2MI_SPARSE_HISTOGRAM) && s != p->V) {
    d2mi_set_error("histogram centroids must have s = V = %d bins", p->V);
    return D2MI_EINVAL;
  }
  p->K = K;
  p->s = s;
  D2MI_TRY(upload_pair(p->cw64, p->cw32, p_w, (size_t)K * s, ctx->stream));
  if (needs_supp) D2MI_TRY(upload_pair(p->cs64, p->cs32, p_supp, (size_t)K * s * p->dim, ctx->stream));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

extern "C" int d2mi_get_centroids(d2mi_ctx *ctx, int phase, double *p_w, double *p_supp) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  PhaseDev *p = d2mi_phase(ctx, phase);
  if (!p) return D2MI_EINVAL;
  if (!p->K) {
    d2mi_set_error("phase %d has no centroids", phase);
    return D2MI_EINVAL;
  }
  if (p_w)
    D2MI_HIP(hipMemcpyAsync(p_w, p->cw64.p, sizeof(double) * p->K * p->s, hipMemcpyDeviceToHost,
                            ctx->stream));
  if (p_supp && p->cs64.p)
    D2MI_HIP(hipMemcpyAsync(p_supp, p->cs64.p, sizeof(double) * p->K * p->s * p->dim,
                            hipMemcpyDeviceToHost, ctx->stream));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

__global__ void k_f32_to_f64(const float *__restrict__ a, double *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

// Two-array read-modify-write streams: the access pattern of a B-ADMM plan
// pass (Z and Y read and written in place) without its arithmetic or
// reductions.  The arrays start at zero and stay there.
__global__ void k_stream2_f1(float *__restrict__ a, float *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float x = a[i], y = b[i];
    a[i] = 0.5f * x + y;
    b[i] = 0.5f * y;
  }
}

__global__ void k_stream2_f4(float4 *__restrict__ a, float4 *__restrict__ b, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x) {
    float4 x = a[i], y = b[i];
    x.x = 0.5f * x.x + y.x; x.y = 0.5f * x.y + y.y;
    x.z = 0.5f * x.z + y.z; x.w = 0.5f * x.w + y.w;
    y.x *= 0.5f; y.y *= 0.5f; y.z *= 0.5f; y.w *= 0.5f;
    a[i] = x;
    b[i] = y;
  }
}

extern "C" int d2mi_stream_probe(d2mi_ctx *ctx, size_t n, int reps, double *gbps,
                                 char *report, int report_len) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (!gbps || n < 1024 || reps < 1) {
    d2mi_set_error("stream probe: n >= 1024 elements and reps >= 1 required");
    return D2MI_EINVAL;
  }
  n &= ~(size_t)3;  // whole float4s
  DevMem a, b;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best = 0;
  std::string rep;
  int rc = a.ensure(n * sizeof(float));
  if (rc == D2MI_OK) rc = b.ensure(n * sizeof(float));
  if (rc == D2MI_OK &&
      (hipMemsetAsync(a.p, 0, n * sizeof(float), ctx->stream) != hipSuccess ||
       hipMemsetAsync(b.p, 0, n * sizeof(float), ctx->stream) != hipSuccess ||
       hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = D2MI_EHIP;
  for (int per_cu : {8, 16, 32}) {
    if (rc != D2MI_OK) break;
    const int grid = ctx->num_cu * per_cu;
    for (int vec = 0; vec < 2 && rc == D2MI_OK; ++vec) {
      float ms_best = 1e30f;
      for (int r = 0; r <= reps; ++r) {  // r = 0 warms up
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = D2MI_EHIP; break; }
        if (vec)
          k_stream2_f4<<<grid, 256, 0, ctx->stream>>>(a.as<float4>(), b.as<float4>(), n / 4);
        else
          k_stream2_f1<<<grid, 256, 0, ctx->stream>>>(a.as<float>(), b.as<float>(), n);
        float ms = 0;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = D2MI_EHIP; break; }
        if (r > 0 && ms < ms_best) ms_best = ms;
      }
      if (rc != D2MI_OK) break;
      const double g = 16.0 * n / (ms_best * 1e-3) / 1e9;
      best = g > best ? g : best;
      char line[128];
      snprintf(line, sizeof line, "%s %d workgroups/CU: %.4f ms, %.1f GB/s\n",
               vec ? "float4" : "dword", per_cu, ms_best, g);
      rep += line;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  a.release();
  b.release();
  if (rc != D2MI_OK) {
    if (rc == D2MI_EHIP) d2mi_set_error("stream probe: HIP call failed");
    return rc;
  }
  *gbps = best;
  if (report && report_len > 0) snprintf(report, report_len, "%s", rep.c_str());
  return D2MI_OK;
}

extern "C" int d2mi_get_plan(d2mi_ctx *ctx, int phase, int precision, int which, double *out) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  PhaseDev *p = d2mi_phase(ctx, phase);
  if (!p) return D2MI_EINVAL;
  const int ps = precision == 64 ? p->plan_s64 : p->plan_s32;
  if (!ps || !out || (which != 0 && which != 1)) {
    d2mi_set_error("no resident plan at precision %d", precision);
    return D2MI_EINVAL;
  }
  const size_t n = (size_t)ps * p->col;
  if (precision == 64) {
    DevMem &src = which ? p->Y64 : p->Z64;
    D2MI_HIP(hipMemcpyAsync(out, src.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  } else {
    DevMem &src = which ? p->Y32 : p->Z32;
    DevMem tmp;
    D2MI_TRY(tmp.ensure(n * sizeof(double)));
    k_f32_to_f64<<<4096, 256, 0, ctx->stream>>>(src.as<float>(), tmp.as<double>(), n);
    D2MI_HIP(hipMemcpyAsync(out, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    D2MI_HIP(hipStreamSynchronize(ctx->stream));
    tmp.release();
  }
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

int d2mi_ensure_plans(d2mi_ctx *ctx, PhaseDev *p, int precision);

extern "C" int d2mi_set_plan(d2mi_ctx *ctx, int phase, int precision, const double *Z) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  PhaseDev *p = d2mi_phase(ctx, phase);
  if (!p || !Z || !p->K) {
    if (p) d2mi_set_error("set centroids before the plan");
    return D2MI_EINVAL;
  }
  const int fresh = d2mi_ensure_plans(ctx, p, precision);
  if (fresh < 0) return fresh;
  const size_t n = (size_t)p->s * p->col;
  if (precision == 64) {
    D2MI_HIP(hipMemcpyAsync(p->Z64.p, Z, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  } else {
    DevMem tmp;
    D2MI_TRY(tmp.assign(Z, n * sizeof(double), ctx->stream));
    k_to_f32<<<4096, 256, 0, ctx->stream>>>(tmp.as<double>(), p->Z32.as<float>(), n);
    D2MI_HIP(hipStreamSynchronize(ctx->stream));
    tmp.release();
  }
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}

// ----------------------------------------------------------------- pdist2

// C[i*n + j] = sum_k (A[j*d+k] - B[i*d+k])^2 by direct differences (no
// ||a||^2 + ||b||^2 - 2ab cancellation), one thread per output, FP64.
__global__ void k_pdist2_f64(int d, size_t n, size_t m, const double *__restrict__ A,
                             const double *__restrict__ B, double *__restrict__ C) {
  const size_t total = n * m;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t / n, j = t - i * n;
    double acc = 0.0;
    for (int k = 0; k < d; ++k) {
      const double u = A[j * d + k] - B[i * d + k];
      acc += u * u;
    }
    C[t] = acc;
  }
}

extern "C" int d2mi_pdist2(d2mi_ctx *ctx, int d, size_t n, size_t m, const double *A,
                           const double *B, double *C) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (d <= 0 || n == 0 || m == 0 || !A || !B || !C) {
    d2mi_set_error("d2mi_pdist2: bad arguments");  // blas_like64.c:201 asserts the same
    return D2MI_EINVAL;
  }
  D2MI_TRY(ctx->pA.assign(A, sizeof(double) * n * d, ctx->stream));
  D2MI_TRY(ctx->pB.assign(B, sizeof(double) * m * d, ctx->stream));
  D2MI_TRY(ctx->pC.ensure(sizeof(double) * n * m));
  const size_t total = n * m;
  unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  k_pdist2_f64<<<blocks, 256, 0, ctx->stream>>>(d, n, m, ctx->pA.as<double>(),
                                                ctx->pB.as<double>(), ctx->pC.as<double>());
  D2MI_HIP(hipGetLastError());
  D2MI_HIP(hipMemcpyAsync(C, ctx->pC.p, sizeof(double) * total, hipMemcpyDeviceToHost,
                          ctx->stream));
  D2MI_HIP(hipStreamSynchronize(ctx->stream));
  return D2MI_OK;
}
