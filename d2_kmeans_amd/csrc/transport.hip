// Batched exact transport LPs on the device: one workgroup per LP, the
// solver in lp_solver.h.  Costs come from supports staged in LDS or read
// from HBM (Euclidean, squared distance by direct differences), from a
// histogram ground-cost matrix, or from a precomputed FP64 cost array.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "d2mi_internal.h"
#include "lp_solver.h"
#include "transport.h"

namespace {

// One LP per workgroup (a grid-stride loop over jobs).  MAXT bounds the
// workgroup size the instantiation accepts: a smaller bound lets the
// compiler keep the solver's state in registers (at 1024 threads it has 128
// VGPRs per lane and spills).
template <int MAXT>
__global__ __launch_bounds__(MAXT) void k_transport(LpArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x, T = blockDim.x;
  // rows carry supports for Euclidean and word phases; columns for Euclidean
  // objects, or for the other centroid table of a centroid pair
  const bool word = a.cost_mode == LP_COST_WORD;
  const bool euclid = a.cost_mode == LP_COST_EUCLID || (word && a.b_is_table);
  const bool rows = euclid || word;
  const int dA = rows && a.stage_a ? a.d : 0, dB = euclid && a.stage_b ? a.d : 0;
  const lp::Layout L(a.n_max, a.m_max, dA, dB, a.table != lp::TAB_NONE,
                     a.cost_mode == LP_COST_SPARSE || word);
  double *A = (double *)(lds + L.A), *B = (double *)(lds + L.B);
  double *wa = (double *)(lds + L.wa), *wb = (double *)(lds + L.wb);
  int *ids = (int *)(lds + L.ids);
  const int njobs = a.njobs_dev ? min(*a.njobs_dev, a.njobs) : a.njobs;
  for (int t = blockIdx.x; t < njobs; t += gridDim.x) {
    const int job = a.job_map ? a.job_map[t] : t;
    if (a.mask && a.mask[job] == 0) continue;
    // row side: centroid `ka` of table A
    const int ka = a.implicit_K > 0 ? job % a.implicit_K : a.job_a[job];
    const int jb = a.implicit_K > 0 ? job / a.implicit_K : a.job_b[job];
    const int n = a.sA;
    const double *Asrc = rows ? a.A_supp + (size_t)ka * n * a.d : nullptr;
    const double *wA = a.A_w + (size_t)ka * n;
    int m;
    const double *Bsrc = nullptr, *wB;
    const int *idsrc = nullptr;
    if (a.b_is_table) {
      m = a.sB;
      Bsrc = euclid ? a.B_supp + (size_t)jb * m * a.d : nullptr;
      wB = a.B_w + (size_t)jb * m;
    } else {
      m = a.p_str[jb];
      const long long off = a.cum[jb];
      Bsrc = euclid ? a.supp + (size_t)off * a.d : (word ? a.vocab : nullptr);
      wB = a.w + off;
      idsrc = a.sym ? a.sym + off : nullptr;
    }
    __syncthreads();
    for (int q = tid; q < n; q += T) wa[q] = wA[q];
    for (int q = tid; q < m; q += T) wb[q] = wB[q];
    if (dA > 0)
      for (int q = tid; q < n * dA; q += T) A[q] = Asrc[q];
    if (dB > 0)
      for (int q = tid; q < m * dB; q += T) B[q] = Bsrc[q];
    int mode = a.cost_mode;
    if (mode == LP_COST_SPARSE || mode == LP_COST_WORD) {
      if (idsrc) {
        for (int q = tid; q < m; q += T) ids[q] = idsrc[q];
      } else {
        // two centroids: full V-bin histograms, or two sets of dense vectors
        mode = mode == LP_COST_SPARSE ? LP_COST_HIST : LP_COST_EUCLID;
      }
    }
    __syncthreads();
    lp::Problem P;
    P.n = n;
    P.m = m;
    P.wa = wa;
    P.wb = wb;
    P.C.mode = mode;
    P.C.d = a.d;
    P.C.n = n;
    P.C.V = a.V;
    P.C.A = dA > 0 ? A : Asrc;
    P.C.B = dB > 0 ? B : Bsrc;
    P.C.ids = ids;
    P.C.dist = a.dist;
    P.C.g = a.gcost ? a.gcost + (size_t)job * a.gstride : nullptr;
    P.table = a.table;
    P.gt = nullptr;
    P.tab_rel = 0x1p-22f;
    P.cpt = a.cpt;
    P.vogel = a.vogel;
    {
      const long long cells = (long long)n * m;
      const lp::DualSchedule ds = lp::dual_schedule(cells);
      const int it = a.dual_iters >= 0 ? a.dual_iters : ds.iters;
      P.dual_iters = a.dual_ok && cells >= a.dual_min && it >= 2 ? it : 0;
      P.dual_eps0 = a.dual_eps0 > 0.f ? a.dual_eps0 : ds.eps0;
      P.dual_eps1 = a.dual_eps1 > 0.f ? a.dual_eps1 : ds.eps1;
      P.dual_reduce = a.dual_reduce >= 0 ? a.dual_reduce : ds.reduce;
    }
    P.prof = a.prof ? a.prof + (size_t)job * lp::PROF_N : nullptr;
    int piv = 0;
    const double v = lp::solve(P, lds, L, &piv, a.status);
    if (tid == 0) {
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
__global__ void k_iota2(int *a, int *b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] = i;
}

}  // namespace

// The device memory of d2mi_transport_batch_noalloc: one fixed array per GPU,
// held until the process ends.  Host code never uses this name as an
// address (it is a host-side stand-in there): the device address comes from
// hipGetSymbolAddress, once per device, in d2mi_noalloc_setup.
static __device__ __align__(256) unsigned char scratch[D2MI_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int d2mi_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    d2mi_set_error("device %d beyond the %d the no-allocation scratch table holds", device,
                   kMaxDevices);
    return D2MI_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return D2MI_OK;
  int cur = -1;
  D2MI_HIP(hipGetDevice(&cur));
  if (cur != device) {
    d2mi_set_error("d2mi_noalloc_setup: device %d is not current (%d is)", device, cur);
    return D2MI_EINVAL;
  }
  void *p = nullptr;
  D2MI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) {
    d2mi_set_error("hipGetSymbolAddress returned no address for the scratch array");
    return D2MI_EHIP;
  }
  g_scratch_ptr[device] = (unsigned char *)p;
  return D2MI_OK;
}

static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

static double env_double(const char *name, double dflt) {
  const char *v = getenv(name);
  return v && *v ? atof(v) : dflt;
}

int d2mi_plan_transport(d2mi_ctx *ctx, LpArgs &a) {
  const bool word = a.cost_mode == LP_COST_WORD;
  const bool euclid = a.cost_mode == LP_COST_EUCLID || (word && a.b_is_table);
  const bool rows = euclid || word;
  const bool sparse = a.cost_mode == LP_COST_SPARSE || word;
  const size_t nn = (size_t)a.n_max + a.m_max;
  if (nn > 32767 || (size_t)a.n_max * a.m_max >= (1u << 24)) {
    d2mi_set_error("transport LP of %d x %d is beyond the solver's size limit", a.n_max, a.m_max);
    return D2MI_EUNSUPPORTED;
  }
  // what goes to LDS, in order of value: the FP32 cost table, then the
  // row supports, then the column supports
  const size_t cap = ctx->lds_max;
  auto bytes = [&](bool tab, bool sa, bool sb) {
    return lp::Layout(a.n_max, a.m_max, rows && sa ? a.d : 0, euclid && sb ? a.d : 0, tab, sparse)
        .total;
  };
  bool tab = bytes(true, false, false) <= cap && env_int("D2MI_LP_TABLE", 1) != 0;
  bool sa = rows && bytes(tab, true, false) <= cap;
  bool sb = euclid && bytes(tab, sa, true) <= cap && (size_t)a.m_max * a.d <= 4096;
  if (bytes(tab, sa, sb) > cap) {
    d2mi_set_error("transport LP of %d x %d needs %zu B of LDS (limit %zu)", a.n_max, a.m_max,
                   bytes(tab, sa, sb), cap);
    return D2MI_EUNSUPPORTED;
  }
  a.table = tab ? lp::TAB_EXACT : lp::TAB_NONE;
  a.stage_a = sa;
  a.stage_b = sb;
  a.lds = bytes(tab, sa, sb);
  // many small LPs: one wave each; few or large LPs: more waves per LP
  const size_t cells = (size_t)a.n_max * a.m_max;
  // few LPs, or LDS for only one or two per CU: more waves per LP
  const size_t per_cu = std::max<size_t>(1, cap / a.lds);
  int T = 64;
  if (a.njobs < ctx->num_cu * 4 || per_cu <= 2) {
    if (cells >= 8192) T = 256;
    else if (cells >= 2048) T = 128;
  } else if (per_cu <= 4 && cells >= 2048) {
    T = 128;
  }
  if (a.threads_hint > 0) T = a.threads_hint;
  T = env_int("D2MI_LP_THREADS", T);
  T = std::max(64, std::min(1024, T / 64 * 64));
  a.threads = T;
  a.cpt = std::max(1, std::min(lp::CPT_MAX, env_int("D2MI_LP_BLOCK", 2048) / T));
  // histogram ground costs on ordered bins (|i-j|, (i-j)^2, ...) usually form
  // a Monge matrix, where the north-west corner is already optimal
  const bool hist = a.cost_mode == LP_COST_HIST || a.cost_mode == LP_COST_SPARSE;
  a.vogel = env_int("D2MI_LP_VOGEL", hist ? 0 : 1) != 0;
  // Vogel on reduced costs from approximate duals, where the pivots it
  // saves outweigh its entropic iterations.  Overrides for tuning:
  // D2MI_LP_DUAL_MIN (cell threshold), D2MI_LP_DUAL_ITERS (0 = plain Vogel),
  // D2MI_LP_DUAL_EPS0 / _EPS1 (first and last temperature), D2MI_LP_DUAL_REDUCE
  a.dual_ok = a.vogel && tab;
  a.dual_min = env_int("D2MI_LP_DUAL_MIN", lp::DUAL_MIN_CELLS);
  a.dual_iters = env_int("D2MI_LP_DUAL_ITERS", -1);
  a.dual_eps0 = (float)env_double("D2MI_LP_DUAL_EPS0", -1.0);
  a.dual_eps1 = (float)env_double("D2MI_LP_DUAL_EPS1", -1.0);
  a.dual_reduce = env_int("D2MI_LP_DUAL_REDUCE", -1);
  return D2MI_OK;
}

int d2mi_launch_transport(d2mi_ctx *ctx, const LpArgs &a0) {
  if (a0.njobs <= 0) return D2MI_OK;
  LpArgs a = a0;
  D2MI_TRY(d2mi_plan_transport(ctx, a));
  // the instantiation with the smallest bound that admits the workgroup
  // (D2MI_LP_BOUND overrides it: 256, 512 or 1024)
  const int want = env_int("D2MI_LP_BOUND", 0);
  const int bound = want >= a.threads && (want == 256 || want == 512 || want == 1024) ? want
                    : a.threads <= 256 ? 256 : a.threads <= 512 ? 512 : 1024;
  auto kern = bound == 256 ? k_transport<256> : bound == 512 ? k_transport<512> : k_transport<1024>;
  int per_cu = 0;
  D2MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, a.threads, a.lds));
  const int grid = std::min(a.njobs, ctx->num_cu * std::max(1, per_cu));
  kern<<<grid, a.threads, a.lds, a.stream ? a.stream : ctx->stream>>>(a);
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

namespace {
// Device staging of one d2mi_transport_batch call, released on every return.
struct BatchStaging {
  DevMem C, A, B, out, ja, jb, piv, st, prof;
  ~BatchStaging() {
    for (DevMem *m : {&C, &A, &B, &out, &ja, &jb, &piv, &st, &prof}) m->release();
  }
};

// d2mi_transport_batch, and with prof_us the per-LP phase times
// (lp::PROF_N per LP, microseconds) of d2mi_transport_profile.
int transport_batch(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C, const double *wX,
                    const double *wY, double *out, int *pivots, double *prof_us) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (n <= 0 || m <= 0 || nbatch <= 0 || !C || !wX || !wY || !out) {
    d2mi_set_error("d2mi_transport_batch: bad arguments");
    return D2MI_EINVAL;
  }
  hipStream_t S = ctx->stream;
  BatchStaging d;
  const size_t nm = (size_t)n * m;
  D2MI_TRY(d.C.assign(C, sizeof(double) * nm * nbatch, S));
  D2MI_TRY(d.A.assign(wX, sizeof(double) * (size_t)n * nbatch, S));
  D2MI_TRY(d.B.assign(wY, sizeof(double) * (size_t)m * nbatch, S));
  D2MI_TRY(d.out.ensure(sizeof(double) * nbatch));
  D2MI_TRY(d.piv.ensure(sizeof(int) * nbatch));
  D2MI_TRY(d.st.ensure(sizeof(int)));
  std::vector<int> idx(nbatch);
  for (int i = 0; i < nbatch; ++i) idx[i] = i;
  D2MI_TRY(d.ja.assign(idx.data(), sizeof(int) * nbatch, S));
  D2MI_TRY(d.jb.assign(idx.data(), sizeof(int) * nbatch, S));
  D2MI_TRY(d2mi_zero_doubles(ctx, d.out.as<double>(), nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, d.piv.as<int>(), nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, d.st.as<int>(), 1));
  const int np = nbatch * lp::PROF_N;
  if (prof_us) {
    D2MI_TRY(d.prof.ensure(sizeof(long long) * np));
    D2MI_TRY(d2mi_zero_doubles(ctx, d.prof.as<double>(), np));  // all-zero bits
  }
  LpArgs a{};
  a.njobs = nbatch;
  a.job_a = d.ja.as<int>();
  a.job_b = d.jb.as<int>();
  a.b_is_table = 1;
  a.cost_mode = LP_COST_GLOBAL;
  a.sA = n;
  a.A_w = d.A.as<double>();
  a.sB = m;
  a.B_w = d.B.as<double>();
  a.gcost = d.C.as<double>();
  a.gstride = (long long)nm;
  a.n_max = n;
  a.m_max = m;
  a.out = d.out.as<double>();
  a.pivots = d.piv.as<int>();
  a.status = d.st.as<int>();
  a.prof = prof_us ? d.prof.as<long long>() : nullptr;
  D2MI_TRY(d2mi_launch_transport(ctx, a));
  D2MI_HIP(hipMemcpyAsync(out, d.out.p, sizeof(double) * nbatch, hipMemcpyDeviceToHost, S));
  if (pivots)
    D2MI_HIP(hipMemcpyAsync(pivots, d.piv.p, sizeof(int) * nbatch, hipMemcpyDeviceToHost, S));
  std::vector<long long> ticks(prof_us ? np : 0);
  if (prof_us)
    D2MI_HIP(hipMemcpyAsync(ticks.data(), d.prof.p, sizeof(long long) * np,
                            hipMemcpyDeviceToHost, S));
  int st = 0;
  D2MI_HIP(hipMemcpyAsync(&st, d.st.p, sizeof(int), hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (prof_us) {
    int khz = 0;  // the wall clock's rate
    D2MI_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    const double us_per_tick = khz > 0 ? 1e3 / khz : 0.0;
    for (int i = 0; i < np; ++i) prof_us[i] = ticks[i] * us_per_tick;
  }
  if (st) {
    d2mi_set_error("a transport LP hit the pivot limit");
    return D2MI_ENUMERIC;
  }
  return D2MI_OK;
}
}  // namespace

extern "C" int d2mi_transport_batch(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                                    const double *wX, const double *wY, double *out,
                                    int *pivots) {
  return transport_batch(ctx, n, m, nbatch, C, wX, wY, out, pivots, nullptr);
}

extern "C" int d2mi_transport_profile(d2mi_ctx *ctx, int n, int m, int nbatch, const double *C,
                                      const double *wX, const double *wY, double *out,
                                      int *pivots, double *phase_us) {
  if (!phase_us) {
    d2mi_set_error("d2mi_transport_profile: null phase_us");
    return D2MI_EINVAL;
  }
  return transport_batch(ctx, n, m, nbatch, C, wX, wY, out, pivots, phase_us);
}

extern "C" int d2mi_transport_batch_noalloc(d2mi_ctx *ctx, int n, int m, int nbatch,
                                            const double *C, const double *wX, const double *wY,
                                            double *out, int *pivots) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (n <= 0 || m <= 0 || nbatch <= 0 || !C || !wX || !wY || !out) {
    d2mi_set_error("d2mi_transport_batch_noalloc: bad arguments");
    return D2MI_EINVAL;
  }
  // staging inside the scratch array, each region 256-byte aligned
  const size_t nm = (size_t)n * m, nb = (size_t)nbatch;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t oC = 0, oA = oC + up(sizeof(double) * nm * nb), oB = oA + up(sizeof(double) * n * nb),
               oO = oB + up(sizeof(double) * m * nb), oP = oO + up(sizeof(double) * nb),
               oJa = oP + up(sizeof(int) * nb), oJb = oJa + up(sizeof(int) * nb),
               oS = oJb + up(sizeof(int) * nb), total = oS + 256;
  if (total > (size_t)D2MI_NOALLOC_SCRATCH_BYTES) {
    d2mi_set_error("d2mi_transport_batch_noalloc: %d LPs of %d x %d need %zu bytes of staging, "
                   "above the fixed %u-byte scratch; split the batch or use "
                   "d2mi_transport_batch", nbatch, n, m, total, D2MI_NOALLOC_SCRATCH_BYTES);
    return D2MI_EINVAL;
  }
  if (!g_scratch_ptr[ctx->device]) D2MI_TRY(d2mi_noalloc_setup(ctx->device));
  unsigned char *base = g_scratch_ptr[ctx->device];
  double *dC = (double *)(base + oC), *dA = (double *)(base + oA), *dB = (double *)(base + oB),
         *dout = (double *)(base + oO);
  int *dpiv = (int *)(base + oP), *dja = (int *)(base + oJa), *djb = (int *)(base + oJb),
      *dst = (int *)(base + oS);
  hipStream_t S = ctx->stream;
  D2MI_HIP(hipMemcpyAsync(dC, C, sizeof(double) * nm * nb, hipMemcpyHostToDevice, S));
  D2MI_HIP(hipMemcpyAsync(dA, wX, sizeof(double) * n * nb, hipMemcpyHostToDevice, S));
  D2MI_HIP(hipMemcpyAsync(dB, wY, sizeof(double) * m * nb, hipMemcpyHostToDevice, S));
  // the array keeps the last call's contents: values, pivot counts and the
  // status flag are accumulated into, so they start from zero (fill kernels)
  k_iota2<<<(nbatch + 255) / 256, 256, 0, S>>>(dja, djb, nbatch);
  D2MI_HIP(hipGetLastError());
  D2MI_TRY(d2mi_zero_doubles(ctx, dout, nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, dpiv, nbatch));
  D2MI_TRY(d2mi_zero_ints(ctx, dst, 1));
  LpArgs a{};
  a.njobs = nbatch;
  a.job_a = dja;
  a.job_b = djb;
  a.b_is_table = 1;
  a.cost_mode = LP_COST_GLOBAL;
  a.sA = n;
  a.A_w = dA;
  a.sB = m;
  a.B_w = dB;
  a.gcost = dC;
  a.gstride = (long long)nm;
  a.n_max = n;
  a.m_max = m;
  a.out = dout;
  a.pivots = dpiv;
  a.status = dst;
  D2MI_TRY(d2mi_launch_transport(ctx, a));
  D2MI_HIP(hipMemcpyAsync(out, dout, sizeof(double) * nbatch, hipMemcpyDeviceToHost, S));
  if (pivots) D2MI_HIP(hipMemcpyAsync(pivots, dpiv, sizeof(int) * nbatch, hipMemcpyDeviceToHost, S));
  int st = 0;
  D2MI_HIP(hipMemcpyAsync(&st, dst, sizeof(int), hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (st) {
    d2mi_set_error("a transport LP hit the pivot limit");
    return D2MI_ENUMERIC;
  }
  return D2MI_OK;
}
