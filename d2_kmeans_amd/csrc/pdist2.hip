// Squared-Euclidean distance tables on the matrix cores: the reference's
// pdist2 (_dpdist2 / _dpdist2_sym, blas_like64.c:199-220) restated as a GEMM,
//     T[p*Q + q] = |x_p|^2 + |y_q|^2 - 2 x_p . y_q      (P x Q, FP32 out),
// with FP32 row norms and FP32 accumulation, the products on MFMA in FP32
// (v_mfma_f32_32x32x2f32) or BF16 (v_mfma_f32_32x32x16_bf16).  Entries whose
// value is small next to |x|^2 + |y|^2 lose digits to cancellation; they are
// recomputed by direct differences in FP32.
//
// Tiling: a workgroup of 4 waves owns a 128 x 128 output tile, each wave a
// 64 x 64 quarter as 2 x 2 blocks of 32 x 32 accumulators; the shared
// dimension is staged through LDS 32 at a time (zero-padded past d).
// MFMA operand layouts (32 x 32 x 2 f32): lane l supplies A[l % 32][l / 32]
// and B[l / 32][l % 32]; result D[i][j] sits in lane j + 32 ((i / 4) % 2),
// register (i % 4) + 4 (i / 8).  The 32 x 32 x 16 bf16 form takes 8
// consecutive k per lane (k = 8 (l / 32) .. +7) with the same result layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "d2mi_internal.h"
#include "transport.h"

namespace {

constexpr int TM = 128, TN = 128, KC = 32, NTH = 256;
constexpr int LDA = KC + 4;  // LDS row stride (floats): rows start on different banks
constexpr float RECOMPUTE = 1e-2f;  // recompute when T < RECOMPUTE * (|x|^2 + |y|^2)

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// FP32 squared row norms
__global__ void k_row_norms(const float *__restrict__ x, int P, int d, float *__restrict__ nx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float *r = x + (size_t)p * d;
  float s = 0.f;
  for (int t = 0; t < d; ++t) s += r[t] * r[t];
  nx[p] = s;
}

// VEC (d % 4 == 0, so rows start 16-byte aligned): each thread fetches its
// share of the next 32-wide slice of the shared dimension as float4 into
// registers right after the barrier, so the loads are in flight while the
// matrix cores work on the current slice; otherwise the slice is fetched
// element by element before each step.
template <bool BF16, bool VEC>
__global__ __launch_bounds__(NTH) void k_pdist2_mfma(const float *__restrict__ x,
                                                     const float *__restrict__ y, int P, int Q,
                                                     int d, const float *__restrict__ nx,
                                                     const float *__restrict__ ny,
                                                     float *__restrict__ T) {
  __shared__ __align__(16) float Xs[TM * LDA];
  __shared__ __align__(16) float Ys[TN * LDA];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  // XCD-aware order: consecutive workgroups land on different XCDs, so give
  // each XCD a contiguous band of row tiles (they share the y tiles in L2)
  const int ntm = (P + TM - 1) / TM, ntn = (Q + TN - 1) / TN, nt = ntm * ntn;
  int b;
  {  // bijection: XCD x (= blockIdx % 8) takes tiles [off(x), off(x) + count(x))
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8, q = nt / 8, r = nt % 8;
    b = xcd * q + min(xcd, r) + slot;
  }
  const int tm = b / ntn, tn = b - (b / ntn) * ntn;
  const int p0 = tm * TM, q0 = tn * TN;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;

  // VEC: 4 float4 of x and 4 of y per thread per slice; group g = tid + i NTH
  // is row g / 8, columns 4 (g % 8) .. +3 of the slice
  constexpr int NV = TM * KC / 4 / NTH;
  float4 px[VEC ? NV : 1], py[VEC ? NV : 1];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int g = tid + i * NTH, r = g >> 3, kk = k0 + (g & 7) * 4;
      const int p = p0 + r, q = q0 + r;
      px[i] = (p < P && kk < d) ? *(const float4 *)(x + (size_t)p * d + kk) : float4{0.f, 0.f, 0.f, 0.f};
      py[i] = (q < Q && kk < d) ? *(const float4 *)(y + (size_t)q * d + kk) : float4{0.f, 0.f, 0.f, 0.f};
    }
  };
  if constexpr (VEC) fetch(0);
  for (int k0 = 0; k0 < d; k0 += KC) {
    if constexpr (VEC) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        const int g = tid + i * NTH, r = g >> 3, k = (g & 7) * 4;
        *(float4 *)(Xs + r * LDA + k) = px[i];
        *(float4 *)(Ys + r * LDA + k) = py[i];
      }
    } else {
      // stage: each thread loads 16 floats of x and 16 of y (rows of 32 k)
      for (int e = tid; e < TM * KC; e += NTH) {
        const int r = e / KC, k = e - (e / KC) * KC;
        const int p = p0 + r, kk = k0 + k;
        Xs[r * LDA + k] = (p < P && kk < d) ? x[(size_t)p * d + kk] : 0.f;
      }
      for (int e = tid; e < TN * KC; e += NTH) {
        const int r = e / KC, k = e - (e / KC) * KC;
        const int q = q0 + r, kk = k0 + k;
        Ys[r * LDA + k] = (q < Q && kk < d) ? y[(size_t)q * d + kk] : 0.f;
      }
    }
    __syncthreads();
    if constexpr (VEC)
      if (k0 + KC < d) fetch(k0 + KC);  // lands while the matrix cores run
    if constexpr (!BF16) {
#pragma unroll
      for (int ks = 0; ks < KC; ks += 2) {
        const int kk = ks + (lane >> 5);
        float av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = Xs[(wr + a * 32 + (lane & 31)) * LDA + kk];
#pragma unroll
        for (int c = 0; c < 2; ++c) bv[c] = Ys[(wc + c * 32 + (lane & 31)) * LDA + kk];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a], bv[c], acc[a][c], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < KC; ks += 16) {
        const int kb = ks + 8 * (lane >> 5);
        v8bf av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const float *src = Xs + (wr + a * 32 + (lane & 31)) * LDA + kb;
#pragma unroll
          for (int u = 0; u < 8; ++u) av[a][u] = (__bf16)src[u];
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const float *src = Ys + (wc + c * 32 + (lane & 31)) * LDA + kb;
#pragma unroll
          for (int u = 0; u < 8; ++u) bv[c][u] = (__bf16)src[u];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < 2; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[a], bv[c], acc[a][c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue: norms, cancellation fix, store (32 consecutive columns per row)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2);
        const int p = p0 + wr + a * 32 + i, q = q0 + wc + c * 32 + (lane & 31);
        if (p < P && q < Q) {
          const float s = nx[p] + ny[q];
          float v = s - 2.f * acc[a][c][r];
          if (v < RECOMPUTE * s) {
            const float *xp = x + (size_t)p * d, *yq = y + (size_t)q * d;
            float t = 0.f;
            for (int k = 0; k < d; ++k) {
              const float u = xp[k] - yq[k];
              t += u * u;
            }
            v = t;
          }
          T[(size_t)p * Q + q] = v < 0.f ? 0.f : v;
        }
      }
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_to_f64(const float *__restrict__ a, double *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

}  // namespace

// T (P x Q, FP32, row-major) for FP32 rows x (P x d) and y (Q x d) on the
// device; nx, ny: FP32 scratch of P and Q floats.
int d2mi_pdist2_mfma_dev(hipStream_t S, int P, int Q, int d, const float *x, const float *y,
                         float *nx, float *ny, float *T, int bf16) {
  if (P <= 0 || Q <= 0 || d <= 0) return D2MI_OK;
  k_row_norms<<<(P + 255) / 256, 256, 0, S>>>(x, P, d, nx);
  k_row_norms<<<(Q + 255) / 256, 256, 0, S>>>(y, Q, d, ny);
  const size_t tiles = (size_t)((P + TM - 1) / TM) * ((Q + TN - 1) / TN);
  if (tiles > 0x7fffffff) {
    d2mi_set_error("pdist2 table of %d x %d too large", P, Q);
    return D2MI_EUNSUPPORTED;
  }
  const bool vec = d % 4 == 0 && ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0);
  if (bf16 && vec) k_pdist2_mfma<true, true><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  else if (bf16) k_pdist2_mfma<true, false><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  else if (vec) k_pdist2_mfma<false, true><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  else k_pdist2_mfma<false, false><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

extern "C" int d2mi_pdist2_mfma(d2mi_ctx *ctx, int d, size_t n, size_t m, const double *A,
                                const double *B, double *C, int bf16, float *ms) {
  D2MI_TRY(d2mi_check_ctx(ctx));
  if (d <= 0 || n == 0 || m == 0 || !A || !B || !C || n > 0x7fffffff || m > 0x7fffffff) {
    d2mi_set_error("d2mi_pdist2_mfma: bad arguments");
    return D2MI_EINVAL;
  }
  hipStream_t S = ctx->stream;
  DevMem a64, b64, a32, b32, na, nb, t32, c64;
  D2MI_TRY(a64.assign(A, sizeof(double) * n * d, S));
  D2MI_TRY(b64.assign(B, sizeof(double) * m * d, S));
  D2MI_TRY(a32.ensure(sizeof(float) * n * d));
  D2MI_TRY(b32.ensure(sizeof(float) * m * d));
  D2MI_TRY(na.ensure(sizeof(float) * n));
  D2MI_TRY(nb.ensure(sizeof(float) * m));
  D2MI_TRY(t32.ensure(sizeof(float) * n * m));
  D2MI_TRY(c64.ensure(sizeof(double) * n * m));
  k_to_f32<<<1024, 256, 0, S>>>(a64.as<double>(), a32.as<float>(), n * d);
  k_to_f32<<<1024, 256, 0, S>>>(b64.as<double>(), b32.as<float>(), m * d);
  D2MI_HIP(hipEventRecord(ctx->ev[0], S));
  // the reference's layout C[i*n + j] (B rows i, A rows j): rows of T are B
  D2MI_TRY(d2mi_pdist2_mfma_dev(S, (int)m, (int)n, d, b32.as<float>(), a32.as<float>(),
                                nb.as<float>(), na.as<float>(), t32.as<float>(), bf16));
  D2MI_HIP(hipEventRecord(ctx->ev[1], S));
  k_to_f64<<<1024, 256, 0, S>>>(t32.as<float>(), c64.as<double>(), n * m);
  D2MI_HIP(hipMemcpyAsync(C, c64.p, sizeof(double) * n * m, hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (ms) D2MI_HIP(hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
  return D2MI_OK;
}
