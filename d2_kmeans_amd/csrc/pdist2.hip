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
//
// BF16 products: both operands are converted to BF16 once, in HBM, rows
// zero-padded to a multiple of 32 (the vocabulary once per phase, the
// centroid supports once per support update), beside the FP32 norms of the
// unrounded rows: T = |x|^2 + |y|^2 - 2 bf16(x).bf16(y) keeps the rounding
// errors zero-mean (the distance of the rounded rows would be biased upward
// by about 2^-16 of |x|^2 + |y|^2, visible in the clustering objective).
// Wide tables (k_pdist2_bf16_dma, 128 x 256 tiles) copy the BF16 slices
// straight from HBM into three LDS stages, two slices in flight while the
// matrix cores work on the third; narrow ones (k_pdist2_bf16, 128 x 128)
// stage them through registers into two LDS buffers.  Each MFMA operand is
// one 128-bit LDS read.  With the products 16 times faster than FP32 ones,
// the 4-byte result table is what bounds these kernels: they are HBM-bound
// on their output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "d2mi_internal.h"
#include "lds_copy.h"
#include "transport.h"

namespace {

constexpr int TM = 128, TN = 128, KC = 32, NTH = 256;
constexpr int LDA = KC + 4;  // LDS row stride (floats): rows start on different banks
constexpr float RECOMPUTE = 1e-2f;  // recompute when T < RECOMPUTE * (|x|^2 + |y|^2)

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));


// Sum over the 64 lanes of a wave (every lane gets it).
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// FP32 squared row norms: one wave per row, coalesced loads.
__global__ __launch_bounds__(256) void k_row_norms(const float *__restrict__ x, int P, int d,
                                                   float *__restrict__ nx) {
  const int lane = threadIdx.x & 63, p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= P) return;
  const float *r = x + (size_t)p * d;
  float s = 0.f;
  for (int t = lane; t < d; t += 64) s += r[t] * r[t];
  s = wave_sum(s);
  if (lane == 0) nx[p] = s;
}

// The tile's row norms in LDS (nxs: the 128 rows of x, nys: of y), read
// once per workgroup; visible after the main loop's first barrier.
__device__ inline void load_norms(const float *nx, const float *ny, int P, int Q, int p0, int q0,
                                  float *nxs, float *nys) {
  const int t = threadIdx.x;
  if (t < TM) nxs[t] = p0 + t < P ? nx[p0 + t] : 0.f;
  else if (t < TM + TN) nys[t - TM] = q0 + t - TM < Q ? ny[q0 + t - TM] : 0.f;
}

// T = |x|^2 + |y|^2 - 2 acc for one wave's 64 x 64 quarter, stored 32
// consecutive columns per row.  A lane's accumulator registers r = 4 rq + r3
// (r3 = 0..3) are rows 4 (lane / 32) + 8 rq + r3 of its 32-row block: four
// consecutive rows, whose norms come in one 16-byte LDS read, tested with
// one ballot.  Entries small next to |x|^2 + |y|^2 are recomputed by direct
// differences from the FP32 rows, each by the whole wave (coalesced loads
// over the shared dimension, a shuffle sum), so a wave that holds one pays a
// few loads rather than a d-step serial loop.
__device__ inline void epilogue(const v16f (&acc)[2][2], const float *nxs, const float *nys,
                                const float *__restrict__ x, const float *__restrict__ y, int P,
                                int Q, int d, int p0, int q0, int wr, int wc, int lane,
                                float *__restrict__ T) {
  const int hi = lane >> 5, lo = lane & 31;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int qc = wc + c * 32 + lo, q = q0 + qc;
    const float nyq = nys[qc];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int rq = 0; rq < 4; ++rq) {
        const int pr = wr + a * 32 + 4 * hi + 8 * rq;  // rows pr .. pr + 3
        const float4 nx4 = *(const float4 *)&nxs[pr];
        const float nxr[4] = {nx4.x, nx4.y, nx4.z, nx4.w};
        float v[4];
        unsigned int flags = 0;  // bit r3: row pr + r3 needs the direct recompute
#pragma unroll
        for (int r3 = 0; r3 < 4; ++r3) {
          const float s = nxr[r3] + nyq;
          v[r3] = s - 2.f * acc[a][c][4 * rq + r3];
          if (p0 + pr + r3 < P && q < Q && v[r3] < RECOMPUTE * s) flags |= 1u << r3;
        }
        if (__ballot(flags != 0)) {
#pragma unroll
          for (int r3 = 0; r3 < 4; ++r3) {
            unsigned long long need = __ballot((flags >> r3) & 1u);
            while (need) {
              const int l = __ffsll((long long)need) - 1;
              need &= need - 1;
              const int pl = __shfl(p0 + pr + r3, l), ql = __shfl(q, l);
              const float *xp = x + (size_t)pl * d, *yq = y + (size_t)ql * d;
              float t = 0.f;
              for (int k = lane; k < d; k += 64) {
                const float u = xp[k] - yq[k];
                t += u * u;
              }
              t = wave_sum(t);
              if (lane == l) v[r3] = t;
            }
          }
        }
#pragma unroll
        for (int r3 = 0; r3 < 4; ++r3) {
          const int p = p0 + pr + r3;
          if (p < P && q < Q) {
            __builtin_nontemporal_store(v[r3] < 0.f ? 0.f : v[r3], &T[(size_t)p * Q + q]);
          }
        }
      }
  }
}

// VEC (d % 4 == 0, so rows start 16-byte aligned): each thread fetches its
// share of the next 32-wide slice of the shared dimension as float4 into
// registers right after the barrier, so the loads are in flight while the
// matrix cores work on the current slice; otherwise the slice is fetched
// element by element before each step.
template <bool VEC>
__global__ __launch_bounds__(NTH, 3) void k_pdist2_mfma(const float *__restrict__ x,
                                                     const float *__restrict__ y, int P, int Q,
                                                     int d, const float *__restrict__ nx,
                                                     const float *__restrict__ ny,
                                                     float *__restrict__ T) {
  __shared__ __align__(16) float Xs[TM * LDA];
  __shared__ __align__(16) float Ys[TN * LDA];
  __shared__ __align__(16) float nxs[TM], nys[TN];
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
  load_norms(nx, ny, P, Q, p0, q0, nxs, nys);
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
    __syncthreads();
  }
  // epilogue: norms, cancellation fix, store (32 consecutive columns per row)
  epilogue(acc, nxs, nys, x, y, P, Q, d, p0, q0, wr, wc, lane, T);
}

// BF16 rows, zero-padded to dp, and FP32 squared norms of the FP32 rows:
// one wave per row, coalesced loads, norms summed by shuffles.
__global__ __launch_bounds__(256) void k_rows_to_bf16(const float *__restrict__ x, int R, int d,
                                                      int dp, __bf16 *__restrict__ xb,
                                                      float *__restrict__ nx) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const float *src = x + (size_t)r * d;
  __bf16 *dst = xb + (size_t)r * dp;
  float acc = 0.f;
  for (int k = lane; k < dp; k += 64) {
    const float f = k < d ? src[k] : 0.f;
    dst[k] = (__bf16)f;
    acc += f * f;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) nx[r] = acc;
}

constexpr int BK = 32;       // shared-dimension slice (bf16 elements)
constexpr int BLD = BK + 8;  // LDS row stride: 80 B, 16-byte aligned, rows rotate banks

// T (P x Q) from BF16 rows xb (P x dp) and yb (Q x dp) with norms nx, ny;
// x, y: the FP32 rows (d wide) for the cancellation recompute.  WN waves
// side by side along Q (two along P): a 128 x 64 WN tile per workgroup.
template <int WN>
__global__ __launch_bounds__(128 * WN, WN == 2 ? 3 : 4) void k_pdist2_bf16(
    const __bf16 *__restrict__ xb, const __bf16 *__restrict__ yb, int P, int Q, int dp,
    const float *__restrict__ nx, const float *__restrict__ ny, const float *__restrict__ x,
    const float *__restrict__ y, int d, float *__restrict__ T) {
  constexpr int NT = 128 * WN, TQ = 64 * WN;
  constexpr int NX = TM * 4 / NT, NY = TQ * 4 / NT;  // 16-byte chunks per thread per slice
  __shared__ __align__(16) __bf16 Xs[2][TM * BLD];
  __shared__ __align__(16) __bf16 Ys[2][TQ * BLD];
  __shared__ __align__(16) float nxs[TM], nys[TQ];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ntm = (P + TM - 1) / TM, ntn = (Q + TQ - 1) / TQ, nt = ntm * ntn;
  int b;
  {  // XCD x (= blockIdx % 8) takes a contiguous band of tiles (shared rows in its L2)
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8, q = nt / 8, r = nt % 8;
    b = xcd * q + min(xcd, r) + slot;
  }
  const int tm = b / ntn, tn = b - (b / ntn) * ntn;
  const int p0 = tm * TM, q0 = tn * TQ;
  const int wr = (w / WN) * 64, wc = (w % WN) * 64;
  if (tid < TM) nxs[tid] = p0 + tid < P ? nx[p0 + tid] : 0.f;
  for (int t = tid; t < TQ; t += NT) nys[t] = q0 + t < Q ? ny[q0 + t] : 0.f;
  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;
  // staging: per slice, rows x 32 bf16 = 4 chunks of 16 B per row; chunk
  // g = tid + i NT is row g / 4, bf16 8 (g % 4) .. +7
  uint4 px[NX], py[NY];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int g = tid + i * NT, r = g >> 2, c = (g & 3) * 8, p = p0 + r;
      px[i] = p < P ? *(const uint4 *)(xb + (size_t)p * dp + k0 + c) : uint4{0, 0, 0, 0};
    }
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      const int g = tid + i * NT, r = g >> 2, c = (g & 3) * 8, q = q0 + r;
      py[i] = q < Q ? *(const uint4 *)(yb + (size_t)q * dp + k0 + c) : uint4{0, 0, 0, 0};
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      const int g = tid + i * NT;
      *(uint4 *)(&Xs[buf][(g >> 2) * BLD + (g & 3) * 8]) = px[i];
    }
#pragma unroll
    for (int i = 0; i < NY; ++i) {
      const int g = tid + i * NT;
      *(uint4 *)(&Ys[buf][(g >> 2) * BLD + (g & 3) * 8]) = py[i];
    }
  };
  const int ns = dp / BK;
  fetch(0);
  stage(0);
  __syncthreads();
  for (int sl = 0; sl < ns; ++sl) {
    const int buf = sl & 1;
    if (sl + 1 < ns) fetch((sl + 1) * BK);  // in flight during the products
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kb = ks + 8 * (lane >> 5);
      v8bf av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        av[a] = *(const v8bf *)(&Xs[buf][(wr + a * 32 + (lane & 31)) * BLD + kb]);
#pragma unroll
      for (int c = 0; c < 2; ++c)
        bv[c] = *(const v8bf *)(&Ys[buf][(wc + c * 32 + (lane & 31)) * BLD + kb]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
    // the other buffer was last read before the previous barrier
    if (sl + 1 < ns) stage(buf ^ 1);
    __syncthreads();
  }
  epilogue(acc, nxs, nys, x, y, P, Q, d, p0, q0, wr, wc, lane, T);
}

// The same GEMM with the operand slices copied from HBM straight into LDS
// (global_load_lds_dwordx4: no staging registers), three LDS stages deep:
// slices sl + 1 and sl + 2 are in flight while the matrix cores work on
// slice sl.  A stage holds the tile's 128 x rows and 64 WN y rows, 32 bf16
// (64 B) each, unpadded; one wave instruction fills 16 rows (1 KB, lane l
// writing bytes 16 l .. 16 l + 15).  Within a row the four 16-byte chunks
// are stored rotated, chunk c at position c ^ ((row / 4) % 4), so the 16
// rows read by one quarter-wave of an MFMA operand load fall on 16
// different bank groups.  Rows past P or Q load row P - 1 or Q - 1 (their
// results are never stored).
constexpr int DMA_STAGES = 3;
// The slices are copied with lds_copy16 (lds_copy.h).
template <int WN>
__global__ __launch_bounds__(128 * WN, 4) void k_pdist2_bf16_dma(
    const __bf16 *__restrict__ xb, const __bf16 *__restrict__ yb, int P, int Q, int dp,
    const float *__restrict__ nx, const float *__restrict__ ny, const float *__restrict__ x,
    const float *__restrict__ y, int d, float *__restrict__ T) {
  constexpr int NT = 128 * WN, TQ = 64 * WN, NW = NT / 64;
  constexpr int XB = TM / 16, YB = TQ / 16, NB = XB + YB;  // 1 KB blocks per stage
  constexpr int BPW = NB / NW;                             // blocks per wave per stage
  static_assert(NB % NW == 0, "whole blocks per wave");
  // one array per stage: the compiler then sees that a stage's reads do not
  // alias the copies in flight into the other two, and waits for nothing
  // but its own stage
  static_assert(DMA_STAGES == 3, "one array per stage below");
  __shared__ __align__(16) __bf16 St0[(TM + TQ) * BK], St1[(TM + TQ) * BK], St2[(TM + TQ) * BK];
  __shared__ __align__(16) float nxs[TM], nys[TQ];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ntm = (P + TM - 1) / TM, ntn = (Q + TQ - 1) / TQ, nt = ntm * ntn;
  int b;
  {  // XCD x (= blockIdx % 8) takes a contiguous band of tiles (shared rows in its L2)
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8, q = nt / 8, r = nt % 8;
    b = xcd * q + min(xcd, r) + slot;
  }
  const int tm = b / ntn, tn = b - (b / ntn) * ntn;
  const int p0 = tm * TM, q0 = tn * TQ;
  const int wr = (w / WN) * 64, wc = (w % WN) * 64;
  if (tid < TM) nxs[tid] = p0 + tid < P ? nx[p0 + tid] : 0.f;
  for (int t = tid; t < TQ; t += NT) nys[t] = q0 + t < Q ? ny[q0 + t] : 0.f;
  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][c][r] = 0.f;
  auto stage_of = [&](auto st_c) -> __bf16 * {
    constexpr int st = decltype(st_c)::value;
    return st == 0 ? St0 : st == 1 ? St1 : St2;
  };
  // this lane's row within a block and the chunk it fetches for its position
  const int lr = lane >> 2, chunk = (lane & 3) ^ ((lane >> 4) & 3);
  auto issue = [&](auto st_c, int sl) {
    __bf16 *St = stage_of(st_c);
    const int k0 = sl * BK + chunk * 8;
#pragma unroll
    for (int i = 0; i < BPW; ++i) {
      const int blk = w + i * NW;  // wave-uniform
      const __bf16 *src;
      if (blk < XB) src = xb + (size_t)min(p0 + blk * 16 + lr, P - 1) * dp + k0;
      else src = yb + (size_t)min(q0 + (blk - XB) * 16 + lr, Q - 1) * dp + k0;
      lds_copy16(src, &St[blk * 16 * BK]);
    }
  };
  auto products = [&](auto st_c) {
    const __bf16 *Xs = stage_of(st_c), *Ys = stage_of(st_c) + TM * BK;
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kc = (ks >> 3) + (lane >> 5);  // this lane's 16-byte chunk of the slice
      v8bf av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int r = wr + a * 32 + (lane & 31);
        av[a] = *(const v8bf *)(&Xs[r * BK + ((kc ^ ((r >> 2) & 3)) << 3)]);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int r = wc + c * 32 + (lane & 31);
        bv[c] = *(const v8bf *)(&Ys[r * BK + ((kc ^ ((r >> 2) & 3)) << 3)]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
  };
  using C0 = std::integral_constant<int, 0>;
  using C1 = std::integral_constant<int, 1>;
  using C2 = std::integral_constant<int, 2>;
  const int ns = dp / BK;
  // slice sl in stage sl % 3
  auto step = [&](auto cur_c, auto nxt2_c, int sl) {
    // slice sl has landed once at most the next slice's BPW copies are
    // outstanding; then a barrier (which also covers the norms' LDS writes
    // and every wave's reads of the stage refilled below, done at sl - 1)
    asm volatile("" ::: "memory");
    if (sl + 1 < ns) __builtin_amdgcn_s_waitcnt(0x0070 | BPW);  // vmcnt(BPW) lgkmcnt(0)
    else __builtin_amdgcn_s_waitcnt(0x0070);                    // vmcnt(0) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (sl + 2 < ns) issue(nxt2_c, sl + 2);
    products(cur_c);
  };
  issue(C0{}, 0);
  if (ns > 1) issue(C1{}, 1);
  for (int sl = 0; sl < ns; sl += 3) {
    step(C0{}, C2{}, sl);
    if (sl + 1 < ns) step(C1{}, C0{}, sl + 1);
    if (sl + 2 < ns) step(C2{}, C1{}, sl + 2);
  }
  epilogue(acc, nxs, nys, x, y, P, Q, d, p0, q0, wr, wc, lane, T);
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
int d2mi_row_norms_dev(hipStream_t S, int R, int d, const float *x, float *nx) {
  if (R <= 0) return D2MI_OK;
  k_row_norms<<<(R + 3) / 4, 256, 0, S>>>(x, R, d, nx);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

// the GEMM alone, norms given
static int pdist2_f32_gemm(hipStream_t S, int P, int Q, int d, const float *x, const float *y,
                           const float *nx, const float *ny, float *T) {
  const size_t tiles = (size_t)((P + TM - 1) / TM) * ((Q + TN - 1) / TN);
  if (tiles > 0x7fffffff) {
    d2mi_set_error("pdist2 table of %d x %d too large", P, Q);
    return D2MI_EUNSUPPORTED;
  }
  const bool vec = d % 4 == 0 && ((uintptr_t)x % 16 == 0) && ((uintptr_t)y % 16 == 0);
  if (vec) k_pdist2_mfma<true><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  else k_pdist2_mfma<false><<<(unsigned)tiles, NTH, 0, S>>>(x, y, P, Q, d, nx, ny, T);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

int d2mi_pdist2_mfma_dev(hipStream_t S, int P, int Q, int d, const float *x, const float *y,
                         float *nx, float *ny, float *T) {
  if (P <= 0 || Q <= 0 || d <= 0) return D2MI_OK;
  D2MI_TRY(d2mi_row_norms_dev(S, P, d, x, nx));
  D2MI_TRY(d2mi_row_norms_dev(S, Q, d, y, ny));
  return pdist2_f32_gemm(S, P, Q, d, x, y, nx, ny, T);
}

int d2mi_pdist2_pad(int d) { return (d + BK - 1) / BK * BK; }

int d2mi_rows_to_bf16_dev(hipStream_t S, int R, int d, const float *x, __bf16 *xb, float *nx) {
  if (R <= 0) return D2MI_OK;
  k_rows_to_bf16<<<(R + 3) / 4, 256, 0, S>>>(x, R, d, d2mi_pdist2_pad(d), xb, nx);
  D2MI_HIP(hipGetLastError());
  return D2MI_OK;
}

int d2mi_pdist2_bf16_dev(hipStream_t S, int P, int Q, int d, const __bf16 *xb, const float *nx,
                         const __bf16 *yb, const float *ny, const float *x, const float *y,
                         float *T) {
  if (P <= 0 || Q <= 0 || d <= 0) return D2MI_OK;
  const size_t tiles = (size_t)((P + TM - 1) / TM) * ((Q + TN - 1) / TN);
  if (tiles > 0x7fffffff) {
    d2mi_set_error("pdist2 table of %d x %d too large", P, Q);
    return D2MI_EUNSUPPORTED;
  }
  // wide tables: 128 x 256 tiles (each x slice staged once serves twice the
  // columns), operands copied into LDS three slices deep; narrow ones:
  // 128 x 128 with register staging
  if (Q >= 192) {
    const size_t t4 = (size_t)((P + TM - 1) / TM) * ((Q + 255) / 256);
    k_pdist2_bf16_dma<4><<<(unsigned)t4, 512, 0, S>>>(xb, yb, P, Q, d2mi_pdist2_pad(d), nx, ny, x,
                                                       y, d, T);
  } else {
    k_pdist2_bf16<2><<<(unsigned)tiles, 256, 0, S>>>(xb, yb, P, Q, d2mi_pdist2_pad(d), nx, ny, x,
                                                      y, d, T);
  }
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
  DevMem a64, b64, a32, b32, na, nb, t32, c64, a16, b16;
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
  // the reference's layout C[i*n + j] (B rows i, A rows j): rows of T are B
  if (bf16) {
    const size_t dp = d2mi_pdist2_pad(d);
    D2MI_TRY(a16.ensure(2 * n * dp));
    D2MI_TRY(b16.ensure(2 * m * dp));
    D2MI_TRY(d2mi_rows_to_bf16_dev(S, (int)n, d, a32.as<float>(), a16.as<__bf16>(), na.as<float>()));
    D2MI_TRY(d2mi_rows_to_bf16_dev(S, (int)m, d, b32.as<float>(), b16.as<__bf16>(), nb.as<float>()));
    D2MI_HIP(hipEventRecord(ctx->ev[0], S));
    D2MI_TRY(d2mi_pdist2_bf16_dev(S, (int)m, (int)n, d, b16.as<__bf16>(), nb.as<float>(),
                                  a16.as<__bf16>(), na.as<float>(), b32.as<float>(),
                                  a32.as<float>(), t32.as<float>()));
  } else {
    D2MI_TRY(d2mi_row_norms_dev(S, (int)n, d, a32.as<float>(), na.as<float>()));
    D2MI_TRY(d2mi_row_norms_dev(S, (int)m, d, b32.as<float>(), nb.as<float>()));
    D2MI_HIP(hipEventRecord(ctx->ev[0], S));
    D2MI_TRY(pdist2_f32_gemm(S, (int)m, (int)n, d, b32.as<float>(), a32.as<float>(),
                             nb.as<float>(), na.as<float>(), t32.as<float>()));
  }
  D2MI_HIP(hipEventRecord(ctx->ev[1], S));
  k_to_f64<<<1024, 256, 0, S>>>(t32.as<float>(), c64.as<double>(), n * m);
  D2MI_HIP(hipMemcpyAsync(C, c64.p, sizeof(double) * n * m, hipMemcpyDeviceToHost, S));
  D2MI_HIP(hipStreamSynchronize(S));
  if (ms) D2MI_HIP(hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
  return D2MI_OK;
}
