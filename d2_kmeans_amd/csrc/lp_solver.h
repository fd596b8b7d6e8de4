// Exact transport LP, one workgroup per problem (device code shared by the
// batched-distance kernels and the assignment server).
//
// The LP is the one the reference hands to Mosek for every distance
// (d2_match_by_distmat, solver_mosek.cc:61-143):
//     min sum_{k,j} C[k + j*n] x_kj,  x >= 0,
//     sum_j x_kj = wA[k] (all n rows),  sum_k x_kj = wB[j] (j < m-1; the last
//     column constraint is dropped, so column m-1 absorbs the remainder).
//
// Transportation simplex in FP64 on a spanning-tree basis (parent pointers,
// root = row 0), run by all waves of the workgroup:
//   * start: Vogel's approximation (wave 0; about half the pivots of the
//     north-west corner), rooted at row 0; north-west corner otherwise;
//   * potentials, depths and the binary-lifting table lvl[b][v] (the 2^b-th
//     ancestor of v) come out of one pointer-jumping pass per pivot: each
//     round doubles the reach of every node, so a tree of depth D costs
//     ceil(log2 D)+1 rounds;
//   * pricing: block search (the next block of T*cpt cells after the last
//     entering cell; the most negative reduced cost in the first block that
//     has one, lowest cell on ties); after 50*(n+m)+1000 pivots, Bland's rule
//     (lowest improving cell, lowest leaving cell on ties), which cannot
//     cycle.  With the same block size the pivot sequence, and so the value,
//     does not depend on the workgroup size;
//   * pivot: every wave finds the apex of the cycle by testing 64 depths at a
//     time against the lifting table; path nodes, the ratio test, the flow
//     update and the re-hanging of the cut subtree (parent pointers reversed
//     along the path) are then parallel over path positions.
// Costs: an FP32 copy of the n*m cost matrix sits in LDS when it fits and
// drives the pivots.  At optimality for the FP32 costs the solver switches to
// the exact FP64 costs: potentials are recomputed from them and every cell
// whose FP32 reduced cost is within the table's error bound of zero is priced
// exactly; pivoting continues on exact costs until no cell improves.  The
// returned value sums exact FP64 costs over the final basis, so it does not
// depend on the FP32 table.
#pragma once

#include <cfloat>
#include <climits>
#include <cmath>
#include <type_traits>

namespace lp {

constexpr int WAVE = 64;
constexpr int MAX_WAVES = 16;
// Vogel start: rows to re-walk after a column is crossed out are walked by
// groups of 64 / cnt lanes (2 to 16) when there are cnt <= 32 of them, by one
// lane each beyond that, and one long row alone by the whole wave
// pointer jumping by wave 0 alone for trees of at most this many nodes
#ifndef D2MI_LP_SOLO_NODES
#define D2MI_LP_SOLO_NODES 256
#endif
// Vogel's start on reduced costs: approximate duals from entropic
// (Sinkhorn) iterations in the log domain, the temperature lowered
// geometrically from eps0 to eps1 times the cost scale, for tables of at
// least DUAL_MIN_CELLS cells.  The schedule depends on the table size
// (measured on MI355X on two LP families: 40 x 20-80 image-like LPs of the
// clustering loop, and 48 x 50-500 assignment LPs in d = 64):
//  - below 4096 cells, 16 iterations from 0.1 to 0.002 of the largest cost.
//    Lower final temperatures need more iterations than the pivots they
//    save are worth, and a cooler first temperature triples the pivots of
//    the image-like LPs, whose costs have many near-ties;
//  - from 4096 cells on, the scale is the largest cost after row, then
//    column minima are subtracted (the spread the LP sees), and the
//    temperature goes from 0.1 to 2e-5 of it in 48 iterations (64 from
//    16384 cells on): a 48 x 500 LP takes 260-310 pivots instead of 1140
//    and 4.3 ms instead of 11.9.
struct DualSchedule {
  int iters;
  float eps0, eps1;
  int reduce;
};
__host__ __device__ inline DualSchedule dual_schedule(long long cells) {
  if (cells < 4096) return {16, 0.1f, 0.002f, 0};
  return {cells >= 16384 ? 64 : 48, 0.1f, 2e-5f, 1};
}
constexpr int DUAL_MIN_CELLS = 512;
// rows of at most DUAL_ROW_THREAD columns: one thread per row in the row
// pass; longer rows: segments of DUAL_SEG columns, one thread each
constexpr int DUAL_ROW_THREAD = 128, DUAL_SEG = 64;
constexpr int JB = 4;        // nodes per thread per batch in the pointer-jumping rounds
constexpr int CPT_MAX = 32;  // cells per thread per pricing block (unrolled)

enum CostMode {
  COST_EUCLID = 0,  // squared Euclidean distance of supports (A rows, B columns), FP64
  COST_HIST = 1,    // dense histogram ground cost dist[k + j*V]
  COST_SPARSE = 2,  // sparse histogram: dist[ids[j]*V + k]
  COST_GLOBAL = 3,  // precomputed FP64 costs g[k + j*n]
  COST_WORD = 4,    // word embeddings: A row k against vocab row ids[j] (B = vocab)
};

enum TableMode {
  TAB_NONE = 0,      // price every cell with the exact cost
  TAB_EXACT = 1,     // LDS table = FP32 rounding of the exact cost
  TAB_GLOBAL32 = 2,  // LDS table copied from a precomputed FP32 table gt[k + j*n]
};

__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline int ceil_log2(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

// LDS layout for problems up to n_max x m_max; offsets in bytes.
struct Layout {
  size_t ct, A, B, wa, wb, ids, accA, accB, flow, par, depA, depB, lvl, path, ctl, total;
  int levels;
  __host__ __device__ __forceinline__ Layout(int n_max, int m_max, int dA, int dB, bool table, bool ids_on) {
    const size_t nn = (size_t)n_max + m_max;
    levels = ceil_log2((int)nn) + 2;
    size_t o = 0;
    ct = o; o = al16(o + (table ? sizeof(float) * (size_t)n_max * m_max : 0));
    A = o; o = al16(o + sizeof(double) * (size_t)n_max * dA);
    B = o; o = al16(o + sizeof(double) * (size_t)m_max * dB);
    wa = o; o = al16(o + sizeof(double) * n_max);
    wb = o; o = al16(o + sizeof(double) * m_max);
    ids = o; o = al16(o + (ids_on ? sizeof(int) * m_max : 0));
    accA = o; o = al16(o + sizeof(double) * nn);
    accB = o; o = al16(o + sizeof(double) * nn);
    flow = o; o = al16(o + sizeof(double) * nn);
    par = o; o = al16(o + sizeof(int) * nn);
    depA = o; o = al16(o + sizeof(int) * nn);
    depB = o; o = al16(o + sizeof(int) * nn);
    lvl = o; o = al16(o + sizeof(short) * nn * levels);
    path = o; o = al16(o + sizeof(int) * (nn + 1));
    ctl = o; o = al16(o + 1024);
    total = o;
  }
};

struct Cost {
  int mode, d, n, V;
  const double *A, *B;  // Euclidean supports: A row k at A + k*d, B column j at B + j*d
  const int *ids;       // sparse-histogram bin ids (LDS)
  const double *__restrict__ dist;
  const double *__restrict__ g;
  __device__ __forceinline__ double operator()(int k, int j) const {
    switch (mode) {
      case COST_EUCLID: {
        const double *a = A + (size_t)k * d, *b = B + (size_t)j * d;
        double acc = 0.0;
        for (int t = 0; t < d; ++t) {
          const double u = a[t] - b[t];
          acc += u * u;
        }
        return acc;
      }
      case COST_WORD: {
        const double *a = A + (size_t)k * d, *b = B + (size_t)ids[j] * d;
        double acc = 0.0;
        for (int t = 0; t < d; ++t) {
          const double u = a[t] - b[t];
          acc += u * u;
        }
        return acc;
      }
      case COST_HIST:
        return dist[k + (size_t)j * V];
      case COST_SPARSE:
        return dist[(size_t)ids[j] * V + k];
      default:
        return g[k + (size_t)j * n];
    }
  }
};

// e = k + j*n  ->  (k, j) without an integer divide (e < 2^24)
__device__ __forceinline__ void split_cell(int e, int n, float inv_n, int &k, int &j) {
  int q = (int)((float)e * inv_n);
  int r = e - q * n;
  if (r < 0) { --q; r += n; }
  else if (r >= n) { ++q; r -= n; }
  k = r;
  j = q;
}

// Cross-lane moves for the selection reductions (min, max, argmin with an
// index tie-break, top-2): DPP, plain VALU moves, where __shfl_xor is an LDS
// round trip (ds_bpermute) per step.  Within each 16-lane row, rotations by
// 1, 2, 4, 8 leave every lane with the row's result; row_bcast:15 into rows 1
// and 3, then row_bcast:31 into rows 2 and 3, carry it to lane 63 (rows the
// row mask leaves out read `id`, the reduction's identity), which is read back
// as a scalar.  Selections are exact, so the order does not matter: results
// equal those of any other order.  All 64 lanes must be active.
enum { DPP_ROR1 = 0x121, DPP_ROR2 = 0x122, DPP_ROR4 = 0x124, DPP_ROR8 = 0x128,
       DPP_BC15 = 0x142, DPP_BC31 = 0x143,
       // lane groups: quad_perm [1,0,3,2] and [2,3,0,1] (lane xor 1, xor 2),
       // lane i <- 7 - i within each 8 lanes, lane i <- 15 - i within each 16
       DPP_QX1 = 0xB1, DPP_QX2 = 0x4E, DPP_HMIRROR = 0x141, DPP_MIRROR = 0x140 };
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_i(int v, int id = 0) {
  return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp_f(float v, float id = 0.f) {
  return __builtin_bit_cast(float, dpp_i<CTRL, ROWS>(__builtin_bit_cast(int, v),
                                                     __builtin_bit_cast(int, id)));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_d(double v, double id = 0.0) {
  typedef int i2 __attribute__((ext_vector_type(2)));
  const i2 u = __builtin_bit_cast(i2, v), w = __builtin_bit_cast(i2, id);
  i2 r;
  r.x = dpp_i<CTRL, ROWS>(u.x, w.x);
  r.y = dpp_i<CTRL, ROWS>(u.y, w.y);
  return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ int lane63_i(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ float lane63_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ double lane63_d(double v) {
  typedef int i2 __attribute__((ext_vector_type(2)));
  i2 u = __builtin_bit_cast(i2, v);
  u.x = __builtin_amdgcn_readlane(u.x, 63);
  u.y = __builtin_amdgcn_readlane(u.y, 63);
  return __builtin_bit_cast(double, u);
}

__device__ __forceinline__ double wave_min_d(double v) {
  v = fmin(v, dpp_d<DPP_ROR1>(v));
  v = fmin(v, dpp_d<DPP_ROR2>(v));
  v = fmin(v, dpp_d<DPP_ROR4>(v));
  v = fmin(v, dpp_d<DPP_ROR8>(v));
  v = fmin(v, dpp_d<DPP_BC15, 0xa>(v, DBL_MAX));
  v = fmin(v, dpp_d<DPP_BC31, 0xc>(v, DBL_MAX));
  return lane63_d(v);
}
// a sum is not a selection: this one keeps its butterfly order (values feed
// the supplies, and a different rounding could change ties in the start)
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  v = min(v, dpp_i<DPP_ROR1>(v));
  v = min(v, dpp_i<DPP_ROR2>(v));
  v = min(v, dpp_i<DPP_ROR4>(v));
  v = min(v, dpp_i<DPP_ROR8>(v));
  v = min(v, dpp_i<DPP_BC15, 0xa>(v, INT_MAX));
  v = min(v, dpp_i<DPP_BC31, 0xc>(v, INT_MAX));
  return lane63_i(v);
}
__device__ __forceinline__ double wave_max_d(double v) {
  v = fmax(v, dpp_d<DPP_ROR1>(v));
  v = fmax(v, dpp_d<DPP_ROR2>(v));
  v = fmax(v, dpp_d<DPP_ROR4>(v));
  v = fmax(v, dpp_d<DPP_ROR8>(v));
  v = fmax(v, dpp_d<DPP_BC15, 0xa>(v, -DBL_MAX));
  v = fmax(v, dpp_d<DPP_BC31, 0xc>(v, -DBL_MAX));
  return lane63_d(v);
}

// Control block in LDS: reduction slots (two banks, alternated per use so a
// slot is never rewritten before every wave has read it) and flags.
struct Ctl {
  double rv[2][MAX_WAVES];
  int ri[2][MAX_WAVES];
  int flag[2][MAX_WAVES];
  int jump_b;  // pointer jumping by one wave: 1 when its results are in the B buffers
};
static_assert(sizeof(Ctl) <= 1024, "the layout reserves 1024 B for the control block");

// Workgroup argmin of (v, i) lexicographically; every thread gets the result.
__device__ __forceinline__ void block_argmin(double &v, int &i, Ctl *ctl, int &bank, int nw) {
  const double wv = wave_min_d(v);
  const int wi = wave_min_i(v == wv ? i : INT_MAX);
  const int wid = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    ctl->rv[bank][wid] = wv;
    ctl->ri[bank][wid] = wi;
  }
  __syncthreads();
  double bv = ctl->rv[bank][0];
  int bi = ctl->ri[bank][0];
  for (int w = 1; w < nw; ++w) {
    const double x = ctl->rv[bank][w];
    const int y = ctl->ri[bank][w];
    if (x < bv || (x == bv && y < bi)) { bv = x; bi = y; }
  }
  bank ^= 1;
  v = bv;
  i = bi;
}

__device__ __forceinline__ double block_sum(double v, Ctl *ctl, int &bank, int nw) {
  v = wave_sum_d(v);
  const int wid = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) ctl->rv[bank][wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += ctl->rv[bank][w];
  bank ^= 1;
  return s;
}

__device__ __forceinline__ double block_max(double v, Ctl *ctl, int &bank, int nw) {
  v = wave_max_d(v);
  const int wid = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) ctl->rv[bank][wid] = v;
  __syncthreads();
  double s = ctl->rv[bank][0];
  for (int w = 1; w < nw; ++w) s = fmax(s, ctl->rv[bank][w]);
  bank ^= 1;
  return s;
}

__device__ __forceinline__ int block_or(int f, Ctl *ctl, int &bank, int nw) {
  f = __any(f) ? 1 : 0;
  const int wid = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) ctl->flag[bank][wid] = f;
  __syncthreads();
  int s = 0;
  for (int w = 0; w < nw; ++w) s |= ctl->flag[bank][w];
  bank ^= 1;
  return s;
}

struct Problem {
  int n, m;
  const double *wa, *wb;  // LDS
  Cost C;
  int table;              // TableMode
  const float *gt;        // TAB_GLOBAL32 source (global)
  float tab_rel;          // |table - exact| <= tab_rel * |table| + 1e-37
  int cpt;                // cells per thread per pricing block
  int vogel;              // start from Vogel's approximation
  int dual_iters;         // > 0: ... on reduced costs from this many entropic iterations
  float dual_eps0;        // first temperature, relative to the cost scale
  float dual_eps1;        // final temperature, relative to the cost scale
  int dual_reduce;        // cost scale: 0 largest cost; 1 largest cost after
                          // subtracting row, then column minima
  long long *prof;        // optional: PROF_N phase times of this LP (wall-clock ticks)
};

// Phases timed when Problem::prof is set (thread 0 reads the wall clock
// right after workgroup barriers; the pivot-loop phases are summed over
// pivots, including the exact-cost verification rounds).
enum ProfPhase {
  PROF_TABLE = 0,  // cost table in LDS and the cost scale
  PROF_DUALS,      // entropic dual iterations and the reduced-cost table
  PROF_START,      // Vogel's approximation (or the north-west corner)
  PROF_ROOT,       // exact table rebuilt, start cells rooted into a tree
  PROF_JUMP,       // pointer jumping: potentials, depths, lifting table
  PROF_PRICE,      // pricing blocks
  PROF_PIVOT,      // apex, ratio test, flows and re-hang
  PROF_VALUE,      // objective over the tree edges
  // inside PROF_START (timed by lane 0 of the wave that runs Vogel's steps)
  PROF_V_ARGMAX,   // the line with the largest penalty
  PROF_V_TAKE,     // its cell: quantity, crossing out
  PROF_V_SCAN,     // the lines whose two cheapest cells were crossed out
  PROF_V_WALK,     // their top-2 minima again
  PROF_N
};

// (value, index) order with the index breaking ties
__device__ __forceinline__ bool lt_vi(float a, int ia, float b, int ib) {
  return a < b || (a == b && ia < ib);
}

// two smallest (value, index) pairs; updated by value with selects only (a
// conditional write through a reference becomes a private-memory store)
struct Top2 {
  float v1, v2;
  int a1, a2;
};

__device__ __forceinline__ Top2 top2_insert(Top2 t, float c, int i) {
  const bool b1 = lt_vi(c, i, t.v1, t.a1), b2 = lt_vi(c, i, t.v2, t.a2);
  Top2 r;
  r.v1 = b1 ? c : t.v1;
  r.a1 = b1 ? i : t.a1;
  r.v2 = b1 ? t.v1 : (b2 ? c : t.v2);
  r.a2 = b1 ? t.a1 : (b2 ? i : t.a2);
  return r;
}

// insert for a walk in increasing index order: the new cell's index is above
// every index held, so it never wins a tie and (value, index) order reduces
// to value order (one compare per slot); empty slots hold +inf
__device__ __forceinline__ Top2 top2_push(Top2 t, float c, int i) {
  const bool b1 = c < t.v1, b2 = c < t.v2;
  Top2 r;
  r.v1 = b1 ? c : t.v1;
  r.a1 = b1 ? i : t.a1;
  r.v2 = b1 ? t.v1 : (b2 ? c : t.v2);
  r.a2 = b1 ? t.a1 : (b2 ? i : t.a2);
  return r;
}

// wave-wide merge, every lane gets the result (DPP as wave_min_d: each lane's
// pair is merged exactly once, and the two smallest (value, index) pairs of
// distinct cells do not depend on the merge order)
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ Top2 top2_merge_dpp(Top2 t) {
  Top2 o;
  o.v1 = dpp_f<CTRL, ROWS>(t.v1, FLT_MAX);
  o.v2 = dpp_f<CTRL, ROWS>(t.v2, FLT_MAX);
  o.a1 = dpp_i<CTRL, ROWS>(t.a1, INT_MAX);
  o.a2 = dpp_i<CTRL, ROWS>(t.a2, INT_MAX);
  return top2_insert(top2_insert(t, o.v1, o.a1), o.v2, o.a2);
}
__device__ __forceinline__ Top2 wave_top2(Top2 t) {
  t = top2_merge_dpp<DPP_ROR1>(t);
  t = top2_merge_dpp<DPP_ROR2>(t);
  t = top2_merge_dpp<DPP_ROR4>(t);
  t = top2_merge_dpp<DPP_ROR8>(t);
  t = top2_merge_dpp<DPP_BC15, 0xa>(t);
  t = top2_merge_dpp<DPP_BC31, 0xc>(t);
  Top2 r;
  r.v1 = lane63_f(t.v1);
  r.v2 = lane63_f(t.v2);
  r.a1 = lane63_i(t.a1);
  r.a2 = lane63_i(t.a2);
  return r;
}

// Vogel's approximation, run by wave 0 alone (no workgroup barrier): each of
// the n+m-1 steps takes the active line (row or column) whose two cheapest
// active cells differ most (lowest line on ties), allocates its cheapest
// cell as far as supply and demand allow and crosses out exactly one line,
// so the cells form a spanning tree.  Line minima are kept per line and
// recomputed only for lines whose two cheapest cells were crossed out.
//
// One wave runs every step, so the cost of a step is its dependency chain:
//   * the lines to redo after a step (those whose two cheapest cells the
//     crossed-out line held) are compacted into a list, and all cnt of them
//     are walked at once: by groups of 64 / cnt lanes each when cnt <= 32
//     (2 to 16 lanes per line, each lane a stride of the line with
//     independent running pairs, the group's pairs merged by DPP), else one
//     lane per line; a lone long row goes to the whole wave.  One VALU pass
//     then covers len / group cells instead of len: a lone wave issues each
//     instruction over 4 cycles, and the walks are most of its work;
//   * each line's penalty is kept in its own array, so the step's argmax
//     reads one value per line.
// The two cheapest (value, index) pairs of a set do not depend on the order
// in which the set is walked, so every form gives the same start.
// Output (LDS arrays idle at this point): cell t = (depA[t], depB[t]) with
// flow accB[t].  Scratch: accA = remaining supply / demand, flow = line
// minima (as float pairs), lvl = their cells and the active flags, par =
// the penalties, path = the list of lines to redo.
template <bool TABLE>
__device__ __attribute__((always_inline)) void vogel_start(const Problem &P, unsigned char *lds, const Layout &L,
                            const float *ct) {
  if (threadIdx.x >= WAVE) return;
  const int lane = threadIdx.x;
  const int n = P.n, m = P.m, nn = n + m;
  double *rem = (double *)(lds + L.accA);
  double *cq = (double *)(lds + L.accB);
  int *ci = (int *)(lds + L.depA), *cj = (int *)(lds + L.depB);
  float *mv1 = (float *)(lds + L.flow), *mv2 = mv1 + nn;
  short *ma1 = (short *)(lds + L.lvl), *ma2 = ma1 + nn, *act = ma1 + 2 * nn;
  float *pen = (float *)(lds + L.par);
  int *redo = (int *)(lds + L.path);
  // TABLE: the FP32 cost table in LDS; a compile-time choice, so the scans
  // below are straight-line code whose loads are all in flight at once
  auto cost = [&](int k, int j) -> float {
    if constexpr (TABLE) return ct[k + j * n];
    else return (float)P.C(k, j);
  };
  const Top2 none = {FLT_MAX, FLT_MAX, INT_MAX, INT_MAX};
  // penalty: gap between the two cheapest active cells (0 with one cell)
  auto store = [&](int v, const Top2 &t) {
    mv1[v] = t.v1;
    ma1[v] = (short)(t.a1 == INT_MAX ? -1 : t.a1);
    mv2[v] = t.v2;
    ma2[v] = (short)(t.a2 == INT_MAX ? -1 : t.a2);
    pen[v] = t.a2 == INT_MAX ? 0.0f : t.v2 - t.v1;
  };
  auto merge = [&](Top2 a, const Top2 &b) {
    return top2_insert(top2_insert(a, b.v1, b.a1), b.v2, b.a2);
  };
  // Scans load unconditionally (addresses stay in range) so the LDS reads
  // of consecutive cells overlap instead of waiting on each other.
  // row k over the active columns, by the whole wave
  auto row_top2_wave = [&](int k) {
    Top2 t = none;
    for (int j = lane; j < m; j += WAVE) {
      const float c = cost(k, j);
      const bool on = act[n + j] != 0;
      t = top2_insert(t, on ? c : FLT_MAX, on ? j : INT_MAX);
    }
    t = wave_top2(t);
    if (lane == 0) store(k, t);
  };
  // a line walked by this lane alone: VB cells per batch (every load of the
  // batch issued before any is used), NP independent running pairs; cells
  // priced on the fly (no table) are computed one at a time
  constexpr int VB = TABLE ? 16 : 1, NP = TABLE ? 4 : 1;
  // Each running pair sees its cells in increasing order (top2_push); an
  // inactive cell costs +inf and is never taken; empty slots read +inf until
  // the end, where they become FLT_MAX as everywhere else.  The result equals
  // top2_insert over the same cells in any order.
  auto walk = [&](auto cell_cost, auto cell_on, int len) -> Top2 {
    const float inf = __builtin_huge_valf();
    Top2 tp[NP];
#pragma unroll
    for (int r = 0; r < NP; ++r) tp[r] = Top2{inf, inf, INT_MAX, INT_MAX};
    int i = 0;
    for (; i + VB <= len; i += VB) {
      float c[VB];
      bool on[VB];
#pragma unroll
      for (int u = 0; u < VB; ++u) {
        c[u] = cell_cost(i + u);
        on[u] = cell_on(i + u);
      }
#pragma unroll
      for (int u = 0; u < VB; ++u)
        tp[u % NP] = top2_push(tp[u % NP], on[u] ? c[u] : inf, i + u);
    }
    for (; i < len; ++i) {
      const float c = cell_cost(i);
      const bool on = cell_on(i);
      tp[0] = top2_push(tp[0], on ? c : inf, i);
    }
#pragma unroll
    for (int r = 1; r < NP; ++r) tp[0] = merge(tp[0], tp[r]);
    if (tp[0].a1 == INT_MAX) tp[0].v1 = FLT_MAX;
    if (tp[0].a2 == INT_MAX) tp[0].v2 = FLT_MAX;
    return tp[0];
  };
  // row k over the active columns, by this lane alone
  auto row_top2_lane = [&](int k) {
    store(k, walk([&](int j) { return cost(k, j); }, [&](int j) { return act[n + j] != 0; }, m));
  };
  // column j over the active rows, by this lane alone
  auto col_top2 = [&](int j) {
    store(n + j, walk([&](int k) { return cost(k, j); }, [&](int k) { return act[k] != 0; }, n));
  };
  // the cnt lines of redo[] (rows, or columns when COL) at once, G lanes per
  // line (G * cnt <= 64): lane l walks cells q, q + G, ... (q = l % G) of line
  // redo[l / G]; the walk's order of (value, position) is that of (value,
  // cell), and the G pairs of a group are merged by DPP within it (quad
  // permutes, then the half-row and row mirrors: every pair merged exactly
  // once).  Lanes past the list repeat its last line and store nothing.
  auto lines_top2_group = [&](auto gc, auto colc, int cnt) {
    constexpr int G = decltype(gc)::value;
    constexpr bool COL = decltype(colc)::value;
    const int q = lane & (G - 1), slot = lane / G;
    const int x = redo[slot < cnt ? slot : cnt - 1];
    const int len = ((COL ? n : m) - q + G - 1) / G;  // 0 when q is past the line
    Top2 t = walk([&](int i) { return COL ? cost(q + i * G, x) : cost(x, q + i * G); },
                  [&](int i) { return act[(COL ? 0 : n) + q + i * G] != 0; }, len);
    if (t.a1 != INT_MAX) t.a1 = q + t.a1 * G;
    if (t.a2 != INT_MAX) t.a2 = q + t.a2 * G;
    if constexpr (G >= 2) t = top2_merge_dpp<DPP_QX1>(t);
    if constexpr (G >= 4) t = top2_merge_dpp<DPP_QX2>(t);
    if constexpr (G >= 8) t = top2_merge_dpp<DPP_HMIRROR>(t);
    if constexpr (G >= 16) t = top2_merge_dpp<DPP_MIRROR>(t);
    if (q == 0 && slot < cnt) store(COL ? n + x : x, t);
  };
  // every line of redo[]: groups of 64 / cnt lanes (16 at most) per line, or
  // one lane per line past 32 lines
  auto lines_top2 = [&](auto colc, int cnt) {
    if (cnt == 0) return;
    if (cnt <= 4) lines_top2_group(std::integral_constant<int, 16>{}, colc, cnt);
    else if (cnt <= 8) lines_top2_group(std::integral_constant<int, 8>{}, colc, cnt);
    else if (cnt <= 16) lines_top2_group(std::integral_constant<int, 4>{}, colc, cnt);
    else if (cnt <= 32) lines_top2_group(std::integral_constant<int, 2>{}, colc, cnt);
    else if constexpr (decltype(colc)::value)
      for (int r = lane; r < cnt; r += WAVE) col_top2(redo[r]);
    else
      for (int r = lane; r < cnt; r += WAVE) row_top2_lane(redo[r]);
  };
  // appends the flagged lines of one 64-line slice to redo[]; returns the new count
  auto append = [&](bool flag, int v, int cnt) {
    const unsigned long long mask = __ballot(flag);
    const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
    if (flag) redo[cnt + below] = v;
    return cnt + __popcll(mask);
  };
  // supplies and demands; the last column takes the remainder (its
  // constraint is the one the LP drops)
  double sbm = 0.0, sa = 0.0;
  for (int k = lane; k < n; k += WAVE) { rem[k] = P.wa[k]; sa += P.wa[k]; }
  for (int j = lane; j < m - 1; j += WAVE) { rem[n + j] = P.wb[j]; sbm += P.wb[j]; }
  sa = wave_sum_d(sa);
  sbm = wave_sum_d(sbm);
  if (lane == 0) rem[n + m - 1] = sa - sbm > 0.0 ? sa - sbm : 0.0;
  for (int v = lane; v < nn; v += WAVE) act[v] = 1;
  __builtin_amdgcn_wave_barrier();
  for (int k = lane; k < n; k += WAVE) row_top2_lane(k);
  for (int j = lane; j < m; j += WAVE) col_top2(j);
  __builtin_amdgcn_wave_barrier();
  int nr = n, nc = m;
  // sub-phase times (lane 0; added to P.prof by solve's thread 0, which is this lane)
  long long vt[4] = {0, 0, 0, 0};
  long long vlast = P.prof ? (long long)wall_clock64() : 0;
  auto vmark = [&](int i) {
    if (P.prof && lane == 0) {
      const long long tnow = (long long)wall_clock64();
      vt[i] += tnow - vlast;
      vlast = tnow;
    }
  };
  for (int t = 0; t < nn - 1; ++t) {
    // line with the largest penalty (crossed-out lines hold -2)
    float bp = -1.0f;
    int bl = INT_MAX;
    for (int v0 = 0; v0 < nn; v0 += 8 * WAVE) {
      float pv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int v = v0 + u * WAVE + lane;
        pv[u] = pen[v < nn ? v : 0];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int v = v0 + u * WAVE + lane;
        const bool better = v < nn && (pv[u] > bp || (pv[u] == bp && v < bl));
        bp = better ? pv[u] : bp;
        bl = better ? v : bl;
      }
    }
    // wave argmax (largest penalty, lowest line on ties), DPP as wave_min_d
    auto pen_merge = [&](float op, int ol) {
      const bool better = op > bp || (op == bp && ol < bl);
      bp = better ? op : bp;
      bl = better ? ol : bl;
    };
    pen_merge(dpp_f<DPP_ROR1>(bp), dpp_i<DPP_ROR1>(bl));
    pen_merge(dpp_f<DPP_ROR2>(bp), dpp_i<DPP_ROR2>(bl));
    pen_merge(dpp_f<DPP_ROR4>(bp), dpp_i<DPP_ROR4>(bl));
    pen_merge(dpp_f<DPP_ROR8>(bp), dpp_i<DPP_ROR8>(bl));
    pen_merge(dpp_f<DPP_BC15, 0xa>(bp, -FLT_MAX), dpp_i<DPP_BC15, 0xa>(bl, INT_MAX));
    pen_merge(dpp_f<DPP_BC31, 0xc>(bp, -FLT_MAX), dpp_i<DPP_BC31, 0xc>(bl, INT_MAX));
    bp = lane63_f(bp);
    bl = lane63_i(bl);
    vmark(0);
    const int is = bl < n ? bl : ma1[bl];
    const int js = bl < n ? ma1[bl] : bl - n;
    const double ra = rem[is], rb = rem[n + js];
    double q = fmin(ra, rb);
    if (q < 0.0) q = 0.0;
    const bool cross_row = (ra - q <= rb - q && nr > 1) || nc == 1;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      const int gone = cross_row ? is : n + js;
      rem[is] = ra - q;
      rem[n + js] = rb - q;
      ci[t] = is;
      cj[t] = js;
      cq[t] = q;
      act[gone] = 0;
      pen[gone] = -2.0f;
    }
    __builtin_amdgcn_wave_barrier();
    vmark(1);
    if (t == nn - 2) break;
    int cnt = 0;
    if (cross_row) {
      --nr;
      // columns whose two cheapest cells included row is
      for (int j0 = 0; j0 < m; j0 += WAVE) {
        const int j = j0 + lane;
        const bool aff = j < m && act[n + j] && (ma1[n + j] == is || ma2[n + j] == is);
        cnt = append(aff, j, cnt);
      }
      __builtin_amdgcn_wave_barrier();
      vmark(2);
      lines_top2(std::true_type{}, cnt);
    } else {
      --nc;
      // rows whose two cheapest cells included column js
      for (int k0 = 0; k0 < n; k0 += WAVE) {
        const int k = k0 + lane;
        const bool aff = k < n && act[k] && (ma1[k] == js || ma2[k] == js);
        cnt = append(aff, k, cnt);
      }
      __builtin_amdgcn_wave_barrier();
      vmark(2);
      // one row: the whole wave if it is long; else every row at once
      if (cnt == 1 && m >= 128) row_top2_wave(redo[0]);
      else lines_top2(std::false_type{}, cnt);
    }
    __builtin_amdgcn_wave_barrier();
    vmark(3);
  }
  if (P.prof && lane == 0)
    for (int i = 0; i < 4; ++i) P.prof[PROF_V_ARGMAX + i] = vt[i];
}

// Solves one LP with the whole workgroup (blockDim.x a multiple of 64, at
// most 64*MAX_WAVES threads); returns the optimal value to all threads.
// status bit 0 is set when the pivot limit is reached.
__device__ __attribute__((always_inline)) double solve(const Problem &P, unsigned char *lds, const Layout &L, int *pivots_out,
                        int *status) {
  const int tid = threadIdx.x, T = blockDim.x, nw = T / WAVE;
  const int lane = tid & (WAVE - 1);
  const int n = P.n, m = P.m, nn = n + m, N = n * m;
  const Cost &C = P.C;
  const bool table = P.table != TAB_NONE;
  float *ct = (float *)(lds + L.ct);
  double *accA = (double *)(lds + L.accA), *accB = (double *)(lds + L.accB);
  double *flow = (double *)(lds + L.flow);
  int *par = (int *)(lds + L.par), *depA = (int *)(lds + L.depA), *depB = (int *)(lds + L.depB);
  short *lvl = (short *)(lds + L.lvl);
  int *path = (int *)(lds + L.path);
  Ctl *ctl = (Ctl *)(lds + L.ctl);
  int bank = 0;
  const float inv_n = 1.0f / (float)n;
  const int Tq = T / n, Tr = T - (T / n) * n;  // a step of T cells = Tq columns + Tr rows
  const int wid = tid / WAVE;
  long long pt[PROF_N] = {};
  long long t_last = P.prof ? (long long)wall_clock64() : 0;
  // adds the time since the last mark to phase ph (thread 0, after a barrier)
  auto mark = [&](int ph) {
    if (P.prof && tid == 0) {
      const long long t = (long long)wall_clock64();
      pt[ph] += t - t_last;
      t_last = t;
    }
  };

  // ---- cost table and the cost scale for the reduced-cost tolerance
  double cmax = 0.0;
  if (table) {
    if (P.table == TAB_GLOBAL32) {
      for (int e = tid; e < N; e += T) {
        const float c = P.gt[e];
        ct[e] = c;
        cmax = fmax(cmax, fabs((double)c));
      }
    } else {
      for (int e = tid; e < N; e += T) {
        int k, j;
        split_cell(e, n, inv_n, k, j);
        const double c = C(k, j);
        ct[e] = (float)c;
        cmax = fmax(cmax, fabs(c));
      }
    }
    cmax *= 1.0 + (double)P.tab_rel;
  } else {
    for (int e = tid; e < N; e += T) {
      int k, j;
      split_cell(e, n, inv_n, k, j);
      cmax = fmax(cmax, fabs(C(k, j)));
    }
  }
  cmax = block_max(cmax, ctl, bank, nw);
  const double eps = 1e-12 * (1.0 + cmax);
  mark(PROF_TABLE);

  // ---- starting basis: Vogel's approximation (about half the pivots of the
  // north-west corner), else the north-west corner.  With the table in LDS,
  // Vogel runs on reduced costs c - u - v from approximate duals: shifting a
  // row or a column by a constant leaves the LP's minimisers unchanged, and
  // with near-optimal duals the cells Vogel picks are mostly optimal ones
  // (3 to 8 times fewer pivots at 48 x 50 ... 48 x 500).  The duals come
  // from entropic iterations on the FP32 table, each column's and each row's
  // log-sum-exp in a fixed order (a column by one thread, a row by one wave),
  // so they, the start and the value do not depend on the workgroup size.
  // The exact table is rebuilt after the start.
  bool have_start = false;
  if (P.vogel && n > 1 && m > 1) {
    const bool duals = table && P.dual_iters > 1;
    if (duals) {
      float *du = (float *)(lds + L.accA), *dv = du + n;  // accA: nn doubles, idle here
      float cm = (float)cmax;
      if (P.dual_reduce) {
        // u = row minima, v = column minima of c - u: the temperature is set
        // against the spread the LP sees (c - u - v >= 0, zero in every line)
        for (int k = tid; k < n; k += T) {
          float mn = FLT_MAX;
          for (int j = 0; j < m; ++j) mn = fminf(mn, ct[k + (size_t)j * n]);
          du[k] = mn;
        }
        __syncthreads();
        float rmx = 0.f;
        for (int j = tid; j < m; j += T) {
          const float *cc = ct + (size_t)j * n;
          float mn = FLT_MAX, mx = -FLT_MAX;
          for (int k = 0; k < n; ++k) {
            const float x = cc[k] - du[k];
            mn = fminf(mn, x);
            mx = fmaxf(mx, x);
          }
          dv[j] = mn;
          rmx = fmaxf(rmx, mx - mn);
        }
        cm = (float)block_max((double)rmx, ctl, bank, nw);
      } else {
        for (int v = tid; v < nn; v += T) du[v] = 0.f;  // u and v
        __syncthreads();
      }
      const int nit = P.dual_iters;
      const float r = exp2f(log2f(P.dual_eps1 / P.dual_eps0) / (float)(nit - 1));
      // long rows: G segments of sg columns, one thread each (segment maxima
      // and sums in LDS that Vogel overwrites later: accB and flow)
      float *smx = (float *)(lds + L.accB);
      const int G0 = m <= DUAL_ROW_THREAD ? 1 : (m + DUAL_SEG - 1) / DUAL_SEG;
      const int G = min(G0, max(1, (2 * nn) / n));  // smx, ssm: n G floats each
      const int sg = (m + G - 1) / G;
      float *ssm = smx + (size_t)n * G;
      float te = P.dual_eps0 * cm;
      // log-sum-exp of a line's cells x(0..len-1) in one sweep: 8-cell
      // batches alternate between two online chains (max and sum of
      // exp(x - max)) merged at the end, a fixed order for every line
      // whatever the workgroup size.  Masked cells are -inf (exp gives 0).
      const float NINF = -__builtin_huge_valf();
      auto lse8 = [](float &mx, float &sm, const float (&x)[8]) {
        float b = x[0];
#pragma unroll
        for (int u = 1; u < 8; ++u) b = fmaxf(b, x[u]);
        const float nm = fmaxf(mx, b);
        float acc = sm * __expf(mx - nm);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += __expf(x[u] - nm);
        mx = nm;
        sm = acc;
      };
      auto lse_line = [&](int len, auto xf, float &mo, float &so) {
        float ma = NINF, sa = 0.f, mb = NINF, sb = 0.f;
        int i = 0;
        for (; i + 16 <= len; i += 16) {
          float xa[8], xb[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            xa[u] = xf(i + u);
            xb[u] = xf(i + 8 + u);
          }
          lse8(ma, sa, xa);
          lse8(mb, sb, xb);
        }
        if (i < len) {  // 1..15 cells left
          float xa[8], xb[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int ia = i + u, ib = i + 8 + u;
            const float va = xf(min(ia, len - 1)), vb = xf(min(ib, len - 1));
            xa[u] = ia < len ? va : NINF;
            xb[u] = ib < len ? vb : NINF;
          }
          lse8(ma, sa, xa);  // holds a cell: ma is finite from here on
          if (i + 8 < len) lse8(mb, sb, xb);
        }
        mo = fmaxf(ma, mb);
        so = sa * __expf(ma - mo) + (sb > 0.f ? sb * __expf(mb - mo) : 0.f);
      };
      for (int it = 0; it < nit && cm > 1e-30f; ++it, te *= r) {
        const float ie = 1.0f / te;
        // v_j = te (log b_j - LSE_k (u_k - c_kj) / te): one thread per
        // column, rows visited from k = j mod n on (column j is contiguous in
        // the table; the rotation spreads the threads over the LDS banks)
        for (int j = tid; j < m; j += T) {
          const float *cc = ct + (size_t)j * n;
          const int k1 = j % n;
          float mx, sm;
          lse_line(n, [&](int i) {
            int k = k1 + i;
            k = k >= n ? k - n : k;
            return (du[k] - cc[k]) * ie;
          }, mx, sm);
          dv[j] = te * (__logf(fmaxf((float)P.wb[j], 1e-30f)) - (mx + __logf(sm)));
        }
        __syncthreads();
        // u_k = te (log a_k - LSE_j (v_j - c_kj) / te): row k's segment g
        // by one thread (consecutive threads, consecutive rows: consecutive
        // addresses); then the segments combined in order
        for (int w = tid; w < n * G; w += T) {
          const int k = w % n, g = w / n;
          const int j0 = g * sg, len = max(0, min(m, j0 + sg) - j0);  // 0: an empty segment
          float mx = NINF, sm = 0.f;
          if (len > 0)
            lse_line(len, [&](int i) {
              const int jj = j0 + i;
              return (dv[jj] - ct[k + (size_t)jj * n]) * ie;
            }, mx, sm);
          smx[w] = mx;
          ssm[w] = sm;
        }
        if (G > 1) __syncthreads();
        for (int k = tid; k < n; k += T) {
          float rm = -FLT_MAX;
          for (int h = 0; h < G; ++h) rm = fmaxf(rm, smx[k + h * n]);
          float sm = 0.f;
          for (int h = 0; h < G; ++h) {
            const float sh = ssm[k + h * n];
            sm += sh > 0.f ? sh * __expf(smx[k + h * n] - rm) : 0.f;
          }
          du[k] = te * (__logf(fmaxf((float)P.wa[k], 1e-30f)) - (rm + __logf(sm)));
        }
        __syncthreads();
      }
      for (int e = tid; e < N; e += T) {
        int k, j;
        split_cell(e, n, inv_n, k, j);
        ct[e] = ct[e] - du[k] - dv[j];
      }
      __syncthreads();
      mark(PROF_DUALS);
    }
    if (table) vogel_start<true>(P, lds, L, ct);
    else vogel_start<false>(P, lds, L, ct);
    __syncthreads();
    mark(PROF_START);
    if (duals) {  // the exact FP32 table again, as built above
      for (int e = tid; e < N; e += T) {
        if (P.table == TAB_GLOBAL32) {
          ct[e] = P.gt[e];
        } else {
          int k, j;
          split_cell(e, n, inv_n, k, j);
          ct[e] = (float)C(k, j);
        }
      }
      __syncthreads();
    }
    // root the basis cells at row 0: a node whose tree neighbour is already
    // rooted hangs under it (each node is reached exactly once in a tree)
    const int *ci = (const int *)(lds + L.depA), *cj = (const int *)(lds + L.depB);
    const double *cq = (const double *)(lds + L.accB);
    for (int v = tid; v < nn; v += T) par[v] = v == 0 ? -1 : -2;
    if (tid == 0) flow[0] = 0.0;
    __syncthreads();
    for (int round = 0; round < nn; ++round) {
      int changed = 0;
      for (int t = tid; t < nn - 1; t += T) {
        const int i = ci[t], v = n + cj[t];
        const int pi = par[i], pv = par[v];
        if (pi != -2 && pv == -2) { par[v] = i; flow[v] = cq[t]; changed = 1; }
        else if (pv != -2 && pi == -2) { par[i] = v; flow[i] = cq[t]; changed = 1; }
      }
      if (!block_or(changed, ctl, bank, nw)) break;
    }
    int missing = 0;
    for (int v = tid; v < nn; v += T) missing |= par[v] == -2;
    have_start = !block_or(missing, ctl, bank, nw);
    mark(PROF_ROOT);
  }
  if (!have_start) {
    // north-west corner: each new cell brings one new node, hung under the
    // endpoint already in the tree; flow[v] = flow on v's parent edge
    if (tid == 0) {
      double sa = 0.0, sbm = 0.0;
      for (int k = 0; k < n; ++k) { accA[k] = P.wa[k]; sa += P.wa[k]; }
      for (int j = 0; j < m - 1; ++j) { accB[j] = P.wb[j]; sbm += P.wb[j]; }
      accB[m - 1] = sa - sbm > 0.0 ? sa - sbm : 0.0;
      par[0] = -1;
      flow[0] = 0.0;
      int i = 0, j = 0, newnode = n;  // the first cell (0,0) brings column 0
      for (int t = 0; t < nn - 1; ++t) {
        double q = fmin(accA[i], accB[j]);
        if (q < 0.0) q = 0.0;
        par[newnode] = newnode >= n ? i : n + j;
        flow[newnode] = q;
        accA[i] -= q;
        accB[j] -= q;
        if (t == nn - 2) break;
        if (i == n - 1) { ++j; newnode = n + j; }
        else if (j == m - 1) { ++i; newnode = i; }
        else if (accA[i] <= accB[j]) { ++i; newnode = i; }
        else { ++j; newnode = n + j; }
      }
    }
    __syncthreads();
    mark(PROF_START);
  }

  // cost of the tree edge at node v (edge to parent p), table or exact
  auto edge_cell = [&](int v, int p) { return v < n ? v + (p - n) * n : p + (v - n) * n; };
  auto edge_k = [&](int v, int p) { return v < n ? v : p; };
  auto edge_j = [&](int v, int p) { return v < n ? p - n : v - n; };

  bool exact_mode = !table;
  int pivots = 0, pos0 = 0;
  const int blk = T * P.cpt;
  const int dantzig_limit = 50 * nn + 1000;
  const int hard_limit = 200000 + 100 * N;
  double *pot = accA;
  int *dep = depA;
  for (;;) {
    // ---- pointer jumping: potentials, depths, lifting table.  A tree of at
    // most D2MI_LP_SOLO_NODES nodes is done by wave 0 alone: its rounds are
    // ordered by the wave itself (LDS operations of one wave complete in
    // order), so a round costs two LDS round trips and no workgroup barrier;
    // the other waves wait at one barrier at the end.  Larger trees use every
    // wave and a workgroup barrier per round.
    const bool solo = nn <= D2MI_LP_SOLO_NODES;
    if (!solo || wid == 0) {
      const int jt = solo ? lane : tid, JT = solo ? WAVE : T;
      auto round_sync = [&]() {
        if (solo) {
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
          asm volatile("" ::: "memory");
        } else {
          __syncthreads();
        }
      };
      // JB nodes per thread at a time: parents, then edge costs, then stores
      for (int v0 = jt; v0 < nn; v0 += JB * JT) {
        int pp[JB];
        double cc[JB];
#pragma unroll
        for (int u = 0; u < JB; ++u) {
          const int v = v0 + u * JT;
          const int p = par[min(v, nn - 1)];  // loads stay in range, no branch
          pp[u] = v < nn ? p : -1;
        }
        if (exact_mode) {
#pragma unroll
          for (int u = 0; u < JB; ++u) {
            const int v = v0 + u * JT;
            cc[u] = pp[u] >= 0 ? C(edge_k(v, pp[u]), edge_j(v, pp[u])) : 0.0;
          }
        } else {
          float cf[JB];
#pragma unroll
          for (int u = 0; u < JB; ++u) {
            const int v = v0 + u * JT;
            cf[u] = ct[pp[u] >= 0 ? edge_cell(v, pp[u]) : 0];
          }
#pragma unroll
          for (int u = 0; u < JB; ++u) cc[u] = pp[u] >= 0 ? (double)cf[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < JB; ++u) {
          const int v = v0 + u * JT;
          if (v < nn) {
            lvl[v] = (short)pp[u];
            accA[v] = cc[u];
            depA[v] = pp[u] >= 0 ? 1 : 0;
          }
        }
      }
      round_sync();
      double *aS = accA, *aD = accB;
      int *dS = depA, *dD = depB;
      for (int r = 0; r + 1 < L.levels; ++r) {
        const short *cur = lvl + (size_t)r * nn;
        short *nxt = lvl + (size_t)(r + 1) * nn;
        const double sg = r == 0 ? -1.0 : 1.0;
        int any = 0;
        // JB nodes per thread at a time: every load of the batch is issued
        // before its stores, so the LDS latencies overlap; loads read clamped
        // indices unconditionally (a guarded load is a branch, and branches,
        // not LDS latency, set the cost of a round on one wave)
        for (int v0 = jt; v0 < nn; v0 += JB * JT) {
          int vv[JB], a[JB], aa[JB], dv[JB], da[JB];
          double av[JB], aav[JB];
#pragma unroll
          for (int u = 0; u < JB; ++u) {
            vv[u] = v0 + u * JT;
            const int vc = min(vv[u], nn - 1);
            a[u] = cur[vc];
            av[u] = aS[vc];
            dv[u] = dS[vc];
          }
#pragma unroll
          for (int u = 0; u < JB; ++u) {
            const int x = a[u] >= 0 ? a[u] : 0;
            const int ax = cur[x];
            aa[u] = a[u] >= 0 ? ax : -1;
            aav[u] = aS[x];
            da[u] = dS[x];
          }
#pragma unroll
          for (int u = 0; u < JB; ++u) {
            if (vv[u] < nn) {
              const bool up = a[u] >= 0;
              aD[vv[u]] = up ? av[u] + sg * aav[u] : av[u];
              dD[vv[u]] = up ? dv[u] + da[u] : dv[u];
              nxt[vv[u]] = (short)aa[u];
              any |= aa[u] >= 0;
            }
          }
        }
        if (solo) {
          any = __any(any) ? 1 : 0;
          round_sync();
        } else {
          any = block_or(any, ctl, bank, nw);
        }
        double *td = aS; aS = aD; aD = td;
        int *ti = dS; dS = dD; dD = ti;
        if (!any) break;
      }
      if (solo && lane == 0) ctl->jump_b = aS == accB ? 1 : 0;
      pot = aS;
      dep = dS;
    }
    if (solo) {
      __syncthreads();
      const bool in_b = ctl->jump_b != 0;
      pot = in_b ? accB : accA;
      dep = in_b ? depB : depA;
    }
    double *pfl = pot == accA ? accB : accA;  // free during the pivot
    mark(PROF_JUMP);  // after the last round's barrier (solo: the one above)

    // ---- pricing
    const bool bland = pivots >= dantzig_limit;
    int ent = INT_MAX;
    double ent_r = 0.0;
    for (int scanned = 0; scanned < N; scanned += blk) {
      int base = (bland ? 0 : pos0) + scanned;
      if (base >= N) base -= N;
      double best = -eps;
      int bestc = INT_MAX;
      // this thread's cells base+tid, base+tid+T, ...: (k, j) stepped without divides
      int e = base + tid;
      if (e >= N) e -= N;
      int k, j;
      split_cell(e, n, inv_n, k, j);
      if (table && !exact_mode) {
        // PB cells at a time: their indices first, then every load of the
        // batch (cells past the block read cell 0), then the same compares in
        // the same order as one cell at a time
        constexpr int PB = 8;
        for (int i0 = 0; i0 < P.cpt; i0 += PB) {
          int ee[PB], kk[PB], jj[PB];
          bool ok[PB];
#pragma unroll
          for (int u = 0; u < PB; ++u) {
            ok[u] = i0 + u < P.cpt && scanned + tid + (i0 + u) * T < N;
            ee[u] = ok[u] ? e : 0;
            kk[u] = ok[u] ? k : 0;
            jj[u] = ok[u] ? j : 0;
            e += T; k += Tr; j += Tq;
            if (k >= n) { k -= n; ++j; }
            if (e >= N) { e -= N; j -= m; }
          }
          float cv[PB];
          double pu[PB], pv[PB];
#pragma unroll
          for (int u = 0; u < PB; ++u) {
            cv[u] = ct[ee[u]];
            pu[u] = pot[kk[u]];
            pv[u] = pot[n + jj[u]];
          }
          // selects, not branches: the same decisions as one cell at a time
          if (bland) {
#pragma unroll
            for (int u = 0; u < PB; ++u) {
              const double r = (double)cv[u] - pu[u] - pv[u];
              const bool take = ok[u] && r < -eps && ee[u] < bestc;
              best = take ? r : best;
              bestc = take ? ee[u] : bestc;
            }
          } else {
#pragma unroll
            for (int u = 0; u < PB; ++u) {
              const double r = (double)cv[u] - pu[u] - pv[u];
              const bool take =
                  ok[u] && (r < best || (r == best && bestc != INT_MAX && ee[u] < bestc));
              best = take ? r : best;
              bestc = take ? ee[u] : bestc;
            }
          }
        }
      } else {
        for (int i = 0; i < P.cpt && scanned + tid + i * T < N; ++i) {
          const double pu = pot[k], pv = pot[n + j];
          double r = 0.0;
          bool priced = true;
          if (!table) {
            r = C(k, j) - pu - pv;
          } else {
            const double c = (double)ct[e];
            r = c - pu - pv;
            if (r < (double)P.tab_rel * fabs(c) + 1e-37 + eps) r = C(k, j) - pu - pv;
            else priced = false;
          }
          if (priced) {
            if (bland) {
              if (r < -eps && e < bestc) { bestc = e; best = r; }
            } else if (r < best || (r == best && bestc != INT_MAX && e < bestc)) {
              best = r;
              bestc = e;
            }
          }
          e += T; k += Tr; j += Tq;
          if (k >= n) { k -= n; ++j; }
          if (e >= N) { e -= N; j -= m; }
        }
      }
      if (bland) {
        double dummy = bestc == INT_MAX ? 0.0 : -1.0;
        block_argmin(dummy, bestc, ctl, bank, nw);
        if (bestc != INT_MAX) { ent = bestc; break; }
      } else {
        block_argmin(best, bestc, ctl, bank, nw);
        if (bestc != INT_MAX && best < -eps) {
          ent = bestc;
          ent_r = best;
          pos0 = base + blk;
          while (pos0 >= N) pos0 -= N;
          break;
        }
      }
    }
    (void)ent_r;
    mark(PROF_PRICE);
    if (ent == INT_MAX) {
      if (!exact_mode) {  // optimal for the FP32 table: verify on exact costs
        exact_mode = true;
        continue;
      }
      break;
    }
    if (pivots >= hard_limit) {
      if (tid == 0) atomicOr(status, 1);
      break;
    }
    int ei, ej;
    split_cell(ent, n, inv_n, ei, ej);
    const int a = n + ej, b = ei;
    const int da = dep[a], db = dep[b];

    // ---- apex of the cycle: largest depth D with anc(a, da-D) == anc(b, db-D)
    auto anc = [&](int v, int k) {
      for (int bit = 0; k; ++bit, k >>= 1)
        if (k & 1) v = lvl[(size_t)bit * nn + v];
      return v;
    };
    // ---- pivot.  Fast path: both arms of the cycle lie within one window of
    // 64 depths below max(da, db), so lane l of wave 0 holds the depth-(top-l)
    // node of each arm and does the ratio test, the flow update and the
    // re-hang with shuffles; other waves go straight to the barrier.
    const int top = max(da, db);
    const int Dl = top - lane;
    const int xa = (Dl >= 0 && Dl <= da) ? anc(a, da - Dl) : -1;
    const int xb = (Dl >= 0 && Dl <= db) ? anc(b, db - Dl) : -2;
    const unsigned long long bal = __ballot(Dl >= 0 && Dl <= min(da, db) && xa == xb);
    if (bal != 0) {
      if (wid == 0) {
        const int Dstar = top - (__ffsll((long long)bal) - 1);
        const int nA = da - Dstar;
        const int tA = da - Dl, tB = db - Dl;
        const bool vA = Dl > Dstar && Dl <= da, vB = Dl > Dstar && Dl <= db;
        const double fA = vA ? flow[xa] : 0.0, fB = vB ? flow[xb] : 0.0;
        const int cA = vA ? edge_cell(xa, par[xa]) : INT_MAX;
        const int cB = vB ? edge_cell(xb, par[xb]) : INT_MAX;
        double th = DBL_MAX;
        int key = INT_MAX;
        if (vA && (tA & 1) == 0) { th = fA; key = bland ? cA : tA; }
        if (vB && (tB & 1) == 0) {
          const int kb = bland ? cB : nA + tB;
          if (fB < th || (fB == th && kb < key)) { th = fB; key = kb; }
        }
        const double theta = wave_min_d(th);
        const int keyw = wave_min_i(th == theta ? key : INT_MAX);
        int qpos = keyw;
        if (bland) {
          int cand = INT_MAX;
          if (vA && (tA & 1) == 0 && fA == theta && cA == keyw) cand = tA;
          if (vB && (tB & 1) == 0 && fB == theta && cB == keyw) cand = min(cand, nA + tB);
          qpos = wave_min_i(cand);
        }
        const int qside = qpos >= nA ? 1 : 0;
        const int qt = qside ? qpos - nA : qpos;
        const int src = lane > 0 ? lane - 1 : 0;  // one step down the arm
        const int xa_p = __shfl(xa, src, WAVE), xb_p = __shfl(xb, src, WAVE);
        const double fA_p = __shfl(fA, src, WAVE), fB_p = __shfl(fB, src, WAVE);
        if (vA) {
          if (qside == 0 && tA <= qt) {
            par[xa] = tA == 0 ? b : xa_p;
            flow[xa] = tA == 0 ? theta : fA_p + (((tA - 1) & 1) ? theta : -theta);
          } else {
            flow[xa] = fA + ((tA & 1) ? theta : -theta);
          }
        }
        if (vB) {
          if (qside == 1 && tB <= qt) {
            par[xb] = tB == 0 ? a : xb_p;
            flow[xb] = tB == 0 ? theta : fB_p + (((tB - 1) & 1) ? theta : -theta);
          } else {
            flow[xb] = fB + ((tB & 1) ? theta : -theta);
          }
        }
      }
    } else {
      int Dstar = 0;
      for (int top = min(da, db); top >= 0; top -= WAVE) {
        const int D = top - lane;
        const bool eq = D >= 0 && anc(a, da - D) == anc(b, db - D);
        const unsigned long long bal = __ballot(eq);
        if (bal) {
          Dstar = top - (__ffsll((long long)bal) - 1);
          break;
        }
      }
      const int nA = da - Dstar, nB = db - Dstar, np = nA + nB;

      // ---- path nodes and the ratio test over the minus edges (even t)
      double th = DBL_MAX;
      int thk = INT_MAX;
      for (int q = tid; q < np; q += T) {
        const bool sideB = q >= nA;
        const int t = sideB ? q - nA : q;
        const int v = anc(sideB ? b : a, t);
        path[q] = v;
        const double f = flow[v];
        pfl[q] = f;
        if ((t & 1) == 0) {
          const int key = bland ? edge_cell(v, par[v]) : q;
          if (f < th || (f == th && key < thk)) { th = f; thk = key; }
        }
      }
      block_argmin(th, thk, ctl, bank, nw);
      // the leaving position: recover it from the key
      int qpos = thk;
      if (bland) {
        // the key is a cell index: find the position holding it
        int found = INT_MAX;
        for (int q = tid; q < np; q += T) {
          const bool sideB = q >= nA;
          const int t = sideB ? q - nA : q;
          if ((t & 1) == 0) {
            const int v = path[q];
            if (pfl[q] == th && edge_cell(v, par[v]) == thk) found = min(found, q);
          }
        }
        double dummy = 0.0;
        block_argmin(dummy, found, ctl, bank, nw);
        qpos = found;
      }
      const double theta = th;
      const int qside = qpos >= nA ? 1 : 0;
      const int qt = qside ? qpos - nA : qpos;

      // ---- flows around the cycle and the re-hang of the cut subtree
      for (int q = tid; q < np; q += T) {
        const int side = q >= nA ? 1 : 0;
        const int t = side ? q - nA : q;
        const int v = path[q];
        if (side == qside && t <= qt) {
          if (t == 0) {
            par[v] = side == 0 ? b : a;
            flow[v] = theta;
          } else {
            par[v] = path[q - 1];
            flow[v] = pfl[q - 1] + (((t - 1) & 1) ? theta : -theta);
          }
        } else {
          flow[v] = pfl[q] + ((t & 1) ? theta : -theta);
        }
      }
    }
    __syncthreads();
    mark(PROF_PIVOT);
    ++pivots;
  }

  // ---- objective on exact costs over the tree edges, summed in a fixed
  // pairwise order so the value does not depend on the workgroup size
  double *sum = pot == accA ? accB : accA;
  for (int v = tid; v < nn; v += T) {
    const int p = par[v];
    sum[v] = p >= 0 ? flow[v] * C(edge_k(v, p), edge_j(v, p)) : 0.0;
  }
  __syncthreads();
  for (int stride = 1; stride < nn; stride <<= 1) {
    for (int i = tid * 2 * stride; i + stride < nn; i += T * 2 * stride) sum[i] += sum[i + stride];
    __syncthreads();
  }
  const double val = sum[0];
  if (pivots_out) *pivots_out = pivots;
  mark(PROF_VALUE);
  if (P.prof && tid == 0)
    for (int i = 0; i <= PROF_VALUE; ++i) P.prof[i] = pt[i];
  __syncthreads();
  return val;
}

}  // namespace lp
