// Device-side pieces of the PCG shared by the single-system kernels
// (sparse.hip) and the continuous-batching kernels (batch.hip): the
// node-block SpMV, fence-free reductions, the solver state and the bodies of
// the two per-iteration kernels.  Each body takes the block's index and the
// block count of its own system, so one launch can carry several systems,
// each on the grid it would get alone (identical reduction order, identical
// results).
//
// Matrix layout.  The global matrix is EasyFEA's scipy-canonical CSR
// (sorted, duplicate-free columns, EasyFEA/Simulations/_simu.py:1062-1102).
// Because every node couples to every DOF of each neighbouring node, row
// (n, a) of that CSR holds, for each neighbour m of n in increasing order,
// the dof_n columns m*dof_n + b.  The kernels therefore index the unchanged
// CSR value array through node-level adjacency (node_ptr, node_cols):
//   value(n, a, pos, b) = dof_n^2 * node_ptr[n] + a * dof_n * cnt(n) + pos * dof_n + b
// which moves one int32 per dof_n^2 values instead of one per value.
#pragma once
#include "efea_common.h"

namespace efea {

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct SpmvArgs {
  int Nnode;
  const int* node_ptr;
  const int* node_cols;
  const void* vals;
  const unsigned char* known;
  // FP32 values only: per node, the dof_n x dof_n block row sum of
  // (K64 - K32), added to the node's own rows so that the rounded operator
  // keeps K's row sums (rigid translations in the null space).  May be null.
  const void* corr = nullptr;
  // most blocks of this system's grids (<= 0: MAXB); part of the reduction
  // order, so two solves agree bitwise only with the same cap
  int cap = 0;
};

// Row-block products of one node against NV vectors.  Lane g of the node's
// G lanes takes positions g, g+G, g+2G, ... (its summation order) and
// accumulates acc[k][a] += K(n, a, pos, b) * X_k(col(pos), b).  The
// positions are visited U at a time with every load of a pass issued before
// the first use (indices past cnt are clamped to the last position and their
// products dropped), so one pass costs one column-index -> vector latency
// chain instead of one per position.  load(c, xv) fills xv[k][b] = X_k(c, b).
template <int G>
struct GatherUnroll {
  static constexpr int U = G <= 2 ? 5 : (G <= 4 ? 3 : 2);
};

template <typename TM, int DN, int G, int NV, typename Load>
__device__ __forceinline__ void gather_rows(const TM* __restrict__ vals, const int* __restrict__ cols, int p0,
                                            int cnt, int g, const Load& load, double (&acc)[NV][DN]) {
  constexpr int U = GatherUnroll<G>::U;
  const int64_t base = (int64_t)DN * DN * p0;
  for (int pos0 = g; pos0 < cnt; pos0 += U * G) {
    int pc[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      pc[u] = min(pos0 + u * G, cnt - 1);
      c[u] = cols[p0 + pc[u]];
    }
    double v[U][DN][DN];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int a = 0; a < DN; ++a)
#pragma unroll
        for (int b = 0; b < DN; ++b) v[u][a][b] = (double)vals[base + (int64_t)a * DN * cnt + (int64_t)pc[u] * DN + b];
    double xv[U][NV][DN];
#pragma unroll
    for (int u = 0; u < U; ++u) load(c[u], xv[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (pos0 + u * G < cnt) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
          for (int a = 0; a < DN; ++a)
#pragma unroll
            for (int b = 0; b < DN; ++b) acc[k][a] += v[u][a][b] * xv[u][k][b];
      }
    }
  }
}

// Node-block SpMV: G lanes cooperate on one node's dof_n rows.
//   MODE 0: y = A x            (rows of known DOFs -> 0 when known != null)
//   MODE 1: y = b - A x        (same masking)
// dot_xy receives x.y (MODE 0) or y.y (MODE 1) over the node's rows.
template <typename TM, int DN, int G, int MODE>
__device__ inline void spmv_node(const SpmvArgs& A, int n, int g, const double* __restrict__ x,
                                 const double* __restrict__ b, double* __restrict__ y, double& dot_xy) {
  const int p0 = A.node_ptr[n], cnt = A.node_ptr[n + 1] - p0;
  double sum[1][DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) sum[0][a] = 0;
  gather_rows<TM, DN, G, 1>((const TM*)A.vals, A.node_cols, p0, cnt, g,
                            [&](int c, double (&xv)[1][DN]) {
#pragma unroll
                              for (int bb = 0; bb < DN; ++bb) xv[0][bb] = x[(int64_t)c * DN + bb];
                            },
                            sum);
#pragma unroll
  for (int a = 0; a < DN; ++a) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) sum[0][a] += __shfl_xor(sum[0][a], o, G);
  }
  dot_xy = 0;
  if (g == 0) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t r = (int64_t)n * DN + a;
      double val = MODE == 1 ? b[r] - sum[0][a] : sum[0][a];
      if (A.known && A.known[r]) val = 0;
      y[r] = val;
      dot_xy += (MODE == 1 ? val : x[r]) * val;
    }
  }
}

// ---------------------------------------------------------------------------
// Fence-free global reductions.
//
// CDNA4 splits the chip into XCDs whose L2 caches are not mutually coherent:
// a device-scope fence is an L2 write-back, so a "last block finishes the
// reduction" pattern (one fence per block) serialises on the write-backs.
// Here every kernel runs a fixed grid of at most MAXB blocks per system
// (grid-stride loops), writes one partial per block with a plain store, and
// the *next* kernel reduces those partials, redundantly in every block and
// in a fixed order (deterministic).  Visibility comes from the kernel
// boundary: one release per dispatch instead of one per block.
constexpr int BS = 256;
constexpr int MAXB = 1024;

// Supported (dof_n, lanes per node) pairs of the node-group kernels; every
// dispatch site expands this list.
#define EFEA_GROUPS(X) X(1, 2) X(1, 4) X(1, 8) X(1, 16) X(2, 4) X(2, 8) X(2, 16) X(3, 8) X(3, 16) X(3, 32)

// Logical block of hardware block b in a grid of nb: the dispatcher deals
// blocks round-robin to the 8 XCDs (b, b+8, b+16, ... on XCD b % 8); this
// bijection gives each XCD one contiguous range of logical blocks, so the
// neighbour rows a block gathers were mostly fetched into the same XCD's L2 by
// its neighbouring blocks.
// Used when the grid covers its rows in one pass (grid-stride passes keep the
// plain order, which measured faster on the 64^3 cube).
__device__ inline int xcd_block(int b, int nb) {
  constexpr int NX = 8;
  const int per = nb / NX, rem = nb % NX, x = b % NX, k = b / NX;
  return x * per + (x < rem ? x : rem) + k;
}

static inline unsigned grid_for(int64_t threads, int cap = 0) {
  const int64_t mx = (cap > 0 && cap < MAXB) ? cap : MAXB;
  int64_t nb = (threads + BS - 1) / BS;
  if (nb > mx) nb = mx;
  if (nb < 1) nb = 1;
  return (unsigned)nb;
}

// Sum of n partials (stride `stride`, offset `off`) in every thread of the block.
__device__ inline double reduce_all(const double* __restrict__ part, int n, int stride, int off, double* sh) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += BS) s += part[stride * i + off];
  s = block_sum<double, BS>(s, sh);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = s;
  __syncthreads();
  const double r = bc;
  __syncthreads();
  return r;
}

// PCG device state.  Scalars stay FP64 in every mode.  The stopping test is
// the Strakos-Tichy delayed estimate of the A-norm error: after k+RING
// iterations ||x* - x_k||_A^2 ~ sum_{j=k}^{k+RING-1} alpha_j (r_j . z_j),
// compared with max(sum_j alpha_j r_j . z_j, x0.A x0).  The estimate is a
// lower bound (the tail beyond the window is left out) and the test runs
// RING iterations behind the iterate it certifies: a window of 8 stops a
// solve 24 iterations earlier than one of 32 (on the notched-square history
// 158 instead of 183 displacement and 50 instead of 72 damage iterations)
// with errors of 1e-9 relative against a sparse direct solve instead of
// 1e-10, for a tolerance of 1e-9.
// A field written by block 0 of a kernel is never read by the other blocks
// of that kernel: rz_prev, rrmax_prev and par_next are per-iteration
// snapshots that break those races.
constexpr int RING = 8;
// This thread's share (stride BS) of n partials of S values each, added in
// the order i = t, t + BS, t + 2 BS, ...  issue() loads the first MAXB / BS
// slots with predicated loads (zero past n), unrolled, so that they are all
// in flight together and with the loads that follow in the kernel; sum()
// adds them where the value is needed, so that the wait for them comes late
// (vector loads complete in order: an early add would also wait for every
// load issued before it).  Grids of more than MAXB blocks (one block per
// aggregate) add the rest in a loop.
template <int S>
struct Partials {
  static constexpr int R = MAXB / BS;
  double v[R][S];
  __device__ __forceinline__ void issue(const double* __restrict__ part, int n) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = (int)threadIdx.x + j * BS;
#pragma unroll
      for (int k = 0; k < S; ++k) v[j][k] = i < n ? part[S * i + k] : 0.0;
    }
  }
  __device__ __forceinline__ void sum(const double* __restrict__ part, int n, double (&out)[S]) const {
#pragma unroll
    for (int k = 0; k < S; ++k) {
      out[k] = v[0][k];
#pragma unroll
      for (int j = 1; j < R; ++j) out[k] += v[j][k];
    }
    for (int i = (int)threadIdx.x + MAXB; i < n; i += BS) {
#pragma unroll
      for (int k = 0; k < S; ++k) out[k] += part[S * i + k];
    }
  }
};

template <int S>
__device__ inline void partial_share(const double* __restrict__ part, int n, double (&out)[S]) {
  Partials<S> P;
  P.issue(part, n);
  P.sum(part, n, out);
}

struct PcgState {
  double rz, rz_prev, alpha, beta, pq;
  double rr, rrmax, rrmax_prev, bb;
  double tolr2, tolA2, delta2;
  double Sall, Sw, xAx0;
  double ring[RING];
  int head, done, iters, nrepl;
  int fresh;     // no iteration to close yet: the first step uses beta = 0
  int justrep;   // the r.z / r.r partials come from a residual replacement
  int par;       // which of the two direction buffers holds p
  int par_next;  // the buffer the current step writes
  int stop;      // pipelined solve: done as seen by the last product pass (read by the next coarse pass)
};
// done: 0 running, 1 converged, 2 breakdown, 3 diverged, 4 paused for a
// residual replacement (mixed mode; the host runs the replacement kernel).

template <typename TP, int DN>
__device__ inline void apply_bj(const TP* __restrict__ Minv, const unsigned char* __restrict__ known, int n,
                                const double (&rv)[DN], double (&zv)[DN]) {
#pragma unroll
  for (int a = 0; a < DN; ++a) {
    double s = 0;
#pragma unroll
    for (int b = 0; b < DN; ++b) s += (double)Minv[(int64_t)n * DN * DN + a * DN + b] * rv[b];
    if (known && known[(int64_t)n * DN + a]) s = 0;
    zv[a] = s;
  }
}

// Iteration bookkeeping, block 0 only.
__device__ inline void pcg_close_iteration(PcgState* st, double rz_new, double rr, bool replaced) {
  const double term = st->alpha * st->rz_prev;  // alpha_j (r_j . z_j)
  st->Sall += term;
  st->ring[st->head] = term;
  st->head = (st->head + 1) % RING;
  double Sw = 0;  // summed afresh: a running add/subtract would cancel catastrophically
  for (int i = 0; i < RING; ++i) Sw += st->ring[i];
  st->Sw = Sw;
  st->iters += 1;
  st->beta = rz_new / st->rz_prev;
  st->rz = rz_new;
  st->rr = rr;
  st->rrmax = replaced ? rr : (rr > st->rrmax_prev ? rr : st->rrmax_prev);
  const double scale = st->Sall > st->xAx0 ? st->Sall : st->xAx0;
  if (st->iters >= RING && Sw <= st->tolA2 * scale) st->done = 1;
  if (rr <= st->tolr2 * st->bb) st->done = 1;
  // non-finite values (NaN, or an overflow that would pass the tests above)
  if (!isfinite(rr) || !isfinite(rz_new) || !isfinite(Sw) || !isfinite(st->Sall)) st->done = 2;
  if (replaced && rr > 1e12 * st->bb) st->done = 3;  // true residual exploded: diverged
}

// Step 1 of an iteration: close the previous iteration (beta from the r.z
// partials of step 2 or of a replacement), form p = z + beta p_old for every
// column it reads, store p for the rows it owns, q = A p (masked rows) and
// per-block partials of p.q.  In mixed mode it pauses the solve (done = 4)
// when the recursive residual has fallen below delta * its running maximum.
// Latency: the partial loads and the first node group's products (A z and
// A p_old, kept apart because beta is not known yet) are issued before the
// block reduction that yields beta.
// TV: the vector type (double; float for the FP32 inner solves of the
// refinement mode, whose products are still formed in double registers).
template <typename TM, int DN, int G, typename TV>
__device__ inline void pcg_step1_body(const SpmvArgs& A, const TV* __restrict__ z,
                                      const TV* __restrict__ P0, const TV* __restrict__ P1,
                                      TV* __restrict__ q, const double* __restrict__ part2, int nb2,
                                      double* __restrict__ part1, PcgState* st, int blk, int nblk) {
  const int done = st->done, par = st->par;
  const bool fresh = st->fresh;
  const int64_t lanes = (int64_t)A.Nnode * G;
  int64_t gid = (int64_t)blk * BS + threadIdx.x;
  const bool have = gid < lanes;
  const int n0 = (int)(gid / G), g0 = (int)(gid % G);
  // Loads that do not depend on the solver state are issued with the state
  // read, so that their latencies overlap: the node's offsets and the r.z /
  // r.r partials (also read on a fresh start, where they are not used).
  int p0 = 0, p1 = 0;
  if (have) {
    p0 = A.node_ptr[n0];
    p1 = A.node_ptr[n0 + 1];
  }
  Partials<2> P2;
  P2.issue(part2, nb2);
  // p1 < p0 never holds; testing it keeps the offset loads ahead of the branch
  if (done || p1 < p0) return;
  __shared__ double sh[BS / 64];
  __shared__ double sh2[BS / 64];
  __shared__ double bc[2];
  const TV* __restrict__ pold = par ? P1 : P0;
  TV* __restrict__ pnew = (TV*)(par ? P0 : P1);
  const TM* __restrict__ vals = (const TM*)A.vals;
  double sz[DN], sp[DN], zo[DN], po[DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) sz[a] = sp[a] = zo[a] = po[a] = 0;
  if (have) {
    const int cnt = p1 - p0;
    double acc[2][DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) acc[0][a] = acc[1][a] = 0;
    gather_rows<TM, DN, G, 2>(vals, A.node_cols, p0, cnt, g0,
                              [&](int c, double (&xv)[2][DN]) {
#pragma unroll
                                for (int bb = 0; bb < DN; ++bb) {
                                  xv[0][bb] = (double)z[(int64_t)c * DN + bb];
                                  xv[1][bb] = (double)pold[(int64_t)c * DN + bb];
                                }
                              },
                              acc);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      sz[a] = acc[0][a];
      sp[a] = acc[1][a];
    }
    // the node's own z and p_old, ahead of the reduction that yields beta
    if (g0 == 0) {
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        zo[a] = (double)z[(int64_t)n0 * DN + a];
        po[a] = (double)pold[(int64_t)n0 * DN + a];
      }
    }
  }
  double beta = 0;
  if (!fresh) {
    double pr2[2];
    P2.sum(part2, nb2, pr2);
    const double prz = block_sum<double, BS>(pr2[0], sh);
    const double prr = block_sum<double, BS>(pr2[1], sh2);
    if (threadIdx.x == 0) {
      bc[0] = prz;
      bc[1] = prr;
    }
    __syncthreads();
    const double rz_new = bc[0], rr = bc[1];
    const bool jr = st->justrep;
    if (!jr && st->delta2 > 0 && rr < st->delta2 * st->rrmax_prev) {
      if (blk == 0 && threadIdx.x == 0) st->done = 4;
      return;
    }
    beta = rz_new / st->rz_prev;
    if (blk == 0 && threadIdx.x == 0) pcg_close_iteration(st, rz_new, rr, jr);
  }
  if (blk == 0 && threadIdx.x == 0) st->par_next = 1 - par;
  double d = 0;
  if (have) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) {
        sz[a] += __shfl_xor(sz[a], o, G);
        sp[a] += __shfl_xor(sp[a], o, G);
      }
    }
    if (g0 == 0) {
      double pn[DN];
#pragma unroll
      for (int a = 0; a < DN; ++a) pn[a] = zo[a] + beta * po[a];
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t r = (int64_t)n0 * DN + a;
        pnew[r] = (TV)pn[a];
        double qv = sz[a] + beta * sp[a];
        if (A.corr) {
          const TM* cr = (const TM*)A.corr + (int64_t)n0 * DN * DN + a * DN;
#pragma unroll
          for (int b = 0; b < DN; ++b) qv += (double)cr[b] * pn[b];
        }
        if (A.known && A.known[r]) qv = 0;
        q[r] = (TV)qv;
        d += pn[a] * qv;
      }
    }
  }
  // later nodes of this block: p is known, one product; the next node's
  // offsets are loaded before this node's gather
  const int64_t stride = (int64_t)nblk * BS;
  gid += stride;
  int np0 = 0, np1 = 0;
  if (gid < lanes) {
    np0 = A.node_ptr[gid / G];
    np1 = A.node_ptr[gid / G + 1];
  }
  for (; gid < lanes; gid += stride) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    const int p0 = np0, cnt = np1 - np0;
    if (gid + stride < lanes) {
      np0 = A.node_ptr[(gid + stride) / G];
      np1 = A.node_ptr[(gid + stride) / G + 1];
    }
    double sum[1][DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) sum[0][a] = 0;
    gather_rows<TM, DN, G, 1>(vals, A.node_cols, p0, cnt, g,
                              [&](int c, double (&xv)[1][DN]) {
#pragma unroll
                                for (int bb = 0; bb < DN; ++bb)
                                  xv[0][bb] = (double)z[(int64_t)c * DN + bb] +
                                              beta * (double)pold[(int64_t)c * DN + bb];
                              },
                              sum);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) sum[0][a] += __shfl_xor(sum[0][a], o, G);
    }
    if (g == 0) {
      double pn[DN];
#pragma unroll
      for (int a = 0; a < DN; ++a)
        pn[a] = (double)z[(int64_t)n * DN + a] + beta * (double)pold[(int64_t)n * DN + a];
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t r = (int64_t)n * DN + a;
        pnew[r] = (TV)pn[a];
        double qv = sum[0][a];
        if (A.corr) {
          const TM* cr = (const TM*)A.corr + (int64_t)n * DN * DN + a * DN;
#pragma unroll
          for (int b = 0; b < DN; ++b) qv += (double)cr[b] * pn[b];
        }
        if (A.known && A.known[r]) qv = 0;
        q[r] = (TV)qv;
        d += pn[a] * qv;
      }
    }
  }
  d = block_sum<double, BS>(d, sh);
  if (threadIdx.x == 0) part1[blk] = d;
}

// Timing harness of the step-1 kernel: scratch layout and the state of a
// running solve whose stopping tests never fire (previous iteration closed
// from nb2 partials of r.z = r.r = 1, beta = 1, tolerances 0, no pause).
struct Step1Bench {
  PcgState* st;
  double* part1;
  double* part2;
  double* P0;
  double* P1;
};

inline size_t step1_bench_bytes(int Nnode, int dof_n, int nb2max) {
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  return align256(sizeof(PcgState)) + align256(8 * (size_t)MAXB) + align256(16 * (size_t)nb2max) + 2 * vb;
}

inline Step1Bench step1_bench_carve(void* scratch, int Nnode, int dof_n, int nb2max) {
  char* c = (char*)scratch;
  Step1Bench b;
  b.st = (PcgState*)c; c += align256(sizeof(PcgState));
  b.part1 = (double*)c; c += align256(8 * (size_t)MAXB);
  b.part2 = (double*)c; c += align256(16 * (size_t)nb2max);
  b.P0 = (double*)c; c += align256(8 * (size_t)Nnode * dof_n);
  b.P1 = (double*)c;
  return b;
}

__device__ inline void step1_bench_state_body(PcgState* st, double* part2, int nb2, double* P0, int64_t n,
                                              int blk, int nblk) {
  for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n; i += (int64_t)nblk * blockDim.x) P0[i] = 0;
  if (blk == 0) {
    for (int i = threadIdx.x; i < nb2; i += blockDim.x) part2[2 * i] = part2[2 * i + 1] = 1.0 / nb2;
    if (threadIdx.x == 0) {
      st->rz = st->rz_prev = st->alpha = st->pq = st->rr = st->rrmax = st->rrmax_prev = 1;
      st->beta = 0;
      st->bb = st->xAx0 = 1;
      st->tolr2 = st->tolA2 = st->delta2 = 0;
      st->Sall = st->Sw = 0;
      for (int i = 0; i < RING; ++i) st->ring[i] = 0;
      st->head = 0;
      st->done = 0;
      st->iters = RING;
      st->nrepl = 0;
      st->fresh = 0;
      st->justrep = 0;
      st->par = 0;
      st->par_next = 1;
    }
  }
}

// Step 2: alpha = rz / pq ; x += alpha p ; r -= alpha q ; z = Minv r ;
// partials of r.z and r.r.  The first node's operands are loaded before the
// reduction that yields alpha.
template <typename TP, int DN>
__device__ inline void pcg_step2_body(int Nnode, const unsigned char* __restrict__ known,
                                      const TP* __restrict__ Minv, double* __restrict__ x, double* __restrict__ r,
                                      double* __restrict__ z, const double* __restrict__ P0,
                                      const double* __restrict__ P1, const double* __restrict__ q,
                                      const double* __restrict__ part1, int nb1, double* __restrict__ part2,
                                      PcgState* st, int blk, int nblk) {
  // Every load is issued before the state test, which follows the block
  // reduction (a launch after convergence does the loads and returns there).
  __shared__ double sh[BS / 64];
  __shared__ double sh2[BS / 64];
  __shared__ double bc;
  Partials<1> P1s;
  P1s.issue(part1, nb1);
  const int pn = st->par_next;
  const double* __restrict__ p = pn ? P1 : P0;
  int n = blk * BS + threadIdx.x;
  double pv[DN], qv[DN], rv[DN], xv[DN], Mv[DN * DN];
  if (n < Nnode) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      pv[a] = p[i];
      qv[a] = q[i];
      rv[a] = r[i];
      xv[a] = x[i];
    }
#pragma unroll
    for (int k = 0; k < DN * DN; ++k) Mv[k] = (double)Minv[(int64_t)n * DN * DN + k];
  }
  double pp1[1];
  P1s.sum(part1, nb1, pp1);
  const double ppq = block_sum<double, BS>(pp1[0], sh);
  if (threadIdx.x == 0) bc = ppq;
  __syncthreads();
  if (st->done) return;
  const double pq = bc;
  const double rz = st->rz;
  if (!(pq > 0) || !(rz > 0)) {
    if (blk == 0 && threadIdx.x == 0) st->done = (rz == 0) ? 1 : 2;  // r = 0 exactly: converged
    return;
  }
  const double al = rz / pq;
  if (blk == 0 && threadIdx.x == 0) {
    st->alpha = al;
    st->pq = pq;
    st->rz_prev = rz;
    st->rrmax_prev = st->rrmax;
    st->fresh = 0;
    st->justrep = 0;
    st->par = pn;
  }
  double s_rz = 0, s_rr = 0;
  if (n < Nnode) {
    double r2[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      x[i] = xv[a] + al * pv[a];
      r2[a] = rv[a] - al * qv[a];
      r[i] = r2[a];
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      double zz = 0;
#pragma unroll
      for (int b = 0; b < DN; ++b) zz += Mv[a * DN + b] * r2[b];
      if (known && known[(int64_t)n * DN + a]) zz = 0;
      z[(int64_t)n * DN + a] = zz;
      s_rz += r2[a] * zz;
      s_rr += r2[a] * r2[a];
    }
  }
  for (n += nblk * BS; n < Nnode; n += nblk * BS) {
    double r2[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      x[i] += al * p[i];
      r2[a] = r[i] - al * q[i];
      r[i] = r2[a];
    }
    apply_bj<TP, DN>(Minv, known, n, r2, zv);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = zv[a];
      s_rz += r2[a] * zv[a];
      s_rr += r2[a] * r2[a];
    }
  }
  s_rz = block_sum<double, BS>(s_rz, sh);
  s_rr = block_sum<double, BS>(s_rr, sh2);
  if (threadIdx.x == 0) {
    part2[2 * blk] = s_rz;
    part2[2 * blk + 1] = s_rr;
  }
}

// Residual replacement (mixed mode), first kernel: r = b - A64 x on the
// node-group grid (coalesced like the SpMV), masked rows.
template <int DN, int G>
__device__ inline void residual_body(const SpmvArgs& A64, const double* __restrict__ x,
                                     const double* __restrict__ b, double* __restrict__ r, int blk, int nblk) {
  const int64_t lanes = (int64_t)A64.Nnode * G;
  for (int64_t gid = (int64_t)blk * BS + threadIdx.x; gid < lanes; gid += (int64_t)nblk * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double dd;
    spmv_node<double, DN, G, 1>(A64, n, g, x, b, r, dd);
  }
}

// Residual replacement, second kernel (node grid, the layout step 1
// reduces): z = Minv r, partials of r.z and r.r; block 0 resumes the solve
// with the replaced residual.
template <typename TP, int DN>
__device__ inline void rebuild_z_body(int Nnode, const unsigned char* __restrict__ known,
                                      const TP* __restrict__ Minv, const double* __restrict__ r,
                                      double* __restrict__ z, double* __restrict__ part2, PcgState* st, int blk,
                                      int nblk) {
  __shared__ double shn[2][BS / 64];
  double red[2] = {0, 0};
  for (int n = blk * BS + threadIdx.x; n < Nnode; n += nblk * BS) {
    double rv[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) rv[a] = r[(int64_t)n * DN + a];
    apply_bj<TP, DN>(Minv, known, n, rv, zv);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = zv[a];
      red[0] += rv[a] * zv[a];
      red[1] += rv[a] * rv[a];
    }
  }
  block_sum_all<2, BS>(red, shn);
  if (threadIdx.x == 0) {
    part2[2 * blk] = red[0];
    part2[2 * blk + 1] = red[1];
  }
  if (blk == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->justrep = 1;
    st->nrepl += 1;
  }
}

// Workspace of one solve: five FP64 vectors, the partial arrays and the state.

struct PcgWork {
  double* r;
  double* z;
  double* P0;
  double* P1;
  double* q;
  double* part1;  // MAXB
  double* partI;  // MAXB
  double* part2;  // 2 MAXB
  PcgState* st;
  double* small;
};

static inline size_t pcg_workspace_size(int Nnode, int dof_n) {
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  return 5 * vb + align256(8 * MAXB) * 2 + align256(16 * MAXB) + align256(sizeof(PcgState)) + 256;
}

static inline PcgWork carve(void* ws, int Nnode, int dof_n) {
  char* c = (char*)ws;
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  PcgWork w;
  w.r = (double*)c; c += vb;
  w.z = (double*)c; c += vb;
  w.P0 = (double*)c; c += vb;
  w.P1 = (double*)c; c += vb;
  w.q = (double*)c; c += vb;
  w.part1 = (double*)c; c += align256(8 * MAXB);
  w.partI = (double*)c; c += align256(8 * MAXB);
  w.part2 = (double*)c; c += align256(16 * MAXB);
  w.st = (PcgState*)c; c += align256(sizeof(PcgState));
  w.small = (double*)c;
  return w;
}

// Host-side helpers shared by the single-system driver and the batch engine:
// initialisation of one solve (state, x0.A x0, r0, z0, p0, scalars) and its
// final true residual.  Defined in sparse.hip.
template <typename TP, int DN, int G>
int pcg_start(const SpmvArgs& A64, const TP* Minv, const double* b, const double* x, const PcgWork& w,
              double tol_r, double tol_A, double delta2, hipStream_t s);
template <int DN, int G>
void pcg_true_residual(const SpmvArgs& A64, const double* b, const double* x, const PcgWork& w, hipStream_t s);
// The same norm with the residual vector written to w.q (dead at a read-back).
template <int DN, int G>
void pcg_spec_residual(const SpmvArgs& A64, const double* b, const double* x, const PcgWork& w, hipStream_t s);
// Fills info[0..3] from a read-back state and the true residual norm^2.
int pcg_finish_info(const void* state, double rr_true, double* info);
// This thread's pinned buffer of the solvers' polls (null if the pinned
// allocation failed): the state and the speculative residual norm are two
// asynchronous copies followed by one synchronisation, where a copy into
// pageable memory would be a blocking staged copy of its own.  Slot 0 serves
// the solves that wait for their polls; slot 1 holds the read-back of a
// deferred pipelined solve between its start and its finish.  Defined in
// sparse.hip.
struct PollBuf {
  PcgState st;
  double rr;
};
PollBuf* poll_buffer(int slot = 0);
// Phases of the pipelined block-Jacobi solve (solve_pcgp): the whole solve,
// or its deferred form split into a start that only queues work and a
// finish that waits for it.
enum { PCGP_WHOLE = 0, PCGP_START = 1, PCGP_FINISH = 2 };

}  // namespace efea
