// Pipelined two-level PCG (Ghysels and Vanroose 2014, preconditioned
// pipelined CG) for latency-bound systems: two kernels per iteration
// instead of the classic solve's three.
//
// The classic two-level iteration needs three global synchronisations (the
// p.q reduction, the restriction P^T r feeding the dense coarse solve, the
// r.z reduction), hence three launches.  The pipelined recurrences
//   gamma = (r, u), delta = (w, u), m = M w, n = A m,
//   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
//   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z
// (u = M r and w = A u kept by recurrence) put both dot products and the
// restriction of w in the same pass as the product n = A m:
//   kernel A (aggregate grid plus one block): the extra block reduces the
//     previous kernel B's partials, forms alpha and beta, closes the
//     previous iteration's stopping-test term and decides whether to stop;
//     every aggregate block sums rc = P^T w from the previous kernel B's
//     per-chunk partials, forms zc = Ac^-1 rc for its aggregate and
//     m = B^-1 w + P zc on its nodes;
//   kernel B (chunks of aggregate-ordered nodes, node-group lanes): reads
//     alpha and beta, forms n = A m, the eight vector updates and the
//     partials of (r, u), (w, u), (r, r) and, per chunk (nodes of one
//     aggregate), of P^T w.
// Kernel B thus starts its product at once: no block of it waits for a
// reduction of the previous pass's partials.
// The stopping test is the classic solve's (delayed A-norm estimate from
// alpha_j (r_j, u_j), residual test), evaluated by kernel A's extra block
// from the partials of the last update: the next kernel B sees the decision
// before it updates anything.
// The recurrences let rounding errors grow (the gap between the recursive and
// the true residual; on the 360 x 39 cantilever the plain pipelined solve
// breaks down with a true relative residual of 1e-3), so every
// PIPE_REPLACE iterations the vectors are recomputed from x and p (residual
// replacement, Cools and Vanroose): r = b - A x, u = M r, w = A u, s = A p,
// q = M s, z = A q, six launches.
// Mixed precision (TM = float): the iterations' products and the block-Jacobi
// inverse use the FP32 copies (with the FP32 values' per-node row-sum
// correction); the start-up and replacement passes, every 32 iterations, use
// the FP64 matrix, so that r = b - A x is the FP64 residual at each
// replacement and the solve converges to the FP64 solution.
// EasyFEA solves directly (Simulations/Solvers.py:225-394); this changes the
// iteration's synchronisation pattern, not the solution (same stopping test).
#include "coarse_device.h"

namespace efea {

// Chunks: runs of at most BS / G nodes of one aggregate, in the aggregate
// order of agg_nodes, one per block of kernel B (nchunk <= MAXB).  Chunk c
// owns the BS / G positions c * BS / G + t; position j holds rows[3j] = its
// node (-1: none), rows[3j + 1], rows[3j + 2] = that node's row range in
// node_cols, so a lane finds its node and row with one load that depends on
// nothing but its block and thread index.  Chunk c writes its restriction
// partial to slot cslot[c] = a * maxch + (its index among aggregate a's
// chunks), so that aggregate a's restriction is the sum of slots a * maxch ..
// a * maxch + maxch - 1 (slots without a chunk stay zero).  first[a * BS + t]:
// node t of aggregate a in agg_nodes order for t < BS (-1 past its end), the
// coarse kernel's node lookup without the agg_ptr load.
struct PipeArgs {
  int nchunk;
  const int* rows;
  const int* cslot;
  int maxch;
  const int* first;
};

constexpr int PIPE_REPLACE = 64;       // FP64 iterations
constexpr int PIPE_REPLACE_MIXED = 32;  // FP32 products (64 let the recurrences drift into breakdowns)

struct PipeWork {
  double *r, *u, *w, *m, *z, *q, *s, *p;
  double* partB;   // 4 per block of kernel B: (r,u), (w,u), (r,r), x.A x; two halves of 4 MAXB
  double* partR;   // 4 MAXB: partials of the residual passes (read-backs, replacements)
  double* rcpart;  // NM per slot: P^T (w or r) of the chunk's nodes
  double* rcpart2; // NM per slot: P^T s (replacement)
  double* small;
  PcgState* st;
};

static inline size_t pcg2p_workspace_size(int Nnode, int dof_n, int nslot) {
  const int nm = dof_n == 1 ? 1 : (dof_n == 2 ? 3 : 6);
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  return 8 * vb + align256(64 * (size_t)MAXB) + align256(32 * (size_t)MAXB) +
         2 * align256(8 * (size_t)nslot * nm) + 256 + align256(sizeof(PcgState));
}

static inline PipeWork carve_pipe(void* ws, int Nnode, int dof_n, int nslot) {
  const int nm = dof_n == 1 ? 1 : (dof_n == 2 ? 3 : 6);
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  char* c = (char*)ws;
  PipeWork w;
  double** v[8] = {&w.r, &w.u, &w.w, &w.m, &w.z, &w.q, &w.s, &w.p};
  for (int k = 0; k < 8; ++k) {
    *v[k] = (double*)c;
    c += vb;
  }
  w.partB = (double*)c; c += align256(64 * (size_t)MAXB);
  w.partR = (double*)c; c += align256(32 * (size_t)MAXB);
  w.rcpart = (double*)c; c += align256(8 * (size_t)nslot * nm);
  w.rcpart2 = (double*)c; c += align256(8 * (size_t)nslot * nm);
  w.small = (double*)c; c += 256;
  w.st = (PcgState*)c;
  return w;
}

// The start-up and residual-replacement passes, grid-stride over chunks.
//   MODE 0: y = A x on the masked rows; r = b - y; partials x.y; rc = P^T r.
//   MODE 1: w = A u (masked); partials (r,u), (w,u), (r,r); rc = P^T w.
//   MODE 3: s = A p (masked); rc = P^T s.
//   MODE 4: z = A q (masked).
template <int DN, int G, int MODE>
__global__ void __launch_bounds__(BS)
k_pp_spmv(SpmvArgs A, PipeArgs PA, const double* __restrict__ rel, const double* __restrict__ vin, double* __restrict__ x, const double* __restrict__ b, PipeWork W) {
  constexpr int NM = Modes<DN>::NM;  // a chunk holds BS / G nodes at most
  __shared__ double shn[NM][BS / 64];
  __shared__ double shp[4][BS / 64];
  double acc[4] = {0, 0, 0, 0};
  for (int c = xcd_block(blockIdx.x, gridDim.x); c < PA.nchunk; c += gridDim.x) {  // XCD-contiguous chunks
    const int g = (int)threadIdx.x % G;
    const int* pos = PA.rows + 3 * ((int64_t)c * (BS / G) + (int)threadIdx.x / G);
    int n = pos[0], p0 = pos[1], p1 = pos[2];
    const bool have = n >= 0;
    if (!have) n = p0 = p1 = 0;
    double sum[1][DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) sum[0][a] = 0;
    // own-row operands, loaded with the gathers
    double own[2][DN];
    if (have && g == 0) {
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t i = (int64_t)n * DN + a;
        if (MODE == 0) {
          own[0][a] = x[i];
          own[1][a] = b[i];
        } else if (MODE == 1) {
          own[0][a] = vin[i];  // u
          own[1][a] = W.r[i];
        }
      }
    }
    if (have) {
      gather_rows<double, DN, G, 1>((const double*)A.vals, A.node_cols, p0, p1 - p0, g,
                                    [&](int col, double (&xv)[1][DN]) {
#pragma unroll
                                      for (int bb = 0; bb < DN; ++bb) xv[0][bb] = vin[(int64_t)col * DN + bb];
                                    },
                                    sum);
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) sum[0][a] += __shfl_xor(sum[0][a], o, G);
    }
    double rcv[NM];
#pragma unroll
    for (int mm = 0; mm < NM; ++mm) rcv[mm] = 0;
    if (have && g == 0) {
      const double* rn = rel + 3 * (int64_t)n;
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t i = (int64_t)n * DN + a;
        const bool kn = A.known && A.known[i];
        const double y = kn ? 0.0 : sum[0][a];
        double rest;  // the vector restricted by P^T in this pass
        if (MODE == 3) {
          W.s[i] = y;
          rest = y;
        } else if (MODE == 4) {
          W.z[i] = y;
          rest = 0;
        } else if (MODE == 0) {
          acc[3] += own[0][a] * y;  // x.A x over the unknown rows
          const double r = kn ? 0.0 : own[1][a] - y;
          W.r[i] = r;
          rest = r;
        } else {
          W.w[i] = y;
          acc[0] += own[1][a] * own[0][a];  // (r, u)
          acc[1] += y * own[0][a];          // (w, u)
          acc[2] += own[1][a] * own[1][a];  // (r, r)
          rest = y;
        }
#pragma unroll
        for (int mm = 0; mm < NM; ++mm) rcv[mm] += pmode<DN>(rn, a, mm) * rest;  // rest = 0 at known DOFs
      }
    }
    if (MODE != 4) {
      block_sum_all<NM, BS>(rcv, shn);
      if (threadIdx.x == 0) {
#pragma unroll
        for (int mm = 0; mm < NM; ++mm) W.rcpart[(int64_t)PA.cslot[c] * NM + mm] = rcv[mm];
      }
    }
  }
  if (MODE <= 1) {
    block_sum_all<4, BS>(acc, shp);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) W.partB[4 * blockIdx.x + q] = acc[q];
    }
  }
}

// Kernel B of an iteration, one chunk per block: alpha, beta and the stop
// decision come from kernel A's extra block (three scalar loads), so the
// chain position -> row -> product starts at once; the own rows' nine
// operands are loaded with the gather.  One block reduction ends it.
template <int DN, int G, typename TM>
__global__ void __launch_bounds__(BS)
k_pp_iter(SpmvArgs A, PipeArgs PA, const double* __restrict__ rel, const double* __restrict__ m,
          double* __restrict__ x, PipeWork W) {
  constexpr int NM = Modes<DN>::NM;
  __shared__ double shr[NM + 3][BS / 64];
  PcgState* st = W.st;
  // logical chunk: each XCD takes one contiguous range of chunks, so the m
  // entries a chunk gathers from its neighbours sit mostly in that XCD's L2
  const int c = xcd_block(blockIdx.x, gridDim.x);
  // the lane's position (node and row range) is read together with the
  // solver state: the exit test uses all three (n < -1 and p1 < p0 never
  // hold), so neither the state read nor the row range adds a round trip
  const int g = (int)threadIdx.x % G;
  const int* pos = PA.rows + 3 * ((int64_t)c * (BS / G) + (int)threadIdx.x / G);
  int n = pos[0], p0 = pos[1], p1 = pos[2];
  const int done = st->done;
  if (done != 0 || n < -1 || p1 < p0) {  // done: decided by kernel A; recorded for the next kernel A
    if (done != 0 && c == 0 && threadIdx.x == 0) st->stop = 1;
    return;
  }
  const bool first = st->justrep != 0;  // the first iteration: beta = 0, no old directions
  const double alpha = st->alpha, beta = st->beta;
  const bool have = n >= 0;
  if (!have) n = p0 = p1 = 0;
  // The own rows' operands, mask and FP32 row-sum correction are loaded here,
  // by every lane (a wave's load costs one instruction whichever lanes are
  // active; the G lanes of a node read one address), with no branch around
  // them: the gather then starts while they are in flight.  Loaded under the
  // leading lane's branch, the gather waited for them where the branch ends;
  // loaded where they are used (the mask after the gather's reduction, the
  // second DOF's after the first DOF's stores), they were dependent memory
  // round trips at the end of every wave.  Without a mask the flag is read
  // from m and ignored; the correction exists only with FP32 values.
  double own[9][DN], relf[3], cor[DN][DN];
  int kb[DN];
  const bool hk = A.known != nullptr;
  const bool lead = have && g == 0;
  {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      own[0][a] = m[i];
      own[1][a] = W.w[i];
      own[2][a] = W.u[i];
      own[3][a] = W.r[i];
      own[4][a] = x[i];
      own[5][a] = W.z[i];
      own[6][a] = W.q[i];
      own[7][a] = W.s[i];
      own[8][a] = W.p[i];
      kb[a] = *(hk ? A.known + i : (const unsigned char*)m);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) relf[d] = rel[3 * (int64_t)n + d];
    if constexpr (sizeof(TM) == 4) {
      const TM* cr = (const TM*)A.corr + (int64_t)n * DN * DN;
#pragma unroll
      for (int a = 0; a < DN; ++a)
#pragma unroll
        for (int bb = 0; bb < DN; ++bb) cor[a][bb] = (double)cr[a * DN + bb];
    }
  }
  double sum[1][DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) sum[0][a] = 0;
  if (have)
    gather_rows<TM, DN, G, 1>((const TM*)A.vals, A.node_cols, p0, p1 - p0, g,
                              [&](int col, double (&xv)[1][DN]) {
#pragma unroll
                                for (int bb = 0; bb < DN; ++bb) xv[0][bb] = m[(int64_t)col * DN + bb];
                              },
                              sum);
#pragma unroll
  for (int a = 0; a < DN; ++a) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) sum[0][a] += __shfl_xor(sum[0][a], o, G);
  }
  if constexpr (sizeof(TM) == 4) {  // FP32 values: the rounding's row sums on the own rows
    if (lead) {
#pragma unroll
      for (int a = 0; a < DN; ++a) {
#pragma unroll
        for (int bb = 0; bb < DN; ++bb) sum[0][a] += cor[a][bb] * own[0][bb];
      }
    }
  }
  double red[NM + 3];
#pragma unroll
  for (int q = 0; q < NM + 3; ++q) red[q] = 0;
  if (lead) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      const double y = (hk && kb[a] != 0) ? 0.0 : sum[0][a];
      const double mv = own[0][a], wv = own[1][a], uv = own[2][a];
      const double zv = y + (first ? 0.0 : beta * own[5][a]);
      const double qv = mv + (first ? 0.0 : beta * own[6][a]);
      const double sv = wv + (first ? 0.0 : beta * own[7][a]);
      const double pv = uv + (first ? 0.0 : beta * own[8][a]);
      const double rnw = own[3][a] - alpha * sv;
      const double un = uv - alpha * qv;
      const double wn = wv - alpha * zv;
      W.z[i] = zv;
      W.q[i] = qv;
      W.s[i] = sv;
      W.p[i] = pv;
      x[i] = own[4][a] + alpha * pv;
      W.r[i] = rnw;
      W.u[i] = un;
      W.w[i] = wn;
      red[0] += rnw * un;
      red[1] += wn * un;
      red[2] += rnw * rnw;
#pragma unroll
      for (int mm = 0; mm < NM; ++mm) red[3 + mm] += pmode<DN>(relf, a, mm) * wn;  // w = 0 at known DOFs
    }
  }
  block_sum_all<NM + 3, BS>(red, shr);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) W.partB[4 * c + q] = red[q];
    W.partB[4 * c + 3] = 0;
#pragma unroll
    for (int mm = 0; mm < NM; ++mm) W.rcpart[(int64_t)PA.cslot[c] * NM + mm] = red[3 + mm];
  }
}

// The scalar step of an iteration (kernel A's extra block): the previous
// kernel B's partials summed in a fixed order give gamma = (r, u),
// delta = (w, u) and (r, r); beta = gamma / gamma_old and
// alpha = gamma / (delta - beta gamma / alpha_old) (alpha = gamma / delta on
// the first iteration); the previous iteration's term alpha_old gamma_old
// enters the delayed A-norm estimate; then the stopping test.
__device__ void pp_scalars(PcgState* st, const double* __restrict__ partPrev, int nbPrev) {
  __shared__ double shp[3][BS / 64];
  Partials<4> P;
  P.issue(partPrev, nbPrev);  // in flight with the state read
  if (st->done) return;
  double pr4[4];
  P.sum(partPrev, nbPrev, pr4);
  double pr[3] = {pr4[0], pr4[1], pr4[2]};
  block_sum_all<3, BS>(pr, shp);
  if (threadIdx.x != 0) return;
  const bool first = st->fresh != 0;
  const double gam = pr[0], del = pr[1], rr = pr[2];
  double alpha, beta;
  if (first) {
    beta = 0;
    alpha = gam / del;
  } else {
    const double a_prev = st->alpha, g_prev = st->rz;
    beta = gam / g_prev;
    alpha = gam / (del - beta * gam / a_prev);
    const double term = a_prev * g_prev;  // closes the previous iteration
    st->Sall += term;
    st->ring[st->head] = term;
    st->head = (st->head + 1) % RING;
    double Sw = 0;
    for (int i = 0; i < RING; ++i) Sw += st->ring[i];
    st->Sw = Sw;
    st->iters += 1;
  }
  st->justrep = first;
  st->fresh = 0;
  st->rr = rr;
  st->rz = gam;
  st->alpha = alpha;
  st->beta = beta;
  st->pq = del;
  const double scale = st->Sall > st->xAx0 ? st->Sall : st->xAx0;
  const bool finite = isfinite(rr) && isfinite(gam) && isfinite(del) && isfinite(st->Sw);
  if (finite && ((st->iters >= RING && st->Sw <= st->tolA2 * scale) || rr <= st->tolr2 * st->bb || gam == 0))
    st->done = 1;  // converged (gam = 0: r = 0 exactly), whatever the next step's alpha
  else if (!finite || !(alpha > 0) || !isfinite(alpha))
    st->done = 2;  // breakdown: the caller redoes the solve with the classic iteration
}

// Kernel A: out = B^-1 vin + P Ac^-1 rc on the block's aggregate.  Every
// thread takes columns c = t, t + BS, ... of the aggregate's NM rows of
// Ac^-1 with rc_c summed from the aggregate slots of c (fixed order); the
// node operands and the rows are loaded before the partials.  MODE 1
// (iterations) runs one more block, the scalar step; its aggregate blocks
// skip the pass once a kernel B has seen the solve stop.
template <typename TP, int DN, int MODE>
__global__ void __launch_bounds__(BS)
k_pp_coarse(CoarseArgs C, PipeArgs PA, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
            const double* __restrict__ vin, double* __restrict__ out, const double* __restrict__ rcpart,
            PcgState* st, const double* __restrict__ partPrev, int nbPrev) {
  constexpr int NM = Modes<DN>::NM;
  constexpr int U = 4;  // columns per pass, loads issued together
  __shared__ double shz[NM][BS / 64];
  const int nc = C.nagg * NM;
  const int blk = blockIdx.x;
  if (MODE == 1 && blk == C.nagg) {
    pp_scalars(st, partPrev, nbPrev);
    return;
  }
  // the thread's node is read together with the stop flag (the test uses
  // both; nf < -1 never holds)
  int nf = PA.first[(int64_t)blk * BS + threadIdx.x];
  if (MODE == 1 && (st->stop != 0 || nf < -1)) return;
  const bool have = nf >= 0;
  if (!have) nf = 0;
  // Every load below is issued by every thread, with clamped indices and no
  // branch around it, so that they are all in flight together: under a
  // branch each group's result was waited for where its branch ended (a
  // dependent round trip per column group and per slot group), and the
  // mask, read where it is used, cost two more after the block reduction.
  double vf[DN], Mf[DN * DN], relf[3];
  int kb[DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) {
    vf[a] = vin[(int64_t)nf * DN + a];
    kb[a] = *(known ? known + (int64_t)nf * DN + a : (const unsigned char*)PA.first);  // no mask: ignored
  }
#pragma unroll
  for (int e = 0; e < DN * DN; ++e) Mf[e] = (double)Minv[(int64_t)nf * DN * DN + e];
#pragma unroll
  for (int d = 0; d < 3; ++d) relf[d] = C.rel[3 * (int64_t)nf + d];
  const int k0 = C.agg_ptr[blk], k1 = C.agg_ptr[blk + 1];
  const double* __restrict__ rows = (const double*)C.Ainv + (int64_t)(blk * NM) * C.ld;
  const int mx = PA.maxch;
  double zc[NM];
#pragma unroll
  for (int mm = 0; mm < NM; ++mm) zc[mm] = 0;
  for (int c0 = threadIdx.x; c0 < nc; c0 += U * BS) {
    double av[U][NM], sl[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + u * BS;
      const int cc = c < nc ? c : 0;
#pragma unroll
      for (int mm = 0; mm < NM; ++mm) av[u][mm] = rows[(int64_t)mm * C.ld + cc];
      const int64_t base = (int64_t)(cc / NM) * mx * NM + (cc % NM);
      // the first four slots (a slot past the aggregate's reads its first)
#pragma unroll
      for (int j = 0; j < 4; ++j) sl[u][j] = rcpart[base + (int64_t)(j < mx ? j : 0) * NM];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + u * BS;
      double rv = 0;
      if (c < nc) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j < mx) rv += sl[u][j];
        const int64_t base = (int64_t)(c / NM) * mx * NM + (c % NM);
        for (int j = 4; j < mx; ++j) rv += rcpart[base + (int64_t)j * NM];
      } else {
#pragma unroll
        for (int mm = 0; mm < NM; ++mm) av[u][mm] = 0.0;
      }
#pragma unroll
      for (int mm = 0; mm < NM; ++mm) zc[mm] += av[u][mm] * rv;
    }
  }
  block_sum_all<NM, BS>(zc, shz);
  if (have) {
    double zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      double s = 0;
#pragma unroll
      for (int b = 0; b < DN; ++b) s += Mf[a * DN + b] * vf[b];
      zv[a] = s;
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)nf * DN + a;
      double cv = 0;
#pragma unroll
      for (int mm = 0; mm < NM; ++mm) cv += pmode<DN>(relf, a, mm) * zc[mm];
      out[i] = (known && kb[a] != 0) ? 0.0 : zv[a] + cv;
    }
  }
  for (int k = k0 + BS + (int)threadIdx.x; k < k1; k += BS) {  // aggregates of more than BS nodes
    const int n = C.agg_nodes[k];
    double v[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) v[a] = vin[(int64_t)n * DN + a];
    apply_bj<TP, DN>(Minv, nullptr, n, v, zv);
    const double* rn = C.rel + 3 * (int64_t)n;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      double cv = 0;
#pragma unroll
      for (int mm = 0; mm < NM; ++mm) cv += pmode<DN>(rn, a, mm) * zc[mm];
      out[i] = (known && known[i]) ? 0.0 : zv[a] + cv;
    }
  }
}

__global__ void __launch_bounds__(BS) k_pp_zero(int64_t n, double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    a[i] = 0;
    b[i] = 0;
  }
}

__global__ void k_pp_state(PcgState* st, double tolr2, double tolA2) {
  if (threadIdx.x == 0) {
    st->rz = st->rz_prev = st->alpha = st->beta = st->pq = 0;
    st->rr = st->rrmax = st->rrmax_prev = st->bb = 0;
    st->tolr2 = tolr2;
    st->tolA2 = tolA2;
    st->delta2 = 0;
    st->Sall = st->Sw = st->xAx0 = 0;
    for (int i = 0; i < RING; ++i) st->ring[i] = 0;
    st->head = st->done = st->iters = st->nrepl = 0;
    st->fresh = 1;
    st->justrep = 0;
    st->par = 0;
    st->par_next = 1;
    st->stop = 0;
  }
}

// After the start-up passes: x0.A x0, the initial residual norm.
__global__ void __launch_bounds__(BS) k_pp_init_scalars(const double* __restrict__ partX, int nbX,
                                                        const double* __restrict__ partW, int nbW, PcgState* st) {
  __shared__ double sh[BS / 64];
  const double xax = reduce_all(partX, nbX, 4, 3, sh);
  const double rr = reduce_all(partW, nbW, 4, 2, sh);
  if (threadIdx.x == 0) {
    st->xAx0 = xax;
    st->bb = rr;
    st->rr = rr;
    st->rrmax = rr;
    st->done = (rr == 0) ? 1 : 0;
  }
}

size_t pcg2p_workspace_bytes(int Nnode, int dof_n, int nslot) { return pcg2p_workspace_size(Nnode, dof_n, nslot); }

// A: the FP64 matrix (start-up, replacements, residuals); Ait: the
// iterations' matrix (A itself, or the FP32 copy with its correction);
// Minv: TP (the job's block-Jacobi inverse, FP32 with FP32 values).
// phase PCGP_WHOLE runs the solve to its end; PCGP_START queues the start-up,
// the first first_check iterations, the residual estimate and the state
// read-back into poll slot 1 (or hs) and returns without waiting;
// PCGP_FINISH waits for that read-back and continues the same loop if the
// stopping test had not been met (*resumed = 1 then).  The stopping test is
// decided on the device and the replacements by the iteration count, so the
// split solve returns what the whole one does.
template <int DN, int G, typename TM, typename TP>
static int pcg2p_run(const SpmvArgs& A, const SpmvArgs& Ait, const PipeArgs& PA, const TP* Minv,
                     const CoarseArgs& C, const double* b, double* x, void* ws, double tol_r, double tol_A,
                     int maxit, int check_every, int first_check, hipStream_t s, double* info, PcgState* hs,
                     int phase, int* resumed) {
  const int Nnode = A.Nnode;
  const int replace_every = sizeof(TM) == 4 ? PIPE_REPLACE_MIXED : PIPE_REPLACE;
  constexpr int NM = Modes<DN>::NM;
  const int nslot = C.nagg * PA.maxch;
  const PipeWork W = carve_pipe(ws, Nnode, DN, nslot);
  const unsigned nbB = (unsigned)PA.nchunk;  // one chunk per block (nchunk <= MAXB)
  const unsigned nbA = (unsigned)C.nagg;
  // partials: start-up pass 0 leaves x.A x in partX; the (r,u), (w,u), (r,r)
  // partials of start-up pass 1, of every kernel B and of the replacements'
  // w = A u pass go to partW, which the next kernel A's extra block reads
  double* partX = W.partB;
  double* partW = W.partB + 4 * MAXB;
  PipeWork WX = W, WW = W;
  WX.partB = partX;
  WW.partB = partW;
  if (phase != PCGP_WHOLE && first_check < 1) {
    set_error("pcg2p: a deferred solve needs first_check >= 1 (got %d)", first_check);
    return EFEA_ERR_ARG;
  }
  PollBuf* pb = poll_buffer(phase == PCGP_WHOLE ? 0 : 1);
  if (!pb) {
    set_error("pcg2p: pinned host allocation failed");
    return EFEA_ERR_HIP;
  }
  if (!hs) hs = &pb->st;
  int launched = 0, pass = 1;
  bool spec = false;
  bool wait_only = false;  // finish: the start's read-back is in flight
  if (phase == PCGP_FINISH) {
    launched = first_check;
    pass = 1 + first_check;
    spec = true;
    wait_only = true;
  } else {
    // slots without a chunk are never written: zero both restriction arrays
    // (at the start only: a finish continues from the partials in them)
    hipLaunchKernelGGL(k_pp_zero, dim3(grid_for((int64_t)nslot * NM)), dim3(BS), 0, s, (int64_t)nslot * NM,
                       W.rcpart, W.rcpart2);
    hipLaunchKernelGGL(k_pp_state, dim3(1), dim3(64), 0, s, W.st, tol_r * tol_r, tol_A * tol_A);
    hipLaunchKernelGGL((k_pp_spmv<DN, G, 0>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                       (const double*)x, x, b, WX);
    hipLaunchKernelGGL((k_pp_coarse<TP, DN, 0>), dim3(nbA), dim3(BS), 0, s, C, PA, A.known, Minv,
                       (const double*)W.r, W.u, (const double*)W.rcpart, W.st, (const double*)nullptr, 0);
    hipLaunchKernelGGL((k_pp_spmv<DN, G, 1>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                       (const double*)W.u, x, b, WW);
    hipLaunchKernelGGL(k_pp_init_scalars, dim3(1), dim3(BS), 0, s, (const double*)partX, (int)nbB,
                       (const double*)partW, (int)nbB, W.st);
    if (int e = check_launch("pcg2p_start")) return e;
  }
  PcgWork fw{};  // residual passes of the read-backs: m is dead there (kernel A rewrites it)
  fw.r = W.m;
  fw.q = W.m;
  fw.part1 = W.partR;
  fw.small = W.small;
  while (true) {
    if (wait_only) {
      wait_only = false;
      if (hipStreamSynchronize(s) != hipSuccess) {
        set_error("pcg2p: state readback failed");
        return EFEA_ERR_HIP;
      }
      if (hs->done || launched >= maxit) break;
      if (resumed) *resumed = 1;
    }
    const int chunk = launched == 0 && first_check > 0 ? first_check : check_every;
    for (int k = 0; k < chunk; ++k) {
      if (pass > 1 && (pass - 1) % replace_every == 0) {
        // residual replacement: r = b - A x, u = M r, w = A u (with the
        // partials the next pass reads), s = A p, q = M s, z = A q
        PipeWork Wr = W;
        Wr.partB = W.partR;
        hipLaunchKernelGGL((k_pp_spmv<DN, G, 0>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                           (const double*)x, x, b, Wr);
        hipLaunchKernelGGL((k_pp_coarse<TP, DN, 0>), dim3(nbA), dim3(BS), 0, s, C, PA, A.known, Minv,
                           (const double*)W.r, W.u, (const double*)W.rcpart, W.st, (const double*)nullptr, 0);
        hipLaunchKernelGGL((k_pp_spmv<DN, G, 1>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                           (const double*)W.u, x, b, WW);
        Wr.rcpart = W.rcpart2;
        hipLaunchKernelGGL((k_pp_spmv<DN, G, 3>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                           (const double*)W.p, x, b, Wr);
        hipLaunchKernelGGL((k_pp_coarse<TP, DN, 0>), dim3(nbA), dim3(BS), 0, s, C, PA, A.known, Minv,
                           (const double*)W.s, W.q, (const double*)W.rcpart2, W.st, (const double*)nullptr, 0);
        hipLaunchKernelGGL((k_pp_spmv<DN, G, 4>), dim3(nbB), dim3(BS), 0, s, A, PA, C.rel,
                           (const double*)W.q, x, b, Wr);
      }
      // kernel A: aggregate blocks plus the scalar step's block
      hipLaunchKernelGGL((k_pp_coarse<TP, DN, 1>), dim3(nbA + 1), dim3(BS), 0, s, C, PA, A.known, Minv,
                         (const double*)W.w, W.m, (const double*)W.rcpart, W.st, (const double*)partW, (int)nbB);
      hipLaunchKernelGGL((k_pp_iter<DN, G, TM>), dim3(nbB), dim3(BS), 0, s, Ait, PA, C.rel,
                         (const double*)W.m, x, WW);
      ++pass;
    }
    launched += chunk;
    spec = first_check > 0;
    if (spec) {
      pcg_spec_residual<DN, G>(A, b, x, fw, s);
      if (hipMemcpyAsync(&pb->rr, W.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("pcg2p: residual readback failed");
        return EFEA_ERR_HIP;
      }
    }
    if (int e = check_launch("pcg2p_iter")) return e;
    if (hipMemcpyAsync(hs, W.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_error("pcg2p: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (phase == PCGP_START) return 0;
    if (hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg2p: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (hs->done || launched >= maxit) break;
  }
  const PcgState fin = *hs;
  if (!spec) {  // else the last read-back carried the residual
    pcg_true_residual<DN, G>(A, b, x, fw, s);
    if (hipMemcpyAsync(&pb->rr, W.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg2p: residual readback failed");
      return EFEA_ERR_HIP;
    }
  }
  return pcg_finish_info(&fin, pb->rr, info);
}

int solve_pcg2p(int precision, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
                const double* vals64, const float* vals32, const float* corr32, const unsigned char* known,
                const void* Minv, int nagg, const int* agg_ptr, const int* agg_nodes, const double* rel,
                const double* Ainv, int ld, int nchunk, const int* rows, const int* cslot, int maxch,
                const int* first, const double* b, double* x, void* ws, double tol_r, double tol_A, int maxit,
                int check_every, int first_check, hipStream_t s, double* info, void* host_state_buf, int phase,
                int* resumed) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals64, known, nullptr, 0};
  SpmvArgs A32{Nnode, node_ptr, node_cols, vals32, known, corr32, 0};
  CoarseArgs C{nagg, agg_ptr, agg_nodes, rel, Ainv, ld, nullptr};
  if (resumed) *resumed = 0;
  PipeArgs PA{nchunk, rows, cslot, maxch, first};
  if (check_every < 1) check_every = 1;
  PcgState* hs = (PcgState*)host_state_buf;  // null: this thread's pinned poll buffer
#define RUNP(D, GV)                                                                                              \
  if (dof_n == D && group == GV) {                                                                              \
    if (precision == 32)                                                                                        \
      return pcg2p_run<D, GV, float, float>(A, A32, PA, (const float*)Minv, C, b, x, ws, tol_r, tol_A, maxit,   \
                                            check_every, first_check, s, info, hs, phase, resumed);             \
    return pcg2p_run<D, GV, double, double>(A, A, PA, (const double*)Minv, C, b, x, ws, tol_r, tol_A, maxit,    \
                                            check_every, first_check, s, info, hs, phase, resumed);             \
  }
  EFEA_GROUPS(RUNP)
#undef RUNP
  set_error("solve_pcg2p: unsupported dof_n %d / group %d", dof_n, group);
  return EFEA_ERR_ARG;
}

// Timing harness of the product pass k_pp_iter (the pipelined solve's
// dominant kernel): its grid, chunks, matrix (FP64, or FP32 with the row-sum
// correction), mask and aggregate coordinates, in the state of a running
// solve (an iteration past the first: the beta terms are formed), on zeroed
// vectors (the timing does not depend on the values).  Launches: four zero
// fills, two state kernels, then reps product passes back to back.
__global__ void k_pp_bench_state(PcgState* st) {
  if (threadIdx.x == 0) {
    st->done = 0;
    st->stop = 0;
    st->justrep = 0;
    st->alpha = 0.5;
    st->beta = 0.25;
  }
}

template <int DN, int G, typename TM>
static int pp_iter_bench_run(const SpmvArgs& Ait, const PipeArgs& PA, const double* rel,
                             int nagg, double* x, void* ws, int reps, hipStream_t s) {
  const PipeWork W = carve_pipe(ws, Ait.Nnode, DN, nagg * PA.maxch);
  const int64_t nd = (int64_t)Ait.Nnode * DN;
  const unsigned gz = grid_for(nd);
  hipLaunchKernelGGL(k_pp_zero, dim3(gz), dim3(BS), 0, s, nd, W.r, W.u);
  hipLaunchKernelGGL(k_pp_zero, dim3(gz), dim3(BS), 0, s, nd, W.w, W.m);
  hipLaunchKernelGGL(k_pp_zero, dim3(gz), dim3(BS), 0, s, nd, W.z, W.q);
  hipLaunchKernelGGL(k_pp_zero, dim3(gz), dim3(BS), 0, s, nd, W.s, W.p);
  hipLaunchKernelGGL(k_pp_zero, dim3(gz), dim3(BS), 0, s, nd, x, W.r);
  hipLaunchKernelGGL(k_pp_state, dim3(1), dim3(64), 0, s, W.st, 0.0, 0.0);
  hipLaunchKernelGGL(k_pp_bench_state, dim3(1), dim3(64), 0, s, W.st);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL((k_pp_iter<DN, G, TM>), dim3((unsigned)PA.nchunk), dim3(BS), 0, s, Ait, PA, rel,
                       (const double*)W.m, x, W);
  return check_launch("pp_iter_bench");
}

int pp_iter_bench(int precision, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
                  const double* vals64, const float* vals32, const float* corr32, const unsigned char* known,
                  int nagg, const double* rel, int nchunk, const int* rows, const int* cslot, int maxch, double* x,
                  void* ws, int reps, hipStream_t s) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals64, known, nullptr, 0};
  SpmvArgs A32{Nnode, node_ptr, node_cols, vals32, known, corr32, 0};
  PipeArgs PA{nchunk, rows, cslot, maxch, nullptr};
#define RUNB(D, GV)                                                                                         \
  if (dof_n == D && group == GV) {                                                                         \
    if (precision == 32) return pp_iter_bench_run<D, GV, float>(A32, PA, rel, nagg, x, ws, reps, s); \
    return pp_iter_bench_run<D, GV, double>(A, PA, rel, nagg, x, ws, reps, s);                  \
  }
  EFEA_GROUPS(RUNB)
#undef RUNB
  set_error("pp_iter_bench: unsupported dof_n %d / group %d", dof_n, group);
  return EFEA_ERR_ARG;
}

}  // namespace efea
