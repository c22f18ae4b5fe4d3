// Pipelined block-Jacobi PCG, one launch per iteration.
//
// Ghysels and Vanroose's pipelined preconditioned CG keeps, besides x, r and
// u = M r, the recurrences w = A u, m = M w, n = A m and the directions
// p, s = A p, q = M s, z = A q, so that an iteration needs one product
// (n = A m) and one global reduction ((r, u), (w, u), (r, r)).  With a
// node-local preconditioner (the nodal block-Jacobi inverse) the new m of a
// node needs only that node's new w, so the launch that updates w also forms
// the m the next launch gathers: the whole iteration is one kernel.  Each
// launch reduces the previous launch's partials redundantly in every block,
// in a fixed order (every block gets the same alpha, beta and stop decision
// bitwise); the solver state and m are double-buffered by launch parity, so
// no block reads what another block of the same launch writes.
//
// The classic iteration (sparse.hip) needs two launches per iteration; the
// solution, the stopping test (Strakos-Tichy delayed A-norm estimate plus
// the residual test) and the reported info are the same.  Rounding errors in
// the recurrences grow, so every 64 iterations (32 with FP32 values; the
// caller may choose) the vectors are recomputed from x and p with the FP64
// matrix (residual replacement, two launches); a solve that still breaks
// down is reported as such and the caller redoes it with the classic
// iteration.  With FP32 values the recursive residual follows the rounded
// operator and falls below the FP64 residual b - A x, which stalls at the
// rounding level: the replaced gamma = (r, u) then jumps (by 300x on a
// damage system of the notched square), and alpha = gamma / (delta - beta
// gamma / alpha_old), which assumes a continuous recurrence, turns negative.
// Those replacements therefore restart the recurrence (beta = 0: p = u,
// s = w, q = m, z = n), as classic CG is restarted from the true residual;
// the A-norm estimate still closes the iteration before the restart.
//
// It replaces the same call as efea_solve_pcg: __Solver_1
// (Simulations/Solvers.py:502-553) with _Solve_Axb's backends
// (Solvers.py:225-394).
#include "pcg_device.h"

namespace efea {

constexpr int PB_REPLACE = 64;        // FP64 values
constexpr int PB_REPLACE_MIXED = 32;  // FP32 values in the iterations' products

struct PbWork {
  double *r, *u, *w, *z, *q, *s, *p;
  double* t;      // scratch of the start-up product and of the read-backs' residual
  double* m[2];   // m = M w, written by launch i into m[(i + 1) & 1]
  double* part[2];  // 4 MAXB each: (r,u), (w,u), (r,r), x.A x per block
  double* partR;  // MAXB: residual partials of the read-backs
  PcgState* st[2];
  double* small;
};

static inline size_t pcgp_workspace_size(int Nnode, int dof_n) {
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  return 10 * vb + 2 * align256(32 * MAXB) + align256(8 * MAXB) + 2 * align256(sizeof(PcgState)) + 256;
}

static inline PbWork carve_pb(void* ws, int Nnode, int dof_n) {
  char* c = (char*)ws;
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  PbWork W;
  double** vecs[] = {&W.r, &W.u, &W.w, &W.z, &W.q, &W.s, &W.p, &W.t, &W.m[0], &W.m[1]};
  for (double** v : vecs) {
    *v = (double*)c;
    c += vb;
  }
  for (int k = 0; k < 2; ++k) {
    W.part[k] = (double*)c;
    c += align256(32 * MAXB);
  }
  W.partR = (double*)c;
  c += align256(8 * MAXB);
  for (int k = 0; k < 2; ++k) {
    W.st[k] = (PcgState*)c;
    c += align256(sizeof(PcgState));
  }
  W.small = (double*)c;
  return W;
}

// The scalar step of an iteration, computed by every thread from the state
// the previous launch left and the reduced partials gamma = (r, u),
// delta = (w, u), rr = (r, r): beta = gamma / gamma_old and
// alpha = gamma / (delta - beta gamma / alpha_old) (alpha = gamma / delta on
// the first iteration); the previous iteration's alpha_old gamma_old enters
// the delayed A-norm estimate; then the stopping test.
struct PbStep {
  double alpha, beta, Sall, Sw, term;
  int iters, head, done;
  bool first;   // beta = 0: the directions start afresh (first iteration or a restart)
  bool closed;  // a previous iteration was closed (its term entered the estimate)
};

__device__ inline PbStep pb_step(const PcgState* si, double gam, double del, double rr) {
  PbStep o;
  o.closed = si->fresh == 0;
  o.first = !o.closed || si->justrep != 0;
  o.Sall = si->Sall;
  o.iters = si->iters;
  o.head = si->head;
  o.term = 0;
  o.Sw = si->Sw;
  if (o.closed) {
    const double a_prev = si->alpha, g_prev = si->rz;
    o.term = a_prev * g_prev;  // closes the previous iteration
    o.Sall += o.term;
    double Sw = 0;  // the window summed afresh: the new term replaces the oldest
    for (int i = 0; i < RING; ++i) Sw += (i == o.head) ? o.term : si->ring[i];
    o.Sw = Sw;
    o.head = (o.head + 1) % RING;
    o.iters += 1;
  }
  if (o.first) {
    o.beta = 0;
    o.alpha = gam / del;
  } else {
    const double a_prev = si->alpha, g_prev = si->rz;
    o.beta = gam / g_prev;
    o.alpha = gam / (del - o.beta * gam / a_prev);
  }
  const double scale = o.Sall > si->xAx0 ? o.Sall : si->xAx0;
  const bool finite = isfinite(rr) && isfinite(gam) && isfinite(del) && isfinite(o.Sw);
  o.done = 0;
  if (finite && ((o.iters >= RING && o.Sw <= si->tolA2 * scale) || rr <= si->tolr2 * si->bb || gam == 0))
    o.done = 1;  // converged (gam = 0: r = 0 exactly), whatever this step's alpha
  else if (!finite || !(o.alpha > 0) || !isfinite(o.alpha))
    o.done = 2;  // breakdown: the caller redoes the solve with the classic iteration
  return o;
}

// Writes the state the next launch reads: a copy of si with this step applied.
__device__ inline void pb_write_state(const PcgState* si, PcgState* so, const PbStep& o, double gam, double del,
                                      double rr) {
  *so = *si;
  if (o.closed) so->ring[si->head] = o.term;
  so->head = o.head;
  so->Sall = o.Sall;
  so->Sw = o.Sw;
  so->iters = o.iters;
  so->fresh = 0;
  so->justrep = 0;
  so->rr = rr;
  so->rz = gam;
  so->pq = del;
  so->alpha = o.alpha;
  so->beta = o.beta;
  so->done = o.done;
}

// One iteration.  par: this launch's parity (it reads st[par], m[par],
// part[par] and writes the other buffers).  TM: the matrix values of the
// product (FP32 with the row-sum correction corr); TP: the block-Jacobi
// inverse.  One lane per thread (the grid covers the Nnode * G lanes in one
// pass): the product n = A m does not need this step's scalars, so the node
// offsets, the gather and the previous launch's partials are all in flight
// before the block reduction that yields alpha and beta.
template <typename TM, typename TP, int DN, int G>
__global__ void __launch_bounds__(BS)
k_pb_iter(SpmvArgs A, const TP* __restrict__ Minv, double* __restrict__ x, PbWork W, int par) {
  __shared__ double shp[3][BS / 64];
  __shared__ double shr[3][BS / 64];
  const PcgState* si = W.st[par];
  PcgState* so = W.st[par ^ 1];
  const double* __restrict__ mi = W.m[par];
  double* __restrict__ mo = W.m[par ^ 1];
  // logical block: each XCD takes one contiguous range of node groups, so the m
  // entries a block gathers from its neighbours sit mostly in that XCD's L2
  const int nb = gridDim.x, blk = xcd_block(blockIdx.x, nb);
  const int64_t gid = (int64_t)blk * BS + threadIdx.x;
  const bool have = gid < (int64_t)A.Nnode * G;
  const int n = have ? (int)(gid / G) : 0, g = (int)(gid % G);
  int p0 = 0, p1 = 0;
  if (have) {
    p0 = A.node_ptr[n];
    p1 = A.node_ptr[n + 1];
  }
  Partials<4> P;
  P.issue(W.part[par], nb);
  // the offsets are read together with the state: the exit test uses both
  // (p1 < p0 never holds)
  const int done_in = si->done;
  if (done_in != 0 || p1 < p0) {  // decided earlier: carry the state over for the next launch and the host
    if (done_in != 0 && blk == 0 && threadIdx.x == 0) *so = *si;
    return;
  }
  const bool lead = have && g == 0;
  // the own rows' operands, block-Jacobi inverse, mask and (FP32 values)
  // row-sum correction, all loaded together ahead of the gather, by every
  // lane and with no branch around them (a conditional load's result is
  // waited for where its branch ends: the mask and correction were two
  // dependent round trips before the gather); without a mask the flag is
  // read from m and ignored
  double own[9][DN], Mown[DN * DN], Cown[DN * DN];
  int kb[DN];
  const bool hk = A.known != nullptr;
  {
#pragma unroll
    for (int e = 0; e < DN * DN; ++e) {
      Mown[e] = (double)Minv[(int64_t)n * DN * DN + e];
      if constexpr (sizeof(TM) == 4)
        Cown[e] = (double)((const TM*)A.corr)[(int64_t)n * DN * DN + e];
      else
        Cown[e] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) kb[a] = *(hk ? A.known + (int64_t)n * DN + a : (const unsigned char*)mi);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      own[0][a] = mi[i];
      own[1][a] = W.w[i];
      own[2][a] = W.u[i];
      own[3][a] = W.r[i];
      own[4][a] = x[i];
      own[5][a] = W.z[i];
      own[6][a] = W.q[i];
      own[7][a] = W.s[i];
      own[8][a] = W.p[i];
    }
  }
  double sum[1][DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) sum[0][a] = 0;
  if (have)
    gather_rows<TM, DN, G, 1>((const TM*)A.vals, A.node_cols, p0, p1 - p0, g,
                              [&](int col, double (&xv)[1][DN]) {
#pragma unroll
                                for (int bb = 0; bb < DN; ++bb) xv[0][bb] = mi[(int64_t)col * DN + bb];
                              },
                              sum);
#pragma unroll
  for (int a = 0; a < DN; ++a) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) sum[0][a] += __shfl_xor(sum[0][a], off, G);
  }
  // the scalar step (every block, same order, same result)
  double pr4[4];
  P.sum(W.part[par], nb, pr4);
  double pr[3] = {pr4[0], pr4[1], pr4[2]};
  block_sum_all<3, BS>(pr, shp);
  const double gam = pr[0], del = pr[1], rr = pr[2];
  const PbStep o = pb_step(si, gam, del, rr);
  if (blk == 0 && threadIdx.x == 0) pb_write_state(si, so, o, gam, del, rr);
  if (o.done) return;  // x is the iterate the test accepted (or the breakdown's start)
  const double alpha = o.alpha, beta = o.beta;
  const bool first = o.first;
  double red[3] = {0, 0, 0};
  if (lead) {
    if constexpr (sizeof(TM) == 4) {  // FP32 values: the rounding's row sums on the own rows
#pragma unroll
      for (int a = 0; a < DN; ++a)
#pragma unroll
        for (int bb = 0; bb < DN; ++bb) sum[0][a] += Cown[a * DN + bb] * own[0][bb];
    }
    bool kn[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) kn[a] = hk && kb[a] != 0;
    double wn[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      const double y = kn[a] ? 0.0 : sum[0][a];
      const double mv = own[0][a], wv = own[1][a], uv = own[2][a];
      const double zv = y + (first ? 0.0 : beta * own[5][a]);
      const double qv = mv + (first ? 0.0 : beta * own[6][a]);
      const double sv = wv + (first ? 0.0 : beta * own[7][a]);
      const double pv = uv + (first ? 0.0 : beta * own[8][a]);
      const double rnw = own[3][a] - alpha * sv;
      const double un = uv - alpha * qv;
      wn[a] = wv - alpha * zv;
      W.z[i] = zv;
      W.q[i] = qv;
      W.s[i] = sv;
      W.p[i] = pv;
      x[i] = own[4][a] + alpha * pv;
      W.r[i] = rnw;
      W.u[i] = un;
      W.w[i] = wn[a];
      red[0] += rnw * un;
      red[1] += wn[a] * un;
      red[2] += rnw * rnw;
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) {  // the next launch's m = M w (0 at known DOFs)
      double mv = 0;
#pragma unroll
      for (int bb = 0; bb < DN; ++bb) mv += Mown[a * DN + bb] * wn[bb];
      mo[(int64_t)n * DN + a] = kn[a] ? 0.0 : mv;
    }
  }
  block_sum_all<3, BS>(red, shr);
  if (threadIdx.x == 0) {
    double* po = W.part[par ^ 1] + 4 * blk;
    po[0] = red[0];
    po[1] = red[1];
    po[2] = red[2];
    po[3] = 0;
  }
}

// Start-up and residual-replacement passes (FP64 values), on the same grid.
//   MODE 0: t = A x; r = b - t (masked); u = M r; partials x.t -> part[3].
//   MODE 1: w = A u (masked); m = M w; partials (r,u), (w,u), (r,r).
//   MODE 2: r = b - A x and s = A p (masked); u = M r, q = M s.
//   MODE 3: w = A u and z = A q (masked); m = M w; partials as MODE 1.
// part: the partial array the pass writes (4 values per block); mo: the m
// buffer the next iteration reads; MODE 3 also counts the replacement in st
// (the state the next iteration reads) and marks a restart when asked.
template <typename TP, int DN, int G, int MODE>
__global__ void __launch_bounds__(BS)
k_pb_pass(SpmvArgs A, const TP* __restrict__ Minv, const double* __restrict__ b, const double* __restrict__ x,
          PbWork W, double* __restrict__ mo, double* __restrict__ part, PcgState* st, int restart) {
  constexpr int NV = (MODE == 0 || MODE == 1) ? 1 : 2;
  __shared__ double shr[4][BS / 64];
  const double* __restrict__ v0 = (MODE == 0 || MODE == 2) ? x : W.u;
  const double* __restrict__ v1 = MODE == 2 ? W.p : W.q;
  double red[4] = {0, 0, 0, 0};
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    const int p0 = A.node_ptr[n], cnt = A.node_ptr[n + 1] - p0;
    double sum[NV][DN];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int a = 0; a < DN; ++a) sum[k][a] = 0;
    gather_rows<double, DN, G, NV>((const double*)A.vals, A.node_cols, p0, cnt, g,
                                   [&](int col, double (&xv)[NV][DN]) {
#pragma unroll
                                     for (int bb = 0; bb < DN; ++bb) {
                                       xv[0][bb] = v0[(int64_t)col * DN + bb];
                                       if (NV == 2) xv[NV - 1][bb] = v1[(int64_t)col * DN + bb];
                                     }
                                   },
                                   sum);
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int a = 0; a < DN; ++a) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) sum[k][a] += __shfl_xor(sum[k][a], off, G);
      }
    if (g != 0) continue;
    double y0[DN], y1[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      const bool kn = A.known && A.known[i];
      y0[a] = kn ? 0.0 : sum[0][a];
      y1[a] = kn ? 0.0 : sum[NV - 1][a];
    }
    if (MODE == 0 || MODE == 2) {
      double rv[DN], uv[DN];
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t i = (int64_t)n * DN + a;
        const bool kn = A.known && A.known[i];
        if (MODE == 0) {
          W.t[i] = y0[a];
          red[3] += x[i] * y0[a];  // x.A x over the unknown rows
        }
        rv[a] = kn ? 0.0 : b[i] - y0[a];
        W.r[i] = rv[a];
      }
      apply_bj<TP, DN>(Minv, A.known, n, rv, uv);
#pragma unroll
      for (int a = 0; a < DN; ++a) W.u[(int64_t)n * DN + a] = uv[a];
      if (MODE == 2) {
        double qv[DN];
        apply_bj<TP, DN>(Minv, A.known, n, y1, qv);
#pragma unroll
        for (int a = 0; a < DN; ++a) {
          W.s[(int64_t)n * DN + a] = y1[a];
          W.q[(int64_t)n * DN + a] = qv[a];
        }
      }
    } else {
      double mv[DN];
      apply_bj<TP, DN>(Minv, A.known, n, y0, mv);
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t i = (int64_t)n * DN + a;
        const double rv = W.r[i], uv = W.u[i];
        W.w[i] = y0[a];
        mo[i] = mv[a];
        if (MODE == 3) W.z[i] = y1[a];
        red[0] += rv * uv;
        red[1] += y0[a] * uv;
        red[2] += rv * rv;
      }
    }
  }
  block_sum_all<4, BS>(red, shr);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = red[k];
  }
  if (MODE == 3 && blockIdx.x == 0 && threadIdx.x == 0 && !st->done) {
    st->nrepl += 1;
    st->justrep = restart;
  }
}

__global__ void k_pb_state(PcgState* st, double tolr2, double tolA2) {
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

// After the start-up passes: x0.A x0 (partials of pass 0) and the initial
// residual norm (partials of pass 1).
__global__ void __launch_bounds__(BS) k_pb_init_scalars(const double* __restrict__ partX,
                                                        const double* __restrict__ partW, int nb, PcgState* st) {
  __shared__ double sh[BS / 64];
  const double xax = reduce_all(partX, nb, 4, 3, sh);
  const double rr = reduce_all(partW, nb, 4, 2, sh);
  if (threadIdx.x == 0) {
    st->xAx0 = xax;
    st->bb = rr;
    st->rr = rr;
    st->rrmax = rr;
    st->done = (rr == 0) ? 1 : 0;
  }
}

size_t pcgp_workspace_bytes(int Nnode, int dof_n) { return pcgp_workspace_size(Nnode, dof_n); }

// A: the FP64 matrix (start-up, replacements, residuals); Ait: the
// iterations' matrix (A itself, or the FP32 copy with its correction).
// phase PCGP_WHOLE runs the solve to its end; PCGP_START queues the start-up
// passes, the first first_check iterations, the residual estimate and the
// state read-back into poll slot 1 (or hs) and returns without waiting;
// PCGP_FINISH waits for that read-back and continues the same loop if the
// stopping test had not been met (*resumed = 1 then).  The iterates do not
// depend on where the polls fall (the stopping test and the replacements are
// decided on the device), so the split solve returns what the whole one does.
template <typename TM, typename TP, int DN, int G>
static int pcgp_run(const SpmvArgs& A, const SpmvArgs& Ait, const TP* Minv, const double* b, double* x, void* ws,
                    double tol_r, double tol_A, int replace_every, int maxit, int check_every, int first_check,
                    hipStream_t s, double* info, PcgState* hs, int phase, int* resumed) {
  const int Nnode = A.Nnode;
  if (replace_every <= 0) replace_every = sizeof(TM) == 4 ? PB_REPLACE_MIXED : PB_REPLACE;
  const int restart = sizeof(TM) == 4 ? 1 : 0;  // FP32 values: replacements restart the recurrence
  const PbWork W = carve_pb(ws, Nnode, DN);
  // one lane per thread in the iterations (every block reduces all partials)
  const int64_t lanes = (int64_t)Nnode * G;
  if (lanes > (int64_t)MAXB * BS) {
    set_error("pcgp: %d nodes of %d lanes exceed one pass of %d blocks", Nnode, G, MAXB);
    return EFEA_ERR_ARG;
  }
  const unsigned nb = grid_for(lanes);
  if (phase != PCGP_WHOLE && first_check < 1) {
    set_error("pcgp: a deferred solve needs first_check >= 1 (got %d)", first_check);
    return EFEA_ERR_ARG;
  }
  PollBuf* pb = poll_buffer(phase == PCGP_WHOLE ? 0 : 1);
  if (!pb) {
    set_error("pcgp: pinned host allocation failed");
    return EFEA_ERR_HIP;
  }
  if (!hs) hs = &pb->st;
  PcgWork fw{};  // residual passes of the read-backs
  fw.q = W.t;
  fw.r = W.t;
  fw.part1 = W.partR;
  fw.small = W.small;
  int launched = 0;
  bool spec = false;
  bool wait_only = false;  // finish: the start's read-back is in flight
  if (phase == PCGP_FINISH) {
    launched = first_check;
    spec = true;
    wait_only = true;
  } else {
    hipLaunchKernelGGL(k_pb_state, dim3(1), dim3(64), 0, s, W.st[0], tol_r * tol_r, tol_A * tol_A);
    hipLaunchKernelGGL((k_pb_pass<TP, DN, G, 0>), dim3(nb), dim3(BS), 0, s, A, Minv, b, (const double*)x, W,
                       W.m[0], W.part[1], W.st[0], 0);
    hipLaunchKernelGGL((k_pb_pass<TP, DN, G, 1>), dim3(nb), dim3(BS), 0, s, A, Minv, b, (const double*)x, W,
                       W.m[0], W.part[0], W.st[0], 0);
    hipLaunchKernelGGL(k_pb_init_scalars, dim3(1), dim3(BS), 0, s, (const double*)W.part[1],
                       (const double*)W.part[0], (int)nb, W.st[0]);
    if (int e = check_launch("pcgp_start")) return e;
  }
  while (true) {
    if (wait_only) {
      wait_only = false;
      if (hipStreamSynchronize(s) != hipSuccess) {
        set_error("pcgp: state readback failed");
        return EFEA_ERR_HIP;
      }
      if (hs->done || launched >= maxit) break;
      if (resumed) *resumed = 1;
    }
    const int chunk = launched == 0 && first_check > 0 ? first_check : check_every;
    for (int k = 0; k < chunk; ++k) {
      const int par = launched & 1;
      if (launched > 0 && launched % replace_every == 0) {
        // residual replacement: r = b - A x, u = M r, s = A p, q = M s, then
        // w = A u, m = M w, z = A q with the partials the next launch reads
        hipLaunchKernelGGL((k_pb_pass<TP, DN, G, 2>), dim3(nb), dim3(BS), 0, s, A, Minv, b, (const double*)x, W,
                           W.m[par], W.part[par], W.st[par], 0);
        hipLaunchKernelGGL((k_pb_pass<TP, DN, G, 3>), dim3(nb), dim3(BS), 0, s, A, Minv, b, (const double*)x, W,
                           W.m[par], W.part[par], W.st[par], restart);
      }
      hipLaunchKernelGGL((k_pb_iter<TM, TP, DN, G>), dim3(nb), dim3(BS), 0, s, Ait, Minv, x, W, par);
      ++launched;
    }
    spec = first_check > 0;
    if (spec) {
      pcg_spec_residual<DN, G>(A, b, x, fw, s);
      if (hipMemcpyAsync(&pb->rr, W.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("pcgp: residual readback failed");
        return EFEA_ERR_HIP;
      }
    }
    if (int e = check_launch("pcgp_iter")) return e;
    if (hipMemcpyAsync(hs, W.st[launched & 1], sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_error("pcgp: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (phase == PCGP_START) return 0;
    if (hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcgp: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (hs->done || launched >= maxit) break;
  }
  const PcgState fin = *hs;
  if (!spec) {  // else the last read-back carried the residual
    pcg_true_residual<DN, G>(A, b, x, fw, s);
    if (hipMemcpyAsync(&pb->rr, W.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcgp: residual readback failed");
      return EFEA_ERR_HIP;
    }
  }
  return pcg_finish_info(&fin, pb->rr, info);
}

int solve_pcgp(int precision, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
               const double* vals64, const float* vals32, const float* corr32, const unsigned char* known,
               const void* Minv, const double* b, double* x, void* ws, double tol_r, double tol_A, int replace_every,
               int maxit, int check_every, int first_check, hipStream_t s, double* info, void* host_state_buf,
               int phase, int* resumed) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals64, known, nullptr, 0};
  if (resumed) *resumed = 0;
  SpmvArgs A32{Nnode, node_ptr, node_cols, vals32, known, corr32, 0};
  if (check_every < 1) check_every = 1;
  PcgState* hs = (PcgState*)host_state_buf;  // null: this thread's pinned poll buffer
#define RUNB(D, GV)                                                                                             \
  if (dof_n == D && group == GV) {                                                                             \
    if (precision == 32)                                                                                       \
      return pcgp_run<float, float, D, GV>(A, A32, (const float*)Minv, b, x, ws, tol_r, tol_A, replace_every,   \
                                           maxit, check_every, first_check, s, info, hs, phase, resumed);      \
    return pcgp_run<double, double, D, GV>(A, A, (const double*)Minv, b, x, ws, tol_r, tol_A, replace_every,    \
                                           maxit, check_every, first_check, s, info, hs, phase, resumed);      \
  }
  EFEA_GROUPS(RUNB)
#undef RUNB
  set_error("solve_pcgp: unsupported dof_n %d / group %d", dof_n, group);
  return EFEA_ERR_ARG;
}

}  // namespace efea
