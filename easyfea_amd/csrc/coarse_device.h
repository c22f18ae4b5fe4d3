// Device-side pieces of the two-level preconditioner shared by the
// single-system kernels (coarse.hip) and the batch engine (batch.hip): the
// rigid-body modes, the per-aggregate bodies of step 2 (with restriction),
// the prolongation, the initial and the replacement residual passes, and the
// workspace layout.  Each body takes the aggregate index `blk` of its block.
#pragma once
#include "pcg_device.h"

namespace efea {

// Rigid-body mode m of DOF component a at a node with coordinates r relative
// to its aggregate centroid.
template <int DN>
__device__ inline double pmode(const double* __restrict__ r, int a, int m) {
  if constexpr (DN == 1) {
    return 1.0;
  } else if constexpr (DN == 2) {
    if (m < 2) return a == m ? 1.0 : 0.0;
    return a == 0 ? -r[1] : r[0];
  } else {
    if (m < 3) return a == m ? 1.0 : 0.0;
    // rotations about x, y, z: w x r
    const int w = m - 3;
    if (a == w) return 0.0;
    if (w == 0) return a == 1 ? -r[2] : r[1];
    if (w == 1) return a == 0 ? r[2] : -r[0];
    return a == 0 ? -r[1] : r[0];
  }
}

template <int DN> struct Modes { static constexpr int NM = DN == 1 ? 1 : (DN == 2 ? 3 : 6); };

struct CoarseArgs {
  int nagg;
  const int* agg_ptr;
  const int* agg_nodes;
  const double* rel;
  const void* Ainv;  // ld x ld, FP64 (FP32 in the refinement mode's inner solves)
  int ld;
  double* rc;  // nagg * NM
};

// Step 2 with restriction: alpha = rz / pq ; x += alpha p ; r -= alpha q ;
// z = B^-1 r ; partials of r.z (block Jacobi part) and r.r ; rc_I = P_I^T r.
// TV: the vector type (double; float in the refinement mode's inner solves,
// with the arithmetic in double registers).
template <typename TP, int DN, typename TV>
__device__ inline void pcg2_step2_body(CoarseArgs C, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
             TV* __restrict__ x, TV* __restrict__ r, TV* __restrict__ z, const TV* __restrict__ P0,
             const TV* __restrict__ P1, const TV* __restrict__ q, const double* __restrict__ part1, int nb1,
             double* __restrict__ part2, PcgState* st, int blk) {
  constexpr int NM = Modes<DN>::NM;
  // the state test follows the loads and the block reduction (a launch after
  // convergence does the loads and returns there)
  __shared__ double sh[BS / 64];
  __shared__ double bc;
  Partials<1> P1s;
  P1s.issue(part1, nb1);
  const int pn = st->par_next;
  const double rz = st->rz;
  const TV* __restrict__ p = pn ? P1 : P0;
  const int k0 = C.agg_ptr[blk], k1 = C.agg_ptr[blk + 1];
  // the first node's operands are loaded before the reduction that yields alpha
  const int kf = k0 + threadIdx.x;
  const bool have = kf < k1;
  const int nf = have ? C.agg_nodes[kf] : 0;
  double xf[DN], pf[DN], rf[DN], qf[DN], Mf[DN * DN], relf[3];
  if (have) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)nf * DN + a;
      xf[a] = (double)x[i];
      pf[a] = (double)p[i];
      rf[a] = (double)r[i];
      qf[a] = (double)q[i];
    }
#pragma unroll
    for (int e = 0; e < DN * DN; ++e) Mf[e] = (double)Minv[(int64_t)nf * DN * DN + e];
#pragma unroll
    for (int d = 0; d < 3; ++d) relf[d] = C.rel[3 * (int64_t)nf + d];
  }
  double pp1[1];
  P1s.sum(part1, nb1, pp1);
  const double ppq = block_sum<double, BS>(pp1[0], sh);
  if (threadIdx.x == 0) bc = ppq;
  __syncthreads();
  if (st->done) return;
  const double pq = bc;
  if (!(pq > 0) || !(rz > 0)) {
    if (blk == 0 && threadIdx.x == 0) st->done = (rz == 0) ? 1 : 2;
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
  double s_rz = 0, s_rr = 0, rcv[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) rcv[m] = 0;
  if (have) {
    double r2[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)nf * DN + a;
      x[i] = (TV)(xf[a] + al * pf[a]);
      r2[a] = rf[a] - al * qf[a];
      r[i] = (TV)r2[a];
    }
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      double zz = 0;
#pragma unroll
      for (int b = 0; b < DN; ++b) zz += Mf[a * DN + b] * r2[b];
      if (known && known[(int64_t)nf * DN + a]) zz = 0;
      z[(int64_t)nf * DN + a] = (TV)zz;
      s_rz += r2[a] * zz;
      s_rr += r2[a] * r2[a];
#pragma unroll
      for (int m = 0; m < NM; ++m) rcv[m] += pmode<DN>(relf, a, m) * r2[a];  // r = 0 at known DOFs
    }
  }
  for (int k = kf + BS; k < k1; k += BS) {
    const int n = C.agg_nodes[k];
    double r2[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      x[i] = (TV)((double)x[i] + al * (double)p[i]);
      r2[a] = (double)r[i] - al * (double)q[i];
      r[i] = (TV)r2[a];
    }
    apply_bj<TP, DN>(Minv, known, n, r2, zv);
    const double* rn = C.rel + 3 * (int64_t)n;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = (TV)zv[a];
      s_rz += r2[a] * zv[a];
      s_rr += r2[a] * r2[a];
#pragma unroll
      for (int m = 0; m < NM; ++m) rcv[m] += pmode<DN>(rn, a, m) * r2[a];
    }
  }
  __shared__ double shn[NM + 2][BS / 64];
  double red[NM + 2];
  red[0] = s_rz;
  red[1] = s_rr;
#pragma unroll
  for (int m = 0; m < NM; ++m) red[2 + m] = rcv[m];
  block_sum_all<NM + 2, BS>(red, shn);
  if (threadIdx.x == 0) {
    part2[2 * blk] = red[0];
    part2[2 * blk + 1] = red[1];
#pragma unroll
    for (int m = 0; m < NM; ++m) C.rc[blk * NM + m] = red[2 + m];
  }
}

// z += P_I zc_I with zc_I = (Ac^-1)_I rc ; r.z partial += rc_I . zc_I.
// TV: vector type; TA: type of the Ac^-1 values (float in the refinement
// mode's inner solves: half the bytes of the dense rows).
template <int DN, typename TV, typename TA>
__device__ inline void pcg2_prolong_body(CoarseArgs C, const unsigned char* __restrict__ known, TV* __restrict__ z,
               double* __restrict__ part2, PcgState* st, int blk) {
  constexpr int NM = Modes<DN>::NM;
  // the state test follows the loads and the coarse rows
  __shared__ double shz[NM][BS / 64];
  const int nc = C.nagg * NM;
  const int k0 = C.agg_ptr[blk], k1 = C.agg_ptr[blk + 1];
  // the first node's data is loaded before the coarse rows
  const int kf = k0 + threadIdx.x;
  const bool have = kf < k1;
  const int nf = have ? C.agg_nodes[kf] : 0;
  double relf[3], zf[DN];
  bool kfree[DN];
  if (have) {
#pragma unroll
    for (int d = 0; d < 3; ++d) relf[d] = C.rel[3 * (int64_t)nf + d];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      zf[a] = (double)z[(int64_t)nf * DN + a];
      kfree[a] = !(known && known[(int64_t)nf * DN + a]);
    }
  }
  // The aggregate's NM rows of Ac^-1 times rc: every thread takes columns
  // c = t, t + BS, ... of all NM rows (one rc load serves NM products), four
  // columns per pass with their loads issued together, then one block
  // reduction of the NM sums (a wavefront per row walking the row alone paid
  // one load latency per 64 columns).
  const TA* __restrict__ rows = (const TA*)C.Ainv + (int64_t)(blk * NM) * C.ld;
  double zc[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) zc[m] = 0;
  int c = threadIdx.x;
  for (; c + 3 * BS < nc; c += 4 * BS) {
    double rv[4], av[4][NM];
#pragma unroll
    for (int u = 0; u < 4; ++u) rv[u] = C.rc[c + u * BS];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int m = 0; m < NM; ++m) av[u][m] = (double)rows[(int64_t)m * C.ld + c + u * BS];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int m = 0; m < NM; ++m) zc[m] += av[u][m] * rv[u];
  }
  for (; c < nc; c += BS) {
    const double rv = C.rc[c];
#pragma unroll
    for (int m = 0; m < NM; ++m) zc[m] += (double)rows[(int64_t)m * C.ld + c] * rv;
  }
  block_sum_all<NM, BS>(zc, shz);
  if (st->done) return;
  if (have) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      if (!kfree[a]) continue;
      double v = 0;
#pragma unroll
      for (int m = 0; m < NM; ++m) v += pmode<DN>(relf, a, m) * zc[m];
      z[(int64_t)nf * DN + a] = (TV)(zf[a] + v);
    }
  }
  for (int k = kf + BS; k < k1; k += BS) {
    const int n = C.agg_nodes[k];
    const double* rn = C.rel + 3 * (int64_t)n;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      if (known && known[i]) continue;
      double v = 0;
#pragma unroll
      for (int m = 0; m < NM; ++m) v += pmode<DN>(rn, a, m) * zc[m];
      z[i] = (TV)((double)z[i] + v);
    }
  }
  if (threadIdx.x == 0) {
    double d = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) d += C.rc[blk * NM + m] * zc[m];
    part2[2 * blk] += d;
  }
}

// Initial z = B^-1 r, P0 = 0, partials of r.z and r.r and the restriction;
// block 0 also folds the x0.A x0 partials into the state.
template <typename TP, int DN, typename TV>
__device__ inline void pcg2_init_body(CoarseArgs C, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
            const TV* __restrict__ r, TV* __restrict__ z, TV* __restrict__ P0,
            const double* __restrict__ partI, int nbI, double* __restrict__ part2, PcgState* st, int blk) {
  constexpr int NM = Modes<DN>::NM;
  __shared__ double sh[BS / 64];
  if (blk == 0 && partI) {
    const double xax = reduce_all(partI, nbI, 1, 0, sh);
    if (threadIdx.x == 0) st->xAx0 = xax;
  }
  double s_rz = 0, s_rr = 0, rcv[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) rcv[m] = 0;
  const int k0 = C.agg_ptr[blk], k1 = C.agg_ptr[blk + 1];
  for (int k = k0 + threadIdx.x; k < k1; k += BS) {
    const int n = C.agg_nodes[k];
    double rv[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) rv[a] = (double)r[(int64_t)n * DN + a];
    apply_bj<TP, DN>(Minv, known, n, rv, zv);
    const double* rn = C.rel + 3 * (int64_t)n;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = (TV)zv[a];
      P0[(int64_t)n * DN + a] = (TV)0;
      s_rz += rv[a] * zv[a];
      s_rr += rv[a] * rv[a];
#pragma unroll
      for (int m = 0; m < NM; ++m) rcv[m] += pmode<DN>(rn, a, m) * rv[a];
    }
  }
  __shared__ double shn[NM + 2][BS / 64];
  double red[NM + 2];
  red[0] = s_rz;
  red[1] = s_rr;
#pragma unroll
  for (int m = 0; m < NM; ++m) red[2 + m] = rcv[m];
  block_sum_all<NM + 2, BS>(red, shn);
  if (threadIdx.x == 0) {
    part2[2 * blk] = red[0];
    part2[2 * blk + 1] = red[1];
#pragma unroll
    for (int m = 0; m < NM; ++m) C.rc[blk * NM + m] = red[2 + m];
  }
}

// Residual replacement, second kernel (aggregate grid, after residual_body):
// z = B^-1 r, partials, restriction; block 0 resumes the solve.  The
// prolongation kernel follows.
template <typename TP, int DN>
__device__ inline void pcg2_rebuild_z_body(CoarseArgs C, const unsigned char* __restrict__ known,
                                           const TP* __restrict__ Minv, const double* __restrict__ r,
                                           double* __restrict__ z, double* __restrict__ part2, PcgState* st,
                                           int blk) {
  constexpr int NM = Modes<DN>::NM;
  double s_rz = 0, s_rr = 0, rcv[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) rcv[m] = 0;
  const int k0 = C.agg_ptr[blk], k1 = C.agg_ptr[blk + 1];
  for (int k = k0 + threadIdx.x; k < k1; k += BS) {
    const int n = C.agg_nodes[k];
    double rv[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) rv[a] = r[(int64_t)n * DN + a];
    apply_bj<TP, DN>(Minv, known, n, rv, zv);
    const double* rn = C.rel + 3 * (int64_t)n;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = zv[a];
      s_rz += rv[a] * zv[a];
      s_rr += rv[a] * rv[a];
#pragma unroll
      for (int m = 0; m < NM; ++m) rcv[m] += pmode<DN>(rn, a, m) * rv[a];
    }
  }
  __shared__ double shn[NM + 2][BS / 64];
  double red[NM + 2];
  red[0] = s_rz;
  red[1] = s_rr;
#pragma unroll
  for (int m = 0; m < NM; ++m) red[2 + m] = rcv[m];
  block_sum_all<NM + 2, BS>(red, shn);
  if (threadIdx.x == 0) {
    part2[2 * blk] = red[0];
    part2[2 * blk + 1] = red[1];
#pragma unroll
    for (int m = 0; m < NM; ++m) C.rc[blk * NM + m] = red[2 + m];
  }
  if (blk == 0 && threadIdx.x == 0) {
    st->done = 0;
    st->justrep = 1;
    st->nrepl += 1;
  }
}

// Workspace of a two-level solve: the one-level workspace (partials sized
// for max(MAXB, nagg) blocks) plus the restriction vector.
static inline size_t pcg2_workspace_size(int Nnode, int dof_n, int nagg) {
  const int nb = nagg > MAXB ? nagg : MAXB;
  const int nm = dof_n == 1 ? 1 : (dof_n == 2 ? 3 : 6);
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  // (+ the refinement mode's three partials per block of its FP64 residual pass)
  return 5 * vb + align256(8 * (size_t)MAXB) * 2 + align256(16 * (size_t)nb) + align256(sizeof(PcgState)) + 256 +
         align256(8 * (size_t)nagg * nm) + align256(24 * (size_t)MAXB);
}

// The one-level work arrays (rc in `small`'s place is separate) of a
// two-level workspace.
struct Pcg2Work {
  PcgWork w;
  double* rc;
};

static inline Pcg2Work carve2(void* ws, int Nnode, int dof_n, int nagg) {
  const int nb = nagg > MAXB ? nagg : MAXB;
  char* c = (char*)ws;
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  Pcg2Work v;
  PcgWork& w = v.w;
  w.r = (double*)c; c += vb;
  w.z = (double*)c; c += vb;
  w.P0 = (double*)c; c += vb;
  w.P1 = (double*)c; c += vb;
  w.q = (double*)c; c += vb;
  w.part1 = (double*)c; c += align256(8 * (size_t)MAXB);
  w.partI = (double*)c; c += align256(8 * (size_t)MAXB);
  w.part2 = (double*)c; c += align256(16 * (size_t)nb);
  w.st = (PcgState*)c; c += align256(sizeof(PcgState));
  w.small = (double*)c; c += 256;
  v.rc = (double*)c;
  return v;
}

// Initialisation of one two-level solve (state, x0.A x0, r0, z0 = M^-1 r0
// with P0 = 0, scalars).  Defined in coarse.hip.
template <typename TP, int DN, int G>
int pcg2_start(const SpmvArgs& A64, const TP* Minv, const CoarseArgs& C, const double* b, const double* x,
               const Pcg2Work& w, double tol_r, double tol_A, double delta2, hipStream_t s);

}  // namespace efea
