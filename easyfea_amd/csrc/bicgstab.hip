// Right-preconditioned BiCGStab (van der Vorst 1992) with the nodal
// block-Jacobi inverse, for the non-symmetric displacement systems of the
// phase-field splits whose c+ / c- are not symmetric (Zhang and the
// AnisotStrain / AnisotStress _PM / _MP variants).  EasyFEA solves those
// directly (Simulations/Solvers.py:225-394, the non-symmetric branch of
// _Solve_Axb); here the Dirichlet-reduced system is solved on the full-length
// vectors with the known-DOF mask, as the PCG does.
//
//   r = b - A x, rh = r, rho_old = alpha = omega = 1, p = v = 0
//   rho = (rh, r); beta = (rho / rho_old) (alpha / omega)
//   p = r + beta (p - omega v); y = M p; v = A y
//   alpha = rho / (rh, v); s = r - alpha v; z = M s; t = A z
//   omega = (t, s) / (t, t); x += alpha y + omega z; r = s - omega t
//
// Five launches per iteration on one node-group grid; each reduction is
// finished redundantly, in a fixed order, by every block of the next launch
// that needs it, and each state field is written by block 0 of one stage
// only (the stage that no other block of the same launch reads it in).
// Stopping test: (r, r) <= tol^2 (r0, r0) on the recursive residual; the
// caller gets the true residual of the returned x.  A stage that finds a
// breakdown stops before writing vectors, in every block alike (each block
// evaluates the same reduced scalars), so x is never left half updated.
#include "pcg_device.h"

namespace efea {

struct BsState {
  double rho, rho_old, alpha, omega, bb, tol2, rr;
  int done, iters;  // done: 1 converged, 2 breakdown
};

struct BsWork {
  double *r, *rh, *p, *v, *y, *s, *z, *t;
  double *P2, *P4, *P5;  // partials: (rh, v); (t, s), (t, t); (rh, r), (r, r)
  BsState* st;
  double* small;
  double* partR;
};

static inline size_t bicgstab_workspace_size(int Nnode, int dof_n) {
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  return 8 * vb + align256(8 * MAXB) + 2 * align256(16 * MAXB) + align256(8 * MAXB) + align256(sizeof(BsState)) +
         256;
}

static inline BsWork carve_bs(void* ws, int Nnode, int dof_n) {
  char* c = (char*)ws;
  const size_t vb = align256(8 * (size_t)Nnode * dof_n);
  BsWork W;
  double** vecs[] = {&W.r, &W.rh, &W.p, &W.v, &W.y, &W.s, &W.z, &W.t};
  for (double** q : vecs) {
    *q = (double*)c;
    c += vb;
  }
  W.P2 = (double*)c;
  c += align256(8 * MAXB);
  W.P4 = (double*)c;
  c += align256(16 * MAXB);
  W.P5 = (double*)c;
  c += align256(16 * MAXB);
  W.partR = (double*)c;
  c += align256(8 * MAXB);
  W.st = (BsState*)c;
  c += align256(sizeof(BsState));
  W.small = (double*)c;
  return W;
}

// The node's G lanes: the product y = A x over the node's rows (masked).
template <int DN, int G>
__device__ inline void bs_product(const SpmvArgs& A, int n, int g, const double* __restrict__ xv,
                                  double (&y)[DN]) {
  const int p0 = A.node_ptr[n], cnt = A.node_ptr[n + 1] - p0;
  double sum[1][DN];
#pragma unroll
  for (int a = 0; a < DN; ++a) sum[0][a] = 0;
  gather_rows<double, DN, G, 1>((const double*)A.vals, A.node_cols, p0, cnt, g,
                                [&](int col, double (&v)[1][DN]) {
#pragma unroll
                                  for (int bb = 0; bb < DN; ++bb) v[0][bb] = xv[(int64_t)col * DN + bb];
                                },
                                sum);
#pragma unroll
  for (int a = 0; a < DN; ++a) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) sum[0][a] += __shfl_xor(sum[0][a], off, G);
    const bool kn = A.known && A.known[(int64_t)n * DN + a];
    y[a] = kn ? 0.0 : sum[0][a];
  }
}

// Start: r = b - A x (masked), rh = r, p = v = 0; partials (rh, r), (r, r).
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_bs_start(SpmvArgs A, const double* __restrict__ b, const double* __restrict__ x, BsWork W) {
  __shared__ double sh[2][BS / 64];
  double red[2] = {0, 0};
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double y[DN];
    bs_product<DN, G>(A, n, g, x, y);
    if (g != 0) continue;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      const bool kn = A.known && A.known[i];
      const double r = kn ? 0.0 : b[i] - y[a];
      W.r[i] = r;
      W.rh[i] = r;
      W.p[i] = 0;
      W.v[i] = 0;
      red[0] += r * r;
      red[1] += r * r;
    }
  }
  block_sum_all<2, BS>(red, sh);
  if (threadIdx.x == 0) {
    W.P5[2 * blockIdx.x] = red[0];
    W.P5[2 * blockIdx.x + 1] = red[1];
  }
}

__global__ void __launch_bounds__(BS) k_bs_state(BsState* st, const double* __restrict__ P5, int nb, double tol2) {
  __shared__ double sh[BS / 64];
  const double rr = reduce_all(P5, nb, 2, 1, sh);
  if (threadIdx.x == 0) {
    st->rho = st->rho_old = st->alpha = st->omega = 1.0;
    st->bb = rr;
    st->rr = rr;
    st->tol2 = tol2;
    st->iters = 0;
    st->done = rr == 0 ? 1 : 0;
  }
}

// Stage 1: rho = (rh, r), the stopping test; p = r + beta (p - omega v), y = M p.
template <typename TP, int DN, int G>
__global__ void __launch_bounds__(BS)
k_bs_stage1(SpmvArgs A, const TP* __restrict__ Minv, BsWork W) {
  __shared__ double sh[2][BS / 64];
  BsState* st = W.st;
  if (st->done) return;
  double pr[2];
  partial_share<2>(W.P5, gridDim.x, pr);
  block_sum_all<2, BS>(pr, sh);
  const double rho = pr[0], rr = pr[1];
  const bool conv = rr <= st->tol2 * st->bb;
  // omega = 0 (t = 0 or t orthogonal to s) would divide the next beta
  const bool broke = !isfinite(rr) || !isfinite(rho) || rho == 0 || !(st->omega != 0) || !isfinite(st->omega);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->rho = rho;
    st->done = conv ? 1 : (broke ? 2 : 0);
  }
  if (conv || broke) return;
  const double beta = (rho / st->rho_old) * (st->alpha / st->omega);
  const double om = st->omega;
  for (int n = blockIdx.x * BS + threadIdx.x; n < A.Nnode; n += gridDim.x * BS) {
    double pv[DN], yv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      pv[a] = W.r[i] + beta * (W.p[i] - om * W.v[i]);
      W.p[i] = pv[a];
    }
    apply_bj<TP, DN>(Minv, A.known, n, pv, yv);
#pragma unroll
    for (int a = 0; a < DN; ++a) W.y[(int64_t)n * DN + a] = yv[a];
  }
}

// Stage 2: v = A y; partials (rh, v).
template <int DN, int G>
__global__ void __launch_bounds__(BS) k_bs_stage2(SpmvArgs A, BsWork W) {
  __shared__ double sh[1][BS / 64];
  if (W.st->done) return;
  double red[1] = {0};
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double y[DN];
    bs_product<DN, G>(A, n, g, W.y, y);
    if (g != 0) continue;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      W.v[i] = y[a];
      red[0] += W.rh[i] * y[a];
    }
  }
  block_sum_all<1, BS>(red, sh);
  if (threadIdx.x == 0) W.P2[blockIdx.x] = red[0];
}

// Stage 3: alpha = rho / (rh, v); s = r - alpha v, z = M s.
template <typename TP, int DN, int G>
__global__ void __launch_bounds__(BS)
k_bs_stage3(SpmvArgs A, const TP* __restrict__ Minv, BsWork W, int nb2) {
  __shared__ double sh[1][BS / 64];
  BsState* st = W.st;
  if (st->done) return;
  double pr[1];
  partial_share<1>(W.P2, nb2, pr);
  block_sum_all<1, BS>(pr, sh);
  const double alpha = st->rho / pr[0];
  const bool broke = !isfinite(alpha) || pr[0] == 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alpha;
    if (broke) st->done = 2;
  }
  if (broke) return;
  for (int n = blockIdx.x * BS + threadIdx.x; n < A.Nnode; n += gridDim.x * BS) {
    double sv[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      sv[a] = W.r[i] - alpha * W.v[i];
      W.s[i] = sv[a];
    }
    apply_bj<TP, DN>(Minv, A.known, n, sv, zv);
#pragma unroll
    for (int a = 0; a < DN; ++a) W.z[(int64_t)n * DN + a] = zv[a];
  }
}

// Stage 4: t = A z; partials (t, s), (t, t).
template <int DN, int G>
__global__ void __launch_bounds__(BS) k_bs_stage4(SpmvArgs A, BsWork W) {
  __shared__ double sh[2][BS / 64];
  if (W.st->done) return;
  double red[2] = {0, 0};
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double y[DN];
    bs_product<DN, G>(A, n, g, W.z, y);
    if (g != 0) continue;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      W.t[i] = y[a];
      red[0] += y[a] * W.s[i];
      red[1] += y[a] * y[a];
    }
  }
  block_sum_all<2, BS>(red, sh);
  if (threadIdx.x == 0) {
    W.P4[2 * blockIdx.x] = red[0];
    W.P4[2 * blockIdx.x + 1] = red[1];
  }
}

// Stage 5: omega = (t, s) / (t, t); x += alpha y + omega z, r = s - omega t;
// partials (rh, r), (r, r) for the next stage 1.
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_bs_stage5(int Nnode, double* __restrict__ x, BsWork W, int nb4) {
  __shared__ double sh[2][BS / 64];
  __shared__ double sh2[2][BS / 64];
  BsState* st = W.st;
  if (st->done) return;
  double pr[2];
  partial_share<2>(W.P4, nb4, pr);
  block_sum_all<2, BS>(pr, sh);
  const double omega = pr[1] > 0 ? pr[0] / pr[1] : 0.0;
  const double alpha = st->alpha;
  double red[2] = {0, 0};
  for (int n = blockIdx.x * BS + threadIdx.x; n < Nnode; n += gridDim.x * BS) {
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      const int64_t i = (int64_t)n * DN + a;
      x[i] += alpha * W.y[i] + omega * W.z[i];
      const double r = W.s[i] - omega * W.t[i];
      W.r[i] = r;
      red[0] += W.rh[i] * r;
      red[1] += r * r;
    }
  }
  block_sum_all<2, BS>(red, sh2);
  if (threadIdx.x == 0) {
    W.P5[2 * blockIdx.x] = red[0];
    W.P5[2 * blockIdx.x + 1] = red[1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // read by the next stage 1 (which tests omega)
    st->omega = omega;
    st->rho_old = st->rho;
    st->iters += 1;
  }
}

size_t bicgstab_workspace_bytes(int Nnode, int dof_n) { return bicgstab_workspace_size(Nnode, dof_n); }

template <typename TP, int DN, int G>
static int bicgstab_run(const SpmvArgs& A, const TP* Minv, const double* b, double* x, void* ws, double tol,
                        int maxit, int check_every, hipStream_t s, double* info) {
  const BsWork W = carve_bs(ws, A.Nnode, DN);
  const unsigned nb = grid_for((int64_t)A.Nnode * G);  // every stage on this grid (<= MAXB blocks)
  hipLaunchKernelGGL((k_bs_start<DN, G>), dim3(nb), dim3(BS), 0, s, A, b, (const double*)x, W);
  hipLaunchKernelGGL(k_bs_state, dim3(1), dim3(BS), 0, s, W.st, (const double*)W.P5, (int)nb, tol * tol);
  if (int e = check_launch("bicgstab_start")) return e;
  PollBuf* pb = poll_buffer();
  if (!pb) {
    set_error("bicgstab: pinned host allocation failed");
    return EFEA_ERR_HIP;
  }
  BsState* hs = (BsState*)&pb->st;
  static_assert(sizeof(BsState) <= sizeof(PcgState), "the poll buffer holds the state");
  int launched = 0;
  while (true) {
    for (int k = 0; k < check_every; ++k) {
      hipLaunchKernelGGL((k_bs_stage1<TP, DN, G>), dim3(nb), dim3(BS), 0, s, A, Minv, W);
      hipLaunchKernelGGL((k_bs_stage2<DN, G>), dim3(nb), dim3(BS), 0, s, A, W);
      hipLaunchKernelGGL((k_bs_stage3<TP, DN, G>), dim3(nb), dim3(BS), 0, s, A, Minv, W, (int)nb);
      hipLaunchKernelGGL((k_bs_stage4<DN, G>), dim3(nb), dim3(BS), 0, s, A, W);
      hipLaunchKernelGGL((k_bs_stage5<DN, G>), dim3(nb), dim3(BS), 0, s, A.Nnode, x, W, (int)nb);
    }
    launched += check_every;
    if (int e = check_launch("bicgstab_iter")) return e;
    if (hipMemcpyAsync(hs, W.st, sizeof(BsState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("bicgstab: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (hs->done || launched >= maxit) break;
  }
  const BsState fin = *hs;
  PcgWork fw{};
  fw.r = W.t;  // dead after the last iteration
  fw.q = W.t;
  fw.part1 = W.partR;
  fw.small = W.small;
  pcg_true_residual<DN, G>(A, b, x, fw, s);
  if (hipMemcpyAsync(&pb->rr, W.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("bicgstab: residual readback failed");
    return EFEA_ERR_HIP;
  }
  info[0] = fin.iters;
  info[1] = fin.bb > 0 ? sqrt(pb->rr / fin.bb) : 0.0;
  info[2] = 0;
  info[3] = fin.done;
  if (fin.done == 2) {
    set_error("bicgstab: breakdown after %d iterations (relative residual %.3e)", fin.iters, info[1]);
    return EFEA_ERR_NOCONV;
  }
  if (fin.done != 1) {
    set_error("bicgstab: not converged after %d iterations (relative residual %.3e)", fin.iters, info[1]);
    return EFEA_ERR_NOCONV;
  }
  return EFEA_OK;
}

int solve_bicgstab(int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols, const double* vals,
                   const unsigned char* known, const double* Minv, const double* b, double* x, void* ws,
                   double tol, int maxit, int check_every, hipStream_t s, double* info) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals, known, nullptr, 0};
  if (check_every < 1) check_every = 1;
#define RUNS(D, GV) \
  if (dof_n == D && group == GV) return bicgstab_run<double, D, GV>(A, Minv, b, x, ws, tol, maxit, check_every, s, info);
  EFEA_GROUPS(RUNS)
#undef RUNS
  set_error("solve_bicgstab: unsupported dof_n %d / group %d", dof_n, group);
  return EFEA_ERR_ARG;
}

}  // namespace efea
