// Bound-constrained linear least squares on the device:
//
//   minimise 1/2 |A x - b|^2  subject to  lb <= x <= ub
//
// for a symmetric sparse A in scalar CSR, the damage problem of the
// phase-field BoundConstrain solver (EasyFEA/Simulations/Solvers.py:376-383
// call scipy.optimize.lsq_linear(A, b, bounds=(lb, ub), tol=1e-10) on the
// damage system; the bounds come from PhaseField.Get_lb_ub,
// Simulations/_phasefield.py:148-163).  The minimiser is unique when A is
// nonsingular; any method that reaches it agrees with lsq_linear to its
// tolerance.
//
// Method (not the trust-region reflective method scipy uses):
//   * primal-dual active set outer loop (Hintermueller, Ito, Kunisch, SIAM
//     J. Optim. 13 (2003) 865-888) on the gradient g = A (A x - b): with the
//     column scaling d_i = sum_j A_ij^2, a variable is held at its lower
//     (upper) bound when x_i - g_i / d_i falls below (above) it, free
//     otherwise;
//   * the free-variable problem min |A_F x_F - (b - A_C x_C)| solved by CGLS
//     (conjugate gradients on the normal equations in the form of Bjoerck,
//     Numerical Methods for Least Squares Problems, 1996, section 7.4) with
//     the same column scaling as preconditioner;
//     each pass stops once the scaled free-set gradient norm has fallen by
//     the factor tol, or once it bounds every free variable's first-order
//     error below kkt_tol / 2;
//   * stop when no variable violates the first-order conditions by more
//     than kkt_tol (in units of x): |x_i - clamp(x_i - g_i/d_i, lb_i, ub_i)|;
//     after max_outer updates without that, the solve reports
//     non-convergence.
//
// Every reduction is a fixed-order two-stage sum, so a solve is bitwise
// repeatable.  Scalars (alpha, beta, gamma) stay on the device; the host
// reads gamma every POLL iterations and the violation count once per outer
// iteration.
#include "efea_common.h"

namespace efea {
namespace {

constexpr int BS = 256;
constexpr int MAXB = 2048;
constexpr int POLL = 16;

enum { SC_GAM = 0, SC_ALPHA, SC_BETA, SC_TMP, SC_N };

unsigned grid(int64_t n) {
  const int64_t nb = (n + BS - 1) / BS;
  return (unsigned)(nb < 1 ? 1 : (nb < MAXB ? nb : MAXB));
}

__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < BS / 64; ++k) s += sh[k];
  return s;
}

// d_i = sum_j A_ij^2 (the squared norm of column i; A is symmetric)
__global__ void __launch_bounds__(BS) k_colnorm2(int N, const int* __restrict__ rp, const double* __restrict__ val,
                                                 double* __restrict__ d) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double s = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * val[k];
    d[i] = s;
  }
}

// y = alpha (A x) + beta c   (c may be null)
__global__ void __launch_bounds__(BS) k_spmv(int N, const int* __restrict__ rp, const int* __restrict__ ci,
                                             const double* __restrict__ val, const double* __restrict__ x,
                                             double alpha, double beta, const double* __restrict__ c,
                                             double* __restrict__ y) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double s = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[ci[k]];
    y[i] = c ? alpha * s + beta * c[i] : alpha * s;
  }
}

// q = (A s) on the free set, 0 elsewhere; z = q / d
__global__ void __launch_bounds__(BS) k_normal_q(int N, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const double* __restrict__ val, const double* __restrict__ s,
                                                 const unsigned char* __restrict__ state,
                                                 const double* __restrict__ d, double* __restrict__ q,
                                                 double* __restrict__ z) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double acc = 0;
    if (state[i] == 0 && d[i] > 0) {
      for (int k = rp[i]; k < rp[i + 1]; ++k) acc += val[k] * s[ci[k]];
      q[i] = acc;
      z[i] = acc / d[i];
    } else {
      q[i] = 0;
      z[i] = 0;
    }
  }
}

// t = A p with the per-block partial sums of t.t (one pass instead of two)
__global__ void __launch_bounds__(BS) k_spmv_tt(int N, const int* __restrict__ rp, const int* __restrict__ ci,
                                                const double* __restrict__ val, const double* __restrict__ p,
                                                double* __restrict__ t, double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double acc = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double v = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) v += val[k] * p[ci[k]];
    t[i] = v;
    acc += v * v;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// k_normal_q with the per-block partial sums of q.z
__global__ void __launch_bounds__(BS) k_normal_qz(int N, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const double* __restrict__ val, const double* __restrict__ s,
                                                  const unsigned char* __restrict__ state,
                                                  const double* __restrict__ d, double* __restrict__ q,
                                                  double* __restrict__ z, double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double acc = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double v = 0, zi = 0;
    if (state[i] == 0 && d[i] > 0) {
      for (int k = rp[i]; k < rp[i + 1]; ++k) v += val[k] * s[ci[k]];
      zi = v / d[i];
    }
    q[i] = v;
    z[i] = zi;
    acc += v * zi;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(BS) k_dot_part(int64_t n, const double* __restrict__ a,
                                                 const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) s += a[i] * b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sums the partials in a fixed order, then: op 0 gam = sum; op 1 alpha =
// gam / sum; op 2 beta = sum / gam, gam = sum.  One block.
__global__ void __launch_bounds__(BS) k_dot_fin(const double* __restrict__ part, int nb, int op,
                                                double* __restrict__ sc) {
  __shared__ double sh[BS / 64];
  double s = 0;
  for (int k = threadIdx.x; k < nb; k += BS) s += part[k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    if (op == 0) {
      sc[SC_GAM] = s;
    } else if (op == 1) {
      sc[SC_ALPHA] = s > 0 ? sc[SC_GAM] / s : 0.0;
    } else {
      sc[SC_BETA] = sc[SC_GAM] > 0 ? s / sc[SC_GAM] : 0.0;
      sc[SC_GAM] = s;
    }
  }
}

// x += alpha p; s -= alpha t
__global__ void __launch_bounds__(BS) k_step_x(int N, const double* __restrict__ sc, const double* __restrict__ p,
                                               const double* __restrict__ t, double* __restrict__ x,
                                               double* __restrict__ s) {
  const double a = sc[SC_ALPHA];
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    x[i] += a * p[i];
    s[i] -= a * t[i];
  }
}

// p = z + beta p; with restart set, p = z
__global__ void __launch_bounds__(BS) k_step_p(int N, const double* __restrict__ sc, const double* __restrict__ z,
                                               double* __restrict__ p, int restart) {
  const double be = restart ? 0.0 : sc[SC_BETA];
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) p[i] = z[i] + be * p[i];
}

__global__ void __launch_bounds__(BS) k_clamp(int N, const double* __restrict__ lb, const double* __restrict__ ub,
                                              double* __restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) x[i] = fmin(fmax(x[i], lb[i]), ub[i]);
}

// Active-set update from the gradient g: state 0 free, 1 at lb, 2 at ub;
// counts the variables whose first-order condition is violated by more
// than kkt_tol (cnt[0]) before the update.
__global__ void __launch_bounds__(BS) k_sets(int N, const double* __restrict__ g, const double* __restrict__ d,
                                             const double* __restrict__ lb, const double* __restrict__ ub,
                                             double kkt_tol, double* __restrict__ x,
                                             unsigned char* __restrict__ state, unsigned* __restrict__ cnt) {
  unsigned viol = 0;
  for (int i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    const double xi = x[i];
    const double tr = d[i] > 0 ? xi - g[i] / d[i] : xi;
    const double pr = fmin(fmax(tr, lb[i]), ub[i]);
    // written so that a NaN counts as a violation
    if (!(fabs(xi - pr) <= kkt_tol) || !(xi >= lb[i] - kkt_tol) || !(xi <= ub[i] + kkt_tol)) ++viol;
    unsigned char st = 0;
    if (tr <= lb[i]) st = 1;
    else if (tr >= ub[i]) st = 2;
    state[i] = st;
    if (st == 1) x[i] = lb[i];
    else if (st == 2) x[i] = ub[i];
  }
  for (int o = 32; o > 0; o >>= 1) viol += __shfl_down(viol, o, 64);
  if ((threadIdx.x & 63) == 0 && viol) atomicAdd(cnt, viol);
}

struct Work {
  double *d, *s, *g, *t, *z, *p, *part, *sc;
  unsigned char* state;
  unsigned* cnt;
};

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

Work carve(void* base, int N) {
  Work w;
  char* c = (char*)base;
  const size_t vec = align_up(sizeof(double) * (size_t)N);
  double** v[6] = {&w.d, &w.s, &w.g, &w.t, &w.z, &w.p};
  for (int k = 0; k < 6; ++k) {
    *v[k] = (double*)c;
    c += vec;
  }
  w.part = (double*)c;
  c += align_up(sizeof(double) * MAXB);
  w.sc = (double*)c;
  c += align_up(sizeof(double) * SC_N);
  w.cnt = (unsigned*)c;
  c += 256;
  w.state = (unsigned char*)c;
  return w;
}

struct Pinned {
  double* gam = nullptr;
  unsigned* cnt = nullptr;
  Pinned() {
    if (hipHostMalloc((void**)&gam, sizeof(double) * 2, hipHostMallocDefault) != hipSuccess) gam = nullptr;
    if (hipHostMalloc((void**)&cnt, sizeof(unsigned) * 2, hipHostMallocDefault) != hipSuccess) cnt = nullptr;
  }
};

Pinned& pinned() {
  static thread_local Pinned p;
  return p;
}

void dot(int N, const double* a, const double* b, int op, const Work& w, hipStream_t s) {
  const unsigned nb = grid(N);
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(BS), 0, s, (int64_t)N, a, b, w.part);
  hipLaunchKernelGGL(k_dot_fin, dim3(1), dim3(BS), 0, s, (const double*)w.part, (int)nb, op, w.sc);
}

}  // namespace

size_t bounded_lsq_workspace_bytes(int N) {
  return 6 * align_up(sizeof(double) * (size_t)(N > 0 ? N : 0)) + align_up(sizeof(double) * MAXB) +
         align_up(sizeof(double) * SC_N) + 256 + align_up((size_t)(N > 0 ? N : 0));
}

int solve_bounded_lsq(int N, const int* rp, const int* ci, const double* val, const double* b, const double* lb,
                      const double* ub, double* x, void* work, double tol, double kkt_tol, int maxit,
                      int max_outer, double* info, hipStream_t s) {
  Pinned& hp = pinned();
  if (!hp.gam || !hp.cnt) {
    set_error("solve_bounded_lsq: pinned host buffer allocation failed");
    return EFEA_ERR_HIP;
  }
  const Work w = carve(work, N);
  const unsigned nb = grid(N);
  int outer = 0, total = 0, status = EFEA_ERR_NOCONV;
  unsigned viol = 0;
  double gref = -1.0;
  hipLaunchKernelGGL(k_colnorm2, dim3(nb), dim3(BS), 0, s, N, rp, val, w.d);
  hipLaunchKernelGGL(k_clamp, dim3(nb), dim3(BS), 0, s, N, lb, ub, x);
  if (hipStreamSynchronize(s) != hipSuccess) return check_launch("solve_bounded_lsq");
  for (;; ++outer) {
    // gradient g = A (A x - b) = -A s with s = b - A x
    hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)x, -1.0, 1.0, b, w.s);
    hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)w.s, -1.0, 0.0,
                       (const double*)nullptr, w.g);
    hipMemsetAsync(w.cnt, 0, sizeof(unsigned), s);
    hipLaunchKernelGGL(k_sets, dim3(nb), dim3(BS), 0, s, N, (const double*)w.g, (const double*)w.d, lb, ub, kkt_tol,
                       x, w.state, w.cnt);
    hipMemcpyAsync(hp.cnt, w.cnt, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return check_launch("solve_bounded_lsq");
    viol = hp.cnt[0];
    if (viol == 0) {
      // free variables may sit up to kkt_tol outside their bounds: clamp
      // them (a move of at most kkt_tol; the first-order conditions then
      // hold within 2 kkt_tol)
      hipLaunchKernelGGL(k_clamp, dim3(nb), dim3(BS), 0, s, N, lb, ub, x);
      status = EFEA_OK;
      break;
    }
    if (outer >= max_outer) break;
    // CGLS on the free set from x (bounded variables now sit on their bound)
    hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)x, -1.0, 1.0, b, w.s);
    hipLaunchKernelGGL(k_normal_q, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)w.s,
                       (const unsigned char*)w.state, (const double*)w.d, w.g, w.z);
    dot(N, w.g, w.z, 0, w, s);
    hipLaunchKernelGGL(k_step_p, dim3(nb), dim3(BS), 0, s, N, (const double*)w.sc, (const double*)w.z, w.p, 1);
    hipMemcpyAsync(hp.gam, w.sc + SC_GAM, sizeof(double), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return check_launch("solve_bounded_lsq");
    // each pass reduces the scaled free-set gradient norm by tol from where it starts
    if (gref < 0) gref = hp.gam[0];
    double gam = hp.gam[0];
    // (a bound on the scaled gradient alone does not bound the distance to
    // the minimiser on ill-conditioned systems, so the pass runs to tol)
    const double stop = tol * tol * gam;
    for (int it = 0; gam > stop && it < maxit;) {
      for (int k = 0; k < POLL; ++k, ++it) {
        hipLaunchKernelGGL(k_spmv_tt, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)w.p, w.t, w.part);
        hipLaunchKernelGGL(k_dot_fin, dim3(1), dim3(BS), 0, s, (const double*)w.part, (int)nb, 1, w.sc);
        hipLaunchKernelGGL(k_step_x, dim3(nb), dim3(BS), 0, s, N, (const double*)w.sc, (const double*)w.p,
                           (const double*)w.t, x, w.s);
        hipLaunchKernelGGL(k_normal_qz, dim3(nb), dim3(BS), 0, s, N, rp, ci, val, (const double*)w.s,
                           (const unsigned char*)w.state, (const double*)w.d, w.g, w.z, w.part);
        hipLaunchKernelGGL(k_dot_fin, dim3(1), dim3(BS), 0, s, (const double*)w.part, (int)nb, 2, w.sc);
        hipLaunchKernelGGL(k_step_p, dim3(nb), dim3(BS), 0, s, N, (const double*)w.sc, (const double*)w.z, w.p, 0);
      }
      hipMemcpyAsync(hp.gam, w.sc + SC_GAM, sizeof(double), hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) return check_launch("solve_bounded_lsq");
      gam = hp.gam[0];
      total += POLL;
      if (!(gam == gam)) break;  // NaN: reported through the violation count
    }
  }
  if (info) {
    info[0] = outer;
    info[1] = total;
    info[2] = viol;
    info[3] = gref;
  }
  if (status != EFEA_OK) {
    set_error("solve_bounded_lsq: %u variables violate the first-order conditions after %d active-set updates",
              viol, outer);
    return EFEA_ERR_NOCONV;
  }
  return check_launch("solve_bounded_lsq");
}

}  // namespace efea
