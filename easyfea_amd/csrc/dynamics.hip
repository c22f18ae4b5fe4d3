// Time-integration kernels: the mass / capacity element operator, linear
// combinations of vectors and of matrix values over one pattern, and the
// per-scheme state update after a time-step solve.  FP64 throughout; this
// file is compiled without FMA contraction so that the state update performs
// the reference's operations one by one, in its order, and matches the
// numpy formulas bit for bit.
#include "efea_common.h"

namespace efea {

// Bilinear.UV (EasyFEA/FEM/Operators/Bilinear.py:42-59 with
// Get_ReactionPart_e_pg, FEM/_group_elem.py:1337-1360):
//   M_e[(a,i),(b,j)] = delta_ij sum_p coef_p wJ_p N_a(p) N_b(p)   (x scale)
// over the rule in `tab` (the mass rule), DOF order node-major
// (a * DOFN + i).  coef: per point (stride NPG), per element (stride 1) or
// the constant coef_const (coef null).
// One workgroup holds EPB = 256 / nPe^2 elements: lanes (element, point)
// put wJ into LDS, lanes (element, a, b) form the scalar entry, and the
// expanded dof_n^2 blocks leave LDS as one coalesced store.
template <typename T, int ET, int NPG, int DOFN>
__global__ void __launch_bounds__(256)
k_uv(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord, const double* __restrict__ tab,
     const double* __restrict__ coef, int coef_stride, double coef_const, double scale, T* __restrict__ Ke) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM;
  constexpr int NPAIR = NPE * NPE;
  constexpr int EPB = 256 / NPAIR;
  constexpr int ND = NPE * DOFN;
  using TB = Tab<NPG, NPE, DIM>;
  __shared__ double s_wJ[EPB][NPG];
  __shared__ double s_M[EPB * NPAIR];

  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  const int t = threadIdx.x;
  if (t < EPB * NPG) {
    const int el = t / NPG, p = t % NPG;
    const int64_t e = e0 + el;
    if (e < Ne) {
      double X[NPE][3];
#pragma unroll
      for (int n = 0; n < NPE; ++n) {
        const int node = connect[e * NPE + n];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
      }
      double dNx[DIM][NPE];
      const double detJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
      const double c = coef ? coef[coef_stride == 1 ? e : e * NPG + p] : coef_const;
      s_wJ[el][p] = c * (detJ * tab[TB::W + p]);
    }
  }
  __syncthreads();
  if (t < EPB * NPAIR) {
    const int el = t / NPAIR, ab = t % NPAIR, a = ab / NPE, b = ab % NPE;
    if (e0 + el < Ne) {
      double m = 0;
#pragma unroll
      for (int p = 0; p < NPG; ++p) m += s_wJ[el][p] * (tab[TB::N + p * NPE + a] * tab[TB::N + p * NPE + b]);
      s_M[t] = m * scale;
    }
  }
  __syncthreads();
  const int64_t nval = (int64_t)min((int64_t)EPB, Ne - e0) * ND * ND;
  T* out = Ke + e0 * ND * ND;
  for (int64_t k = t; k < nval; k += 256) {
    const int el = (int)(k / (ND * ND)), r = (int)(k % (ND * ND));
    const int A = r / ND, B = r % ND;
    double v = 0;
    if (A % DOFN == B % DOFN) v = s_M[el * NPAIR + (A / DOFN) * NPE + B / DOFN];
    out[k] = (T)v;
  }
}

// out = c0 x0 + c1 x1 + c2 x2 + c3 x3 (terms with a null x are absent).
// Every lane reads its entries before writing, so out may be one of the x.
__global__ void __launch_bounds__(256)
k_lincomb(int64_t n, double c0, const double* x0, double c1, const double* x1, double c2, const double* x2,
          double c3, const double* x3, double* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double s = c0 * x0[i];
    if (x1) s += c1 * x1[i];
    if (x2) s += c2 * x2[i];
    if (x3) s += c3 * x3[i];
    out[i] = s;
  }
}

// _Solver_Update_solutions (EasyFEA/Simulations/_simu.py:1552-1640): the new
// (u, v, a) from the previous state and the solve's unknown x.  Each scheme
// evaluates the reference's expressions operation by operation, scalars
// first where Python evaluates them first.  Outputs are written after all
// inputs of the lane are read, so they may alias the inputs.
__global__ void __launch_bounds__(256)
k_time_update(int scheme, int64_t n, double dt, double beta, double gamma, double alpha, const double* u_n,
              const double* v_n, const double* a_n, const double* x, double* u_out, double* v_out, double* a_out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double un = u_n[i], vn = v_n[i], an = a_n[i], u1 = x[i];
    double u = u1, v = vn, a = an;
    switch (scheme) {
      case EFEA_SCHEME_PARABOLIC: {  // :1583-1590
        const double vt = un + ((1 - alpha) * dt) * vn;
        v = (u1 - vt) / (alpha * dt);
        break;
      }
      case EFEA_SCHEME_NEWMARK:
      case EFEA_SCHEME_HHT_NEWMARK: {  // :1592-1604
        const double ut = (un + dt * vn) + ((dt * dt) / 2 * (1 - 2 * beta)) * an;
        const double vt = vn + (dt * (1 - gamma)) * an;
        a = (u1 - ut) / (beta * (dt * dt));
        v = vt + (gamma * dt) * a;
        break;
      }
      case EFEA_SCHEME_MIDPOINT: {  // :1606-1613
        v = (2 / dt) * (u1 - un) - vn;
        a = (2 / dt) * (v - vn) - an;
        break;
      }
      case EFEA_SCHEME_HHT: {  // :1615-1624
        a = (1 / (beta * dt)) * ((u1 - un) / dt - vn) + (1 - 1 / (2 * beta)) * an;
        v = dt * ((1 - gamma) * an + gamma * a) + vn;
        break;
      }
      case EFEA_SCHEME_EULER_IMPLICIT: {  // :1626-1630
        v = (u1 - un) / dt;
        a = (v - vn) / dt;
        break;
      }
      case EFEA_SCHEME_EULER_EXPLICIT: {  // :1632-1638 (x is the acceleration a^n)
        a = u1;
        u = un + dt * vn;
        v = vn + dt * a;
        break;
      }
      default:  // elliptic: u only
        break;
    }
    u_out[i] = u;
    if (v_out) v_out[i] = v;
    if (a_out) a_out[i] = a;
  }
}

template <typename T, int ET, int NPG, int DOFN>
static int launch_uv(int64_t Ne, const int* connect, const double* coord, const double* tab, const double* coef,
                     int cs, double cc, double scale, void* Ke, hipStream_t s) {
  constexpr int NPAIR = Elem<ET>::NPE * Elem<ET>::NPE;
  constexpr int EPB = 256 / NPAIR;
  const int64_t nb = (Ne + EPB - 1) / EPB;
  if (nb == 0) return EFEA_OK;
  hipLaunchKernelGGL((k_uv<T, ET, NPG, DOFN>), dim3((unsigned)nb), dim3(256), 0, s, Ne, connect, coord, tab, coef,
                     cs, cc, scale, (T*)Ke);
  return check_launch("uv_ke");
}

template <typename T, int ET, int NPG>
static int uv_dofn(int dof_n, int64_t Ne, const int* connect, const double* coord, const double* tab,
                   const double* coef, int cs, double cc, double scale, void* Ke, hipStream_t s) {
  if (dof_n == 1) return launch_uv<T, ET, NPG, 1>(Ne, connect, coord, tab, coef, cs, cc, scale, Ke, s);
  if (dof_n == Elem<ET>::DIM) {
    return launch_uv<T, ET, NPG, Elem<ET>::DIM>(Ne, connect, coord, tab, coef, cs, cc, scale, Ke, s);
  }
  set_error("uv_ke: dof_n %d must be 1 or the element dimension", dof_n);
  return EFEA_ERR_ARG;
}

int uv_ke(int elem, int prec, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
          const double* coef, int coef_stride, double coef_const, int dof_n, double scale, void* Ke,
          hipStream_t s) {
  if (coef && coef_stride != 1 && coef_stride != nPg) {
    set_error("uv_ke: coef_stride %d must be 1 (per element) or nPg = %d (per point)", coef_stride, nPg);
    return EFEA_ERR_ARG;
  }
#define UV(T, ET, NPG) return uv_dofn<T, ET, NPG>(dof_n, Ne, connect, coord, tab, coef, coef_stride, coef_const, \
                                                  scale, Ke, s)
  if (prec == 64) {
    if (elem == ET_TRI3 && nPg == 3) UV(double, ET_TRI3, 3);
    if (elem == ET_QUAD4 && nPg == 4) UV(double, ET_QUAD4, 4);
    if (elem == ET_HEXA8 && nPg == 8) UV(double, ET_HEXA8, 8);
    if (elem == ET_TETRA4 && nPg == 4) UV(double, ET_TETRA4, 4);
    if (elem == ET_TETRA10 && nPg == 4) UV(double, ET_TETRA10, 4);
  }
#undef UV
  set_error("uv_ke: unsupported element %d / nPg %d / precision %d", elem, nPg, prec);
  return EFEA_ERR_ARG;
}

static unsigned grid_for(int64_t n) {
  const int64_t nb = (n + 255) / 256;
  return (unsigned)(nb < 4096 ? nb : 4096);
}

int lincomb(int64_t n, double c0, const double* x0, double c1, const double* x1, double c2, const double* x2,
            double c3, const double* x3, double* out, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(256), 0, s, n, c0, x0, c1, x1, c2, x2, c3, x3, out);
  return check_launch("lincomb");
}

int time_update(int scheme, int64_t n, double dt, double beta, double gamma, double alpha, const double* u_n,
                const double* v_n, const double* a_n, const double* x, double* u_out, double* v_out,
                double* a_out, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_time_update, dim3(grid_for(n)), dim3(256), 0, s, scheme, n, dt, beta, gamma, alpha, u_n,
                       v_n, a_n, x, u_out, v_out, a_out);
  return check_launch("time_update");
}

}  // namespace efea
