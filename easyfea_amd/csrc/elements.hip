// Element kernels: isotropic linear-elastic stiffness, reaction/diffusion
// (UV, GradU_A_GradV) and source (V) element operators.
//
// Work decomposition (one workgroup of 256 lanes holds EPB elements):
//   phase 1  one lane per (element, integration point) computes the
//            Jacobian, |det| * w and the physical derivatives into LDS;
//   phase 2  one lane per (element, node pair a,b) accumulates the
//            dim x dim block K_ab over the integration points;
//   phase 3  the staged element matrices leave LDS as one contiguous,
//            coalesced store (the EPB matrices are adjacent in HBM).
// The element stiffness uses the isotropic closed form
//   K_ab,ij = sum_p wJ (lam dN_a,i dN_b,j + mu dN_a,j dN_b,i + mu d_ij dN_a.dN_b)
// which equals wJ * B^T C B for the Kelvin-Mandel C of
// EasyFEA/Models/Elastic/_laws.py:407-482 (shear rows carry 1/sqrt(2), the
// shear diagonal of C carries 2 mu), at a tenth of the dense flops.
#include "efea_common.h"

namespace efea {

template <typename T, int ET, int NPG>
__global__ void __launch_bounds__(256)
k_elastic_iso(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
              const double* __restrict__ tab, const double* __restrict__ E, int E_stride,
              double nu, int planeStress, double thickness, T* __restrict__ Ke) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, ND = NPE * DIM;
  constexpr int NPAIR = NPE * NPE;
  constexpr int EPB = 256 / NPAIR;
  using TB = Tab<NPG, NPE, DIM>;
  __shared__ T s_dNx[EPB][NPG][DIM][NPE];
  __shared__ T s_wJ[EPB][NPG];
  __shared__ T s_K[EPB * ND * ND];

  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  const int t = threadIdx.x;

  if (t < EPB * NPG) {
    const int el = t / NPG, p = t % NPG;
    const int64_t e = e0 + el;
    if (e < Ne) {
      T X[NPE][3];
#pragma unroll
      for (int n = 0; n < NPE; ++n) {
        const int node = connect[e * NPE + n];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[n][k] = (T)coord[(int64_t)node * 3 + k];
      }
      T dNx[DIM][NPE];
      const T detJ = point_geometry<T, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
      s_wJ[el][p] = detJ * (T)tab[TB::W + p];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int n = 0; n < NPE; ++n) s_dNx[el][p][i][n] = dNx[i][n];
    }
  }
  __syncthreads();

  if (t < EPB * NPAIR) {
    const int el = t / NPAIR, ab = t % NPAIR, a = ab / NPE, b = ab % NPE;
    const int64_t e = e0 + el;
    if (e < Ne) {
      const double Ee = E[E_stride ? e : 0];
      double lam = Ee * nu / ((1 + nu) * (1 - 2 * nu));
      if (DIM == 2 && planeStress) lam = Ee * nu / (1 - nu * nu);
      const double mu = Ee / (2 * (1 + nu));
      const T tl = (T)lam, tm = (T)mu;
      T acc[DIM][DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) acc[i][j] = 0;
#pragma unroll
      for (int p = 0; p < NPG; ++p) {
        const T w = s_wJ[el][p];
        T ga[DIM], gb[DIM], dot = 0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          ga[i] = s_dNx[el][p][i][a];
          gb[i] = s_dNx[el][p][i][b];
          dot += ga[i] * gb[i];
        }
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int j = 0; j < DIM; ++j) {
            T v = tl * ga[i] * gb[j] + tm * ga[j] * gb[i];
            if (i == j) v += tm * dot;
            acc[i][j] += w * v;
          }
      }
      const T th = (T)(DIM == 2 ? thickness : 1.0);
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j)
          s_K[el * ND * ND + (a * DIM + i) * ND + b * DIM + j] = acc[i][j] * th;
    }
  }
  __syncthreads();

  const int64_t nval = (int64_t)min((int64_t)EPB, Ne - e0) * ND * ND;
  T* out = Ke + e0 * ND * ND;
  for (int64_t k = t; k < nval; k += 256) out[k] = s_K[k];
}

// Reaction + diffusion + source for scalar fields:
//   K_e = sum_pm r_p wJ N^T N + k sum_pr wJ dN^T dN ,  F_e = sum_pm f_p wJ N
// (Operators/Bilinear.py:42-59, 229-249 with A = I; Operators/Linear.py:18-35).
// r, f are per mass point (Ne, nPg_m) or null; kdiff per element or scalar.
template <typename T, int ET, int NPGM, int NPGR>
__global__ void __launch_bounds__(256)
k_reaction_diffusion(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
                     const double* __restrict__ tab_m, const double* __restrict__ tab_r,
                     const double* __restrict__ r, double r_const, const double* __restrict__ kdiff,
                     int k_stride, const double* __restrict__ f, double scale,
                     T* __restrict__ Ke, T* __restrict__ Fe) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM;
  constexpr int NPAIR = NPE * NPE;
  constexpr int EPB = 256 / NPAIR;
  constexpr int NPGT = NPGM + NPGR;
  using TM = Tab<NPGM, NPE, DIM>;
  using TR = Tab<NPGR, NPE, DIM>;
  __shared__ double s_dNx[EPB][NPGR][DIM][NPE];
  __shared__ double s_wJr[EPB][NPGR];
  __shared__ double s_wJm[EPB][NPGM];
  __shared__ T s_K[EPB * NPE * NPE];

  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  const int t = threadIdx.x;
  if (t < EPB * NPGT) {
    const int el = t / NPGT, q = t % NPGT;
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
      if (q < NPGM) {
        const double detJ = point_geometry<double, NPE, DIM>(tab_m + TM::DN + q * DIM * NPE, X, dNx);
        s_wJm[el][q] = detJ * tab_m[TM::W + q];
      } else {
        const int p = q - NPGM;
        const double detJ = point_geometry<double, NPE, DIM>(tab_r + TR::DN + p * DIM * NPE, X, dNx);
        s_wJr[el][p] = detJ * tab_r[TR::W + p];
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int n = 0; n < NPE; ++n) s_dNx[el][p][i][n] = dNx[i][n];
      }
    }
  }
  __syncthreads();

  if (t < EPB * NPAIR) {
    const int el = t / NPAIR, ab = t % NPAIR, a = ab / NPE, b = ab % NPE;
    const int64_t e = e0 + el;
    if (e < Ne) {
      double R = 0;
#pragma unroll
      for (int p = 0; p < NPGM; ++p) {
        const double rp = r ? r[e * NPGM + p] : r_const;
        R += rp * s_wJm[el][p] * tab_m[TM::N + p * NPE + a] * tab_m[TM::N + p * NPE + b];
      }
      double D = 0;
#pragma unroll
      for (int p = 0; p < NPGR; ++p) {
        double dot = 0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) dot += s_dNx[el][p][i][a] * s_dNx[el][p][i][b];
        D += s_wJr[el][p] * dot;
      }
      const double kd = kdiff ? kdiff[k_stride ? e : 0] : 0.0;
      s_K[el * NPAIR + ab] = (T)((R + kd * D) * scale);
      if (Fe && b == 0) {
        double F = 0;
#pragma unroll
        for (int p = 0; p < NPGM; ++p) F += f[e * NPGM + p] * s_wJm[el][p] * tab_m[TM::N + p * NPE + a];
        Fe[e * NPE + a] = (T)(F * scale);
      }
    }
  }
  __syncthreads();
  const int64_t nval = (int64_t)min((int64_t)EPB, Ne - e0) * NPAIR;
  T* out = Ke + e0 * NPAIR;
  for (int64_t k = t; k < nval; k += 256) out[k] = s_K[k];
}

// Load integral sum_p wJ f_p N_a (EasyFEA/Simulations/_simu.py:2770-2842)
// on line, surface or volume elements.  Coordinates arrive already expressed
// in each element's own frame (Ne, nPe, dim) so that the Jacobian is square.
template <int NPE, int DIM, int NPG>
__global__ void k_boundary_load(int64_t Ne, const double* __restrict__ Xloc, const double* __restrict__ tab,
                                const double* __restrict__ f, double* __restrict__ out) {
  using TB = Tab<NPG, NPE, DIM>;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ne) return;
  double X[NPE][3];
#pragma unroll
  for (int n = 0; n < NPE; ++n)
#pragma unroll
    for (int k = 0; k < 3; ++k) X[n][k] = k < DIM ? Xloc[(e * NPE + n) * DIM + k] : 0.0;
  double acc[NPE];
#pragma unroll
  for (int n = 0; n < NPE; ++n) acc[n] = 0;
#pragma unroll
  for (int p = 0; p < NPG; ++p) {
    double dNx[DIM][NPE];
    const double wJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx) * tab[TB::W + p];
    const double fp = f[e * NPG + p];
#pragma unroll
    for (int n = 0; n < NPE; ++n) acc[n] += wJ * fp * tab[TB::N + p * NPE + n];
  }
#pragma unroll
  for (int n = 0; n < NPE; ++n) out[e * NPE + n] = acc[n];
}

// ---------------------------------------------------------------------------
// launchers

template <typename T, int ET, int NPG>
static int launch_elastic(int64_t Ne, const int* connect, const double* coord, const double* tab,
                          const double* E, int E_stride, double nu, int ps, double th, void* Ke,
                          hipStream_t s) {
  constexpr int NPAIR = Elem<ET>::NPE * Elem<ET>::NPE;
  constexpr int EPB = 256 / NPAIR;
  const int64_t nb = (Ne + EPB - 1) / EPB;
  if (nb == 0) return EFEA_OK;
  hipLaunchKernelGGL((k_elastic_iso<T, ET, NPG>), dim3((unsigned)nb), dim3(256), 0, s, Ne, connect, coord,
                     tab, E, E_stride, nu, ps, th, (T*)Ke);
  return check_launch("elastic_ke");
}

int elastic_ke(int elem, int prec, int64_t Ne, const int* connect, const double* coord, const double* tab,
               int nPg, const double* E, int E_stride, double nu, int ps, double th, void* Ke, hipStream_t s) {
#define EL(T, ET, NPG) return launch_elastic<T, ET, NPG>(Ne, connect, coord, tab, E, E_stride, nu, ps, th, Ke, s)
  if (prec == 64) {
    if (elem == ET_TRI3 && nPg == 1) EL(double, ET_TRI3, 1);
    if (elem == ET_QUAD4 && nPg == 4) EL(double, ET_QUAD4, 4);
    if (elem == ET_HEXA8 && nPg == 8) EL(double, ET_HEXA8, 8);
    if (elem == ET_TETRA4 && nPg == 1) EL(double, ET_TETRA4, 1);
    if (elem == ET_TETRA10 && nPg == 4) EL(double, ET_TETRA10, 4);
  } else if (prec == 32) {
    if (elem == ET_TRI3 && nPg == 1) EL(float, ET_TRI3, 1);
    if (elem == ET_QUAD4 && nPg == 4) EL(float, ET_QUAD4, 4);
    if (elem == ET_HEXA8 && nPg == 8) EL(float, ET_HEXA8, 8);
    if (elem == ET_TETRA4 && nPg == 1) EL(float, ET_TETRA4, 1);
    if (elem == ET_TETRA10 && nPg == 4) EL(float, ET_TETRA10, 4);
  }
#undef EL
  set_error("elastic_ke: unsupported element %d / nPg %d / precision %d", elem, nPg, prec);
  return EFEA_ERR_ARG;
}

template <typename T, int ET, int NPGM, int NPGR>
static int launch_rd(int64_t Ne, const int* connect, const double* coord, const double* tm, const double* tr,
                     const double* r, double rc, const double* kd, int ks, const double* f, double scale,
                     void* Ke, void* Fe, hipStream_t s) {
  constexpr int NPAIR = Elem<ET>::NPE * Elem<ET>::NPE;
  constexpr int EPB = 256 / NPAIR;
  const int64_t nb = (Ne + EPB - 1) / EPB;
  if (nb == 0) return EFEA_OK;
  hipLaunchKernelGGL((k_reaction_diffusion<T, ET, NPGM, NPGR>), dim3((unsigned)nb), dim3(256), 0, s, Ne,
                     connect, coord, tm, tr, r, rc, kd, ks, f, scale, (T*)Ke, (T*)Fe);
  return check_launch("reaction_diffusion");
}

int reaction_diffusion(int elem, int prec, int64_t Ne, const int* connect, const double* coord,
                       const double* tab_m, int nPgm, const double* tab_r, int nPgr, const double* r,
                       double r_const, const double* kdiff, int k_stride, const double* f, double scale,
                       void* Ke, void* Fe, hipStream_t s) {
#define RD(T, ET, M, R) \
  return launch_rd<T, ET, M, R>(Ne, connect, coord, tab_m, tab_r, r, r_const, kdiff, k_stride, f, scale, Ke, Fe, s)
  if (prec == 64) {
    if (elem == ET_TRI3 && nPgm == 3 && nPgr == 1) RD(double, ET_TRI3, 3, 1);
    if (elem == ET_QUAD4 && nPgm == 4 && nPgr == 4) RD(double, ET_QUAD4, 4, 4);
    if (elem == ET_HEXA8 && nPgm == 8 && nPgr == 8) RD(double, ET_HEXA8, 8, 8);
    if (elem == ET_TETRA4 && nPgm == 4 && nPgr == 1) RD(double, ET_TETRA4, 4, 1);
    if (elem == ET_TETRA10 && nPgm == 4 && nPgr == 4) RD(double, ET_TETRA10, 4, 4);
  } else if (prec == 32) {
    if (elem == ET_TRI3 && nPgm == 3 && nPgr == 1) RD(float, ET_TRI3, 3, 1);
    if (elem == ET_QUAD4 && nPgm == 4 && nPgr == 4) RD(float, ET_QUAD4, 4, 4);
    if (elem == ET_HEXA8 && nPgm == 8 && nPgr == 8) RD(float, ET_HEXA8, 8, 8);
    if (elem == ET_TETRA4 && nPgm == 4 && nPgr == 1) RD(float, ET_TETRA4, 4, 1);
    if (elem == ET_TETRA10 && nPgm == 4 && nPgr == 4) RD(float, ET_TETRA10, 4, 4);
  }
#undef RD
  set_error("reaction_diffusion: unsupported element %d / nPg %d,%d / precision %d", elem, nPgm, nPgr, prec);
  return EFEA_ERR_ARG;
}

int boundary_load(int elem, int64_t Ne, const double* Xloc, const double* tab, int nPg, const double* f,
                  double* out, hipStream_t s) {
  const int64_t nb = (Ne + 127) / 128;
  if (nb == 0) return EFEA_OK;
  if (elem == ET_SEG2 && nPg == 2) {
    hipLaunchKernelGGL((k_boundary_load<2, 1, 2>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  if (elem == ET_QUAD4 && nPg == 4) {
    hipLaunchKernelGGL((k_boundary_load<4, 2, 4>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  if (elem == ET_TRI3 && nPg == 3) {
    hipLaunchKernelGGL((k_boundary_load<3, 2, 3>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  // body loads of 3D meshes integrate over the volume elements themselves
  if (elem == ET_TETRA4 && nPg == 4) {
    hipLaunchKernelGGL((k_boundary_load<4, 3, 4>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  if (elem == ET_TETRA10 && nPg == 4) {
    hipLaunchKernelGGL((k_boundary_load<10, 3, 4>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  if (elem == ET_HEXA8 && nPg == 8) {
    hipLaunchKernelGGL((k_boundary_load<8, 3, 8>), dim3((unsigned)nb), dim3(128), 0, s, Ne, Xloc, tab, f, out);
    return check_launch("boundary_load");
  }
  set_error("boundary_load: unsupported element %d / nPg %d", elem, nPg);
  return EFEA_ERR_ARG;
}

}  // namespace efea
