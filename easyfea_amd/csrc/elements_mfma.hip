// Dense 3D element stiffness on the FP64 matrix cores:
//   K_e = sum_p wJ_p B_p^T D B_p        (Operators/Bilinear.py:62-79)
// for HEXA8 (24 x 24, 8 points) and TETRA10 (30 x 30, 4 points), with any
// Kelvin-Mandel constitutive matrix D (per element or shared) or the
// isotropic one formed from E and nu.
//
// One wavefront per element, four per workgroup.  Stacking the points, the
// element matrix is one product K = Bs^T Ys with Bs = [B_0; ...; B_(P-1)]
// (6P x ND) and Ys = [wJ_0 D B_0; ...] (6P x ND):
//   phase 1  lane (p, a) forms the Jacobian at point p, the physical
//            derivatives of node a and node a's three columns of B_p and of
//            wJ_p D B_p, into the wave's LDS tiles;
//   phase 2  the 2 x 2 output tiles of 16 x 16 (ND padded to 32) accumulate
//            6P / 4 steps of v_mfma_f64_16x16x4_f64 each (A = Bs^T, B = Ys,
//            operands read from LDS; padding lanes feed zeros);
//   phase 3  each lane stores its four values of every tile, rows of 16
//            consecutive columns per quarter-wave.
// Lane layout of the 16x16x4 FP64 instruction (checked on the device): lane
// l holds A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register j
// holds D[l / 16 + 4 j][l % 16].
#include "efea_common.h"

namespace efea {

typedef double v4d __attribute__((ext_vector_type(4)));

template <int ET, int NPG>
__global__ void __launch_bounds__(256)
k_elastic_mfma(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
               const double* __restrict__ tab, const double* __restrict__ Dmat, int D_stride,
               const double* __restrict__ E, int E_stride, double nu, double* __restrict__ Ke) {
  constexpr int NPE = Elem<ET>::NPE, ND = NPE * 3, KD = 6 * NPG, WAVES = 4;
  static_assert(Elem<ET>::DIM == 3 && KD % 4 == 0 && ND <= 32, "3D elements of at most 32 DOFs");
  using TB = Tab<NPG, NPE, 3>;
  __shared__ double sB[WAVES][KD][ND];
  __shared__ double sY[WAVES][KD][ND];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int64_t e = (int64_t)blockIdx.x * WAVES + w;
  const bool live = e < Ne;

  // phase 1: geometry, B and wJ D B at (point p, node a)
  if (live && l < NPG * NPE) {
    const int p = l / NPE, a = l % NPE;
    double X[NPE][3];
#pragma unroll
    for (int n = 0; n < NPE; ++n) {
      const int node = connect[e * NPE + n];
#pragma unroll
      for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
    }
    const double* dN = tab + TB::DN + p * 3 * NPE;
    double F[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double s = 0;
#pragma unroll
        for (int n = 0; n < NPE; ++n) s += dN[d * NPE + n] * X[n][k];
        F[d][k] = s;
      }
    const double det = F[0][0] * (F[1][1] * F[2][2] - F[2][1] * F[1][2]) -
                       F[0][1] * (F[1][0] * F[2][2] - F[2][0] * F[1][2]) +
                       F[0][2] * (F[1][0] * F[2][1] - F[2][0] * F[1][1]);
    const double r = 1.0 / det;
    double inv[3][3];
    inv[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * r;
    inv[0][1] = -(F[0][1] * F[2][2] - F[0][2] * F[2][1]) * r;
    inv[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * r;
    inv[1][0] = -(F[1][0] * F[2][2] - F[1][2] * F[2][0]) * r;
    inv[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * r;
    inv[1][2] = -(F[0][0] * F[1][2] - F[0][2] * F[1][0]) * r;
    inv[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * r;
    inv[2][1] = -(F[0][0] * F[2][1] - F[0][1] * F[2][0]) * r;
    inv[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * r;
    double g[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = inv[i][0] * dN[a] + inv[i][1] * dN[NPE + a] + inv[i][2] * dN[2 * NPE + a];
    const double wJ = (det < 0 ? -det : det) * tab[TB::W + p];
    // the constitutive matrix (Kelvin-Mandel, rows xx yy zz yz xz xy)
    double D[6][6];
    if (Dmat) {
      const double* Dp = Dmat + (D_stride ? e * 36 : 0);
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) D[i][j] = Dp[i * 6 + j];
    } else {
      const double Ee = E[E_stride ? e : 0];
      const double lam = Ee * nu / ((1 + nu) * (1 - 2 * nu)), mu = Ee / (2 * (1 + nu));
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) D[i][j] = (i < 3 && j < 3) ? lam + (i == j ? 2 * mu : 0.0) : (i == j ? 2 * mu : 0.0);
    }
    // node a's columns of B_p (shear rows carry 1/sqrt(2))
    const double c = 0.70710678118654752440;
    double Bc[3][6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Bc[0][k] = Bc[1][k] = Bc[2][k] = 0;
    Bc[0][0] = g[0]; Bc[0][4] = g[2] * c; Bc[0][5] = g[1] * c;
    Bc[1][1] = g[1]; Bc[1][3] = g[2] * c; Bc[1][5] = g[0] * c;
    Bc[2][2] = g[2]; Bc[2][3] = g[1] * c; Bc[2][4] = g[0] * c;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double y = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) y += D[i][k] * Bc[cc][k];
        sB[w][6 * p + i][3 * a + cc] = Bc[cc][i];
        sY[w][6 * p + i][3 * a + cc] = wJ * y;
      }
  }
  __syncthreads();

  // phases 2 and 3: K = Bs^T Ys on the matrix cores
  if (live) {
    const int q = l / 16, m = l % 16;
    double* out = Ke + e * ND * ND;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const int ca = mt * 16 + m, cb = nt * 16 + m;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KD / 4; ++s) {
          const int k = 4 * s + q;
          const double av = ca < ND ? sB[w][k][ca] : 0.0;
          const double bv = cb < ND ? sY[w][k][cb] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = mt * 16 + q + 4 * j;
          if (row < ND && cb < ND) out[row * ND + cb] = acc[j];
        }
      }
  }
}

template <int ET, int NPG>
static int launch_elastic_mfma(int64_t Ne, const int* connect, const double* coord, const double* tab,
                               const double* D, int D_stride, const double* E, int E_stride, double nu, double* Ke,
                               hipStream_t s) {
  const int64_t nb = (Ne + 3) / 4;
  if (nb == 0) return EFEA_OK;
  hipLaunchKernelGGL((k_elastic_mfma<ET, NPG>), dim3((unsigned)nb), dim3(256), 0, s, Ne, connect, coord, tab, D,
                     D_stride, E, E_stride, nu, Ke);
  return check_launch("elastic_ke_dense");
}

int elastic_ke_dense(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
                     const double* D, int D_stride, const double* E, int E_stride, double nu, double* Ke,
                     hipStream_t s) {
  if (elem == ET_HEXA8 && nPg == 8)
    return launch_elastic_mfma<ET_HEXA8, 8>(Ne, connect, coord, tab, D, D_stride, E, E_stride, nu, Ke, s);
  if (elem == ET_TETRA10 && nPg == 4)
    return launch_elastic_mfma<ET_TETRA10, 4>(Ne, connect, coord, tab, D, D_stride, E, E_stride, nu, Ke, s);
  set_error("elastic_ke_dense: element %d with %d points (HEXA8 / 8 and TETRA10 / 4 only)", elem, nPg);
  return EFEA_ERR_ARG;
}

}  // namespace efea
