// Device post-processing of linear-elastic results: per element, the strain
// and stress averaged over the stiffness integration points, their von Mises
// measures and the deformation energy; and element -> node averaging.
//
// Per element e and integration point p (EasyFEA/Models/Elastic/_laws.py:127-185):
//   eps_p = B_p u_e        Kelvin-Mandel [xx, yy, (zz), sqrt2 yz, sqrt2 xz, sqrt2 xy]
//   sig_p = C eps_p        isotropic: sig = lam tr(eps) I + 2 mu eps (componentwise in KM)
// Reported per element (Models/_utils.py:302-434): the tensor components
// (shear divided by sqrt 2), averaged over the points with equal weights, the
// average of the pointwise von Mises measures
//   2D: sqrt(xx^2 + yy^2 - xx yy + 3 xy^2)
//   3D: sqrt(((xx-yy)^2 + (yy-zz)^2 + (zz-xx)^2 + 6 (xy^2 + yz^2 + xz^2)) / 2)
// and W_e = 1/2 sum_p wJ_p sig_p . eps_p (x thickness in 2D)
// (Simulations/_elastic.py:323-380).  Output row of element e:
//   [eps comps (ns), Evm, sig comps (ns), Svm, W_e]  (2 ns + 3 values).
// One lane per element; the element's nodes are gathered once.
#include "efea_common.h"

namespace efea {

template <int DIM>
struct KM {
  static constexpr int NS = DIM == 2 ? 3 : 6;
};

template <int DIM>
__device__ inline double von_mises(const double (&t)[KM<DIM>::NS]) {
  if (DIM == 2) {
    const double xx = t[0], yy = t[1], xy = t[2];
    return sqrt(xx * xx + yy * yy - xx * yy + 3 * xy * xy);
  } else {
    const double xx = t[0], yy = t[1], zz = t[2], yz = t[3], xz = t[4], xy = t[5];
    return sqrt(0.5 * ((xx - yy) * (xx - yy) + (yy - zz) * (yy - zz) + (zz - xx) * (zz - xx) +
                       6 * (xy * xy + yz * yz + xz * xz)));
  }
}

template <int ET, int NPG>
__global__ void __launch_bounds__(256)
k_strain_stress(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
                const double* __restrict__ tab, const double* __restrict__ u, const double* __restrict__ E,
                int E_stride, double nu, int planeStress, double thickness, double* __restrict__ out) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, NS = KM<DIM>::NS, NOUT = 2 * NS + 3;
  using TB = Tab<NPG, NPE, DIM>;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ne) return;
  const double isq2 = 0.70710678118654752440;  // 1 / sqrt(2)
  double X[NPE][3], ue[NPE][DIM];
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const int node = connect[e * NPE + n];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
#pragma unroll
    for (int k = 0; k < DIM; ++k) ue[n][k] = u[(int64_t)node * DIM + k];
  }
  const double Ee = E[E_stride ? e : 0];
  double lam = Ee * nu / ((1 + nu) * (1 - 2 * nu));
  if (DIM == 2 && planeStress) lam = Ee * nu / (1 - nu * nu);
  const double mu = Ee / (2 * (1 + nu));
  double se[NS], ss[NS], evm = 0, svm = 0, w = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) se[k] = ss[k] = 0;
#pragma unroll
  for (int p = 0; p < NPG; ++p) {
    double dNx[DIM][NPE];
    const double detJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
    // displacement gradient g[i][j] = d u_i / d x_j
    double g[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double s = 0;
#pragma unroll
        for (int n = 0; n < NPE; ++n) s += dNx[j][n] * ue[n][i];
        g[i][j] = s;
      }
    double eps[NS];  // Kelvin-Mandel
    if (DIM == 2) {
      eps[0] = g[0][0];
      eps[1] = g[1][1];
      eps[2] = (g[0][1] + g[1][0]) * isq2;
    } else {
      eps[0] = g[0][0];
      eps[1] = g[1][1];
      eps[2] = g[2][2];
      eps[3] = (g[1][2] + g[2][1]) * isq2;
      eps[4] = (g[0][2] + g[2][0]) * isq2;
      eps[5] = (g[0][1] + g[1][0]) * isq2;
    }
    double tr = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) tr += eps[i];
    double sig[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) sig[k] = (k < DIM ? lam * tr : 0.0) + 2 * mu * eps[k];
    double dot = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) dot += sig[k] * eps[k];
    w += detJ * tab[TB::W + p] * dot;
    // tensor components (shear / sqrt 2)
    double te[NS], ts[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      te[k] = k < DIM ? eps[k] : eps[k] * isq2;
      ts[k] = k < DIM ? sig[k] : sig[k] * isq2;
      se[k] += te[k];
      ss[k] += ts[k];
    }
    evm += von_mises<DIM>(te);
    svm += von_mises<DIM>(ts);
  }
  double* o = out + e * NOUT;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    o[k] = se[k] / NPG;
    o[NS + 1 + k] = ss[k] / NPG;
  }
  o[NS] = evm / NPG;
  o[2 * NS + 1] = svm / NPG;
  o[2 * NS + 2] = 0.5 * w * (DIM == 2 ? thickness : 1.0);
}

// ZZ1 smoothed energy (Simulations/_elastic.py:323-418): at each stiffness
// point the stress interpolated from the nodal stresses sig_n (Nn, ns; tensor
// components, shear not scaled) with N, dotted with the point's strain in
// Kelvin-Mandel form: Ws_e = 1/2 sum_p wJ sig_s . eps (x thickness in 2D).
template <int ET, int NPG>
__global__ void __launch_bounds__(256)
k_zz1_energy(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
             const double* __restrict__ tab, const double* __restrict__ u, const double* __restrict__ sig_n,
             double thickness, double* __restrict__ out) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, NS = KM<DIM>::NS;
  using TB = Tab<NPG, NPE, DIM>;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ne) return;
  const double isq2 = 0.70710678118654752440, sq2 = 1.41421356237309504880;
  double X[NPE][3], ue[NPE][DIM], sn[NPE][NS];
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const int node = connect[e * NPE + n];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
#pragma unroll
    for (int k = 0; k < DIM; ++k) ue[n][k] = u[(int64_t)node * DIM + k];
#pragma unroll
    for (int k = 0; k < NS; ++k) sn[n][k] = sig_n[(int64_t)node * NS + k];
  }
  double w = 0;
#pragma unroll
  for (int p = 0; p < NPG; ++p) {
    double dNx[DIM][NPE];
    const double detJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
    double g[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        double s = 0;
#pragma unroll
        for (int n = 0; n < NPE; ++n) s += dNx[j][n] * ue[n][i];
        g[i][j] = s;
      }
    double eps[NS];
    if (DIM == 2) {
      eps[0] = g[0][0];
      eps[1] = g[1][1];
      eps[2] = (g[0][1] + g[1][0]) * isq2;
    } else {
      eps[0] = g[0][0];
      eps[1] = g[1][1];
      eps[2] = g[2][2];
      eps[3] = (g[1][2] + g[2][1]) * isq2;
      eps[4] = (g[0][2] + g[2][0]) * isq2;
      eps[5] = (g[0][1] + g[1][0]) * isq2;
    }
    double dot = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double sk = 0;
#pragma unroll
      for (int n = 0; n < NPE; ++n) sk += tab[TB::N + p * NPE + n] * sn[n][k];
      dot += (k < DIM ? sk : sk * sq2) * eps[k];
    }
    w += detJ * tab[TB::W + p] * dot;
  }
  out[e] = 0.5 * w * (DIM == 2 ? thickness : 1.0);
}

template <int ET, int NPG>
static int launch_zz1(int64_t Ne, const int* connect, const double* coord, const double* tab, const double* u,
                      const double* sig_n, double th, double* out, hipStream_t s) {
  const int64_t nb = (Ne + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL((k_zz1_energy<ET, NPG>), dim3((unsigned)nb), dim3(256), 0, s, Ne, connect, coord, tab, u,
                       sig_n, th, out);
  return check_launch("zz1_energy");
}

int zz1_energy(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
               const double* u, const double* sig_n, double th, double* out, hipStream_t s) {
#define ZZ(ET, NPG) return launch_zz1<ET, NPG>(Ne, connect, coord, tab, u, sig_n, th, out, s)
  if (elem == ET_TRI3 && nPg == 1) ZZ(ET_TRI3, 1);
  if (elem == ET_QUAD4 && nPg == 4) ZZ(ET_QUAD4, 4);
  if (elem == ET_HEXA8 && nPg == 8) ZZ(ET_HEXA8, 8);
  if (elem == ET_TETRA4 && nPg == 1) ZZ(ET_TETRA4, 1);
  if (elem == ET_TETRA10 && nPg == 4) ZZ(ET_TETRA10, 4);
#undef ZZ
  set_error("zz1_energy: unsupported element %d with %d points", elem, nPg);
  return EFEA_ERR_ARG;
}

// values_n[n] = (sum of values_e over the elements holding node n, in
// increasing element order) / their count; 0 for nodes of no element
// (FEM/_mesh.py:822-873 with the node -> element incidence of
// FEM/_group_elem.py:322-344).  One lane per (node, column).
__global__ void __launch_bounds__(256)
k_node_average(int Nn, int ncol, const int* __restrict__ n_ptr, const int* __restrict__ n_elems,
               const double* __restrict__ values_e, double* __restrict__ values_n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)Nn * ncol) return;
  const int n = (int)(t / ncol), c = (int)(t % ncol);
  const int k0 = n_ptr[n], k1 = n_ptr[n + 1];
  double s = 0;
  for (int k = k0; k < k1; ++k) s += values_e[(int64_t)n_elems[k] * ncol + c];
  values_n[t] = k1 > k0 ? s / (k1 - k0) : 0.0;
}

template <int ET, int NPG>
static int launch_ss(int64_t Ne, const int* connect, const double* coord, const double* tab, const double* u,
                     const double* E, int E_stride, double nu, int ps, double th, double* out, hipStream_t s) {
  const int64_t nb = (Ne + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL((k_strain_stress<ET, NPG>), dim3((unsigned)nb), dim3(256), 0, s, Ne, connect, coord, tab, u,
                       E, E_stride, nu, ps, th, out);
  return check_launch("strain_stress");
}

int strain_stress(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
                  const double* u, const double* E, int E_stride, double nu, int ps, double th, double* out,
                  hipStream_t s) {
#define SS(ET, NPG) return launch_ss<ET, NPG>(Ne, connect, coord, tab, u, E, E_stride, nu, ps, th, out, s)
  if (elem == ET_TRI3 && nPg == 1) SS(ET_TRI3, 1);
  if (elem == ET_QUAD4 && nPg == 4) SS(ET_QUAD4, 4);
  if (elem == ET_HEXA8 && nPg == 8) SS(ET_HEXA8, 8);
  if (elem == ET_TETRA4 && nPg == 1) SS(ET_TETRA4, 1);
  if (elem == ET_TETRA10 && nPg == 4) SS(ET_TETRA10, 4);
#undef SS
  set_error("strain_stress: unsupported element %d with %d points", elem, nPg);
  return EFEA_ERR_ARG;
}

int node_average(int Nn, int ncol, const int* n_ptr, const int* n_elems, const double* values_e, double* values_n,
                 hipStream_t s) {
  const int64_t nt = (int64_t)Nn * ncol;
  if (nt > 0)
    hipLaunchKernelGGL(k_node_average, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, Nn, ncol, n_ptr, n_elems,
                       values_e, values_n);
  return check_launch("node_average");
}

}  // namespace efea
