// Phase-field element kernels (2D: TRI3, QUAD4; 3D: TETRA4, TETRA10, HEXA8),
// FP64 throughout.
//
//  * k_pf_damage: at the mass points, strain from u, energy split psi+,
//    history psi+ <- old where psi+ < old (Simulations/_phasefield.py:494-538),
//    then the damage operators
//      Kd_e = sum_m r wJ N^T N + k sum_r wJ dN^T dN ,   Fd_e = sum_m f wJ N
//    (_phasefield.py:540-573; Models/_phasefield.py:236-293), x thickness.
//  * k_pf_elastic: at the rigi points, strain from the current u, split
//    c+ / c-, degradation g = (1 - N.d)^2 + 1e-12, c = g c+ + c-, and
//    K_e = sum_p wJ B^T c B x thickness (_phasefield.py:444-482).
// The split follows Models/_phasefield.py: Amor :451-483, Miehe :503-537 with
// the 2D closed-form eigen decomposition :751-802 and the projector build
// :1043-1128; Rp/Rm :485-501.  In 3D the splits without a spectral
// decomposition (Amor, Bourdin) are provided.  Element matrices carry the
// thickness in 2D only (the reference scales them when its dimension is 2).
#include "efea_common.h"

namespace efea {

struct PfParams {
  int split;  // EFEA_SPLIT_*
  int regu;   // 0 AT1, 1 AT2
  int planeStress;
  double nu;
  double l0;
  double thickness;
  double k_res;
};

__device__ inline void lame2d(double E, double nu, int planeStress, double& lam, double& mu) {
  lam = E * nu / ((1 + nu) * (1 - 2 * nu));
  if (planeStress) lam = E * nu / (1 - nu * nu);
  mu = E / (2 * (1 + nu));
}

__device__ inline double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : (x == 0 ? 0.0 : x)); }
__device__ inline double heav(double x) { return x < 0 ? 0.0 : (x > 0 ? 1.0 : (x == 0 ? 0.5 : x)); }

// Projector of the positive part of a 2D Kelvin-Mandel tensor t = [txx, tyy,
// sqrt(2) txy] (__Spectral_Decomposition, Models/_phasefield.py:1043-1128,
// with the 2D eigen decomposition :751-802): P = betaP I + g0 m1 m1^T + g1
// m2 m2^T; the negative projector is I - P.
__device__ inline void proj2d(const double (&t)[3], double (&P)[3][3]) {
  const double r2 = sqrt(2.0);
  const double m00 = t[0], m11 = t[1], m01 = t[2] / r2;
  const double det = m00 * m11 - m01 * m01;
  const double tr = m00 + m11;
  const double dl = tr * tr - 4 * det;
  const double sq = sqrt(dl);
  const double v0 = (tr - sq) / 2, v1 = (tr + sq) / 2;
  double A00 = 1, A01 = 0, A11 = 0;  // M1
  if (v0 != v1) {
    const double d = v0 - v1;
    A00 = (m00 - v1) / d;
    A01 = (m01 - 0.0) / d;
    A11 = (m11 - v1) / d;
  }
  const double B00 = 1 - A00, B01 = 0 - A01, B11 = 1 - A11;  // M2 = I - M1
  const double m1[3] = {A00, A11, A01 * r2};
  const double m2[3] = {B00, B11, B01 * r2};
  const double vp0 = (v0 + fabs(v0)) / 2, vp1 = (v1 + fabs(v1)) / 2;
  const double dv0 = heav(v0), dv1 = heav(v1);
  double vd = v0 - v1;
  if (vd == 0) vd = 1;
  const double betaP = (vp0 - vp1) / vd;
  const double g0 = dv0 - betaP, g1 = dv1 - betaP;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double Iij = (i == j) ? 1.0 : 0.0;
      P[i][j] = ((betaP * Iij) + (g0 * (m1[i] * m1[j]))) + (g1 * (m2[i] * m2[j]));
    }
}

template <int DIM>
struct PfDim {
  static constexpr int NS = DIM == 2 ? 3 : 6;  // Kelvin-Mandel components
};

template <int N>
__device__ inline void mmN(const double (&A)[N][N], const double (&B)[N][N], double (&C)[N][N]) {  // C = A B
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += A[i][k] * B[k][j];
      C[i][j] = s;
    }
}
template <int N>
__device__ inline void mtmN(const double (&A)[N][N], const double (&B)[N][N], double (&C)[N][N]) {  // C = A^T B
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) s += A[k][i] * B[k][j];
      C[i][j] = s;
    }
}

template <int DIM>
__device__ inline void proj_km(const double (&t)[PfDim<DIM>::NS], double (&P)[PfDim<DIM>::NS][PfDim<DIM>::NS]) {
  static_assert(DIM == 2, "spectral projectors are provided in 2D");
  proj2d(t, P);
}

// The cross terms of the anisotropic splits (Models/_phasefield.py:548-566,
// 660-676): 0 both to c+, 1 (_NoCross) both to c-, 2 (_PM) Cpm to c+ and Cmp
// to c-, 3 (_MP) Cmp to c+ and Cpm to c-.
template <int NS>
__device__ inline void cross_assign(int how, const double (&Cpp)[NS][NS], const double (&Cpm)[NS][NS],
                                    const double (&Cmm)[NS][NS], const double (&Cmp)[NS][NS],
                                    double (&cP)[NS][NS], double (&cM)[NS][NS]) {
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      if (how == 0) {
        cP[i][j] = (Cpp[i][j] + Cpm[i][j]) + Cmp[i][j];
        cM[i][j] = Cmm[i][j];
      } else if (how == 1) {
        cP[i][j] = Cpp[i][j];
        cM[i][j] = (Cmm[i][j] + Cpm[i][j]) + Cmp[i][j];
      } else if (how == 2) {
        cP[i][j] = Cpp[i][j] + Cpm[i][j];
        cM[i][j] = Cmm[i][j] + Cmp[i][j];
      } else {
        cP[i][j] = Cpp[i][j] + Cmp[i][j];
        cM[i][j] = Cmm[i][j] + Cpm[i][j];
      }
    }
}

// cP, cM of the 2D split (PhaseField.Calc_C, Models/_phasefield.py:396-431) for
// the Kelvin-Mandel strain e of an isotropic material (Lame constants lam,
// mu of the model; E, nu for the stress split), codes EFEA_SPLIT_*:
//   Amor (:451-483), Miehe (:503-537), Bourdin (:433-449), Stress (:573-634),
//   He (:680-749), AnisotStrain[_NoCross, _PM, _MP] (:539-566),
//   Zhang (:636-658), AnisotStress[_NoCross, _PM, _MP] (:636-676).
// Zhang and the _PM / _MP cross assignments give a non-symmetric c+ / c-.
// The reference's evaluation order of each product and sum is kept.
template <int DIM>
__device__ inline void split_t(int split, const double (&e)[PfDim<DIM>::NS], double lam, double mu, double E,
                               double nu, int planeStress, double (&cP)[PfDim<DIM>::NS][PfDim<DIM>::NS],
                               double (&cM)[PfDim<DIM>::NS][PfDim<DIM>::NS]) {
  constexpr int NS = PfDim<DIM>::NS;
  double IxI[NS][NS], C[NS][NS];
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      IxI[i][j] = (i < DIM && j < DIM) ? 1.0 : 0.0;
      C[i][j] = lam * IxI[i][j] + (i == j ? 2 * mu : 0.0);  // lam IxI + 2 mu I
    }
  double tr_e = 0;
#pragma unroll
  for (int i = 0; i < DIM; ++i) tr_e += e[i];
  if (split == EFEA_SPLIT_AMOR) {
    const double Rp = (1 + sgn(tr_e)) / 2, Rm = (1 + sgn(-tr_e)) / 2;
    const double bulk = lam + 2 * mu / DIM;
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double Iij = (i == j) ? 1.0 : 0.0;
        cP[i][j] = bulk * (Rp * IxI[i][j]) + 2 * mu * (Iij - 1.0 / DIM * IxI[i][j]);
        cM[i][j] = bulk * (Rm * IxI[i][j]);
      }
    return;
  }
  if (split == EFEA_SPLIT_BOURDIN) {
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        cP[i][j] = C[i][j];
        cM[i][j] = 0;
      }
    return;
  }
  double P[NS][NS], M[NS][NS];
  if (split == EFEA_SPLIT_MIEHE || split == EFEA_SPLIT_ANISOT_STRAIN || split == EFEA_SPLIT_ANISOT_STRAIN_NOCROSS ||
      split == EFEA_SPLIT_ANISOT_STRAIN_PM || split == EFEA_SPLIT_ANISOT_STRAIN_MP) {
    proj_km<DIM>(e, P);
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) M[i][j] = (i == j ? 1.0 : 0.0) - P[i][j];
    if (split == EFEA_SPLIT_MIEHE) {
      const double Rp = (1 + sgn(tr_e)) / 2, Rm = (1 + sgn(-tr_e)) / 2;
#pragma unroll
      for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          cP[i][j] = lam * (Rp * IxI[i][j]) + 2 * mu * P[i][j];
          cM[i][j] = lam * (Rm * IxI[i][j]) + 2 * mu * M[i][j];
        }
      return;
    }
    double PTC[NS][NS], MTC[NS][NS], Cpp[NS][NS], Cpm[NS][NS], Cmm[NS][NS], Cmp[NS][NS];
    mtmN<NS>(P, C, PTC);
    mtmN<NS>(M, C, MTC);
    mmN<NS>(PTC, P, Cpp);
    mmN<NS>(PTC, M, Cpm);
    mmN<NS>(MTC, M, Cmm);
    mmN<NS>(MTC, P, Cmp);
    cross_assign<NS>(split == EFEA_SPLIT_ANISOT_STRAIN ? 0 : split == EFEA_SPLIT_ANISOT_STRAIN_PM ? 2
                     : split == EFEA_SPLIT_ANISOT_STRAIN_MP ? 3 : 1, Cpp, Cpm, Cmm, Cmp, cP, cM);
    return;
  }
  if (split == EFEA_SPLIT_HE) {
    // C^(1/2) and C^(-1/2) of the isotropic C: eigenvalue 2 mu + DIM lam along
    // IxI / DIM, 2 mu on its complement
    const double a = sqrt(2 * mu), b = sqrt(2 * mu + DIM * lam);
    double sqC[NS][NS], isqC[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double Iij = (i == j) ? 1.0 : 0.0;
        sqC[i][j] = a * Iij + (b - a) * (IxI[i][j] / DIM);
        isqC[i][j] = Iij / a + (1 / b - 1 / a) * (IxI[i][j] / DIM);
      }
    double et[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double s = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) s += sqC[i][j] * e[j];
      et[i] = s;
    }
    double Pt[NS][NS], Mt[NS][NS], T[NS][NS];
    proj_km<DIM>(et, Pt);
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) Mt[i][j] = (i == j ? 1.0 : 0.0) - Pt[i][j];
    mmN<NS>(isqC, Pt, T);
    mmN<NS>(T, sqC, P);
    mmN<NS>(isqC, Mt, T);
    mmN<NS>(T, sqC, M);
    mmN<NS>(C, P, cP);
    mmN<NS>(C, M, cM);
    return;
  }
  // stress-based splits: the projectors of sigma = C e
  double sg[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) s += C[i][j] * e[j];
    sg[i] = s;
  }
  proj_km<DIM>(sg, P);
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) M[i][j] = (i == j ? 1.0 : 0.0) - P[i][j];
  if (split == EFEA_SPLIT_STRESS) {
    double tr_s = 0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) tr_s += sg[i];
    const double Rp = (1 + sgn(tr_s)) / 2, Rm = (1 + sgn(-tr_s)) / 2;
    // 2D: (1+nu)/E P - c R IxI, c = nu/E (plane stress) or nu(1+nu)/E;
    // 3D: 1/(2 mu) P - nu/E R IxI
    const double cp = DIM == 2 ? (1 + nu) / E : 1 / (2 * mu);
    const double cv = (DIM == 2 && !planeStress) ? nu * (1 + nu) / E : nu / E;
    double sP[NS][NS], sM[NS][NS], T[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        sP[i][j] = (cp * P[i][j]) - (cv * Rp * IxI[i][j]);
        sM[i][j] = (cp * M[i][j]) - (cv * Rm * IxI[i][j]);
      }
    mtmN<NS>(C, sP, T);
    mmN<NS>(T, C, cP);
    mtmN<NS>(C, sM, T);
    mmN<NS>(T, C, cM);
    return;
  }
  // Zhang: c+ = P C, c- = M C.  AnisotStress[_NoCross, _PM, _MP]: Cp = P C,
  // Cm = M C and the compliance S = C^-1 = I / (2 mu) - lam / (2 mu (2 mu +
  // DIM lam)) IxI
  double S[NS][NS], Cp[NS][NS], Cm[NS][NS], PS[NS][NS], MS[NS][NS], Cpp[NS][NS], Cpm[NS][NS], Cmm[NS][NS],
      Cmp[NS][NS];
  if (split == EFEA_SPLIT_ZHANG) {
    mmN<NS>(P, C, cP);
    mmN<NS>(M, C, cM);
    return;
  }
  const double s2 = lam / (2 * mu * (2 * mu + DIM * lam));
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) S[i][j] = (i == j ? 1 / (2 * mu) : 0.0) - s2 * IxI[i][j];
  mmN<NS>(P, C, Cp);
  mmN<NS>(M, C, Cm);
  mtmN<NS>(Cp, S, PS);
  mtmN<NS>(Cm, S, MS);
  mmN<NS>(PS, Cp, Cpp);
  mmN<NS>(PS, Cm, Cpm);
  mmN<NS>(MS, Cm, Cmm);
  mmN<NS>(MS, Cp, Cmp);
  cross_assign<NS>(split == EFEA_SPLIT_ANISOT_STRESS ? 0 : split == EFEA_SPLIT_ANISOT_STRESS_PM ? 2
                   : split == EFEA_SPLIT_ANISOT_STRESS_MP ? 3 : 1, Cpp, Cpm, Cmm, Cmp, cP, cM);
}
template <int NPE>
__device__ inline void strain2d(const double (&dNx)[2][NPE], const double* ue, double e[3]) {
  const double c = 1 / sqrt(2.0);
  e[0] = e[1] = e[2] = 0;
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    e[0] += dNx[0][n] * ue[2 * n];
    e[1] += dNx[1][n] * ue[2 * n + 1];
    e[2] += (dNx[1][n] * c) * ue[2 * n] + (dNx[0][n] * c) * ue[2 * n + 1];
  }
}

// 3D Kelvin-Mandel strain [xx, yy, zz, sqrt2 yz, sqrt2 xz, sqrt2 xy].
template <int NPE>
__device__ inline void strain3d(const double (&dNx)[3][NPE], const double* ue, double e[6]) {
  const double c = 1 / sqrt(2.0);
#pragma unroll
  for (int k = 0; k < 6; ++k) e[k] = 0;
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const double ux = ue[3 * n], uy = ue[3 * n + 1], uz = ue[3 * n + 2];
    e[0] += dNx[0][n] * ux;
    e[1] += dNx[1][n] * uy;
    e[2] += dNx[2][n] * uz;
    e[3] += (dNx[2][n] * c) * uy + (dNx[1][n] * c) * uz;
    e[4] += (dNx[2][n] * c) * ux + (dNx[0][n] * c) * uz;
    e[5] += (dNx[1][n] * c) * ux + (dNx[0][n] * c) * uy;
  }
}

// 3D splits without a spectral decomposition: Amor (bulk lam + 2 mu / 3 on
// the volumetric part, Isym - IxI / 3 deviatoric) and Bourdin.
__device__ inline void split3d(int split, const double e[6], double lam, double mu, double cP[6][6],
                               double cM[6][6]) {
  const double tr_e = e[0] + e[1] + e[2];
  const double Rp = (1 + sgn(tr_e)) / 2, Rm = (1 + sgn(-tr_e)) / 2;
  const double bulk = lam + 2 * mu / 3;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double IxI = (i < 3 && j < 3) ? 1.0 : 0.0;
      const double Iij = (i == j) ? 1.0 : 0.0;
      if (split == EFEA_SPLIT_AMOR) {
        cP[i][j] = bulk * (Rp * IxI) + 2 * mu * (Iij - 1.0 / 3 * IxI);
        cM[i][j] = bulk * (Rm * IxI);
      } else {  // Bourdin: c+ = C
        cP[i][j] = lam * IxI + (i == j ? 2 * mu : 0.0);
        cM[i][j] = 0;
      }
    }
}

template <int DIM, int NPE>
__device__ inline void strain_km(const double (&dNx)[DIM][NPE], const double* ue, double (&e)[PfDim<DIM>::NS]) {
  if constexpr (DIM == 2)
    strain2d<NPE>(dNx, ue, e);
  else
    strain3d<NPE>(dNx, ue, e);
}

// c+ / c- of the element's material at one point (E, nu, plane stress of the model)
template <int DIM>
__device__ inline void split_km(int split, const double (&e)[PfDim<DIM>::NS], double E, double nu, int planeStress,
                                double (&cP)[PfDim<DIM>::NS][PfDim<DIM>::NS],
                                double (&cM)[PfDim<DIM>::NS][PfDim<DIM>::NS]) {
  double lam, mu;
  if constexpr (DIM == 2) {
    lame2d(E, nu, planeStress, lam, mu);
    split_t<2>(split, e, lam, mu, E, nu, planeStress, cP, cM);
  } else {
    lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    mu = E / (2 * (1 + nu));
    split3d(split, e, lam, mu, cP, cM);
  }
}

// B columns of node a (derivatives dx, dy[, dz]) in Kelvin-Mandel rows.
template <int DIM>
__device__ inline void bcols(const double* dn, double (&B)[DIM][PfDim<DIM>::NS]) {
  const double c = 1 / sqrt(2.0);
  if constexpr (DIM == 2) {
    const double ax = dn[0], ay = dn[1];
    B[0][0] = ax; B[0][1] = 0; B[0][2] = ay * c;
    B[1][0] = 0; B[1][1] = ay; B[1][2] = ax * c;
  } else {
    const double ax = dn[0], ay = dn[1], az = dn[2];
    B[0][0] = ax; B[0][1] = 0; B[0][2] = 0; B[0][3] = 0; B[0][4] = az * c; B[0][5] = ay * c;
    B[1][0] = 0; B[1][1] = ay; B[1][2] = 0; B[1][3] = az * c; B[1][4] = 0; B[1][5] = ax * c;
    B[2][0] = 0; B[2][1] = 0; B[2][2] = az; B[2][3] = ay * c; B[2][4] = ax * c; B[2][5] = 0;
  }
}

template <int ET, int NPGM, int NPGR>
__global__ void __launch_bounds__(256)
k_pf_damage(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
            const double* __restrict__ tab_m, const double* __restrict__ tab_r, const double* __restrict__ u,
            const double* __restrict__ H_old, const double* __restrict__ E, int E_stride,
            const double* __restrict__ Gc, int Gc_stride, PfParams P, double* __restrict__ psiP,
            double* __restrict__ Kd, double* __restrict__ Fd) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, NS = PfDim<DIM>::NS;
  constexpr int NPAIR = NPE * NPE;
  constexpr int EPB = 256 / NPAIR;
  constexpr int NPGT = NPGM + NPGR;
  using TM = Tab<NPGM, NPE, DIM>;
  using TR = Tab<NPGR, NPE, DIM>;
  __shared__ double s_dNx[EPB][NPGR][DIM][NPE];
  __shared__ double s_wJr[EPB][NPGR];
  __shared__ double s_wJm[EPB][NPGM];
  __shared__ double s_r[EPB][NPGM];
  __shared__ double s_f[EPB][NPGM];
  __shared__ double s_K[EPB * NPAIR];

  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  const int t = threadIdx.x;
  if (t < EPB * NPGT) {
    const int el = t / NPGT, q = t % NPGT;
    const int64_t e = e0 + el;
    if (e < Ne) {
      double X[NPE][3];
      double ue[DIM * NPE];
#pragma unroll
      for (int n = 0; n < NPE; ++n) {
        const int node = connect[e * NPE + n];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
#pragma unroll
        for (int k = 0; k < DIM; ++k) ue[DIM * n + k] = u[(int64_t)node * DIM + k];
      }
      double dNx[DIM][NPE];
      if (q < NPGM) {
        const double detJ = point_geometry<double, NPE, DIM>(tab_m + TM::DN + q * DIM * NPE, X, dNx);
        s_wJm[el][q] = detJ * tab_m[TM::W + q];
        double eps[NS];
        strain_km<DIM, NPE>(dNx, ue, eps);
        double cP[NS][NS], cM[NS][NS];
        split_km<DIM>(P.split, eps, E[E_stride ? e : 0], P.nu, P.planeStress, cP, cM);
        double psi = 0;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          double sig = 0;
#pragma unroll
          for (int j = 0; j < NS; ++j) sig += cP[i][j] * eps[j];
          psi += 1.0 / 2 * eps[i] * sig;
        }
        if (H_old) {
          const double old = H_old[e * NPGM + q];
          if (psi - old < 0) psi = old;
        }
        psiP[e * NPGM + q] = psi;
        const double gc = Gc[Gc_stride ? e : 0];
        double r, f;
        if (P.regu == 1) {
          r = 2 * psi + gc / P.l0;
          f = 2 * psi;
        } else {
          r = 2 * psi;
          f = 2 * psi - (3 * gc) / (8 * P.l0);
          f = (f + fabs(f)) / 2;
        }
        s_r[el][q] = r;
        s_f[el][q] = f;
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
      for (int p = 0; p < NPGM; ++p)
        R += s_r[el][p] * s_wJm[el][p] * tab_m[TM::N + p * NPE + a] * tab_m[TM::N + p * NPE + b];
      double D = 0;
#pragma unroll
      for (int p = 0; p < NPGR; ++p) {
        double dot = 0;
#pragma unroll
        for (int i = 0; i < DIM; ++i) dot += s_dNx[el][p][i][a] * s_dNx[el][p][i][b];
        D += s_wJr[el][p] * dot;
      }
      const double gc = Gc[Gc_stride ? e : 0];
      const double kd = P.regu == 1 ? gc * P.l0 : 3.0 / 4 * gc * P.l0;
      const double th = DIM == 2 ? P.thickness : 1.0;
      s_K[el * NPAIR + ab] = (R + kd * D) * th;
      if (b == 0) {
        double F = 0;
#pragma unroll
        for (int p = 0; p < NPGM; ++p) F += s_f[el][p] * s_wJm[el][p] * tab_m[TM::N + p * NPE + a];
        Fd[e * NPE + a] = F * th;
      }
    }
  }
  __syncthreads();
  const int64_t nval = (int64_t)min((int64_t)EPB, Ne - e0) * NPAIR;
  double* out = Kd + e0 * NPAIR;
  for (int64_t k = t; k < nval; k += 256) out[k] = s_K[k];
}

template <int ET, int NPG>
__global__ void __launch_bounds__(256)
k_pf_elastic(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
             const double* __restrict__ tab, const double* __restrict__ u, const double* __restrict__ d,
             const double* __restrict__ E, int E_stride, PfParams P, double* __restrict__ Ke) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, ND = DIM * NPE, NS = PfDim<DIM>::NS;
  constexpr int NPAIR = NPE * NPE;
  constexpr int EPB = 256 / NPAIR;
  using TB = Tab<NPG, NPE, DIM>;
  __shared__ double s_dNx[EPB][NPG][DIM][NPE];
  __shared__ double s_wJ[EPB][NPG];
  __shared__ double s_c[EPB][NPG][NS][NS];
  __shared__ double s_K[EPB * ND * ND];

  const int64_t e0 = (int64_t)blockIdx.x * EPB;
  const int t = threadIdx.x;
  if (t < EPB * NPG) {
    const int el = t / NPG, p = t % NPG;
    const int64_t e = e0 + el;
    if (e < Ne) {
      double X[NPE][3];
      double ue[DIM * NPE];
      double de[NPE];
#pragma unroll
      for (int n = 0; n < NPE; ++n) {
        const int node = connect[e * NPE + n];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
#pragma unroll
        for (int k = 0; k < DIM; ++k) ue[DIM * n + k] = u[(int64_t)node * DIM + k];
        de[n] = d[node];
      }
      double dNx[DIM][NPE];
      const double detJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
      s_wJ[el][p] = detJ * tab[TB::W + p];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int n = 0; n < NPE; ++n) s_dNx[el][p][i][n] = dNx[i][n];
      double eps[NS];
      strain_km<DIM, NPE>(dNx, ue, eps);
      double cP[NS][NS], cM[NS][NS];
      split_km<DIM>(P.split, eps, E[E_stride ? e : 0], P.nu, P.planeStress, cP, cM);
      double dp = 0;
#pragma unroll
      for (int n = 0; n < NPE; ++n) dp += tab[TB::N + p * NPE + n] * de[n];
      const double g = (1 - dp) * (1 - dp) + P.k_res;
#pragma unroll
      for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) s_c[el][p][i][j] = g * cP[i][j] + cM[i][j];
    }
  }
  __syncthreads();

  if (t < EPB * NPAIR) {
    const int el = t / NPAIR, ab = t % NPAIR, a = ab / NPE, b = ab % NPE;
    const int64_t e = e0 + el;
    if (e < Ne) {
      double acc[DIM][DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) acc[i][j] = 0;
#pragma unroll
      for (int p = 0; p < NPG; ++p) {
        double da[DIM], db[DIM];
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
          da[i] = s_dNx[el][p][i][a];
          db[i] = s_dNx[el][p][i][b];
        }
        double Ba[DIM][NS], Bb[DIM][NS];
        bcols<DIM>(da, Ba);
        bcols<DIM>(db, Bb);
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int j = 0; j < DIM; ++j) {
            double s = 0;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
              double cb = 0;
#pragma unroll
              for (int l = 0; l < NS; ++l) cb += s_c[el][p][k][l] * Bb[j][l];
              s += Ba[i][k] * cb;
            }
            acc[i][j] += s_wJ[el][p] * s;
          }
      }
      const double th = DIM == 2 ? P.thickness : 1.0;
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) s_K[el * ND * ND + (DIM * a + i) * ND + DIM * b + j] = acc[i][j] * th;
    }
  }
  __syncthreads();
  const int64_t nval = (int64_t)min((int64_t)EPB, Ne - e0) * ND * ND;
  double* out = Ke + e0 * ND * ND;
  for (int64_t k = t; k < nval; k += 256) out[k] = s_K[k];
}

// Phase-field results per element (Simulations/_phasefield.py:682-760,
// 838-916): at each rigi point eps = B u_e, the split c+ / c-, and the damaged
// stress sig = g(d) (c+ eps) + c- eps with g = (1 - N.d)^2 + 1e-12.  Output row
// as the elastic post-processing: [eps xx, yy, xy (tensor), Evm, sig xx, yy,
// xy, Svm, 1/2 sum_p wJ sig.eps x thickness], components and von Mises
// measures averaged over the points.  One lane per element.
template <int ET, int NPG>
__global__ void __launch_bounds__(256)
k_pf_strain_stress(int64_t Ne, const int* __restrict__ connect, const double* __restrict__ coord,
                   const double* __restrict__ tab, const double* __restrict__ u, const double* __restrict__ d,
                   const double* __restrict__ E, int E_stride, PfParams P, double* __restrict__ out) {
  constexpr int NPE = Elem<ET>::NPE, DIM = Elem<ET>::DIM, NS = PfDim<DIM>::NS, NOUT = 2 * NS + 3;
  using TB = Tab<NPG, NPE, DIM>;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ne) return;
  double X[NPE][3], ue[DIM * NPE], de[NPE];
#pragma unroll
  for (int n = 0; n < NPE; ++n) {
    const int node = connect[e * NPE + n];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[n][k] = coord[(int64_t)node * 3 + k];
#pragma unroll
    for (int k = 0; k < DIM; ++k) ue[DIM * n + k] = u[(int64_t)node * DIM + k];
    de[n] = d[node];
  }
  const double Ee = E[E_stride ? e : 0];
  const double isq2 = 1 / sqrt(2.0);
  double se[NS], ss[NS], evm = 0, svm = 0, w = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) se[k] = ss[k] = 0;
#pragma unroll
  for (int p = 0; p < NPG; ++p) {
    double dNx[DIM][NPE];
    const double detJ = point_geometry<double, NPE, DIM>(tab + TB::DN + p * DIM * NPE, X, dNx);
    double eps[NS];
    strain_km<DIM, NPE>(dNx, ue, eps);
    double cP[NS][NS], cM[NS][NS];
    split_km<DIM>(P.split, eps, Ee, P.nu, P.planeStress, cP, cM);
    double dp = 0;
#pragma unroll
    for (int n = 0; n < NPE; ++n) dp += tab[TB::N + p * NPE + n] * de[n];
    const double g = (1 - dp) * (1 - dp) + P.k_res;
    double sig[NS];
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      double sp = 0, sm = 0;
#pragma unroll
      for (int b = 0; b < NS; ++b) {
        sp += cP[a][b] * eps[b];
        sm += cM[a][b] * eps[b];
      }
      sig[a] = g * sp + sm;
    }
    double dot = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) dot += sig[k] * eps[k];
    w += detJ * tab[TB::W + p] * dot;
    double te[NS], ts[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      te[k] = k < DIM ? eps[k] : eps[k] * isq2;
      ts[k] = k < DIM ? sig[k] : sig[k] * isq2;
      se[k] += te[k];
      ss[k] += ts[k];
    }
    if constexpr (DIM == 2) {
      evm += sqrt(te[0] * te[0] + te[1] * te[1] - te[0] * te[1] + 3 * te[2] * te[2]);
      svm += sqrt(ts[0] * ts[0] + ts[1] * ts[1] - ts[0] * ts[1] + 3 * ts[2] * ts[2]);
    } else {
      evm += sqrt(0.5 * ((te[0] - te[1]) * (te[0] - te[1]) + (te[1] - te[2]) * (te[1] - te[2]) +
                         (te[2] - te[0]) * (te[2] - te[0]) + 6 * (te[5] * te[5] + te[3] * te[3] + te[4] * te[4])));
      svm += sqrt(0.5 * ((ts[0] - ts[1]) * (ts[0] - ts[1]) + (ts[1] - ts[2]) * (ts[1] - ts[2]) +
                         (ts[2] - ts[0]) * (ts[2] - ts[0]) + 6 * (ts[5] * ts[5] + ts[3] * ts[3] + ts[4] * ts[4])));
    }
  }
  double* o = out + e * NOUT;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    o[k] = se[k] / NPG;
    o[NS + 1 + k] = ss[k] / NPG;
  }
  o[NS] = evm / NPG;
  o[2 * NS + 1] = svm / NPG;
  o[2 * NS + 2] = 0.5 * w * (DIM == 2 ? P.thickness : 1.0);
}

// Gauss-point split evaluation for testing and post-processing:
// eps (n, 3) -> cP, cM (n, 3, 3), psiP, psiM (n).
__global__ void k_pf_split_points(int64_t n, int split, const double* __restrict__ eps,
                                  const double* __restrict__ E, int E_stride, double nu, int planeStress,
                                  double* __restrict__ cPo, double* __restrict__ cMo, double* __restrict__ psiP,
                                  double* __restrict__ psiM) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e[3] = {eps[3 * i], eps[3 * i + 1], eps[3 * i + 2]};
  double cP[3][3], cM[3][3];
  split_km<2>(split, e, E[E_stride ? i : 0], nu, planeStress, cP, cM);
  double pp = 0, pm = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double sp = 0, sm = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      sp += cP[a][b] * e[b];
      sm += cM[a][b] * e[b];
      cPo[9 * i + 3 * a + b] = cP[a][b];
      cMo[9 * i + 3 * a + b] = cM[a][b];
    }
    pp += 1.0 / 2 * e[a] * sp;
    pm += 1.0 / 2 * e[a] * sm;
  }
  psiP[i] = pp;
  psiM[i] = pm;
}

static PfParams make_params(int split, int regu, int ps, double nu, double l0, double th) {
  PfParams P;
  P.split = split;
  P.regu = regu;
  P.planeStress = ps;
  P.nu = nu;
  P.l0 = l0;
  P.thickness = th;
  P.k_res = 1e-12;
  return P;
}

template <int ET>
static unsigned pf_blocks(int64_t Ne) {
  constexpr int EPB = 256 / (Elem<ET>::NPE * Elem<ET>::NPE);
  return (unsigned)((Ne + EPB - 1) / EPB);
}

// In 3D only the splits without a spectral decomposition are provided.
static bool pf_split_ok(int elem, int split, const char* what) {
  if (split < EFEA_SPLIT_AMOR || split > EFEA_SPLIT_ANISOT_STRESS_MP) {
    set_error("%s: unknown split %d", what, split);
    return false;
  }
  const bool three_d = elem == ET_TETRA4 || elem == ET_TETRA10 || elem == ET_HEXA8;
  if (three_d && split != EFEA_SPLIT_AMOR && split != EFEA_SPLIT_BOURDIN) {
    set_error("%s: split %d is not provided in 3D (Amor and Bourdin are)", what, split);
    return false;
  }
  return true;
}

int pf_damage(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab_m, int nPgm,
              const double* tab_r, int nPgr, const double* u, const double* H_old, const double* E,
              int E_stride, double nu, int ps, const double* Gc, int Gc_stride, double l0, int split, int regu,
              double thickness, double* psiP, double* Kd, double* Fd, hipStream_t s) {
  const PfParams P = make_params(split, regu, ps, nu, l0, thickness);
  if (!pf_split_ok(elem, split, "pf_damage")) return EFEA_ERR_ARG;
  if (Ne == 0) return EFEA_OK;
#define PFD(ET, M, R)                                                                                           \
  if (elem == ET && nPgm == M && nPgr == R) {                                                                   \
    hipLaunchKernelGGL((k_pf_damage<ET, M, R>), dim3(pf_blocks<ET>(Ne)), dim3(256), 0, s, Ne, connect, coord,   \
                       tab_m, tab_r, u, H_old, E, E_stride, Gc, Gc_stride, P, psiP, Kd, Fd);                     \
    return check_launch("pf_damage");                                                                           \
  }
  PFD(ET_TRI3, 3, 1)
  PFD(ET_QUAD4, 4, 4)
  PFD(ET_TETRA4, 4, 1)
  PFD(ET_TETRA10, 4, 4)
  PFD(ET_HEXA8, 8, 8)
#undef PFD
  set_error("pf_damage: unsupported element %d / nPg %d,%d", elem, nPgm, nPgr);
  return EFEA_ERR_ARG;
}

int pf_elastic(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
               const double* u, const double* d, const double* E, int E_stride, double nu, int ps, int split,
               double thickness, double* Ke, hipStream_t s) {
  const PfParams P = make_params(split, 1, ps, nu, 1.0, thickness);
  if (!pf_split_ok(elem, split, "pf_elastic")) return EFEA_ERR_ARG;
  if (Ne == 0) return EFEA_OK;
#define PFE(ET, G)                                                                                              \
  if (elem == ET && nPg == G) {                                                                                 \
    hipLaunchKernelGGL((k_pf_elastic<ET, G>), dim3(pf_blocks<ET>(Ne)), dim3(256), 0, s, Ne, connect, coord, tab, \
                       u, d, E, E_stride, P, Ke);                                                               \
    return check_launch("pf_elastic");                                                                          \
  }
  PFE(ET_TRI3, 1)
  PFE(ET_QUAD4, 4)
  PFE(ET_TETRA4, 1)
  PFE(ET_TETRA10, 4)
  PFE(ET_HEXA8, 8)
#undef PFE
  set_error("pf_elastic: unsupported element %d / nPg %d", elem, nPg);
  return EFEA_ERR_ARG;
}

int pf_strain_stress(int elem, int64_t Ne, const int* connect, const double* coord, const double* tab, int nPg,
                     const double* u, const double* d, const double* E, int E_stride, double nu, int ps, int split,
                     double thickness, double* out, hipStream_t s) {
  const PfParams P = make_params(split, 1, ps, nu, 1.0, thickness);
  if (!pf_split_ok(elem, split, "pf_strain_stress")) return EFEA_ERR_ARG;
  if (Ne == 0) return EFEA_OK;
  const unsigned nb = (unsigned)((Ne + 255) / 256);
#define PFS(ET, G)                                                                                              \
  if (elem == ET && nPg == G) {                                                                                 \
    hipLaunchKernelGGL((k_pf_strain_stress<ET, G>), dim3(nb), dim3(256), 0, s, Ne, connect, coord, tab, u, d, E, \
                       E_stride, P, out);                                                                       \
    return check_launch("pf_strain_stress");                                                                    \
  }
  PFS(ET_TRI3, 1)
  PFS(ET_QUAD4, 4)
  PFS(ET_TETRA4, 1)
  PFS(ET_TETRA10, 4)
  PFS(ET_HEXA8, 8)
#undef PFS
  set_error("pf_strain_stress: unsupported element %d / nPg %d", elem, nPg);
  return EFEA_ERR_ARG;
}

int pf_split_points(int64_t n, int split, const double* eps, const double* E, int E_stride, double nu, int ps,
                    double* cP, double* cM, double* psiP, double* psiM, hipStream_t s) {
  if (n == 0) return EFEA_OK;
  hipLaunchKernelGGL(k_pf_split_points, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, n, split, eps, E,
                     E_stride, nu, ps, cP, cM, psiP, psiM);
  return check_launch("pf_split_points");
}

}  // namespace efea
