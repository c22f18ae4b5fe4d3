// Shared definitions for the efea HIP library (gfx950 / CDNA4).
//
// Conventions used by every kernel:
//   * connectivity is int32 (Ne, nPe), node coordinates are FP64 (Nn, 3);
//   * DOF numbering is dof = node * dof_n + component
//     (EasyFEA/FEM/_boundary_conditions.py:160-201);
//   * element matrices are element-major (Ne, nd, nd), row i then column j,
//     the order EasyFEA ravels K_e before its bincount scatter
//     (EasyFEA/Simulations/_simu.py:1037);
//   * reference-element tables arrive as one FP64 block per integration
//     rule: [w (nPg)] [N (nPg*nPe)] [dN (nPg*dim*nPe)].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "efea.h"

#define EFEA_OK 0

namespace efea {

void set_error(const char* fmt, ...);
int check_launch(const char* what);

enum ElemCode { ET_TRI3 = 0, ET_QUAD4 = 1, ET_HEXA8 = 2, ET_TETRA4 = 3, ET_TETRA10 = 4, ET_SEG2 = 5 };

template <int ET> struct Elem;
template <> struct Elem<ET_TRI3> { static constexpr int NPE = 3, DIM = 2; };
template <> struct Elem<ET_QUAD4> { static constexpr int NPE = 4, DIM = 2; };
template <> struct Elem<ET_HEXA8> { static constexpr int NPE = 8, DIM = 3; };
template <> struct Elem<ET_TETRA4> { static constexpr int NPE = 4, DIM = 3; };
template <> struct Elem<ET_TETRA10> { static constexpr int NPE = 10, DIM = 3; };
template <> struct Elem<ET_SEG2> { static constexpr int NPE = 2, DIM = 1; };

// Table offsets inside one rule block.
template <int NPG, int NPE, int DIM> struct Tab {
  static constexpr int W = 0;
  static constexpr int N = NPG;
  static constexpr int DN = NPG + NPG * NPE;
  static constexpr int SIZE = NPG + NPG * NPE + NPG * DIM * NPE;
};

// Jacobian (transposed, F[d][k] = sum_n dN[d][n] X[n][k]), |det|, inverse
// and physical derivatives at one integration point
// (EasyFEA/FEM/_group_elem.py:832-915, 1083-1105; FEM/_linalg.py:533-656).
template <typename T, int NPE, int DIM>
__device__ inline T point_geometry(const double* __restrict__ dN, const T (&X)[NPE][3],
                                   T (&dNx)[DIM][NPE]) {
  T F[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      T s = 0;
#pragma unroll
      for (int n = 0; n < NPE; ++n) s += (T)dN[d * NPE + n] * X[n][k];
      F[d][k] = s;
    }
  T inv[DIM][DIM];
  T det;
  if constexpr (DIM == 1) {
    det = F[0][0];
    inv[0][0] = (T)1 / det;
  } else if constexpr (DIM == 2) {
    det = F[0][0] * F[1][1] - F[1][0] * F[0][1];
    T r = (T)1 / det;
    inv[0][0] = F[1][1] * r;
    inv[0][1] = -F[0][1] * r;
    inv[1][0] = -F[1][0] * r;
    inv[1][1] = F[0][0] * r;
  } else {
    det = F[0][0] * (F[1][1] * F[2][2] - F[2][1] * F[1][2]) -
          F[0][1] * (F[1][0] * F[2][2] - F[2][0] * F[1][2]) +
          F[0][2] * (F[1][0] * F[2][1] - F[2][0] * F[1][1]);
    T r = (T)1 / det;
    inv[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * r;
    inv[0][1] = -(F[0][1] * F[2][2] - F[0][2] * F[2][1]) * r;
    inv[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * r;
    inv[1][0] = -(F[1][0] * F[2][2] - F[1][2] * F[2][0]) * r;
    inv[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * r;
    inv[1][2] = -(F[0][0] * F[1][2] - F[0][2] * F[1][0]) * r;
    inv[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * r;
    inv[2][1] = -(F[0][0] * F[2][1] - F[0][1] * F[2][0]) * r;
    inv[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * r;
  }
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int n = 0; n < NPE; ++n) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < DIM; ++j) s += inv[i][j] * (T)dN[j * NPE + n];
      dNx[i][n] = s;
    }
  return det < 0 ? -det : det;
}

// Block-wide helpers -------------------------------------------------------

template <typename T>
__device__ inline T warp_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block reduction (fixed tree order), result valid in thread 0.
template <typename T, int BS>
__device__ inline T block_sum(T v, T* sh) {
  v = warp_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += sh[i];
  }
  __syncthreads();
  return s;
}

// Deterministic block reduction of N values at once (each value summed in the
// order of block_sum), result valid in every thread; one barrier pair.
template <int N, int BS>
__device__ inline void block_sum_all(double (&v)[N], double (*sh)[BS / 64]) {
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = warp_sum(v[q]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) sh[q][w] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    double s = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += sh[q][i];
    v[q] = s;
  }
  __syncthreads();
}

}  // namespace efea
