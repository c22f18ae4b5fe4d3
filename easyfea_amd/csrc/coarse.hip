// Two-level preconditioner: block Jacobi plus an additive coarse correction
// over node aggregates, M^-1 r = B^-1 r + P Ac^-1 P^T r.
//
// P holds per aggregate the rigid-body modes of its nodes (2D: ux, uy and the
// rotation about the aggregate centroid; 3D: three translations and three
// rotations; scalar problems: the constant), zero at known DOFs.  Per matrix:
//   * k_coarse_assemble: Ac = P^T A P, one wavefront per nonzero
//     aggregate-pair block (I, J), summing over a host-sorted list of node
//     pairs (deterministic, no atomics);
//   * blocked Gauss-Jordan inversion of the SPD matrix Ac in place (pivot
//     block in LDS, then row panel, trailing update and column panel), then
//     symmetrisation.
// Per PCG iteration, on a grid of one block per aggregate:
//   * step 2 (alpha, x, r, z = B^-1 r, partials of r.z and r.r) also forms
//     the aggregate's restriction rc_I = P_I^T r;
//   * the prolongation kernel computes its rows zc_I = (Ac^-1)_I rc, adds
//     P_I zc_I to z and rc_I . zc_I to the block's r.z partial.
// Step 1 (SpMV and p update) is the one-level kernel unchanged.
#include <rocblas/rocblas.h>

#include <map>
#include <mutex>

#include "coarse_device.h"

namespace efea {

int pcg_finish_info(const void* state, double rr_true, double* info);
template <int DN, int G>
void pcg_spec_residual(const SpmvArgs& A64, const double* b, const double* x, const PcgWork& w, hipStream_t s);
template <int DN, int G>
void launch_replace_residual(const SpmvArgs& A64, const double* x, const double* b, double* r, hipStream_t s);

constexpr int GB = 32;  // Gauss-Jordan block
// Up to this many blocks per dimension the two-launch step (update forming
// P A_kj per tile) wins: the launches, not the flops, bound it.
constexpr int GJ_FUSED_MAX = 32;
// Above this order the inverse takes NB-block steps with DGEMM panels and
// updates (the GB-tile kernels re-stream the whole matrix once per 32 columns).
constexpr int GJ_BLAS_MIN = 1024;
constexpr int GJ_NB = 256;

// Doubles of scratch efea_coarse_setup needs for an ld x ld coarse matrix.
size_t coarse_scratch_doubles(int ld) {
  size_t n = 2 * (size_t)GB * GB;
  if (ld > GJ_BLAS_MIN) n += (size_t)GJ_NB * GJ_NB + 2 * (size_t)ld * GJ_NB;
  return n;
}

// One rocBLAS handle per device, bound to the caller's stream at each use.
static rocblas_handle blas_handle(hipStream_t s) {
  static std::mutex mu;
  static std::map<int, rocblas_handle> handles;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("coarse_setup: no device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  rocblas_handle& h = handles[dev];
  if (!h && rocblas_create_handle(&h) != rocblas_status_success) {
    h = nullptr;
    set_error("coarse_setup: rocblas_create_handle failed");
    return nullptr;
  }
  if (rocblas_set_stream(h, s) != rocblas_status_success) {
    set_error("coarse_setup: rocblas_set_stream failed");
    return nullptr;
  }
  return h;
}

__global__ void k_fill_f64(int64_t n, double* __restrict__ a, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Coarse matrix.  pair_ptr[cb]..pair_ptr[cb+1] lists the (node, neighbour
// position) pairs whose nodes lie in aggregates (blk_I[cb], blk_J[cb]).
template <int DN>
__global__ void __launch_bounds__(256)
k_coarse_assemble(int ncb, const int* __restrict__ pair_ptr, const int* __restrict__ pair_node,
                  const int* __restrict__ pair_pos, const int* __restrict__ blk_I, const int* __restrict__ blk_J,
                  const int* __restrict__ node_ptr, const int* __restrict__ node_cols,
                  const double* __restrict__ vals, const unsigned char* __restrict__ known,
                  const double* __restrict__ rel, double* __restrict__ Ac, int ld) {
  constexpr int NM = Modes<DN>::NM;
  const int cb = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (cb >= ncb) return;  // whole wavefront exits together
  double acc[NM][NM];
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int l = 0; l < NM; ++l) acc[m][l] = 0;
  for (int k = pair_ptr[cb] + lane; k < pair_ptr[cb + 1]; k += 64) {
    const int i = pair_node[k], pos = pair_pos[k];
    const int p0 = node_ptr[i], cnt = node_ptr[i + 1] - p0;
    const int j = node_cols[p0 + pos];
    const int64_t base = (int64_t)DN * DN * p0;
    const double* ri = rel + 3 * (int64_t)i;
    const double* rj = rel + 3 * (int64_t)j;
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      if (known && known[(int64_t)i * DN + a]) continue;
      double T[NM];
#pragma unroll
      for (int l = 0; l < NM; ++l) T[l] = 0;
#pragma unroll
      for (int b = 0; b < DN; ++b) {
        if (known && known[(int64_t)j * DN + b]) continue;
        const double v = vals[base + (int64_t)a * DN * cnt + (int64_t)pos * DN + b];
#pragma unroll
        for (int l = 0; l < NM; ++l) T[l] += v * pmode<DN>(rj, b, l);
      }
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double pm = pmode<DN>(ri, a, m);
#pragma unroll
        for (int l = 0; l < NM; ++l) acc[m][l] += pm * T[l];
      }
    }
  }
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int l = 0; l < NM; ++l) {
      double v = acc[m][l];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      acc[m][l] = v;
    }
  if (lane == 0) {
    const int I = blk_I[cb], J = blk_J[cb];
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int l = 0; l < NM; ++l) Ac[(int64_t)(I * NM + m) * ld + J * NM + l] = acc[m][l];
  }
}

// Unit diagonal where a mode vanishes on the unknown DOFs (and in the padding);
// the other diagonal entries scaled by 1 + shift (shift = 0: unchanged).
__global__ void k_coarse_guard(int ld, double* __restrict__ Ac, double shift) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ld) {
    const double v = Ac[(int64_t)k * ld + k];
    Ac[(int64_t)k * ld + k] = v == 0.0 ? 1.0 : v * (1.0 + shift);
  }
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion of the SPD matrix in place (no pivoting),
// GB x GB blocks.  Block step k with P = inv(A_kk):
//   A_ij -= A_ik P A_kj   (i, j != k)      k_gj_update
//   A_kj <- P A_kj, A_ik <- -A_ik P         k_gj_panels (+ the next pivot)
//   A_kk <- P                               the pivot inversion itself
// The trailing update reads the panels before they are scaled (it forms
// P A_kj itself), so the panels and the inversion of the next pivot block
// (final once the update of step k is done) share one launch: two launches
// per block step.

// Wavefront broadcast of lane `lane`'s value (lane is a compile-time constant
// after unrolling).
__device__ inline double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inverts pivot block k with the first wavefront of the block: lane l < GB
// holds row l in registers; each elimination step broadcasts the pivot row
// with readlane (no LDS, no barrier).  Writes P to Pb and in place.
__device__ inline void gj_pivot_wave(double* __restrict__ A, int ld, int k, double* __restrict__ Pb) {
  const int l = threadIdx.x;
  if (l >= 64) return;
  const int k0 = k * GB;
  const bool own = l < GB;
  double row[GB];
#pragma unroll
  for (int j = 0; j < GB; ++j) row[j] = own ? A[(int64_t)(k0 + l) * ld + k0 + j] : 0.0;
#pragma unroll
  for (int s = 0; s < GB; ++s) {
    double rs[GB];
#pragma unroll
    for (int j = 0; j < GB; ++j) rs[j] = readlane_f64(row[j], s);
    const double p = 1.0 / rs[s];
    const double ap = row[s] * p;
    const bool piv = l == s;
#pragma unroll
    for (int j = 0; j < GB; ++j) {
      const double upd = (j == s) ? -ap : row[j] - ap * rs[j];
      const double prow = (j == s) ? p : rs[j] * p;
      row[j] = piv ? prow : upd;
    }
  }
  if (own) {
#pragma unroll
    for (int j = 0; j < GB; ++j) {
      Pb[l * GB + j] = row[j];
      A[(int64_t)(k0 + l) * ld + k0 + j] = row[j];
    }
  }
}

__global__ void __launch_bounds__(64) k_gj_pivot(double* __restrict__ A, int ld, int k, double* __restrict__ Pb) {
  gj_pivot_wave(A, ld, k, Pb);
}

// Trailing update of tile (i, j), i, j != k: A_ij -= A_ik (P A_kj) with the
// row panel unscaled (PRESCALED false: the tile forms P A_kj itself) or
// A_ij -= A_ik A_kj with the row panel already scaled (PRESCALED true).
template <bool PRESCALED>
__global__ void __launch_bounds__(256) k_gj_update(double* __restrict__ A, int ld, int n, int k,
                                                   const double* __restrict__ Pb) {
  __shared__ double Ps[GB][GB + 1];
  __shared__ double L[GB][GB + 1];
  __shared__ double R[GB][GB + 1];
  __shared__ double T[GB][GB + 1];
  const int nbk = n / GB;
  int ib = blockIdx.y, jb = blockIdx.x;
  if (ib >= k) ++ib;
  if (jb >= k) ++jb;
  if (ib >= nbk || jb >= nbk) return;
  const int k0 = k * GB, i0 = ib * GB, j0 = jb * GB;
  for (int e = threadIdx.x; e < GB * GB; e += 256) {
    const int r = e / GB, c = e % GB;
    if (!PRESCALED) Ps[r][c] = Pb[e];
    L[r][c] = A[(int64_t)(i0 + r) * ld + k0 + c];
    R[r][c] = A[(int64_t)(k0 + r) * ld + j0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < GB * GB; e += 256) {
    const int r = e / GB, c = e % GB;
    double s = 0;
    if (PRESCALED) {
      s = R[r][c];
    } else {
#pragma unroll 8
      for (int t = 0; t < GB; ++t) s += Ps[r][t] * R[t][c];
    }
    T[r][c] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < GB * GB; e += 256) {
    const int r = e / GB, c = e % GB;
    double s = 0;
#pragma unroll 8
    for (int t = 0; t < GB; ++t) s += L[r][t] * T[t][c];
    A[(int64_t)(i0 + r) * ld + j0 + c] -= s;
  }
}

// Panels of step k and the inversion of pivot block k+1.  WHICH 0: blocks
// 0 .. nbk-2 row tiles, nbk-1 .. 2nbk-3 column tiles, then the pivot block;
// WHICH 1: row tiles only; WHICH 2: column tiles, then the pivot block.
template <int WHICH>
__global__ void __launch_bounds__(256) k_gj_panels(double* __restrict__ A, int ld, int n, int k,
                                                   const double* __restrict__ Pb, double* __restrict__ Pnext) {
  __shared__ double Ps[GB][GB + 1];
  __shared__ double T[GB][GB + 1];
  const int nbk = n / GB;
  const int nrow = WHICH == 2 ? 0 : nbk - 1, ncol = WHICH == 1 ? 0 : nbk - 1;
  const int b = blockIdx.x;
  if (b >= nrow + ncol) {
    if (WHICH != 1 && k + 1 < nbk) gj_pivot_wave(A, ld, k + 1, Pnext);
    return;
  }
  const bool rowp = b < nrow;
  int t = rowp ? b : b - nrow;
  if (t >= k) ++t;
  const int k0 = k * GB, t0 = t * GB;
  for (int e = threadIdx.x; e < GB * GB; e += 256) {
    const int r = e / GB, c = e % GB;
    Ps[r][c] = Pb[e];
    T[r][c] = rowp ? A[(int64_t)(k0 + r) * ld + t0 + c] : A[(int64_t)(t0 + r) * ld + k0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < GB * GB; e += 256) {
    const int r = e / GB, c = e % GB;
    double s = 0;
    if (rowp) {
#pragma unroll 8
      for (int q = 0; q < GB; ++q) s += Ps[r][q] * T[q][c];
      A[(int64_t)(k0 + r) * ld + t0 + c] = s;
    } else {
#pragma unroll 8
      for (int q = 0; q < GB; ++q) s += T[r][q] * Ps[q][c];
      A[(int64_t)(t0 + r) * ld + k0 + c] = -s;
    }
  }
}

// X <- (X + X^T) / 2 (upper triangle pairs).
__global__ void k_symmetrize(double* __restrict__ A, int ld) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)ld * ld) return;
  const int i = (int)(e / ld), j = (int)(e % ld);
  if (j <= i) return;
  const double v = 0.5 * (A[(int64_t)i * ld + j] + A[(int64_t)j * ld + i]);
  A[(int64_t)i * ld + j] = v;
  A[(int64_t)j * ld + i] = v;
}

// In-place inverse of the n x n SPD block at A (leading dimension ld, n a
// multiple of GB) with the GB x GB kernels.  Pb: 2 GB^2 doubles (two pivot
// inverses used in turn).
static void gj_invert(double* A, int ld, int n, double* Pb, hipStream_t s) {
  const int nbk = n / GB;
  hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(64), 0, s, A, ld, 0, Pb);
  for (int k = 0; k < nbk; ++k) {
    double* Pk = Pb + (k & 1) * GB * GB;
    double* Pn = Pb + ((k + 1) & 1) * GB * GB;
    if (nbk > 1 && nbk <= GJ_FUSED_MAX) {
      // two launches per step: the update forms P A_kj itself
      hipLaunchKernelGGL(k_gj_update<false>, dim3(nbk - 1, nbk - 1), dim3(256), 0, s, A, ld, n, k, (const double*)Pk);
      hipLaunchKernelGGL(k_gj_panels<0>, dim3(2 * (nbk - 1) + 1), dim3(256), 0, s, A, ld, n, k, (const double*)Pk,
                         Pn);
    } else if (nbk > 1) {
      // large matrices: scale the row panel once (half the update flops)
      hipLaunchKernelGGL(k_gj_panels<1>, dim3(nbk - 1), dim3(256), 0, s, A, ld, n, k, (const double*)Pk, Pn);
      hipLaunchKernelGGL(k_gj_update<true>, dim3(nbk - 1, nbk - 1), dim3(256), 0, s, A, ld, n, k, (const double*)Pk);
      hipLaunchKernelGGL(k_gj_panels<2>, dim3(nbk), dim3(256), 0, s, A, ld, n, k, (const double*)Pk, Pn);
    }
  }
}

// Large coarse matrices: Gauss-Jordan with NB x NB block steps whose panel
// products and trailing update are rocBLAS DGEMMs.  The buffer is read as a
// column-major matrix B; the matrix is symmetric, so inverting B inverts A.
// Step k (rows/columns kb of size nb): P = inv(B_kk) (GB kernels on a copy),
// C = B(:, kb), W = P B(kb, :), B -= C W, then B(kb, :) = W,
// B(:, kb) = -C P, B_kk = P.  Scratch: Pb (2 GB^2), P (NB^2), C (ld x NB),
// W (NB x ld) doubles.
static int gj_invert_blocked(double* B, int ld, double* scratch, hipStream_t s) {
  rocblas_handle h = blas_handle(s);
  if (!h) return EFEA_ERR_HIP;
  double* Pb = scratch;
  double* P = Pb + 2 * GB * GB;
  double* Cw = P + (size_t)GJ_NB * GJ_NB;
  double* Ww = Cw + (size_t)ld * GJ_NB;
  const double one = 1.0, minus = -1.0, zero = 0.0;
  for (int k0 = 0; k0 < ld; k0 += GJ_NB) {
    const int nb = ld - k0 < GJ_NB ? ld - k0 : GJ_NB;  // a multiple of GB
    double* Bkk = B + (size_t)k0 * ld + k0;
    if (hipMemcpy2DAsync(P, nb * sizeof(double), Bkk, ld * sizeof(double), nb * sizeof(double), nb,
                         hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpy2DAsync(Cw, ld * sizeof(double), B + (size_t)k0 * ld, ld * sizeof(double), ld * sizeof(double), nb,
                         hipMemcpyDeviceToDevice, s) != hipSuccess) {
      set_error("coarse_setup: panel copy failed");
      return EFEA_ERR_HIP;
    }
    gj_invert(P, nb, nb, Pb, s);
    // W = P B(kb, :)                     (nb x ld)
    // B = B - C W                        (ld x ld)
    // B(:, kb) = -C P                    (ld x nb)
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nb, ld, nb, &one, P, nb, B + k0, ld, &zero,
                      Ww, nb) != rocblas_status_success ||
        rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ld, ld, nb, &minus, Cw, ld, Ww, nb, &one, B,
                      ld) != rocblas_status_success ||
        rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ld, nb, nb, &minus, Cw, ld, P, nb, &zero,
                      B + (size_t)k0 * ld, ld) != rocblas_status_success) {
      set_error("coarse_setup: rocblas_dgemm failed");
      return EFEA_ERR_HIP;
    }
    // B(kb, :) = W, then B_kk = P
    if (hipMemcpy2DAsync(B + k0, ld * sizeof(double), Ww, nb * sizeof(double), nb * sizeof(double), ld,
                         hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpy2DAsync(Bkk, ld * sizeof(double), P, nb * sizeof(double), nb * sizeof(double), nb,
                         hipMemcpyDeviceToDevice, s) != hipSuccess) {
      set_error("coarse_setup: panel copy failed");
      return EFEA_ERR_HIP;
    }
  }
  return EFEA_OK;
}

template <int DN>
static int coarse_setup_t(int ncb, const int* pair_ptr, const int* pair_node, const int* pair_pos,
                          const int* blk_I, const int* blk_J, const int* node_ptr, const int* node_cols,
                          const double* vals, const unsigned char* known, const double* rel, int ld,
                          double shift, double* Ac, double* Pb, hipStream_t s) {
  const int64_t n2 = (int64_t)ld * ld;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, n2, Ac, 0.0);
  if (ncb > 0)
    hipLaunchKernelGGL((k_coarse_assemble<DN>), dim3((unsigned)((ncb + 3) / 4)), dim3(256), 0, s, ncb, pair_ptr,
                       pair_node, pair_pos, blk_I, blk_J, node_ptr, node_cols, vals, known, rel, Ac, ld);
  hipLaunchKernelGGL(k_coarse_guard, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, s, ld, Ac, shift);
  if (ld > GJ_BLAS_MIN) {
    if (int e = gj_invert_blocked(Ac, ld, Pb, s)) return e;
  } else {
    gj_invert(Ac, ld, ld, Pb, s);
  }
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, Ac, ld);
  return check_launch("coarse_setup");
}

int coarse_setup(int dof_n, int ncb, const int* pair_ptr, const int* pair_node, const int* pair_pos,
                 const int* blk_I, const int* blk_J, const int* node_ptr, const int* node_cols, const double* vals,
                 const unsigned char* known, const double* rel, int ld, double shift, double* Ac, double* Pb,
                 hipStream_t s) {
  if (ld % GB != 0) {
    set_error("coarse_setup: leading dimension %d not a multiple of %d", ld, GB);
    return EFEA_ERR_ARG;
  }
  if (dof_n == 1)
    return coarse_setup_t<1>(ncb, pair_ptr, pair_node, pair_pos, blk_I, blk_J, node_ptr, node_cols, vals, known,
                             rel, ld, shift, Ac, Pb, s);
  if (dof_n == 2)
    return coarse_setup_t<2>(ncb, pair_ptr, pair_node, pair_pos, blk_I, blk_J, node_ptr, node_cols, vals, known,
                             rel, ld, shift, Ac, Pb, s);
  if (dof_n == 3)
    return coarse_setup_t<3>(ncb, pair_ptr, pair_node, pair_pos, blk_I, blk_J, node_ptr, node_cols, vals, known,
                             rel, ld, shift, Ac, Pb, s);
  set_error("coarse_setup: dof_n %d", dof_n);
  return EFEA_ERR_ARG;
}

int coarse_ld(int nagg, int dof_n) {
  const int nc = nagg * (dof_n == 1 ? 1 : (dof_n == 2 ? 3 : 6));
  return (nc + GB - 1) / GB * GB;
}

// ---------------------------------------------------------------------------
// Per-iteration kernels on the aggregate grid (block = aggregate).

// TV: vector type (float in the refinement mode's inner solves); TA: type of
// the coarse inverse's values.
template <typename TP, int DN, typename TV = double>
__global__ void __launch_bounds__(BS)
k_pcg2_step2(CoarseArgs C, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
             TV* __restrict__ x, TV* __restrict__ r, TV* __restrict__ z, const TV* __restrict__ P0,
             const TV* __restrict__ P1, const TV* __restrict__ q, const double* __restrict__ part1, int nb1,
             double* __restrict__ part2, PcgState* st) {
  pcg2_step2_body<TP, DN>(C, known, Minv, x, r, z, P0, P1, q, part1, nb1, part2, st, blockIdx.x);
}

template <int DN, typename TV = double, typename TA = double>
__global__ void __launch_bounds__(BS)
k_pcg2_prolong(CoarseArgs C, const unsigned char* __restrict__ known, TV* __restrict__ z,
               double* __restrict__ part2, PcgState* st) {
  pcg2_prolong_body<DN, TV, TA>(C, known, z, part2, st, blockIdx.x);
}

template <typename TP, int DN, typename TV = double>
__global__ void __launch_bounds__(BS)
k_pcg2_init(CoarseArgs C, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
            const TV* __restrict__ r, TV* __restrict__ z, TV* __restrict__ P0,
            const double* __restrict__ partI, int nbI, double* __restrict__ part2, PcgState* st) {
  pcg2_init_body<TP, DN>(C, known, Minv, r, z, P0, partI, nbI, part2, st, blockIdx.x);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS)
k_pcg2_replace_z(CoarseArgs C, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
                 const double* __restrict__ r, double* __restrict__ z, double* __restrict__ part2, PcgState* st) {
  pcg2_rebuild_z_body<TP, DN>(C, known, Minv, r, z, part2, st, blockIdx.x);
}

// Step 1 on a two-level solve: the one-level kernel (partials of step 2 now
// come from nagg blocks).
template <typename TM, int DN, int G, typename TV = double>
__global__ void __launch_bounds__(BS)
k_pcg2_step1(SpmvArgs A, const TV* __restrict__ z, const TV* __restrict__ P0, const TV* __restrict__ P1,
             TV* __restrict__ q, const double* __restrict__ part2, int nb2, double* __restrict__ part1,
             PcgState* st) {
  const bool one_pass = (int64_t)gridDim.x * BS >= (int64_t)A.Nnode * G;
  pcg_step1_body<TM, DN, G>(A, z, P0, P1, q, part2, nb2, part1, st,
                            one_pass ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x, gridDim.x);
}

__global__ void k_state_reset2(PcgState* st, double tolr2, double tolA2, double delta2) {
  if (threadIdx.x == 0) {
    st->rz = st->rz_prev = st->alpha = st->beta = st->pq = 0;
    st->rr = st->rrmax = st->rrmax_prev = st->bb = 0;
    st->tolr2 = tolr2;
    st->tolA2 = tolA2;
    st->delta2 = delta2;
    st->Sall = st->Sw = st->xAx0 = 0;
    for (int i = 0; i < RING; ++i) st->ring[i] = 0;
    st->head = st->done = st->iters = st->nrepl = 0;
    st->fresh = 1;
    st->justrep = 0;
    st->par = 0;
    st->par_next = 1;
  }
}

__global__ void __launch_bounds__(BS) k_pcg2_init_scalars(const double* __restrict__ part2, int nb2, PcgState* st) {
  __shared__ double sh[BS / 64];
  const double rz = reduce_all(part2, nb2, 2, 0, sh);
  const double rr = reduce_all(part2, nb2, 2, 1, sh);
  if (threadIdx.x == 0) {
    st->rz = rz;
    st->rr = rr;
    st->rrmax = rr;
    st->bb = rr;
    st->done = (rr == 0) ? 1 : 0;
  }
}

// x0.A x0 and r0 = b - A x0 on the node-group grid.
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_xax_residual(SpmvArgs A, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
               double* __restrict__ y, double* __restrict__ partI) {
  __shared__ double sh[BS / 64];
  double d = 0;
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double dd, d2;
    spmv_node<double, DN, G, 0>(A, n, g, x, nullptr, y, dd);
    spmv_node<double, DN, G, 1>(A, n, g, x, b, r, d2);
    d += dd;
  }
  d = block_sum<double, BS>(d, sh);
  if (threadIdx.x == 0) partI[blockIdx.x] = d;
}

size_t pcg2_workspace_bytes(int Nnode, int dof_n, int nagg) { return pcg2_workspace_size(Nnode, dof_n, nagg); }

template <typename TP, int DN, int G>
int pcg2_start(const SpmvArgs& A64, const TP* Minv, const CoarseArgs& C, const double* b, const double* x,
               const Pcg2Work& v, double tol_r, double tol_A, double delta2, hipStream_t s) {
  const PcgWork& w = v.w;
  CoarseArgs Cw = C;
  Cw.rc = v.rc;
  const unsigned nbS = grid_for((int64_t)A64.Nnode * G, A64.cap);
  const unsigned nbA = (unsigned)C.nagg;
  hipLaunchKernelGGL(k_state_reset2, dim3(1), dim3(64), 0, s, w.st, tol_r * tol_r, tol_A * tol_A, delta2);
  hipLaunchKernelGGL((k_xax_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, x, b, w.r, w.q, w.partI);
  hipLaunchKernelGGL((k_pcg2_init<TP, DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, Minv, (const double*)w.r, w.z,
                     w.P0, (const double*)w.partI, (int)nbS, w.part2, w.st);
  hipLaunchKernelGGL((k_pcg2_prolong<DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, w.z, w.part2, w.st);
  hipLaunchKernelGGL(k_pcg2_init_scalars, dim3(1), dim3(BS), 0, s, (const double*)w.part2, (int)nbA, w.st);
  return check_launch("pcg2_init");
}

#define EFEA_PCG2_START(DN, G)                                                                             \
  template int pcg2_start<double, DN, G>(const SpmvArgs&, const double*, const CoarseArgs&, const double*,   \
                                         const double*, const Pcg2Work&, double, double, double, hipStream_t); \
  template int pcg2_start<float, DN, G>(const SpmvArgs&, const float*, const CoarseArgs&, const double*,     \
                                        const double*, const Pcg2Work&, double, double, double, hipStream_t);
EFEA_GROUPS(EFEA_PCG2_START)
#undef EFEA_PCG2_START

template <typename TM, typename TP, int DN, int G>
static int pcg2_run(const SpmvArgs& A64, const SpmvArgs& AK, const TP* Minv, const CoarseArgs& C, const double* b,
                    double* x, void* ws, double tol_r, double tol_A, double delta, int maxit, int check_every,
                    int first_check, hipStream_t s, double* info) {
  const int Nnode = A64.Nnode;
  const Pcg2Work v = carve2(ws, Nnode, DN, C.nagg);
  const PcgWork& w = v.w;
  CoarseArgs Cw = C;
  Cw.rc = v.rc;
  const unsigned nbS = grid_for((int64_t)Nnode * G, A64.cap);
  const unsigned nbA = (unsigned)C.nagg;
  const bool mixed = (void*)AK.vals != (void*)A64.vals;
  if (int e = pcg2_start<TP, DN, G>(A64, Minv, C, b, x, v, tol_r, tol_A, mixed ? delta * delta : 0.0, s)) return e;
  PollBuf* pb = poll_buffer();
  if (!pb) {
    set_error("pcg2: pinned host allocation failed");
    return EFEA_ERR_HIP;
  }
  PcgState* hs = &pb->st;
  int launched = 0;
  bool spec = false;
  while (true) {
    const int chunk = launched == 0 && first_check > 0 ? first_check : check_every;
    for (int k = 0; k < chunk; ++k) {
      hipLaunchKernelGGL((k_pcg2_step1<TM, DN, G>), dim3(nbS), dim3(BS), 0, s, AK, (const double*)w.z,
                         (const double*)w.P0, (const double*)w.P1, w.q, (const double*)w.part2, (int)nbA, w.part1,
                         w.st);
      hipLaunchKernelGGL((k_pcg2_step2<TP, DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, Minv, x, w.r, w.z,
                         (const double*)w.P0, (const double*)w.P1, (const double*)w.q, (const double*)w.part1,
                         (int)nbS, w.part2, w.st);
      hipLaunchKernelGGL((k_pcg2_prolong<DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, w.z, w.part2, w.st);
    }
    launched += chunk;
    // FP64 solve with predicted read-backs: this read-back is likely the last
    // (the residual goes to q and part1, dead at a read-back)
    spec = !mixed && first_check > 0;
    if (spec) {
      pcg_spec_residual<DN, G>(A64, b, x, w, s);
      if (hipMemcpyAsync(&pb->rr, w.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("pcg2: residual readback failed");
        return EFEA_ERR_HIP;
      }
    }
    if (int e = check_launch("pcg2_iter")) return e;
    if (hipMemcpyAsync(hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg2: state readback failed");
      return EFEA_ERR_HIP;
    }
    if (hs->done == 4 || (mixed && delta <= 0 && hs->done == 0 && launched < maxit)) {
      launch_replace_residual<DN, G>(A64, (const double*)x, b, w.r, s);
      hipLaunchKernelGGL((k_pcg2_replace_z<TP, DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, Minv,
                         (const double*)w.r, w.z, w.part2, w.st);
      hipLaunchKernelGGL((k_pcg2_prolong<DN>), dim3(nbA), dim3(BS), 0, s, Cw, A64.known, w.z, w.part2, w.st);
      if (int e = check_launch("pcg2_replace")) return e;
      if (launched >= maxit) break;
      continue;
    }
    if (hs->done || launched >= maxit) break;
  }
  const PcgState fin = *hs;
  if (!spec) {  // else the last read-back carried the residual
    pcg_true_residual<DN, G>(A64, b, x, w, s);
    if (hipMemcpyAsync(&pb->rr, w.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg2: residual readback failed");
      return EFEA_ERR_HIP;
    }
  }
  return pcg_finish_info(&fin, pb->rr, info);
}

// ---------------------------------------------------------------------------
// Refinement mode (mode 33): FP64 iterative refinement around an FP32 inner
// two-level PCG.  Each outer step forms r = b - A64 x in FP64 (masked rows)
// and its FP32 copy, solves A32 c = r from c = 0 with FP32 vectors, FP32
// matrix values (row-sum corrected), FP32 block Jacobi and FP32 Ac^-1, to a
// relative accuracy tol_in, and updates x += c in FP64.  It stops when the
// FP64 residual is below tol_r of the first one, or when the error estimate
// after the update - the inner solve's own A-norm error estimate plus the
// FP32 operator's share, taken as 1e-3 of the correction (measured: the
// refinement contracts like the FP64 solve) - is below tol_A of x's energy
// (x.A64 x from the residual pass).  A residual that grows between outer
// steps reports divergence (status 3), for the caller's FP64 fallback.

// r64 = b - A64 x on the masked rows -> r32; partials of r.r and x.A x.
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_ref_residual(SpmvArgs A64, const double* __restrict__ x, const double* __restrict__ b, float* __restrict__ r32,
               double* __restrict__ part) {
  __shared__ double shn[2][BS / 64];
  double red[2] = {0, 0};
  const int64_t lanes = (int64_t)A64.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    const int p0 = A64.node_ptr[n], cnt = A64.node_ptr[n + 1] - p0;
    double sum[1][DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) sum[0][a] = 0;
    gather_rows<double, DN, G, 1>((const double*)A64.vals, A64.node_cols, p0, cnt, g,
                                  [&](int c, double (&xv)[1][DN]) {
#pragma unroll
                                    for (int bb = 0; bb < DN; ++bb) xv[0][bb] = x[(int64_t)c * DN + bb];
                                  },
                                  sum);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) sum[0][a] += __shfl_xor(sum[0][a], o, G);
    }
    if (g == 0) {
#pragma unroll
      for (int a = 0; a < DN; ++a) {
        const int64_t i = (int64_t)n * DN + a;
        const double xi = x[i];
        red[1] += xi * sum[0][a];
        double ri = b[i] - sum[0][a];
        if (A64.known && A64.known[i]) ri = 0;
        r32[i] = (float)ri;
        red[0] += ri * ri;
      }
    }
  }
  block_sum_all<2, BS>(red, shn);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0];
    part[2 * blockIdx.x + 1] = red[1];
  }
}

__global__ void __launch_bounds__(BS) k_ref_sums(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[BS / 64];
  const double rr = reduce_all(part, nb, 2, 0, sh);
  const double xax = reduce_all(part, nb, 2, 1, sh);
  if (threadIdx.x == 0) {
    out[0] = rr;
    out[1] = xax;
  }
}

// x += c (c = 0 at known DOFs); c is cleared for the next inner solve.
__global__ void __launch_bounds__(BS) k_ref_update(int64_t n, float* __restrict__ c, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    x[i] += (double)c[i];
    c[i] = 0.0f;
  }
}

__global__ void __launch_bounds__(BS) k_fill_f32(int64_t n, float* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) a[i] = 0.0f;
}

// Inner state: relative tolerances, and x0.A x0 standing for the absolute
// A-norm target, so the stopping test max(Sall, xAx0) * tol_in^2 stops the
// inner solve at the outer target as well.
__global__ void k_ref_state(PcgState* st, double tol_in2, double xax0) {
  if (threadIdx.x == 0) {
    st->rz = st->rz_prev = st->alpha = st->beta = st->pq = 0;
    st->rr = st->rrmax = st->rrmax_prev = st->bb = 0;
    st->tolr2 = tol_in2;
    st->tolA2 = tol_in2;
    st->delta2 = 0;
    st->Sall = st->Sw = 0;
    st->xAx0 = xax0;
    for (int i = 0; i < RING; ++i) st->ring[i] = 0;
    st->head = st->done = st->iters = st->nrepl = 0;
    st->fresh = 1;
    st->justrep = 0;
    st->par = 0;
    st->par_next = 1;
  }
}

struct RefWork {
  float *r, *c, *z, *q, *P0, *P1;
  double *part1, *part2, *partR, *small, *rc;
  PcgState* st;
};

// The two-level workspace with each FP64 vector slot holding two FP32 vectors.
static RefWork carve_ref(void* ws, int Nnode, int dof_n, int nagg) {
  const Pcg2Work v = carve2(ws, Nnode, dof_n, nagg);
  const int64_t n = (int64_t)Nnode * dof_n;
  RefWork w;
  w.r = (float*)v.w.r;
  w.c = w.r + n;
  w.z = (float*)v.w.z;
  w.q = w.z + n;
  w.P0 = (float*)v.w.P0;
  w.P1 = w.P0 + n;
  w.part1 = v.w.part1;
  w.part2 = v.w.part2;
  w.small = v.w.small;
  w.st = v.w.st;
  w.rc = v.rc;
  const int nm = dof_n == 1 ? 1 : (dof_n == 2 ? 3 : 6);
  w.partR = (double*)((char*)v.rc + align256(8 * (size_t)nagg * nm));
  return w;
}

template <int DN, int G>
static int pcg2_refine(const SpmvArgs& A64, const SpmvArgs& A32, const float* Minv32, const CoarseArgs& C32,
                       const double* b, double* x, void* ws, double tol_r, double tol_A, double tol_in, int maxit,
                       int check_every, hipStream_t s, double* info, PcgState* hs) {
  const int Nnode = A64.Nnode;
  const int64_t n = (int64_t)Nnode * DN;
  const RefWork w = carve_ref(ws, Nnode, DN, C32.nagg);
  CoarseArgs C = C32;
  C.rc = w.rc;
  const unsigned nbS = grid_for((int64_t)Nnode * G, A64.cap);
  const unsigned nbA = (unsigned)C.nagg;
  const unsigned nbV = grid_for(n, A64.cap);
  constexpr double MU2 = 1e-6;  // (FP32 operator's share of the error per step)^2
  double hsum[2] = {0, 0}, rr0 = -1, rr_prev = -1;
  int total = 0, outer = 0, status = 0;
  double est = INFINITY;  // estimated squared A-norm error after the last update
  hipLaunchKernelGGL(k_fill_f32, dim3(nbV), dim3(BS), 0, s, n, w.c);
  while (true) {
    hipLaunchKernelGGL((k_ref_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, (const double*)x, b, w.r, w.partR);
    hipLaunchKernelGGL(k_ref_sums, dim3(1), dim3(BS), 0, s, (const double*)w.partR, (int)nbS, w.small);
    if (hipMemcpyAsync(hsum, w.small, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg2 refinement: residual readback failed");
      return EFEA_ERR_HIP;
    }
    const double rr = hsum[0], xax = hsum[1];
    if (rr0 < 0) rr0 = rr;
    if (!isfinite(rr) || !isfinite(xax)) {
      status = 2;
      break;
    }
    if (rr <= tol_r * tol_r * rr0 || rr == 0 || (outer > 0 && est <= tol_A * tol_A * xax)) {
      status = 1;
      break;
    }
    if (outer > 0 && rr > 4.0 * rr_prev) {  // the correction made the residual grow: diverged
      status = 3;
      break;
    }
    if (total >= maxit) break;
    rr_prev = rr;
    // inner solve A32 c = r, c from 0; its stopping test relative to
    // max(its Sall, (tol_A / tol_in)^2 x.A x), i.e. also at the outer target
    const double xax0 = xax > 0 ? xax * (tol_A / tol_in) * (tol_A / tol_in) : 0.0;
    hipLaunchKernelGGL(k_ref_state, dim3(1), dim3(64), 0, s, w.st, tol_in * tol_in, xax0);
    hipLaunchKernelGGL((k_pcg2_init<float, DN, float>), dim3(nbA), dim3(BS), 0, s, C, A64.known, Minv32,
                       (const float*)w.r, w.z, w.P0, (const double*)nullptr, 0, w.part2, w.st);
    hipLaunchKernelGGL((k_pcg2_prolong<DN, float, float>), dim3(nbA), dim3(BS), 0, s, C, A64.known, w.z, w.part2,
                       w.st);
    hipLaunchKernelGGL(k_pcg2_init_scalars, dim3(1), dim3(BS), 0, s, (const double*)w.part2, (int)nbA, w.st);
    int launched = 0;
    while (true) {
      for (int k = 0; k < check_every; ++k) {
        hipLaunchKernelGGL((k_pcg2_step1<float, DN, G, float>), dim3(nbS), dim3(BS), 0, s, A32, (const float*)w.z,
                           (const float*)w.P0, (const float*)w.P1, w.q, (const double*)w.part2, (int)nbA, w.part1,
                           w.st);
        hipLaunchKernelGGL((k_pcg2_step2<float, DN, float>), dim3(nbA), dim3(BS), 0, s, C, A64.known, Minv32, w.c,
                           w.r, w.z, (const float*)w.P0, (const float*)w.P1, (const float*)w.q,
                           (const double*)w.part1, (int)nbS, w.part2, w.st);
        hipLaunchKernelGGL((k_pcg2_prolong<DN, float, float>), dim3(nbA), dim3(BS), 0, s, C, A64.known, w.z,
                           w.part2, w.st);
      }
      launched += check_every;
      if (int e = check_launch("pcg2_refine_inner")) return e;
      if (hipMemcpyAsync(hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("pcg2 refinement: state readback failed");
        return EFEA_ERR_HIP;
      }
      if (hs->done || total + launched >= maxit) break;
    }
    total += hs->iters;
    if (hs->done == 2) {  // breakdown of the FP32 inner solve
      status = 2;
      break;
    }
    // error after the update: the inner's own estimate (its window sum once
    // it has one, else everything it did) plus the FP32 operator's share
    est = (hs->iters >= RING ? hs->Sw : hs->Sall) + MU2 * hs->Sall;
    hipLaunchKernelGGL(k_ref_update, dim3(nbV), dim3(BS), 0, s, n, w.c, x);
    ++outer;
  }
  info[0] = total;
  info[1] = rr0 > 0 ? sqrt(hsum[0] / rr0) : 0.0;
  info[2] = outer;
  info[3] = status;
  if (status == 2 || status == 3) {
    set_error("pcg2 refinement: %s after %d outer steps (%d inner iterations)",
              status == 2 ? "breakdown or non-finite values" : "divergence", outer, total);
    return EFEA_ERR_NOCONV;
  }
  if (status != 1) {
    set_error("pcg2 refinement: not converged after %d inner iterations (relative residual %.3e)", total, info[1]);
    return EFEA_ERR_NOCONV;
  }
  return check_launch("pcg2_refine");
}

int step1_bench_init(const Step1Bench& w, int nb2, int64_t ndof, hipStream_t s);

// `reps` launches of the two-level solve's step-1 kernel (k_pcg2_step1: r.z
// partials from the nagg aggregate blocks) on its solve grid, in the state of
// a running solve (timing harness).
int pcg2_step1_bench(int prec, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
                     const void* vals, const float* corr32, const unsigned char* known, int nagg, const double* z,
                     double* q, void* scratch, int reps, hipStream_t s) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals, known, prec == 32 ? (const void*)corr32 : nullptr, 0};
  const unsigned nbS = grid_for((int64_t)Nnode * group);
  const Step1Bench w = step1_bench_carve(scratch, Nnode, dof_n, nagg > MAXB ? nagg : MAXB);
  if (int e = step1_bench_init(w, nagg, (int64_t)Nnode * dof_n, s)) return e;
#define SB(TM, D, GV)                                                                                        \
  if (dof_n == D && group == GV) {                                                                          \
    for (int k = 0; k < reps; ++k)                                                                          \
      hipLaunchKernelGGL((k_pcg2_step1<TM, D, GV>), dim3(nbS), dim3(BS), 0, s, A, z, (const double*)w.P0,     \
                         (const double*)w.P1, q, (const double*)w.part2, nagg, w.part1, w.st);               \
    return check_launch("pcg2_step1_bench");                                                                \
  }
  if (prec == 64) {
#define SB64(D, GV) SB(double, D, GV)
    EFEA_GROUPS(SB64)
#undef SB64
  } else if (prec == 32) {
#define SB32(D, GV) SB(float, D, GV)
    EFEA_GROUPS(SB32)
#undef SB32
  }
#undef SB
  set_error("pcg2_step1_bench: unsupported precision %d / dof_n %d / group %d", prec, dof_n, group);
  return EFEA_ERR_ARG;
}

int solve_pcg2(int mode, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
               const double* vals64, const float* vals32, const float* corr32, const unsigned char* known,
               const void* Minv, int nagg,
               const int* agg_ptr, const int* agg_nodes, const double* rel, const double* Ainv, int ld,
               const double* b, double* x, void* ws, double tol_r, double tol_A, double delta, int maxit,
               int check_every, int first_check, int max_blocks, hipStream_t s, double* info) {
  SpmvArgs A64{Nnode, node_ptr, node_cols, vals64, known, nullptr, max_blocks};
  SpmvArgs A32{Nnode, node_ptr, node_cols, vals32, known, corr32, max_blocks};
  CoarseArgs C{nagg, agg_ptr, agg_nodes, rel, Ainv, ld, nullptr};
  if (check_every < 1) check_every = 1;
  if (mode == 33) {
    // refinement: Minv and Ainv are the FP32 block-Jacobi and coarse inverses;
    // delta (replace_ratio) is the inner solves' relative accuracy
    CoarseArgs C32{nagg, agg_ptr, agg_nodes, rel, Ainv, ld, nullptr};
    PcgState hs;
    const double tol_in = delta > 0 ? delta : 1e-5;
#define RUNR(D, GV)                                                                                          \
    if (dof_n == D && group == GV && vals32)                                                                \
      return pcg2_refine<D, GV>(A64, A32, (const float*)Minv, C32, b, x, ws, tol_r, tol_A, tol_in, maxit,    \
                                check_every, s, info, &hs);
    EFEA_GROUPS(RUNR)
#undef RUNR
  }
#define RUN2(D, GV)                                                                                            \
  if (dof_n == D && group == GV) {                                                                            \
    if (mode == 64)                                                                                           \
      return pcg2_run<double, double, D, GV>(A64, A64, (const double*)Minv, C, b, x, ws, tol_r, tol_A, delta, \
                                             maxit, check_every, first_check, s, info);                       \
    return pcg2_run<float, float, D, GV>(A64, A32, (const float*)Minv, C, b, x, ws, tol_r, tol_A, delta,      \
                                         maxit, check_every, first_check, s, info);                           \
  }
  if (mode == 64 || (mode == 32 && vals32)) {
    EFEA_GROUPS(RUN2)
  }
#undef RUN2
  set_error("solve_pcg2: unsupported mode %d / dof_n %d / group %d", mode, dof_n, group);
  return EFEA_ERR_ARG;
}

}  // namespace efea
