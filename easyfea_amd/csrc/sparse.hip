// Sparse kernels: deterministic CSR scatter, block-Jacobi setup, the
// single-system PCG driver and its kernels (thin wrappers over the bodies in
// pcg_device.h, which the batch engine shares).
#include "pcg_device.h"

namespace efea {

// ---------------------------------------------------------------------------
// Deterministic scatter: out[s] = sum_{k in seg[s]..seg[s+1]} vals[perm[k]],
// summed sequentially from 0 in increasing element-entry order, which is the
// order np.bincount accumulates (_simu.py:1051-1055).  No atomics.
// Latency-bound by its chain of dependent loads (segment bounds -> entry
// indices -> entry values), so each thread keeps SPT slots in flight (slots
// t, t + 256, ... of its block's range, coalesced across the wave) and takes
// a slot's entries eight at a time, the eight index loads and then the eight
// value loads issued together (a slot of a HEXA8 mesh gathers up to 8
// entries, of a TRI3 mesh up to 6).  The adds stay in entry order.
template <typename T, int SPT>
__global__ void __launch_bounds__(256) k_scatter(int64_t nslot, const int* __restrict__ seg,
                                                 const int* __restrict__ perm, const T* __restrict__ vals,
                                                 T* __restrict__ out) {
  constexpr int U = 8;
  const int64_t s0 = (int64_t)blockIdx.x * (256 * SPT) + threadIdx.x;
  int k0[SPT], k1[SPT];
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    const int64_t s = s0 + 256 * j;
    k0[j] = s < nslot ? seg[s] : 0;
    k1[j] = s < nslot ? seg[s + 1] : 0;
  }
  T acc[SPT];
  int kb[SPT];
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    acc[j] = 0;
    kb[j] = k0[j];
  }
  while (true) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < SPT; ++j) any |= kb[j] < k1[j];
    if (!any) break;
    int idx[SPT][U];
#pragma unroll
    for (int j = 0; j < SPT; ++j)
#pragma unroll
      for (int u = 0; u < U; ++u) idx[j][u] = kb[j] + u < k1[j] ? perm[kb[j] + u] : -1;
    T v[SPT][U];
#pragma unroll
    for (int j = 0; j < SPT; ++j)
#pragma unroll
      for (int u = 0; u < U; ++u) v[j][u] = idx[j][u] >= 0 ? vals[idx[j][u]] : T(0);
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (kb[j] + u < k1[j]) acc[j] += v[j][u];
      kb[j] += U;
    }
  }
#pragma unroll
  for (int j = 0; j < SPT; ++j) {
    const int64_t s = s0 + 256 * j;
    if (s < nslot) out[s] = acc[j];
  }
}

int scatter(int prec, int64_t nslot, const int* seg, const int* perm, const void* vals, void* out,
            hipStream_t s) {
  if (nslot == 0) return EFEA_OK;
  constexpr int SPT = 2;
  const unsigned nb = (unsigned)((nslot + 256 * SPT - 1) / (256 * SPT));
  if (prec == 64)
    hipLaunchKernelGGL((k_scatter<double, SPT>), dim3(nb), dim3(256), 0, s, nslot, seg, perm, (const double*)vals,
                       (double*)out);
  else if (prec == 32)
    hipLaunchKernelGGL((k_scatter<float, SPT>), dim3(nb), dim3(256), 0, s, nslot, seg, perm, (const float*)vals,
                       (float*)out);
  else {
    set_error("scatter: precision %d", prec);
    return EFEA_ERR_ARG;
  }
  return check_launch("scatter");
}

__global__ void k_cast64to32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// FP32 copy of the values and, per node, the dof_n x dof_n block row sums
// of the rounding differences (K64 - K32), accumulated in FP64 and stored in
// FP32 (see SpmvArgs::corr).  16 lanes per node walk its positions (coalesced
// like the SpMV) and reduce with shuffles.
template <int DN>
__global__ void __launch_bounds__(256) k_split_f32(int Nnode, const int* __restrict__ node_ptr,
                                                   const double* __restrict__ v64, float* __restrict__ v32,
                                                   float* __restrict__ corr) {
  constexpr int G = 16;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int n = (int)(gid / G), g = (int)(gid % G);
  const bool have = n < Nnode;
  double c[DN][DN];
#pragma unroll
  for (int a = 0; a < DN; ++a)
#pragma unroll
    for (int b = 0; b < DN; ++b) c[a][b] = 0;
  if (have) {
    const int p0 = node_ptr[n], cnt = node_ptr[n + 1] - p0;
    const int64_t base = (int64_t)DN * DN * p0;
    for (int pos = g; pos < cnt; pos += G) {
#pragma unroll
      for (int a = 0; a < DN; ++a)
#pragma unroll
        for (int b = 0; b < DN; ++b) {
          const int64_t k = base + (int64_t)a * DN * cnt + (int64_t)pos * DN + b;
          const double v = v64[k];
          const float f = (float)v;
          v32[k] = f;
          c[a][b] += v - (double)f;
        }
    }
  }
#pragma unroll
  for (int a = 0; a < DN; ++a)
#pragma unroll
    for (int b = 0; b < DN; ++b) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) c[a][b] += __shfl_xor(c[a][b], o, G);
    }
  if (have && g == 0) {
#pragma unroll
    for (int a = 0; a < DN; ++a)
#pragma unroll
      for (int b = 0; b < DN; ++b) corr[(int64_t)n * DN * DN + a * DN + b] = (float)c[a][b];
  }
}

int split_f32(int Nnode, int dof_n, const int* node_ptr, const double* v64, float* v32, float* corr,
              hipStream_t s) {
  if (Nnode == 0) return EFEA_OK;
  const unsigned nb = (unsigned)(((int64_t)Nnode * 16 + 255) / 256);
  if (dof_n == 1) hipLaunchKernelGGL(k_split_f32<1>, dim3(nb), dim3(256), 0, s, Nnode, node_ptr, v64, v32, corr);
  else if (dof_n == 2) hipLaunchKernelGGL(k_split_f32<2>, dim3(nb), dim3(256), 0, s, Nnode, node_ptr, v64, v32, corr);
  else if (dof_n == 3) hipLaunchKernelGGL(k_split_f32<3>, dim3(nb), dim3(256), 0, s, Nnode, node_ptr, v64, v32, corr);
  else {
    set_error("split_f32: dof_n %d", dof_n);
    return EFEA_ERR_ARG;
  }
  return check_launch("split_f32");
}

int cast_f64_f32(int64_t n, const double* a, float* b, hipStream_t s) {
  if (n == 0) return EFEA_OK;
  hipLaunchKernelGGL(k_cast64to32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, b);
  return check_launch("cast");
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ a, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Single-system kernels: one system on the whole grid.

__global__ void k_state_reset(PcgState* st, double tolr2, double tolA2, double delta2) {
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

template <typename TM, int DN, int G>
__global__ void __launch_bounds__(BS)
k_pcg_step1(SpmvArgs A, const double* __restrict__ z, const double* __restrict__ P0, const double* __restrict__ P1,
            double* __restrict__ q, const double* __restrict__ part2, int nb2, double* __restrict__ part1,
            PcgState* st) {
  const bool one_pass = (int64_t)gridDim.x * BS >= (int64_t)A.Nnode * G;
  pcg_step1_body<TM, DN, G>(A, z, P0, P1, q, part2, nb2, part1, st,
                            one_pass ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x, gridDim.x);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS)
k_pcg_step2(int Nnode, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
            double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, const double* __restrict__ P0,
            const double* __restrict__ P1, const double* __restrict__ q, const double* __restrict__ part1, int nb1,
            double* __restrict__ part2, PcgState* st) {
  // the same node -> XCD ranges as step 1 (whose p and q it reads)
  const bool one_pass = (int64_t)gridDim.x * BS >= Nnode;
  pcg_step2_body<TP, DN>(Nnode, known, Minv, x, r, z, P0, P1, q, part1, nb1, part2, st,
                         one_pass ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x, gridDim.x);
}

template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_replace_residual(SpmvArgs A64, const double* __restrict__ x, const double* __restrict__ b,
                   double* __restrict__ r) {
  residual_body<DN, G>(A64, x, b, r, blockIdx.x, gridDim.x);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS)
k_replace_z(int Nnode, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
            const double* __restrict__ r, double* __restrict__ z, double* __restrict__ part2, PcgState* st) {
  rebuild_z_body<TP, DN>(Nnode, known, Minv, r, z, part2, st, blockIdx.x, gridDim.x);
}

// Residual r = b - A x (FP64), masked; per-block partials of r.r.
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_residual(SpmvArgs A, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r,
           double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double d = 0;
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double dd;
    spmv_node<double, DN, G, 1>(A, n, g, x, b, r, dd);
    d += dd;
  }
  d = block_sum<double, BS>(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

// x.A x over unknown rows (y = A x as scratch), per-block partials.
template <int DN, int G>
__global__ void __launch_bounds__(BS)
k_xax(SpmvArgs A, const double* __restrict__ x, double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double d = 0;
  const int64_t lanes = (int64_t)A.Nnode * G;
  for (int64_t gid = (int64_t)blockIdx.x * BS + threadIdx.x; gid < lanes; gid += (int64_t)gridDim.x * BS) {
    const int n = (int)(gid / G), g = (int)(gid % G);
    double dd;
    spmv_node<double, DN, G, 0>(A, n, g, x, nullptr, y, dd);
    d += dd;
  }
  d = block_sum<double, BS>(d, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

// Initial z = Minv r, p = z (buffer P0), partials of r.z and r.r; block 0
// also folds the x0.A x0 partials into the state.
template <typename TP, int DN>
__global__ void __launch_bounds__(BS)
k_pcg_init(int Nnode, const unsigned char* __restrict__ known, const TP* __restrict__ Minv,
           const double* __restrict__ r, double* __restrict__ z, double* __restrict__ p,
           const double* __restrict__ partI, int nbI, double* __restrict__ part2, PcgState* st) {
  __shared__ double sh[BS / 64];
  __shared__ double sh2[BS / 64];
  if (blockIdx.x == 0) {
    const double xax = reduce_all(partI, nbI, 1, 0, sh);
    if (threadIdx.x == 0) st->xAx0 = xax;
  }
  double s_rz = 0, s_rr = 0;
  for (int n = blockIdx.x * BS + threadIdx.x; n < Nnode; n += gridDim.x * BS) {
    double rv[DN], zv[DN];
#pragma unroll
    for (int a = 0; a < DN; ++a) rv[a] = r[(int64_t)n * DN + a];
    apply_bj<TP, DN>(Minv, known, n, rv, zv);
#pragma unroll
    for (int a = 0; a < DN; ++a) {
      z[(int64_t)n * DN + a] = zv[a];
      p[(int64_t)n * DN + a] = zv[a];
      s_rz += rv[a] * zv[a];
      s_rr += rv[a] * rv[a];
    }
  }
  s_rz = block_sum<double, BS>(s_rz, sh);
  s_rr = block_sum<double, BS>(s_rr, sh2);
  if (threadIdx.x == 0) {
    part2[2 * blockIdx.x] = s_rz;
    part2[2 * blockIdx.x + 1] = s_rr;
  }
}

__global__ void __launch_bounds__(BS) k_pcg_init_scalars(const double* __restrict__ part2, int nb2, PcgState* st) {
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

// One-block finalize of a partial array into *out.
__global__ void __launch_bounds__(BS) k_sum_partials(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[BS / 64];
  const double s = reduce_all(part, nb, 1, 0, sh);
  if (threadIdx.x == 0) *out = s;
}

// Block-Jacobi: invert each node's diagonal dof_n x dof_n block, with the
// rows/columns of known DOFs replaced by identity (the reduced system's
// diagonal block), in FP64, stored in T.
template <typename T, int DN>
__global__ void k_bj_setup(int Nnode, const int* __restrict__ node_ptr, const int* __restrict__ node_cols,
                           const double* __restrict__ vals, const unsigned char* __restrict__ known,
                           T* __restrict__ Minv) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= Nnode) return;
  const int p0 = node_ptr[n], cnt = node_ptr[n + 1] - p0;
  int pos = -1;
  for (int k = 0; k < cnt; ++k)
    if (node_cols[p0 + k] == n) pos = k;
  double Bm[DN][DN];
  const int64_t base = (int64_t)DN * DN * p0;
#pragma unroll
  for (int a = 0; a < DN; ++a)
#pragma unroll
    for (int b = 0; b < DN; ++b) Bm[a][b] = pos >= 0 ? vals[base + (int64_t)a * DN * cnt + pos * DN + b] : 0.0;
#pragma unroll
  for (int a = 0; a < DN; ++a) {
    const bool ka = known ? known[(int64_t)n * DN + a] : false;
#pragma unroll
    for (int b = 0; b < DN; ++b) {
      const bool kb = known ? known[(int64_t)n * DN + b] : false;
      if (ka || kb) Bm[a][b] = (a == b) ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int a = 0; a < DN; ++a)
    if (Bm[a][a] == 0.0) {  // orphan / empty row: identity keeps z finite
#pragma unroll
      for (int b = 0; b < DN; ++b) Bm[a][b] = Bm[b][a] = (a == b) ? 1.0 : 0.0;
    }
  double I[DN][DN];
  if constexpr (DN == 1) {
    I[0][0] = 1.0 / Bm[0][0];
  } else if constexpr (DN == 2) {
    const double det = Bm[0][0] * Bm[1][1] - Bm[1][0] * Bm[0][1];
    I[0][0] = Bm[1][1] / det;
    I[0][1] = -Bm[0][1] / det;
    I[1][0] = -Bm[1][0] / det;
    I[1][1] = Bm[0][0] / det;
  } else {
    const double(&F)[3][3] = Bm;
    const double det = F[0][0] * (F[1][1] * F[2][2] - F[2][1] * F[1][2]) -
                       F[0][1] * (F[1][0] * F[2][2] - F[2][0] * F[1][2]) +
                       F[0][2] * (F[1][0] * F[2][1] - F[2][0] * F[1][1]);
    const double rd = 1.0 / det;
    I[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * rd;
    I[0][1] = -(F[0][1] * F[2][2] - F[0][2] * F[2][1]) * rd;
    I[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * rd;
    I[1][0] = -(F[1][0] * F[2][2] - F[1][2] * F[2][0]) * rd;
    I[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * rd;
    I[1][2] = -(F[0][0] * F[1][2] - F[0][2] * F[1][0]) * rd;
    I[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * rd;
    I[2][1] = -(F[0][0] * F[2][1] - F[0][1] * F[2][0]) * rd;
    I[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * rd;
  }
#pragma unroll
  for (int a = 0; a < DN; ++a)
#pragma unroll
    for (int b = 0; b < DN; ++b) Minv[(int64_t)n * DN * DN + a * DN + b] = (T)I[a][b];
}

// Deterministic dot product (b may be null: sum of a): partials + finalize.
__global__ void __launch_bounds__(BS) k_dot_part(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    s += a[i] * (b ? b[i] : 1.0);
  s = block_sum<double, BS>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_gather_sum(int64_t n, const int* __restrict__ idx, const double* __restrict__ y,
                             double* __restrict__ out) {
  __shared__ double sh[BS / 64];
  double s = 0;
  for (int64_t i = threadIdx.x; i < n; i += BS) s += y[idx[i]];
  s = block_sum<double, BS>(s, sh);
  if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
// Host-side drivers

size_t pcg_workspace_bytes(int Nnode, int dof_n, int group) {
  (void)group;
  return pcg_workspace_size(Nnode, dof_n);
}

PollBuf* poll_buffer(int slot) {
  static thread_local PollBuf* h[2] = {nullptr, nullptr};
  PollBuf*& p = h[slot != 0 ? 1 : 0];
  if (!p) {
    if (hipHostMalloc((void**)&p, sizeof(PollBuf), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

// Initialisation of one solve: state, x0.A x0, r0 = b - A x0, z0 = p0 =
// Minv r0, r.z and r.r.  Asynchronous on s.
template <typename TP, int DN, int G>
int pcg_start(const SpmvArgs& A64, const TP* Minv, const double* b, const double* x, const PcgWork& w,
              double tol_r, double tol_A, double delta2, hipStream_t s) {
  const int Nnode = A64.Nnode;
  const unsigned nbS = grid_for((int64_t)Nnode * G, A64.cap);
  const unsigned nbN = grid_for(Nnode, A64.cap);
  hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, w.st, tol_r * tol_r, tol_A * tol_A, delta2);
  hipLaunchKernelGGL((k_xax<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, x, w.q, w.partI);
  hipLaunchKernelGGL((k_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, x, b, w.r, w.part1);
  hipLaunchKernelGGL((k_pcg_init<TP, DN>), dim3(nbN), dim3(BS), 0, s, Nnode, A64.known, Minv, (const double*)w.r,
                     w.z, w.P0, (const double*)w.partI, (int)nbS, w.part2, w.st);
  hipLaunchKernelGGL(k_pcg_init_scalars, dim3(1), dim3(BS), 0, s, (const double*)w.part2, (int)nbN, w.st);
  return check_launch("pcg_init");
}

// Final true residual r.r = |b - A x|^2 into w.small[0].  Asynchronous on s.
template <int DN, int G>
void pcg_true_residual(const SpmvArgs& A64, const double* b, const double* x, const PcgWork& w, hipStream_t s) {
  const unsigned nbS = grid_for((int64_t)A64.Nnode * G, A64.cap);
  hipLaunchKernelGGL((k_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, x, b, w.r, w.part1);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, (const double*)w.part1, (int)nbS, w.small);
}

// The same norm with the residual vector written to q, which holds nothing
// live at a state read-back (the next step 1 rewrites it before any read):
// computed with a read-back that may be the last, it spares the solve a
// second host round trip and leaves the iterates untouched.
template <int DN, int G>
void pcg_spec_residual(const SpmvArgs& A64, const double* b, const double* x, const PcgWork& w, hipStream_t s) {
  const unsigned nbS = grid_for((int64_t)A64.Nnode * G, A64.cap);
  hipLaunchKernelGGL((k_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, x, b, w.q, w.part1);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, (const double*)w.part1, (int)nbS, w.small);
}

#define EFEA_PCG_START(DN, G)                                                                              \
  template int pcg_start<double, DN, G>(const SpmvArgs&, const double*, const double*, const double*,        \
                                        const PcgWork&, double, double, double, hipStream_t);                \
  template int pcg_start<float, DN, G>(const SpmvArgs&, const float*, const double*, const double*,          \
                                       const PcgWork&, double, double, double, hipStream_t);                 \
  template void pcg_true_residual<DN, G>(const SpmvArgs&, const double*, const double*, const PcgWork&,      \
                                         hipStream_t);                                                       \
  template void pcg_spec_residual<DN, G>(const SpmvArgs&, const double*, const double*, const PcgWork&,      \
                                         hipStream_t);
EFEA_GROUPS(EFEA_PCG_START)
#undef EFEA_PCG_START

// r = b - A64 x (masked rows) on the node-group grid; used by the two-level
// driver's residual replacement.
template <int DN, int G>
void launch_replace_residual(const SpmvArgs& A64, const double* x, const double* b, double* r, hipStream_t s) {
  hipLaunchKernelGGL((k_replace_residual<DN, G>), dim3(grid_for((int64_t)A64.Nnode * G, A64.cap)), dim3(BS), 0, s, A64, x, b,
                     r);
}
#define EFEA_REPLACE_RESIDUAL(DN, G) \
  template void launch_replace_residual<DN, G>(const SpmvArgs&, const double*, const double*, double*, hipStream_t);
EFEA_GROUPS(EFEA_REPLACE_RESIDUAL)
#undef EFEA_REPLACE_RESIDUAL

// Reads back the state and the true residual and fills info[0..3]
// (iterations, relative residual, replacements, status).
static int pcg_finish(const PcgState& hs, double rr_true, double* info) {
  int done = hs.done == 4 ? 0 : hs.done;
  const int iters = hs.iters;
  info[0] = iters;
  info[1] = hs.bb > 0 ? sqrt(rr_true / hs.bb) : 0.0;
  // A "converged" iterate whose true residual is not small is a divergence,
  // unless the iterations changed x by a negligible energy (sum alpha r.z
  // within the A-norm tolerance of x0.A x0): a start at the solution (a
  // steady state carried into a transient step, a repeated solve) has an
  // initial residual at round-off that no iteration reduces further.
  const bool moved = !(hs.Sall <= hs.tolA2 * hs.xAx0);
  if (done == 1 && !(info[1] <= 1e-4) && moved) done = 3;
  info[2] = hs.nrepl;
  info[3] = done;
  if (done == 2 || done == 3) {
    set_error("pcg: %s after %d iterations", done == 2 ? "breakdown (p.Ap <= 0, r.z <= 0 or NaN)" : "divergence",
              iters);
    return EFEA_ERR_NOCONV;
  }
  if (done != 1) {
    set_error("pcg: not converged after %d iterations (relative residual %.3e)", iters, info[1]);
    return EFEA_ERR_NOCONV;
  }
  return EFEA_OK;
}

int pcg_finish_info(const void* state, double rr_true, double* info) {
  return pcg_finish(*(const PcgState*)state, rr_true, info);
}

template <typename TM, typename TP, int DN, int G>
static int pcg_run(const SpmvArgs& A64, const SpmvArgs& AK, const TP* Minv, const double* b, double* x,
                   void* ws, double tol_r, double tol_A, double delta, int maxit, int check_every,
                   int first_check, hipStream_t s, double* info, PcgState* hs_buf) {
  const int Nnode = A64.Nnode;
  PcgWork w = carve(ws, Nnode, DN);
  const unsigned nbS = grid_for((int64_t)Nnode * G, A64.cap);  // node-group kernels
  const unsigned nbN = grid_for(Nnode, A64.cap);               // node kernels
  const bool mixed = (void*)AK.vals != (void*)A64.vals;
  if (int e = pcg_start<TP, DN, G>(A64, Minv, b, x, w, tol_r, tol_A, mixed ? delta * delta : 0.0, s)) return e;
  // the caller's host buffer for the state read-backs, else this thread's pinned one
  PollBuf* pb = poll_buffer();
  if (!pb) {
    set_error("pcg: pinned host allocation failed");
    return EFEA_ERR_HIP;
  }
  PcgState* hs = hs_buf ? hs_buf : &pb->st;
  int launched = 0;
  bool spec = false;
  while (true) {
    const int chunk = launched == 0 && first_check > 0 ? first_check : check_every;
    for (int k = 0; k < chunk; ++k) {
      hipLaunchKernelGGL((k_pcg_step1<TM, DN, G>), dim3(nbS), dim3(BS), 0, s, AK, (const double*)w.z,
                         (const double*)w.P0, (const double*)w.P1, w.q, (const double*)w.part2, (int)nbN, w.part1,
                         w.st);
      hipLaunchKernelGGL((k_pcg_step2<TP, DN>), dim3(nbN), dim3(BS), 0, s, Nnode, A64.known, Minv, x, w.r, w.z,
                         (const double*)w.P0, (const double*)w.P1, (const double*)w.q, (const double*)w.part1,
                         (int)nbS, w.part2, w.st);
    }
    launched += chunk;
    // FP64 solve with predicted read-backs: this read-back is likely the last
    spec = !mixed && first_check > 0;
    if (spec) {
      pcg_spec_residual<DN, G>(A64, b, x, w, s);
      if (hipMemcpyAsync(&pb->rr, w.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("pcg: residual readback failed");
        return EFEA_ERR_HIP;
      }
    }
    if (int e = check_launch("pcg_iter")) return e;
    if (hipMemcpyAsync(hs, w.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("pcg: state readback failed");
      return EFEA_ERR_HIP;
    }
    // mixed mode: the residual is replaced by the FP64 b - K x after a pause
    // (replace_ratio > 0) or at every poll (replace_ratio = 0: periodic)
    if (hs->done == 4 || (mixed && delta <= 0 && hs->done == 0 && launched < maxit)) {
      hipLaunchKernelGGL((k_replace_residual<DN, G>), dim3(nbS), dim3(BS), 0, s, A64, (const double*)x, b, w.r);
      hipLaunchKernelGGL((k_replace_z<TP, DN>), dim3(nbN), dim3(BS), 0, s, Nnode, A64.known, Minv,
                         (const double*)w.r, w.z, w.part2, w.st);
      if (int e = check_launch("pcg_replace")) return e;
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
      set_error("pcg: residual readback failed");
      return EFEA_ERR_HIP;
    }
  }
  return pcg_finish(fin, pb->rr, info);
}

// mode 64: FP64 matrix and preconditioner.  mode 32: FP32 matrix values and
// preconditioner for the Krylov products, FP64 vectors, reliable FP64
// residual replacement.  host_state: a host PcgState the state read-backs
// land in (pageable memory is fine); null: this thread's pinned buffer,
// allocated on its first use.
int solve_pcg(int mode, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
              const double* vals64, const float* vals32, const float* corr32, const unsigned char* known,
              const void* Minv, const double* b, double* x, void* ws, double tol_r, double tol_A, double delta,
              int maxit, int check_every, int first_check, int max_blocks, hipStream_t s, double* info,
              void* host_state_buf) {
  PcgState* hsb = (PcgState*)host_state_buf;
  SpmvArgs A64{Nnode, node_ptr, node_cols, vals64, known, nullptr, max_blocks};
  SpmvArgs A32{Nnode, node_ptr, node_cols, vals32, known, corr32, max_blocks};
  if (check_every < 1) check_every = 1;
#define RUN(D, GV)                                                                                          \
  if (dof_n == D && group == GV) {                                                                         \
    if (mode == 64)                                                                                        \
      return pcg_run<double, double, D, GV>(A64, A64, (const double*)Minv, b, x, ws, tol_r, tol_A, delta,  \
                                            maxit, check_every, first_check, s, info, hsb);                \
    return pcg_run<float, float, D, GV>(A64, A32, (const float*)Minv, b, x, ws, tol_r, tol_A, delta, maxit, \
                                        check_every, first_check, s, info, hsb);                           \
  }
  if (mode == 64 || (mode == 32 && vals32)) {
    EFEA_GROUPS(RUN)
  }
#undef RUN
  set_error("solve_pcg: unsupported mode %d / dof_n %d / group %d", mode, dof_n, group);
  return EFEA_ERR_ARG;
}

// Standalone y = A x (masked rows), FP64 vectors, FP64 or FP32 values.  It
// runs the solver's own step-1 kernel with beta = 0 (p = x), so a timing of
// this entry point measures the PCG's dominant kernel.  scratch:
// spmv_scratch_bytes(Nnode, dof_n) bytes (state, partials, p copy).
size_t spmv_scratch_bytes(int Nnode, int dof_n) {
  return align256(sizeof(PcgState)) + align256(8 * MAXB) + align256(8 * (size_t)Nnode * dof_n);
}

__global__ void k_spmv_state_init(PcgState* st) {
  if (threadIdx.x == 0) {
    st->done = 0;
    st->fresh = 1;
    st->par = 0;
  }
}

int spmv(int prec, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols, const void* vals,
         const unsigned char* known, const double* x, double* y, void* scratch, int reps, hipStream_t s) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals, known};
  if (reps < 1) reps = 1;
  if (Nnode == 0) return EFEA_OK;
  const unsigned nb = grid_for((int64_t)Nnode * group);
  char* c = (char*)scratch;
  PcgState* st = (PcgState*)c;
  double* part = (double*)(c + align256(sizeof(PcgState)));
  double* pcopy = (double*)(c + align256(sizeof(PcgState)) + align256(8 * MAXB));
  hipLaunchKernelGGL(k_spmv_state_init, dim3(1), dim3(64), 0, s, st);
#define SP(TM, D, GV)                                                                                      \
  if (dof_n == D && group == GV) {                                                                        \
    for (int k = 0; k < reps; ++k)                                                                        \
      hipLaunchKernelGGL((k_pcg_step1<TM, D, GV>), dim3(nb), dim3(BS), 0, s, A, x, x, (const double*)pcopy, y, \
                         (const double*)part, 1, part, st);                                               \
    return check_launch("spmv");                                                                          \
  }
  if (prec == 64) {
#define SP64(D, GV) SP(double, D, GV)
    EFEA_GROUPS(SP64)
#undef SP64
  } else if (prec == 32) {
#define SP32(D, GV) SP(float, D, GV)
    EFEA_GROUPS(SP32)
#undef SP32
  }
#undef SP
  set_error("spmv: unsupported precision %d / dof_n %d / group %d", prec, dof_n, group);
  return EFEA_ERR_ARG;
}

__global__ void __launch_bounds__(BS) k_step1_bench_state(PcgState* st, double* part2, int nb2, double* P0,
                                                          int64_t n) {
  step1_bench_state_body(st, part2, nb2, P0, n, blockIdx.x, gridDim.x);
}

int step1_bench_init(const Step1Bench& w, int nb2, int64_t ndof, hipStream_t s) {
  hipLaunchKernelGGL(k_step1_bench_state, dim3(grid_for(ndof)), dim3(BS), 0, s, w.st, w.part2, nb2, w.P0, ndof);
  return check_launch("step1_bench_init");
}

// `reps` launches of the block-Jacobi solve's step-1 kernel (k_pcg_step1) on
// its solve grid, in the state of a running solve (timing harness).
int step1_bench(int prec, int Nnode, int dof_n, int group, const int* node_ptr, const int* node_cols,
                const void* vals, const float* corr32, const unsigned char* known, const double* z, double* q,
                void* scratch, int reps, hipStream_t s) {
  SpmvArgs A{Nnode, node_ptr, node_cols, vals, known, prec == 32 ? (const void*)corr32 : nullptr, 0};
  const unsigned nbS = grid_for((int64_t)Nnode * group), nbN = grid_for(Nnode);
  const Step1Bench w = step1_bench_carve(scratch, Nnode, dof_n, MAXB);
  if (int e = step1_bench_init(w, (int)nbN, (int64_t)Nnode * dof_n, s)) return e;
#define SB(TM, D, GV)                                                                                        \
  if (dof_n == D && group == GV) {                                                                          \
    for (int k = 0; k < reps; ++k)                                                                          \
      hipLaunchKernelGGL((k_pcg_step1<TM, D, GV>), dim3(nbS), dim3(BS), 0, s, A, z, (const double*)w.P0,      \
                         (const double*)w.P1, q, (const double*)w.part2, (int)nbN, w.part1, w.st);           \
    return check_launch("step1_bench");                                                                     \
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
  set_error("step1_bench: unsupported precision %d / dof_n %d / group %d", prec, dof_n, group);
  return EFEA_ERR_ARG;
}

int bj_setup(int prec, int Nnode, int dof_n, const int* node_ptr, const int* node_cols, const double* vals,
             const unsigned char* known, void* Minv, hipStream_t s) {
  const unsigned nb = (unsigned)((Nnode + 127) / 128);
  if (nb == 0) return EFEA_OK;
#define BJ(T, D) \
  hipLaunchKernelGGL((k_bj_setup<T, D>), dim3(nb), dim3(128), 0, s, Nnode, node_ptr, node_cols, vals, known, (T*)Minv)
  if (prec == 64) {
    if (dof_n == 1) BJ(double, 1);
    else if (dof_n == 2) BJ(double, 2);
    else if (dof_n == 3) BJ(double, 3);
  } else {
    if (dof_n == 1) BJ(float, 1);
    else if (dof_n == 2) BJ(float, 2);
    else if (dof_n == 3) BJ(float, 3);
  }
#undef BJ
  return check_launch("bj_setup");
}

// scratch: MAXB doubles of partials.
int dot(int64_t n, const double* a, const double* b, double* scratch, double* out, hipStream_t s) {
  const unsigned nb = grid_for(n);
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(BS), 0, s, n, a, b, scratch);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(BS), 0, s, (const double*)scratch, (int)nb, out);
  return check_launch("dot");
}

int gather_sum(int64_t n, const int* idx, const double* y, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_sum, dim3(1), dim3(BS), 0, s, n, idx, y, out);
  return check_launch("gather_sum");
}

// Lanes per node from the largest neighbour count, as measured on MI355X
// (PCG step-1 kernel, FP64 and FP32 values, TRI3/QUAD4 meshes of 10^4..10^6
// nodes, HEXA8/TETRA4/TETRA10 of 2*10^3..3*10^5): up to 12 neighbours of
// 1-2 DOF blocks, 4 lanes (three positions per lane in one pass) beat 8 by
// 10-25% and 16 by up to 2.6x on large meshes; 3-DOF blocks take 8 lanes up
// to 16 neighbours (TETRA4) and 16 beyond (HEXA8: 27, TETRA10: 65; 32 lanes
// lose up to 25% in FP32).  Counts between the measured families interpolate.
// Scalar patterns of at most 12 neighbours take 2 lanes (five positions per
// lane in one pass) from 5*10^4 nodes: 23% faster than 4 at 10^6 nodes
// (QUAD4 1024^2: 30.3 vs 39.6 us FP64), 8% slower at 10^4 (latency floor).
int default_group(int dof_n, int max_cnt, int Nnode) {
  if (dof_n == 3) return max_cnt > 16 ? 16 : 8;
  if (dof_n == 1 && max_cnt <= 12 && Nnode >= 50000) return 2;
  return max_cnt > 24 ? 16 : max_cnt > 12 ? 8 : 4;
}

}  // namespace efea
