// Continuous-batching PCG: many independent systems advance together, one
// launch per PCG step for all of them.
//
// Each system keeps the grid it would get alone (grid_for(Nnode*G, cap) blocks
// for the node-group kernel, grid_for(Nnode, cap) for the node kernel); a launch
// concatenates those grids and a block->system map sends every block to its
// system's descriptor.  The per-block code is the single-system code
// (pcg_device.h), so every system's reductions run in the same order as in
// a solo solve and the iterates are bitwise the ones efea_solve_pcg computes.
// Systems join (efea_batch_add) and leave (efea_batch_pop) between chunks of
// iterations; a system that has converged idles (its blocks return at once)
// until the next poll retires it.  A batch runs either the one-level (block
// Jacobi) or the two-level iteration for all its systems; the two-level one
// adds a grid of one block per aggregate for step 2 and the prolongation.
#include "coarse_device.h"

#include <vector>

namespace efea {

int pcg_finish_info(const void* state, double rr_true, double* info);

struct BatchSys {
  SpmvArgs A64;         // FP64 values: residuals and replacements
  const void* valsK;    // values of the Krylov products (FP32 copy in mode 32)
  const void* corrK;    // FP32 row-sum correction (mode 32) or null
  const void* Minv;
  const double* b;
  double* x;
  double* r;
  double* z;
  double* P0;
  double* P1;
  double* q;
  double* part1;
  double* part2;
  PcgState* st;
  CoarseArgs C;    // two-level batches
  int blkS0, nbS;  // block range in the node-group grid
  int blkN0, nbN;  // block range in the node grid
  int blkA0;       // first block in the aggregate grid (C.nagg blocks)
  int nb2;         // blocks whose r.z / r.r partials step 1 reduces
  int replace;     // set by the host at a poll: run a residual replacement
  int pad_;
};

template <typename TM, int DN, int G>
__global__ void __launch_bounds__(BS) kb_step1(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  SpmvArgs AK = S.A64;
  AK.vals = S.valsK;
  AK.corr = S.corrK;
  pcg_step1_body<TM, DN, G>(AK, S.z, S.P0, S.P1, S.q, S.part2, S.nb2, S.part1, S.st, (int)blockIdx.x - S.blkS0,
                            S.nbS);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS) kb_step2(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  pcg_step2_body<TP, DN>(S.A64.Nnode, S.A64.known, (const TP*)S.Minv, S.x, S.r, S.z, S.P0, S.P1, S.q, S.part1,
                         S.nbS, S.part2, S.st, (int)blockIdx.x - S.blkN0, S.nbN);
}

template <int DN, int G>
__global__ void __launch_bounds__(BS) kb_replace_residual(const BatchSys* __restrict__ sys,
                                                          const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  if (!S.replace) return;
  residual_body<DN, G>(S.A64, S.x, S.b, S.r, (int)blockIdx.x - S.blkS0, S.nbS);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS) kb_replace_z(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  if (!S.replace) return;
  rebuild_z_body<TP, DN>(S.A64.Nnode, S.A64.known, (const TP*)S.Minv, S.r, S.z, S.part2, S.st,
                         (int)blockIdx.x - S.blkN0, S.nbN);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS) kb2_step2(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  pcg2_step2_body<TP, DN>(S.C, S.A64.known, (const TP*)S.Minv, S.x, S.r, S.z, S.P0, S.P1, S.q, S.part1, S.nbS,
                          S.part2, S.st, (int)blockIdx.x - S.blkA0);
}

// ONLY_REPLACED: after a replacement, only the replaced systems get the
// coarse correction of their new residual.
template <int DN, bool ONLY_REPLACED>
__global__ void __launch_bounds__(BS) kb2_prolong(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  if (ONLY_REPLACED && !S.replace) return;
  pcg2_prolong_body<DN, double, double>(S.C, S.A64.known, S.z, S.part2, S.st, (int)blockIdx.x - S.blkA0);
}

template <typename TP, int DN>
__global__ void __launch_bounds__(BS) kb2_replace_z(const BatchSys* __restrict__ sys, const int* __restrict__ map) {
  const BatchSys& S = sys[map[blockIdx.x]];
  if (!S.replace) return;
  pcg2_rebuild_z_body<TP, DN>(S.C, S.A64.known, (const TP*)S.Minv, S.r, S.z, S.part2, S.st,
                              (int)blockIdx.x - S.blkA0);
}

// One block per system: copies its state into a contiguous array for a
// single device-to-host read.
__global__ void kb_gather_state(const BatchSys* __restrict__ sys, PcgState* __restrict__ out) {
  const int* src = (const int*)sys[blockIdx.x].st;
  int* dst = (int*)(out + blockIdx.x);
  for (int k = threadIdx.x; k < (int)(sizeof(PcgState) / 4); k += blockDim.x) dst[k] = src[k];
}

struct BatchEntry {
  int id;
  BatchSys d;
  PcgWork w;
  int launched, maxit;
  bool finished;
  bool periodic;  // mixed mode, replace_ratio = 0: residual replaced at every poll
  double info[4];
};

constexpr int MAXA = 4096;  // aggregates per system in a two-level batch

struct Batch {
  int mode, dof_n, group, cap;
  int two_level;
  int next_id = 0;
  std::vector<BatchEntry> active;
  std::vector<BatchEntry> done;
  BatchSys* h_sys[2] = {nullptr, nullptr};
  int flip = 0;
  BatchSys* d_sys = nullptr;
  int* h_map = nullptr;  // [cap*MAXB node-group | cap*MAXB node | cap*MAXA aggregate blocks]
  int* d_map = nullptr;
  PcgState* h_st = nullptr;
  PcgState* d_st = nullptr;
  double* h_rr = nullptr;
  int totS = 0, totN = 0, totA = 0;
  bool dirty = true;
  // a launched chunk whose states have not been read yet: the first
  // `inflight` entries of `active` ran `inflight_iters` iterations
  int inflight = 0, inflight_iters = 0;
  hipStream_t inflight_stream = nullptr;
};

static void batch_free(Batch* b) {
  if (!b) return;
  for (int i = 0; i < 2; ++i)
    if (b->h_sys[i]) (void)hipHostFree(b->h_sys[i]);
  if (b->h_map) (void)hipHostFree(b->h_map);
  if (b->h_st) (void)hipHostFree(b->h_st);
  if (b->h_rr) (void)hipHostFree(b->h_rr);
  if (b->d_sys) (void)hipFree(b->d_sys);
  if (b->d_map) (void)hipFree(b->d_map);
  if (b->d_st) (void)hipFree(b->d_st);
  delete b;
}

static bool supported(int dof_n, int group) {
#define SUP(D, GV) if (dof_n == D && group == GV) return true;
  EFEA_GROUPS(SUP)
#undef SUP
  return false;
}

static size_t map_size(int cap) { return 2 * (size_t)cap * MAXB + (size_t)cap * MAXA; }

int batch_create(int mode, int dof_n, int group, int two_level, int capacity, void** out) {
  if (mode != 64 && mode != 32) {
    set_error("batch_create: mode %d (64 or 32)", mode);
    return EFEA_ERR_ARG;
  }
  if (!supported(dof_n, group)) {
    set_error("batch_create: unsupported dof_n %d / group %d", dof_n, group);
    return EFEA_ERR_ARG;
  }
  if (capacity < 1 || capacity > 1024) {
    set_error("batch_create: capacity %d outside [1, 1024]", capacity);
    return EFEA_ERR_ARG;
  }
  Batch* b = new Batch();
  b->mode = mode;
  b->dof_n = dof_n;
  b->group = group;
  b->cap = capacity;
  b->two_level = two_level ? 1 : 0;
  const size_t nmap = map_size(capacity);
  bool ok = hipHostMalloc((void**)&b->h_sys[0], capacity * sizeof(BatchSys), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void**)&b->h_sys[1], capacity * sizeof(BatchSys), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void**)&b->h_map, nmap * sizeof(int), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void**)&b->h_st, capacity * sizeof(PcgState), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void**)&b->h_rr, capacity * sizeof(double), hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void**)&b->d_sys, capacity * sizeof(BatchSys)) == hipSuccess &&
            hipMalloc((void**)&b->d_map, nmap * sizeof(int)) == hipSuccess &&
            hipMalloc((void**)&b->d_st, capacity * sizeof(PcgState)) == hipSuccess;
  if (!ok) {
    batch_free(b);
    set_error("batch_create: allocation failed (capacity %d)", capacity);
    return EFEA_ERR_HIP;
  }
  *out = b;
  return EFEA_OK;
}

int batch_destroy(void* h) {
  Batch* B = (Batch*)h;
  if (B && B->inflight) (void)hipStreamSynchronize(B->inflight_stream);  // kernels still read its buffers
  batch_free(B);
  return EFEA_OK;
}

template <typename TP, int DN, int G>
static int start_dispatch(const SpmvArgs& A64, const void* Minv, const CoarseArgs* C, const double* b,
                          const double* x, const Pcg2Work& v, double tol_r, double tol_A, double delta2,
                          hipStream_t s) {
  if (C) return pcg2_start<TP, DN, G>(A64, (const TP*)Minv, *C, b, x, v, tol_r, tol_A, delta2, s);
  return pcg_start<TP, DN, G>(A64, (const TP*)Minv, b, x, v.w, tol_r, tol_A, delta2, s);
}

int batch_add(void* h, int Nnode, const int* node_ptr, const int* node_cols, const double* vals64,
              const float* vals32, const float* corr32, const unsigned char* known, const void* Minv, int nagg,
              const int* agg_ptr,
              const int* agg_nodes, const double* rel, const double* Ainv, int ld, const double* b, double* x,
              void* ws, size_t ws_bytes, double tol_r, double tol_A, double delta, int maxit, int max_blocks,
              hipStream_t s) {
  Batch* B = (Batch*)h;
  if (B->two_level) {
    if (nagg < 1 || nagg > MAXA || !agg_ptr || !agg_nodes || !rel || !Ainv || ld < 1) {
      set_error("batch_add: two-level batch needs 1..%d aggregates and the coarse arrays", MAXA);
      return EFEA_ERR_ARG;
    }
  } else {
    nagg = 0;
  }
  const size_t need = B->two_level ? pcg2_workspace_size(Nnode, B->dof_n, nagg) : pcg_workspace_size(Nnode, B->dof_n);
  if (ws_bytes < need) {
    set_error("batch_add: workspace of %zu bytes, %zu needed", ws_bytes, need);
    return EFEA_ERR_SIZE;
  }
  if ((int)B->active.size() >= B->cap) {
    set_error("batch_add: batch full (%d systems)", B->cap);
    return EFEA_ERR_SIZE;
  }
  if (B->mode == 32 && !vals32) {
    set_error("batch_add: mode 32 needs the FP32 value copy");
    return EFEA_ERR_ARG;
  }
  BatchEntry e{};
  e.id = B->next_id++;
  const Pcg2Work v = B->two_level ? carve2(ws, Nnode, B->dof_n, nagg) : Pcg2Work{carve(ws, Nnode, B->dof_n), nullptr};
  e.w = v.w;
  e.launched = 0;
  e.maxit = maxit;
  e.finished = false;
  e.periodic = delta <= 0;
  BatchSys& d = e.d;
  d.A64 = SpmvArgs{Nnode, node_ptr, node_cols, vals64, known, nullptr, max_blocks};
  d.valsK = B->mode == 32 ? (const void*)vals32 : (const void*)vals64;
  d.corrK = B->mode == 32 ? (const void*)corr32 : nullptr;
  d.Minv = Minv;
  d.b = b;
  d.x = x;
  d.r = e.w.r;
  d.z = e.w.z;
  d.P0 = e.w.P0;
  d.P1 = e.w.P1;
  d.q = e.w.q;
  d.part1 = e.w.part1;
  d.part2 = e.w.part2;
  d.st = e.w.st;
  d.nbS = (int)grid_for((int64_t)Nnode * B->group, max_blocks);
  d.nbN = (int)grid_for(Nnode, max_blocks);
  d.C = CoarseArgs{nagg, agg_ptr, agg_nodes, rel, Ainv, ld, v.rc};
  d.nb2 = B->two_level ? nagg : d.nbN;
  d.replace = 0;
  const double delta2 = B->mode == 32 ? delta * delta : 0.0;
  int rc = EFEA_ERR_ARG;
  const CoarseArgs* Cp = B->two_level ? &d.C : nullptr;
#define ST(D, GV)                                                                                          \
  if (B->dof_n == D && B->group == GV)                                                                    \
    rc = B->mode == 64 ? start_dispatch<double, D, GV>(d.A64, Minv, Cp, b, x, v, tol_r, tol_A, delta2, s) \
                       : start_dispatch<float, D, GV>(d.A64, Minv, Cp, b, x, v, tol_r, tol_A, delta2, s);
  EFEA_GROUPS(ST)
#undef ST
  if (rc) return rc;
  B->active.push_back(e);
  B->dirty = true;
  return e.id;
}

static int upload_sys(Batch* B, hipStream_t s) {
  // two pinned buffers used in turn: the previous upload from this buffer
  // completed before the stream synchronisation of the last poll
  BatchSys* hs = B->h_sys[B->flip];
  B->flip ^= 1;
  for (size_t i = 0; i < B->active.size(); ++i) hs[i] = B->active[i].d;
  if (hipMemcpyAsync(B->d_sys, hs, B->active.size() * sizeof(BatchSys), hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("batch: descriptor upload failed");
    return EFEA_ERR_HIP;
  }
  return EFEA_OK;
}

static int rebuild(Batch* B, hipStream_t s) {
  int* mapS = B->h_map;
  int* mapN = B->h_map + (size_t)B->cap * MAXB;
  int* mapA = B->h_map + 2 * (size_t)B->cap * MAXB;
  int tS = 0, tN = 0, tA = 0;
  for (size_t i = 0; i < B->active.size(); ++i) {
    BatchSys& d = B->active[i].d;
    d.blkS0 = tS;
    d.blkN0 = tN;
    d.blkA0 = tA;
    for (int k = 0; k < d.nbS; ++k) mapS[tS + k] = (int)i;
    for (int k = 0; k < d.nbN; ++k) mapN[tN + k] = (int)i;
    for (int k = 0; k < d.C.nagg; ++k) mapA[tA + k] = (int)i;
    tS += d.nbS;
    tN += d.nbN;
    tA += d.C.nagg;
  }
  B->totS = tS;
  B->totN = tN;
  B->totA = tA;
  const size_t offN = (size_t)B->cap * MAXB, offA = 2 * (size_t)B->cap * MAXB;
  if (!B->active.empty() &&
      (hipMemcpyAsync(B->d_map, mapS, tS * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(B->d_map + offN, mapN, tN * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
       (tA > 0 &&
        hipMemcpyAsync(B->d_map + offA, mapA, tA * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess))) {
    set_error("batch: map upload failed");
    return EFEA_ERR_HIP;
  }
  B->dirty = false;
  return upload_sys(B, s);
}

template <typename TM, typename TP, int DN, int G>
static int batch_launch_t(Batch* B, int iters, hipStream_t s) {
  const int n = (int)B->active.size();
  const int* mapS = B->d_map;
  const int* mapN = B->d_map + (size_t)B->cap * MAXB;
  const int* mapA = B->d_map + 2 * (size_t)B->cap * MAXB;
  const BatchSys* dsys = B->d_sys;
  for (int k = 0; k < iters; ++k) {
    hipLaunchKernelGGL((kb_step1<TM, DN, G>), dim3(B->totS), dim3(BS), 0, s, dsys, mapS);
    if (B->two_level) {
      hipLaunchKernelGGL((kb2_step2<TP, DN>), dim3(B->totA), dim3(BS), 0, s, dsys, mapA);
      hipLaunchKernelGGL((kb2_prolong<DN, false>), dim3(B->totA), dim3(BS), 0, s, dsys, mapA);
    } else {
      hipLaunchKernelGGL((kb_step2<TP, DN>), dim3(B->totN), dim3(BS), 0, s, dsys, mapN);
    }
  }
  hipLaunchKernelGGL(kb_gather_state, dim3(n), dim3(64), 0, s, dsys, B->d_st);
  if (int e = check_launch("batch_iter")) return e;
  if (hipMemcpyAsync(B->h_st, B->d_st, n * sizeof(PcgState), hipMemcpyDeviceToHost, s) != hipSuccess) {
    set_error("batch: state readback failed");
    return EFEA_ERR_HIP;
  }
  B->inflight = n;
  B->inflight_iters = iters;
  B->inflight_stream = s;
  return EFEA_OK;
}

// Reads the states of the launched chunk (synchronising the stream), queues
// the residual replacements, finishes the systems that stopped.  Systems
// added while the chunk ran (entries past `inflight`) are left as they are.
template <typename TM, typename TP, int DN, int G>
static int batch_wait_t(Batch* B, hipStream_t s) {
  const int n = B->inflight, iters = B->inflight_iters;
  B->inflight = 0;
  const int* mapS = B->d_map;
  const int* mapN = B->d_map + (size_t)B->cap * MAXB;
  const int* mapA = B->d_map + 2 * (size_t)B->cap * MAXB;
  const BatchSys* dsys = B->d_sys;
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("batch: state readback failed");
    return EFEA_ERR_HIP;
  }
  bool any_replace = false;
  int nfin = 0;
  for (int i = 0; i < n; ++i) {
    BatchEntry& e = B->active[i];
    const PcgState& st = B->h_st[i];
    e.launched += iters;
    e.finished = (st.done != 0 && st.done != 4) || e.launched >= e.maxit;
    // mixed mode: replacement after a pause, or at every poll (periodic)
    e.d.replace = st.done == 4 || (B->mode == 32 && e.periodic && st.done == 0 && !e.finished);
    any_replace |= e.d.replace != 0;
    nfin += e.finished;
  }
  if (any_replace) {
    // the kernels address systems through the maps of the chunk, which cover
    // the first n descriptors only
    if (int e = upload_sys(B, s)) return e;
    hipLaunchKernelGGL((kb_replace_residual<DN, G>), dim3(B->totS), dim3(BS), 0, s, dsys, mapS);
    if (B->two_level) {
      hipLaunchKernelGGL((kb2_replace_z<TP, DN>), dim3(B->totA), dim3(BS), 0, s, dsys, mapA);
      hipLaunchKernelGGL((kb2_prolong<DN, true>), dim3(B->totA), dim3(BS), 0, s, dsys, mapA);
    } else {
      hipLaunchKernelGGL((kb_replace_z<TP, DN>), dim3(B->totN), dim3(BS), 0, s, dsys, mapN);
    }
    if (int e = check_launch("batch_replace")) return e;
    for (int i = 0; i < n; ++i) B->active[i].d.replace = 0;
  }
  if (nfin == 0) return 0;
  int j = 0;
  for (int i = 0; i < n; ++i) {
    BatchEntry& e = B->active[i];
    if (!e.finished) continue;
    pcg_true_residual<DN, G>(e.d.A64, e.d.b, e.d.x, e.w, s);
    if (hipMemcpyAsync(B->h_rr + j, e.w.small, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_error("batch: residual readback failed");
      return EFEA_ERR_HIP;
    }
    ++j;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("batch: residual readback failed");
    return EFEA_ERR_HIP;
  }
  std::vector<BatchEntry> keep;
  keep.reserve(B->active.size());
  j = 0;
  for (int i = 0; i < (int)B->active.size(); ++i) {
    BatchEntry& e = B->active[i];
    if (i >= n || !e.finished) {
      keep.push_back(e);
      continue;
    }
    (void)pcg_finish_info(&B->h_st[i], B->h_rr[j++], e.info);
    B->done.push_back(e);
  }
  B->active.swap(keep);
  B->dirty = true;
  return nfin;
}

// Launches `iters` iterations of every active system and the read-back of
// their states, without waiting (at most one chunk in flight per batch).
int batch_launch(void* h, int iters, hipStream_t s) {
  Batch* B = (Batch*)h;
  if (B->inflight) {
    set_error("batch_launch: a chunk is already in flight (call batch_wait first)");
    return EFEA_ERR_ARG;
  }
  if (B->active.empty()) return 0;
  if (iters < 1) iters = 1;
  if (B->dirty) {
    if (int e = rebuild(B, s)) return e;
  }
#define LN(D, GV)                                                                             \
  if (B->dof_n == D && B->group == GV)                                                       \
    return B->mode == 64 ? batch_launch_t<double, double, D, GV>(B, iters, s)                \
                         : batch_launch_t<float, float, D, GV>(B, iters, s);
  EFEA_GROUPS(LN)
#undef LN
  set_error("batch_launch: unsupported configuration");
  return EFEA_ERR_ARG;
}

// Waits for the chunk in flight; systems that converged, failed or reached
// maxit move to the finished list.  Returns how many finished (0 when no
// chunk was in flight).
int batch_wait(void* h, hipStream_t s) {
  Batch* B = (Batch*)h;
  if (!B->inflight) return 0;
#define WT(D, GV)                                                                             \
  if (B->dof_n == D && B->group == GV)                                                       \
    return B->mode == 64 ? batch_wait_t<double, double, D, GV>(B, s)                         \
                         : batch_wait_t<float, float, D, GV>(B, s);
  EFEA_GROUPS(WT)
#undef WT
  set_error("batch_wait: unsupported configuration");
  return EFEA_ERR_ARG;
}

// Runs `iters` iterations of every active system, then polls: systems that
// converged, failed or reached maxit move to the finished list.  Returns the
// number of systems that finished in this call.
int batch_run(void* h, int iters, hipStream_t s) {
  if (int e = batch_launch(h, iters, s)) return e;
  return batch_wait(h, s);
}

// Pops one finished system: its id and info[0..3] (iterations, relative
// residual, replacements, status: 1 converged, 0 stopped at maxit, 2
// breakdown, 3 diverged).  Returns 0, or 1 with *id = -1 when nothing has
// finished.
int batch_pop(void* h, int* id, double* info) {
  Batch* B = (Batch*)h;
  if (B->done.empty()) {
    *id = -1;
    return 1;
  }
  const BatchEntry e = B->done.front();
  B->done.erase(B->done.begin());
  *id = e.id;
  for (int k = 0; k < 4; ++k) info[k] = e.info[k];
  return EFEA_OK;
}

int batch_active(void* h) { return (int)((Batch*)h)->active.size(); }

}  // namespace efea
