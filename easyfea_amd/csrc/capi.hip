// C ABI of the efea HIP library: plain pointers, sizes and a hipStream_t,
// int status returns (0 ok, negative error), a thread-local message getter.
// Declarations and their reference counterparts: include/efea.h.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "efea_common.h"
#include "pcg_device.h"

namespace efea {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return EFEA_ERR_HIP;
  }
  return EFEA_OK;
}

// implemented in elements.hip / sparse.hip / phasefield.hip
int elastic_ke(int, int, int64_t, const int*, const double*, const double*, int, const double*, int, double, int,
               double, void*, hipStream_t);
int reaction_diffusion(int, int, int64_t, const int*, const double*, const double*, int, const double*, int,
                       const double*, double, const double*, int, const double*, double, void*, void*,
                       hipStream_t);
int boundary_load(int, int64_t, const double*, const double*, int, const double*, double*, hipStream_t);
int scatter(int, int64_t, const int*, const int*, const void*, void*, hipStream_t);
int cast_f64_f32(int64_t, const double*, float*, hipStream_t);
int spmv(int, int, int, int, const int*, const int*, const void*, const unsigned char*, const double*, double*,
         void*, int, hipStream_t);
size_t spmv_scratch_bytes(int, int);
int bj_setup(int, int, int, const int*, const int*, const double*, const unsigned char*, void*, hipStream_t);
size_t pcg_workspace_bytes(int, int, int);
int solve_pcg(int, int, int, int, const int*, const int*, const double*, const float*, const float*,
              const unsigned char*, const void*, const double*, double*, void*, double, double, double, int, int,
              int, int, hipStream_t, double*, void*);
int split_f32(int, int, const int*, const double*, float*, float*, hipStream_t);
int strain_stress(int, int64_t, const int*, const double*, const double*, int, const double*, const double*, int,
                  double, int, double, double*, hipStream_t);
int node_average(int, int, const int*, const int*, const double*, double*, hipStream_t);
int step1_bench(int, int, int, int, const int*, const int*, const void*, const float*, const unsigned char*,
                const double*, double*, void*, int, hipStream_t);
int pcg2_step1_bench(int, int, int, int, const int*, const int*, const void*, const float*, const unsigned char*, int,
                     const double*, double*, void*, int, hipStream_t);
int dot(int64_t, const double*, const double*, double*, double*, hipStream_t);
int gather_sum(int64_t, const int*, const double*, double*, hipStream_t);
int default_group(int, int, int);
int pf_damage(int, int64_t, const int*, const double*, const double*, int, const double*, int, const double*,
              const double*, const double*, int, double, int, const double*, int, double, int, int, double,
              double*, double*, double*, hipStream_t);
int pf_elastic(int, int64_t, const int*, const double*, const double*, int, const double*, const double*,
               const double*, int, double, int, int, double, double*, hipStream_t);
int pf_strain_stress(int, int64_t, const int*, const double*, const double*, int, const double*, const double*,
                     const double*, int, double, int, int, double, double*, hipStream_t);
int pf_split_points(int64_t, int, const double*, const double*, int, double, int, double*, double*, double*,
                    double*, hipStream_t);
int uv_ke(int, int, int64_t, const int*, const double*, const double*, int, const double*, int, double, int, double,
          void*, hipStream_t);
int lincomb(int64_t, double, const double*, double, const double*, double, const double*, double, const double*,
            double*, hipStream_t);
int time_update(int, int64_t, double, double, double, double, const double*, const double*, const double*,
                const double*, double*, double*, double*, hipStream_t);
int zz1_energy(int, int64_t, const int*, const double*, const double*, int, const double*, const double*, double,
               double*, hipStream_t);
size_t bounded_lsq_workspace_bytes(int);
int solve_bounded_lsq(int, const int*, const int*, const double*, const double*, const double*, const double*,
                      double*, void*, double, double, int, int, double*, hipStream_t);
int batch_create(int, int, int, int, int, void**);
int batch_destroy(void*);
int batch_add(void*, int, const int*, const int*, const double*, const float*, const float*, const unsigned char*,
              const void*, int, const int*, const int*, const double*, const double*, int, const double*, double*,
              void*, size_t, double, double, double, int, int, hipStream_t);
int batch_run(void*, int, hipStream_t);
int batch_launch(void*, int, hipStream_t);
int batch_wait(void*, hipStream_t);
int batch_pop(void*, int*, double*);
int batch_active(void*);
int coarse_ld(int, int);
size_t coarse_scratch_doubles(int);
int coarse_setup(int, int, const int*, const int*, const int*, const int*, const int*, const int*, const int*,
                 const double*, const unsigned char*, const double*, int, double, double*, double*, hipStream_t);
size_t pcg2_workspace_bytes(int, int, int);
int solve_pcg2(int, int, int, int, const int*, const int*, const double*, const float*, const float*,
               const unsigned char*, const void*, int, const int*, const int*, const double*, const double*, int,
               const double*, double*, void*, double, double, double, int, int, int, int, hipStream_t, double*);

size_t pcg2p_workspace_bytes(int, int, int);
size_t pcgp_workspace_bytes(int, int);
size_t bicgstab_workspace_bytes(int, int);
int solve_bicgstab(int, int, int, const int*, const int*, const double*, const unsigned char*, const double*,
                   const double*, double*, void*, double, int, int, hipStream_t, double*);
int solve_pcgp(int, int, int, int, const int*, const int*, const double*, const float*, const float*,
               const unsigned char*, const void*, const double*, double*, void*, double, double, int, int, int, int,
               hipStream_t, double*, void*, int, int*);
int solve_pcg2p(int, int, int, int, const int*, const int*, const double*, const float*, const float*,
                const unsigned char*, const void*, int, const int*, const int*, const double*, const double*, int, int,
                const int*, const int*, int, const int*, const double*, double*, void*, double, double, int, int, int,
                hipStream_t, double*, void*, int, int*);
int pp_iter_bench(int, int, int, int, const int*, const int*, const double*, const float*, const float*,
                  const unsigned char*, int, const double*, int, const int*, const int*, int, double*, void*, int,
                  hipStream_t);

int elastic_ke_dense(int, int64_t, const int*, const double*, const double*, int, const double*, int, const double*,
                     int, double, double*, hipStream_t);
}  // namespace efea

using namespace efea;

#define STREAM(s) ((hipStream_t)(s))

// Scratch of efea_solve_pcg_noalloc: the PCG workspace (vectors, partials,
// state) in a file-scope device array, one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: it is a host-side stand-in there.  Its device address, per device,
// comes from hipGetSymbolAddress at the entry point's first call on that
// device (with the device current) and is kept in noalloc_ptr.  Nothing in
// it has to start at zero: the solve's start kernels write the state, the
// partials and the vectors before any read.  Calls that overlap on one GPU
// share the array and race.
static __device__ __align__(256) unsigned char scratch[EFEA_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device.
static int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("efea_solve_pcg_noalloc: no current device (or device id %d above %d)", dev, kMaxDevices - 1);
    return EFEA_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("efea_solve_pcg_noalloc: hipGetSymbolAddress on device %d: %s", dev, hipGetErrorString(e));
      return EFEA_ERR_HIP;
    }
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return EFEA_OK;
}

extern "C" {

int efea_version(void) { return EFEA_ABI_VERSION; }

const char* efea_last_error(void) { return g_err; }

int efea_elastic_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect, const double* coord,
                    const double* tables, int nPg, const double* E, int E_stride, double nu, int plane_stress,
                    double thickness, void* Ke, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables || !E || !Ke))) {
    set_error("efea_elastic_ke: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return elastic_ke(elem_type, precision, Ne, connect, coord, tables, nPg, E, E_stride, nu, plane_stress,
                    thickness, Ke, STREAM(stream));
}

int efea_elastic_ke_dense(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                          const double* tables, int nPg, const double* D, int D_stride, const double* E, int E_stride,
                          double nu, double* Ke, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables || !Ke || (!D && !E)))) {
    set_error("efea_elastic_ke_dense: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return elastic_ke_dense(elem_type, Ne, connect, coord, tables, nPg, D, D_stride, E, E_stride, nu, Ke,
                          STREAM(stream));
}

int efea_elastic_strain_stress(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                               const double* tables, int nPg, const double* u, const double* E, int E_stride,
                               double nu, int plane_stress, double thickness, double* out, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables || !u || !E || !out))) {
    set_error("efea_elastic_strain_stress: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return strain_stress(elem_type, Ne, connect, coord, tables, nPg, u, E, E_stride, nu, plane_stress, thickness, out,
                       STREAM(stream));
}

int efea_node_average(int Nn, int ncol, const int32_t* node_ptr, const int32_t* node_elems, const double* values_e,
                      double* values_n, void* stream) {
  if (Nn < 0 || ncol < 1 || (Nn > 0 && (!node_ptr || !node_elems || !values_e || !values_n))) {
    set_error("efea_node_average: null pointer, negative size or ncol < 1");
    return EFEA_ERR_ARG;
  }
  return node_average(Nn, ncol, node_ptr, node_elems, values_e, values_n, STREAM(stream));
}

int efea_reaction_diffusion_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect,
                               const double* coord, const double* tables_mass, int nPg_mass,
                               const double* tables_rigi, int nPg_rigi, const double* r, double r_const,
                               const double* k, int k_stride, const double* f, double scale, void* Ke, void* Fe,
                               void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables_mass || !tables_rigi || !Ke || (Fe && !f)))) {
    set_error("efea_reaction_diffusion_ke: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return reaction_diffusion(elem_type, precision, Ne, connect, coord, tables_mass, nPg_mass, tables_rigi,
                            nPg_rigi, r, r_const, k, k_stride, f, scale, Ke, Fe, STREAM(stream));
}

int efea_boundary_load(int elem_type, int64_t Ne, const double* X_local, const double* tables, int nPg,
                       const double* f, double* out, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!X_local || !tables || !f || !out))) {
    set_error("efea_boundary_load: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return boundary_load(elem_type, Ne, X_local, tables, nPg, f, out, STREAM(stream));
}

int efea_csr_scatter(int precision, int64_t nslot, const int32_t* seg_ptr, const int32_t* perm,
                     const void* entries, void* out_values, void* stream) {
  if (nslot < 0 || (nslot > 0 && (!seg_ptr || !perm || !entries || !out_values))) {
    set_error("efea_csr_scatter: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return scatter(precision, nslot, seg_ptr, perm, entries, out_values, STREAM(stream));
}

int efea_split_f32(int Nnode, int dof_n, const int32_t* node_ptr, const double* values64, float* values32,
                   float* corr32, void* stream) {
  if (Nnode < 0 || (Nnode > 0 && (!node_ptr || !values64 || !values32 || !corr32))) {
    set_error("efea_split_f32: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return split_f32(Nnode, dof_n, node_ptr, values64, values32, corr32, STREAM(stream));
}

int efea_cast_f64_f32(int64_t n, const double* src, float* dst, void* stream) {
  if (n < 0 || (n > 0 && (!src || !dst))) {
    set_error("efea_cast_f64_f32: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return cast_f64_f32(n, src, dst, STREAM(stream));
}

size_t efea_spmv_scratch_bytes(int Nnode, int dof_n) { return spmv_scratch_bytes(Nnode, dof_n); }

int efea_spmv(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
              const void* values, const uint8_t* known, const double* x, double* y, void* scratch, void* stream) {
  if (Nnode < 0 || (Nnode > 0 && (!node_ptr || !node_cols || !values || !x || !y || !scratch))) {
    set_error("efea_spmv: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return spmv(precision, Nnode, dof_n, group, node_ptr, node_cols, values, known, x, y, scratch, 1, STREAM(stream));
}

int efea_spmv_repeat(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const void* values, const uint8_t* known, const double* x, double* y,
                     void* scratch, int reps, void* stream) {
  if (Nnode < 0 || reps < 1 || (Nnode > 0 && (!node_ptr || !node_cols || !values || !x || !y || !scratch))) {
    set_error("efea_spmv_repeat: null pointer, negative size or reps < 1");
    return EFEA_ERR_ARG;
  }
  return spmv(precision, Nnode, dof_n, group, node_ptr, node_cols, values, known, x, y, scratch, reps,
              STREAM(stream));
}

size_t efea_step1_bench_scratch_bytes(int Nnode, int dof_n, int nagg) {
  if (Nnode < 0 || dof_n < 1 || nagg < 0) return 0;
  return step1_bench_bytes(Nnode, dof_n, nagg > MAXB ? nagg : MAXB);
}

int efea_step1_bench(int two_level, int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const void* values, const float* corr32, const uint8_t* known,
                     int nagg, const double* z, double* q, void* scratch, int reps, void* stream) {
  if (Nnode <= 0 || reps < 1 || !node_ptr || !node_cols || !values || !z || !q || !scratch ||
      (two_level && nagg <= 0)) {
    set_error("efea_step1_bench: null pointer, empty system, reps < 1 or nagg <= 0");
    return EFEA_ERR_ARG;
  }
  if (two_level)
    return pcg2_step1_bench(precision, Nnode, dof_n, group, node_ptr, node_cols, values, corr32, known, nagg, z, q,
                            scratch, reps, STREAM(stream));
  return step1_bench(precision, Nnode, dof_n, group, node_ptr, node_cols, values, corr32, known, z, q, scratch,
                     reps, STREAM(stream));
}

int efea_bj_setup(int precision, int Nnode, int dof_n, const int32_t* node_ptr, const int32_t* node_cols,
                  const double* values, const uint8_t* known, void* Minv, void* stream) {
  if (Nnode < 0 || (Nnode > 0 && (!node_ptr || !node_cols || !values || !Minv))) {
    set_error("efea_bj_setup: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return bj_setup(precision, Nnode, dof_n, node_ptr, node_cols, values, known, Minv, STREAM(stream));
}

size_t efea_pcg_workspace_bytes(int Nnode, int dof_n, int group) {
  return pcg_workspace_bytes(Nnode, dof_n, group);
}

int efea_default_group(int dof_n, int max_neighbours, int Nnode) {
  return default_group(dof_n, max_neighbours, Nnode);
}

int efea_solve_pcg(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                   const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                   const void* Minv,
                   const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                   double tol_anorm, double replace_ratio, int maxit, int check_every, int first_check,
                   int max_blocks, void* stream, double* info) {
  if (Nnode <= 0 || !node_ptr || !node_cols || !values64 || !Minv || !b || !x || !workspace || !info) {
    set_error("efea_solve_pcg: null pointer or empty system");
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < pcg_workspace_bytes(Nnode, dof_n, group)) {
    set_error("efea_solve_pcg: workspace of %zu bytes, %zu needed", workspace_bytes,
              pcg_workspace_bytes(Nnode, dof_n, group));
    return EFEA_ERR_SIZE;
  }
  return solve_pcg(mode, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, b, x,
                   workspace, tol_residual, tol_anorm, replace_ratio, maxit, check_every, first_check, max_blocks,
                   STREAM(stream), info, nullptr);
}

int efea_solve_pcg_noalloc(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                           const int32_t* node_cols, const double* values64, const float* values32,
                           const float* corr32, const uint8_t* known, const void* Minv, const double* b, double* x,
                           double tol_residual, double tol_anorm, double replace_ratio, int maxit, int check_every,
                           int first_check, int max_blocks, void* stream, double* info) {
  if (Nnode <= 0 || !node_ptr || !node_cols || !values64 || !Minv || !b || !x || !info) {
    set_error("efea_solve_pcg_noalloc: null pointer or empty system");
    return EFEA_ERR_ARG;
  }
  const size_t need = pcg_workspace_bytes(Nnode, dof_n, group);
  if (need > (size_t)EFEA_NOALLOC_SCRATCH_BYTES) {
    set_error("efea_solve_pcg_noalloc: %d nodes x %d DOF need %zu bytes of scratch, more than the %zu of the "
              "library's static scratch; use efea_solve_pcg with a workspace", Nnode, dof_n, need,
              (size_t)EFEA_NOALLOC_SCRATCH_BYTES);
    return EFEA_ERR_SIZE;
  }
  void* ws = nullptr;
  if (int e = noalloc_scratch(&ws)) return e;
  // the state read-backs land in pageable memory on this stack: no pinned allocation
  PcgState hs;
  return solve_pcg(mode, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, b, x, ws,
                   tol_residual, tol_anorm, replace_ratio, maxit, check_every, first_check, max_blocks,
                   STREAM(stream), info, &hs);
}

size_t efea_pcgp_workspace_bytes(int Nnode, int dof_n) { return pcgp_workspace_bytes(Nnode, dof_n); }

// Argument checks shared by efea_solve_pcgp and its deferred form.
static int pcgp_check(const char* fn, int precision, int Nnode, int group, const int32_t* node_ptr,
                      const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                      const void* Minv, const double* b, double* x, void* workspace, size_t workspace_bytes,
                      int dof_n) {
  if (Nnode <= 0 || !node_ptr || !node_cols || !values64 || !Minv || !b || !x || !workspace) {
    set_error("%s: null pointer or empty system", fn);
    return EFEA_ERR_ARG;
  }
  if ((precision != 64 && precision != 32) || (precision == 32 && (!values32 || !corr32))) {
    set_error("%s: precision %d (64, or 32 with FP32 values and their row-sum correction)", fn, precision);
    return EFEA_ERR_ARG;
  }
  if (group <= 0 || (int64_t)Nnode * group > (int64_t)efea_pcgp_max_lanes()) {
    set_error("%s: %d nodes x %d lanes exceed %d lanes (one pass of the grid)", fn, Nnode, group,
              efea_pcgp_max_lanes());
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < pcgp_workspace_bytes(Nnode, dof_n)) {
    set_error("%s: workspace of %zu bytes, %zu needed", fn, workspace_bytes, pcgp_workspace_bytes(Nnode, dof_n));
    return EFEA_ERR_SIZE;
  }
  return EFEA_OK;
}

int efea_solve_pcgp(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                    const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                    const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                    size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                    int check_every, int first_check, void* stream, double* info) {
  if (!info) {
    set_error("efea_solve_pcgp: null info");
    return EFEA_ERR_ARG;
  }
  if (int e = pcgp_check("efea_solve_pcgp", precision, Nnode, group, node_ptr, node_cols, values64, values32, corr32, Minv,
                         b, x, workspace, workspace_bytes, dof_n))
    return e;
  return solve_pcgp(precision, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, b,
                    x, workspace, tol_residual, tol_anorm, replace_every, maxit, check_every, first_check,
                    STREAM(stream), info, nullptr, PCGP_WHOLE, nullptr);
}

int efea_pcgp_start(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                    const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                    const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                    size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                    int check_every, int first_check, void* stream) {
  if (int e = pcgp_check("efea_pcgp_start", precision, Nnode, group, node_ptr, node_cols, values64, values32, corr32, Minv,
                         b, x, workspace, workspace_bytes, dof_n))
    return e;
  if (first_check < 1) {
    set_error("efea_pcgp_start: first_check %d (the iterations to queue, >= 1)", first_check);
    return EFEA_ERR_ARG;
  }
  return solve_pcgp(precision, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, b,
                    x, workspace, tol_residual, tol_anorm, replace_every, maxit, check_every, first_check,
                    STREAM(stream), nullptr, nullptr, PCGP_START, nullptr);
}

int efea_pcgp_finish(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                     const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                     size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                     int check_every, int first_check, void* stream, double* info, int* resumed) {
  if (!info) {
    set_error("efea_pcgp_finish: null info");
    return EFEA_ERR_ARG;
  }
  if (int e = pcgp_check("efea_pcgp_finish", precision, Nnode, group, node_ptr, node_cols, values64, values32, corr32, Minv,
                         b, x, workspace, workspace_bytes, dof_n))
    return e;
  if (first_check < 1) {
    set_error("efea_pcgp_finish: first_check %d (the start's, >= 1)", first_check);
    return EFEA_ERR_ARG;
  }
  if (check_every < 1) check_every = 1;
  return solve_pcgp(precision, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, b,
                    x, workspace, tol_residual, tol_anorm, replace_every, maxit, check_every, first_check,
                    STREAM(stream), info, nullptr, PCGP_FINISH, resumed);
}

int efea_pcgp_max_lanes(void) { return 1024 * 256; }

size_t efea_bicgstab_workspace_bytes(int Nnode, int dof_n) { return bicgstab_workspace_bytes(Nnode, dof_n); }

int efea_solve_bicgstab(int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                        const double* values, const uint8_t* known, const double* Minv, const double* b, double* x,
                        void* workspace, size_t workspace_bytes, double tol_residual, int maxit, int check_every,
                        void* stream, double* info) {
  if (Nnode <= 0 || !node_ptr || !node_cols || !values || !Minv || !b || !x || !workspace || !info) {
    set_error("efea_solve_bicgstab: null pointer or empty system");
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < bicgstab_workspace_bytes(Nnode, dof_n)) {
    set_error("efea_solve_bicgstab: workspace of %zu bytes, %zu needed", workspace_bytes,
              bicgstab_workspace_bytes(Nnode, dof_n));
    return EFEA_ERR_SIZE;
  }
  return solve_bicgstab(Nnode, dof_n, group, node_ptr, node_cols, values, known, Minv, b, x, workspace,
                        tol_residual, maxit, check_every, STREAM(stream), info);
}

size_t efea_pcg2p_workspace_bytes(int Nnode, int dof_n, int nslot) {
  return pcg2p_workspace_bytes(Nnode, dof_n, nslot);
}

int efea_pp_iter_bench(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                       const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                       const uint8_t* known, int nagg, const double* rel, int nchunk, const int32_t* chunk_rows,
                       const int32_t* chunk_slot, int chunks_per_agg, double* x, void* workspace,
                       size_t workspace_bytes, int reps, void* stream) {
  if (Nnode <= 0 || nagg <= 0 || nchunk <= 0 || nchunk > 1024 || chunks_per_agg <= 0 || reps < 1 || !node_ptr ||
      !node_cols || !rel || !chunk_rows || !chunk_slot || !x ||
      !workspace || (precision == 64 && !values64) || (precision == 32 && (!values32 || !corr32)) ||
      (precision != 64 && precision != 32)) {
    set_error("efea_pp_iter_bench: null pointer, empty system, bad precision, reps < 1 or over 1024 chunks");
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < pcg2p_workspace_bytes(Nnode, dof_n, nagg * chunks_per_agg)) {
    set_error("efea_pp_iter_bench: workspace of %zu bytes, %zu needed", workspace_bytes,
              pcg2p_workspace_bytes(Nnode, dof_n, nagg * chunks_per_agg));
    return EFEA_ERR_SIZE;
  }
  return pp_iter_bench(precision, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, nagg,
                       rel, nchunk, chunk_rows, chunk_slot, chunks_per_agg, x, workspace, reps, STREAM(stream));
}

// Argument checks shared by efea_solve_pcg2p and its deferred form.
static int pcg2p_check(const char* fn, int precision, int Nnode, int dof_n, const int32_t* node_ptr,
                       const int32_t* node_cols, const double* values64, const float* values32,
                       const float* corr32, const void* Minv,
                       int nagg, const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel,
                       const double* Ainv, int ld, int nchunk, const int32_t* chunk_rows, const int32_t* chunk_slot,
                       int chunks_per_agg, const int32_t* agg_first, const double* b, double* x, void* workspace,
                       size_t workspace_bytes) {
  if (Nnode <= 0 || nagg <= 0 || nchunk <= 0 || chunks_per_agg <= 0 || !node_ptr || !node_cols || !values64 ||
      !Minv || !agg_ptr || !agg_nodes || !rel || !Ainv || !chunk_rows || !chunk_slot || !agg_first || !b || !x ||
      !workspace) {
    set_error("%s: null pointer or empty system", fn);
    return EFEA_ERR_ARG;
  }
  if ((precision != 64 && precision != 32) || (precision == 32 && (!values32 || !corr32))) {
    set_error("%s: precision %d (64, or 32 with FP32 values and their row-sum correction)", fn, precision);
    return EFEA_ERR_ARG;
  }
  if (nchunk > 1024) {
    set_error("%s: %d node chunks, at most 1024 (one block each)", fn, nchunk);
    return EFEA_ERR_ARG;
  }
  if (ld < coarse_ld(nagg, dof_n)) {
    set_error("%s: leading dimension %d < %d", fn, ld, coarse_ld(nagg, dof_n));
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < pcg2p_workspace_bytes(Nnode, dof_n, nagg * chunks_per_agg)) {
    set_error("%s: workspace of %zu bytes, %zu needed", fn, workspace_bytes,
              pcg2p_workspace_bytes(Nnode, dof_n, nagg * chunks_per_agg));
    return EFEA_ERR_SIZE;
  }
  return EFEA_OK;
}

// The three pipelined two-level entry points share their argument list up
// to the stream (PCG2P_ARGS) and differ in the phase they run.
#define PCG2P_PARAMS                                                                                               \
  int precision, int Nnode, int dof_n, int group, const int32_t *node_ptr, const int32_t *node_cols,              \
      const double *values64, const float *values32, const float *corr32, const uint8_t *known, const void *Minv, \
      int nagg, const int32_t *agg_ptr, const int32_t *agg_nodes, const double *rel, const double *Ainv, int ld,  \
      int nchunk, const int32_t *chunk_rows, const int32_t *chunk_slot, int chunks_per_agg,                      \
      const int32_t *agg_first, const double *b, double *x, void *workspace, size_t workspace_bytes,             \
      double tol_residual, double tol_anorm, int maxit, int check_every, int first_check, void *stream
#define PCG2P_CHECK(fn)                                                                                           \
  pcg2p_check(fn, precision, Nnode, dof_n, node_ptr, node_cols, values64, values32, corr32, Minv, nagg, agg_ptr,  \
              agg_nodes, rel, Ainv, ld, nchunk, chunk_rows, chunk_slot, chunks_per_agg, agg_first, b, x,          \
              workspace, workspace_bytes)
#define PCG2P_RUN(info, phase, resumed)                                                                           \
  solve_pcg2p(precision, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, nagg, \
              agg_ptr, agg_nodes, rel, Ainv, ld, nchunk, chunk_rows, chunk_slot, chunks_per_agg, agg_first, b, x, \
              workspace, tol_residual, tol_anorm, maxit, check_every, first_check, STREAM(stream), info, nullptr, \
              phase, resumed)

int efea_solve_pcg2p(PCG2P_PARAMS, double* info) {
  if (!info) {
    set_error("efea_solve_pcg2p: null info");
    return EFEA_ERR_ARG;
  }
  if (int e = PCG2P_CHECK("efea_solve_pcg2p")) return e;
  return PCG2P_RUN(info, PCGP_WHOLE, nullptr);
}

int efea_pcg2p_start(PCG2P_PARAMS) {
  if (int e = PCG2P_CHECK("efea_pcg2p_start")) return e;
  if (first_check < 1) {
    set_error("efea_pcg2p_start: first_check %d (the iterations to queue, >= 1)", first_check);
    return EFEA_ERR_ARG;
  }
  return PCG2P_RUN(nullptr, PCGP_START, nullptr);
}

int efea_pcg2p_finish(PCG2P_PARAMS, double* info, int* resumed) {
  if (!info) {
    set_error("efea_pcg2p_finish: null info");
    return EFEA_ERR_ARG;
  }
  if (int e = PCG2P_CHECK("efea_pcg2p_finish")) return e;
  if (first_check < 1) {
    set_error("efea_pcg2p_finish: first_check %d (the start's, >= 1)", first_check);
    return EFEA_ERR_ARG;
  }
  return PCG2P_RUN(info, PCGP_FINISH, resumed);
}

#undef PCG2P_RUN
#undef PCG2P_CHECK
#undef PCG2P_PARAMS

int efea_batch_create(int mode, int dof_n, int group, int two_level, int capacity, void** handle) {
  if (!handle) {
    set_error("efea_batch_create: null handle pointer");
    return EFEA_ERR_ARG;
  }
  *handle = nullptr;
  return batch_create(mode, dof_n, group, two_level, capacity, handle);
}

int efea_batch_destroy(void* handle) {
  if (!handle) return EFEA_OK;
  return batch_destroy(handle);
}

int efea_batch_add(void* handle, int Nnode, const int32_t* node_ptr, const int32_t* node_cols,
                   const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                   const void* Minv, int nagg,
                   const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                   const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                   double tol_anorm, double replace_ratio, int maxit, int max_blocks, void* stream) {
  if (!handle || Nnode <= 0 || !node_ptr || !node_cols || !values64 || !Minv || !b || !x || !workspace) {
    set_error("efea_batch_add: null pointer or empty system");
    return EFEA_ERR_ARG;
  }
  if (maxit < 1) {
    set_error("efea_batch_add: maxit %d", maxit);
    return EFEA_ERR_ARG;
  }
  return batch_add(handle, Nnode, node_ptr, node_cols, values64, values32, corr32, known, Minv, nagg, agg_ptr,
                   agg_nodes,
                   rel, Ainv, ld, b, x, workspace, workspace_bytes, tol_residual, tol_anorm, replace_ratio, maxit,
                   max_blocks, STREAM(stream));
}

int efea_batch_run(void* handle, int iterations, void* stream) {
  if (!handle) {
    set_error("efea_batch_run: null handle");
    return EFEA_ERR_ARG;
  }
  return batch_run(handle, iterations, STREAM(stream));
}

int efea_batch_launch(void* handle, int iterations, void* stream) {
  if (!handle) {
    set_error("efea_batch_launch: null handle");
    return EFEA_ERR_ARG;
  }
  return batch_launch(handle, iterations, STREAM(stream));
}

int efea_batch_wait(void* handle, void* stream) {
  if (!handle) {
    set_error("efea_batch_wait: null handle");
    return EFEA_ERR_ARG;
  }
  return batch_wait(handle, STREAM(stream));
}

int efea_batch_pop(void* handle, int* id, double* info) {
  if (!handle || !id || !info) {
    set_error("efea_batch_pop: null pointer");
    return EFEA_ERR_ARG;
  }
  return batch_pop(handle, id, info);
}

int efea_batch_active(void* handle) {
  if (!handle) {
    set_error("efea_batch_active: null handle");
    return EFEA_ERR_ARG;
  }
  return batch_active(handle);
}

int efea_coarse_ld(int nagg, int dof_n) {
  if (nagg < 1 || dof_n < 1 || dof_n > 3) {
    set_error("efea_coarse_ld: nagg %d / dof_n %d", nagg, dof_n);
    return EFEA_ERR_ARG;
  }
  return coarse_ld(nagg, dof_n);
}

size_t efea_coarse_scratch_bytes(int ld) { return ld > 0 ? 8 * coarse_scratch_doubles(ld) : 0; }

int efea_coarse_setup(int dof_n, int nblocks, const int32_t* pair_ptr, const int32_t* pair_node,
                      const int32_t* pair_pos, const int32_t* blk_I, const int32_t* blk_J, const int32_t* node_ptr,
                      const int32_t* node_cols, const double* values64, const uint8_t* known, const double* rel,
                      int ld, double diag_shift, double* Ainv, double* scratch, void* stream) {
  if (nblocks < 0 || ld <= 0 || !Ainv || !scratch || !(diag_shift >= 0.0) ||
      (nblocks > 0 && (!pair_ptr || !pair_node || !pair_pos || !blk_I || !blk_J || !node_ptr || !node_cols ||
                       !values64 || !rel))) {
    set_error("efea_coarse_setup: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return coarse_setup(dof_n, nblocks, pair_ptr, pair_node, pair_pos, blk_I, blk_J, node_ptr, node_cols, values64,
                      known, rel, ld, diag_shift, Ainv, scratch, STREAM(stream));
}

size_t efea_pcg2_workspace_bytes(int Nnode, int dof_n, int nagg) { return pcg2_workspace_bytes(Nnode, dof_n, nagg); }

int efea_solve_pcg2(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                    const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                    const void* Minv, int nagg,
                    const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                    const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                    double tol_anorm, double replace_ratio, int maxit, int check_every, int first_check,
                    int max_blocks, void* stream, double* info) {
  if (Nnode <= 0 || nagg <= 0 || !node_ptr || !node_cols || !values64 || !Minv || !agg_ptr || !agg_nodes || !rel ||
      !Ainv || !b || !x || !workspace || !info) {
    set_error("efea_solve_pcg2: null pointer or empty system");
    return EFEA_ERR_ARG;
  }
  if (ld < coarse_ld(nagg, dof_n)) {
    set_error("efea_solve_pcg2: leading dimension %d < %d", ld, coarse_ld(nagg, dof_n));
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < pcg2_workspace_bytes(Nnode, dof_n, nagg)) {
    set_error("efea_solve_pcg2: workspace of %zu bytes, %zu needed", workspace_bytes,
              pcg2_workspace_bytes(Nnode, dof_n, nagg));
    return EFEA_ERR_SIZE;
  }
  return solve_pcg2(mode, Nnode, dof_n, group, node_ptr, node_cols, values64, values32, corr32, known, Minv, nagg,
                    agg_ptr,
                    agg_nodes, rel, Ainv, ld, b, x, workspace, tol_residual, tol_anorm, replace_ratio, maxit,
                    check_every, first_check, max_blocks, STREAM(stream), info);
}

int efea_dot(int64_t n, const double* a, const double* b, double* scratch, double* out, void* stream) {
  if (n < 0 || !a || !scratch || !out) {
    set_error("efea_dot: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return dot(n, a, b, scratch, out, STREAM(stream));
}

int efea_gather_sum(int64_t n, const int32_t* idx, const double* y, double* out, void* stream) {
  if (n < 0 || (n > 0 && (!idx || !y)) || !out) {
    set_error("efea_gather_sum: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return gather_sum(n, idx, y, out, STREAM(stream));
}

int efea_pf_damage_ke(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                      const double* tables_mass, int nPg_mass, const double* tables_rigi, int nPg_rigi,
                      const double* u, const double* H_old, const double* E, int E_stride, double nu,
                      int plane_stress, const double* Gc, int Gc_stride, double l0, int split, int regu,
                      double thickness, double* psiP, double* Kd, double* Fd, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables_mass || !tables_rigi || !u || !E || !Gc || !psiP ||
                            !Kd || !Fd))) {
    set_error("efea_pf_damage_ke: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (split < EFEA_SPLIT_AMOR || split > EFEA_SPLIT_ANISOT_STRESS_MP || regu < 0 || regu > 1) {
    set_error("efea_pf_damage_ke: split %d / regularisation %d not supported", split, regu);
    return EFEA_ERR_ARG;
  }
  return pf_damage(elem_type, Ne, connect, coord, tables_mass, nPg_mass, tables_rigi, nPg_rigi, u, H_old, E,
                   E_stride, nu, plane_stress, Gc, Gc_stride, l0, split, regu, thickness, psiP, Kd, Fd,
                   STREAM(stream));
}

int efea_pf_strain_stress(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                          const double* tables_rigi, int nPg_rigi, const double* u, const double* d, const double* E,
                          int E_stride, double nu, int plane_stress, int split, double thickness, double* out,
                          void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables_rigi || !u || !d || !E || !out))) {
    set_error("efea_pf_strain_stress: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (split < EFEA_SPLIT_AMOR || split > EFEA_SPLIT_ANISOT_STRESS_MP) {
    set_error("efea_pf_strain_stress: split %d not supported", split);
    return EFEA_ERR_ARG;
  }
  return pf_strain_stress(elem_type, Ne, connect, coord, tables_rigi, nPg_rigi, u, d, E, E_stride, nu, plane_stress,
                          split, thickness, out, STREAM(stream));
}

int efea_pf_elastic_ke(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                       const double* tables_rigi, int nPg_rigi, const double* u, const double* d,
                       const double* E, int E_stride, double nu, int plane_stress, int split, double thickness,
                       double* Ke, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables_rigi || !u || !d || !E || !Ke))) {
    set_error("efea_pf_elastic_ke: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (split < EFEA_SPLIT_AMOR || split > EFEA_SPLIT_ANISOT_STRESS_MP) {
    set_error("efea_pf_elastic_ke: split %d not supported", split);
    return EFEA_ERR_ARG;
  }
  return pf_elastic(elem_type, Ne, connect, coord, tables_rigi, nPg_rigi, u, d, E, E_stride, nu, plane_stress,
                    split, thickness, Ke, STREAM(stream));
}

int efea_pf_split_points(int64_t n, int split, const double* eps, const double* E, int E_stride, double nu,
                         int plane_stress, double* cP, double* cM, double* psiP, double* psiM, void* stream) {
  if (n < 0 || (n > 0 && (!eps || !E || !cP || !cM || !psiP || !psiM))) {
    set_error("efea_pf_split_points: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (split < EFEA_SPLIT_AMOR || split > EFEA_SPLIT_ANISOT_STRESS_MP) {
    set_error("efea_pf_split_points: split %d not supported", split);
    return EFEA_ERR_ARG;
  }
  return pf_split_points(n, split, eps, E, E_stride, nu, plane_stress, cP, cM, psiP, psiM, STREAM(stream));
}

int efea_uv_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect, const double* coord,
               const double* tables_mass, int nPg_mass, const double* coef, int coef_stride, double coef_const,
               int dof_n, double scale, void* Ke, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables_mass || !Ke))) {
    set_error("efea_uv_ke: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return uv_ke(elem_type, precision, Ne, connect, coord, tables_mass, nPg_mass, coef, coef_stride, coef_const, dof_n,
               scale, Ke, STREAM(stream));
}

int efea_lincomb(int64_t n, double c0, const double* x0, double c1, const double* x1, double c2, const double* x2,
                 double c3, const double* x3, double* out, void* stream) {
  if (n < 0 || (n > 0 && (!x0 || !out))) {
    set_error("efea_lincomb: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return lincomb(n, c0, x0, c1, x1, c2, x2, c3, x3, out, STREAM(stream));
}

int efea_time_update(int scheme, int64_t n, double dt, double beta, double gamma, double alpha, const double* u_n,
                     const double* v_n, const double* a_n, const double* x, double* u_out, double* v_out,
                     double* a_out, void* stream) {
  if (n < 0 || (n > 0 && (!u_n || !v_n || !a_n || !x || !u_out))) {
    set_error("efea_time_update: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (scheme < EFEA_SCHEME_ELLIPTIC || scheme > EFEA_SCHEME_EULER_EXPLICIT) {
    set_error("efea_time_update: unknown scheme %d", scheme);
    return EFEA_ERR_ARG;
  }
  if (scheme != EFEA_SCHEME_ELLIPTIC && !(dt > 0)) {
    set_error("efea_time_update: dt must be > 0");
    return EFEA_ERR_ARG;
  }
  return time_update(scheme, n, dt, beta, gamma, alpha, u_n, v_n, a_n, x, u_out, v_out, a_out, STREAM(stream));
}

int efea_zz1_energy(int elem_type, int64_t Ne, const int32_t* connect, const double* coord, const double* tables,
                    int nPg, const double* u, const double* sig_n, double thickness, double* out, void* stream) {
  if (Ne < 0 || (Ne > 0 && (!connect || !coord || !tables || !u || !sig_n || !out))) {
    set_error("efea_zz1_energy: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  return zz1_energy(elem_type, Ne, connect, coord, tables, nPg, u, sig_n, thickness, out, STREAM(stream));
}

size_t efea_bounded_lsq_workspace_bytes(int N) { return bounded_lsq_workspace_bytes(N); }

int efea_solve_bounded_lsq(int N, const int32_t* row_ptr, const int32_t* cols, const double* values,
                           const double* b, const double* lb, const double* ub, double* x, void* workspace,
                           size_t workspace_bytes, double tol, double kkt_tol, int maxit, int max_outer,
                           double* info, void* stream) {
  if (N < 0 || (N > 0 && (!row_ptr || !cols || !values || !b || !lb || !ub || !x || !workspace))) {
    set_error("efea_solve_bounded_lsq: null pointer or negative size");
    return EFEA_ERR_ARG;
  }
  if (workspace_bytes < bounded_lsq_workspace_bytes(N)) {
    set_error("efea_solve_bounded_lsq: workspace of %zu bytes, %zu needed", workspace_bytes,
              bounded_lsq_workspace_bytes(N));
    return EFEA_ERR_SIZE;
  }
  if (!(tol > 0 && tol < 1) || !(kkt_tol > 0) || maxit < 1 || max_outer < 1) {
    set_error("efea_solve_bounded_lsq: need 0 < tol < 1, kkt_tol > 0, maxit >= 1, max_outer >= 1");
    return EFEA_ERR_ARG;
  }
  return solve_bounded_lsq(N, row_ptr, cols, values, b, lb, ub, x, workspace, tol, kkt_tol, maxit, max_outer, info,
                           STREAM(stream));
}

}  // extern "C"
