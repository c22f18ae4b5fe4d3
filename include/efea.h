/* efea: MI355X (gfx950) HIP library for EasyFEA's finite-element solve path.
 *
 * C ABI: plain pointers and sizes, no C++ or torch types.  Every pointer
 * argument except those documented as host pointers is a device pointer;
 * `stream` is a hipStream_t (NULL = default stream).  Every int function
 * returns 0 on success or a negative code (EFEA_ERR_*); efea_last_error()
 * then returns a thread-local message.  Calls are asynchronous on `stream`
 * unless stated otherwise, and reentrant per stream.
 *
 * Conventions shared with the reference (EasyFEA v3.5.1):
 *   dof = node * dof_n + component      (FEM/_boundary_conditions.py:160-201)
 *   element DOF order [n0x, n0y, n1x, ...]   (FEM/_group_elem.py:358-380)
 *   element matrices (Ne, nd, nd) row-major, the order of
 *   K_e.ravel() in Simulations/_simu.py:1037
 *   Kelvin-Mandel strains [xx, yy, sqrt2 xy] / [xx, yy, zz, sqrt2 yz, sqrt2 xz, sqrt2 xy]
 *   (FEM/_group_elem.py:1241-1312, Models/_utils.py:111-152)
 * Reference-element tables: one FP64 block per integration rule,
 *   [w (nPg)] [N (nPg*nPe)] [dN (nPg*dim*nPe)]  (FEM/_gauss.py:364-516 and the FEM/Elems modules)
 * Element codes: 0 TRI3, 1 QUAD4, 2 HEXA8, 3 TETRA4, 4 TETRA10, 5 SEG2.
 */
#ifndef EFEA_H
#define EFEA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFEA_ABI_VERSION 1

#define EFEA_ERR_ARG (-1)    /* invalid argument / unsupported combination */
#define EFEA_ERR_HIP (-2)    /* HIP runtime error */
#define EFEA_ERR_NOCONV (-3) /* iterative solve did not converge (info filled) */
#define EFEA_ERR_SIZE (-4)   /* workspace too small */

int efea_version(void);
const char* efea_last_error(void);

/* Isotropic linear-elastic element stiffness K_e (Ne, nd, nd), precision 64
 * or 32, Young's modulus E[e*E_stride] (E_stride 0: one value), Poisson nu,
 * plane stress flag (2D), thickness multiplies 2D matrices.
 * Replaces Operators.Bilinear.LinearizedElasticity (FEM/Operators/Bilinear.py:62-79)
 * called from Simulations.Elastic.Construct_local_matrix_system
 * (Simulations/_elastic.py:123-152) with Isotropic._Behavior (Models/Elastic/_laws.py:407-482). */
int efea_elastic_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect, const double* coord,
                    const double* tables, int nPg, const double* E, int E_stride, double nu, int plane_stress,
                    double thickness, void* Ke, void* stream);

/* Dense 3D element stiffness on the FP64 matrix cores,
 * K_e = sum_p wJ_p B_p^T D B_p, for HEXA8 (nPg 8) and TETRA10 (nPg 4): D is a
 * Kelvin-Mandel 6 x 6 (row-major; rows xx yy zz yz xz xy) per element
 * (D_stride 1) or shared (0); with D null, the isotropic matrix of E
 * (E_stride 0 / 1) and nu.  Output (Ne, nd, nd) FP64 as efea_elastic_ke.
 * Replaces Operators.Bilinear.LinearizedElasticity (FEM/Operators/
 * Bilinear.py:62-79, the einsum of :79) for constitutive matrices of any
 * form (Models/Elastic/_laws.py: Orthotropic, TransverselyIsotropic,
 * Anisotropic). */
int efea_elastic_ke_dense(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                          const double* tables, int nPg, const double* D, int D_stride, const double* E, int E_stride,
                          double nu, double* Ke, void* stream);

/* Linear-elastic post-processing per element (Simulations/_elastic.py:229-380,
 * Models/Elastic/_laws.py:127-185, Models/_utils.py:302-434): at each stiffness
 * integration point eps = B u_e (Kelvin-Mandel) and sig = C eps (isotropic,
 * E per element when E_stride = 1); out (Ne, 2 ns + 3) FP64 holds per element
 * [strain tensor components (ns, shear without the sqrt 2), Evm, stress
 * components (ns), Svm, W_e] - components and von Mises measures averaged over
 * the points with equal weights, W_e = 1/2 sum_p wJ sig.eps (x thickness in
 * 2D).  ns = 3 in 2D [xx, yy, xy], 6 in 3D [xx, yy, zz, yz, xz, xy]. */
int efea_elastic_strain_stress(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                               const double* tables, int nPg, const double* u, const double* E, int E_stride,
                               double nu, int plane_stress, double thickness, double* out, void* stream);
/* Element -> node values (Mesh.Get_Node_Values, FEM/_mesh.py:822-873): for node n,
 * the mean of values_e (Ne, ncol) over the elements listed in
 * node_elems[node_ptr[n] .. node_ptr[n+1]) (increasing element order); 0 for a
 * node of no element. */
int efea_node_average(int Nn, int ncol, const int32_t* node_ptr, const int32_t* node_elems, const double* values_e,
                      double* values_n, void* stream);

/* Scalar reaction-diffusion-source element operators:
 * K_e = sum_mass r wJ N^T N + k sum_rigi wJ dN^T dN, F_e = sum_mass f wJ N, times scale.
 * r: (Ne, nPg_mass) or NULL (then r_const); k: k[e*k_stride] or NULL (no diffusion);
 * Fe/f may be NULL.  Replaces Operators.Bilinear.UV (Bilinear.py:42-59),
 * Operators.Bilinear.GradU_A_GradV with A = I (Bilinear.py:229-249) and
 * Operators.Linear.V (Operators/Linear.py:18-35). */
int efea_reaction_diffusion_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect,
                               const double* coord, const double* tables_mass, int nPg_mass,
                               const double* tables_rigi, int nPg_rigi, const double* r, double r_const,
                               const double* k, int k_stride, const double* f, double scale, void* Ke, void* Fe,
                               void* stream);

/* Load integral out[e, a] = sum_p wJ f[e, p] N_a on coordinates given in each element's own
 * frame X_local (Ne, nPe, dim): line and surface loads on SEG2 (nPg 2), TRI3 (3), QUAD4 (4);
 * body loads (add_volumeLoad, _simu.py:2683-2741) also on TETRA4 (4), TETRA10 (4), HEXA8 (8).
 * Replaces the integration of _Simu.__Bc_Integration_Dim (Simulations/_simu.py:2770-2842). */
int efea_boundary_load(int elem_type, int64_t Ne, const double* X_local, const double* tables, int nPg,
                       const double* f, double* out, void* stream);

/* Deterministic CSR scatter: out[s] = sum_{k=seg_ptr[s]}^{seg_ptr[s+1]-1} entries[perm[k]],
 * accumulated sequentially from zero in increasing element-entry order, with no
 * atomics (bit-identical to np.bincount given identical entries).
 * Replaces _Simu.__Assemble_csr's bincount (Simulations/_simu.py:1051-1055)
 * over the map of _Simu.__Get_csr_map (_simu.py:1062-1102). */
int efea_csr_scatter(int precision, int64_t nslot, const int32_t* seg_ptr, const int32_t* perm,
                     const void* entries, void* out_values, void* stream);

int efea_cast_f64_f32(int64_t n, const double* src, float* dst, void* stream);

/* FP32 copy of the CSR values plus, per node, the dof_n x dof_n block row sums
 * of the rounding differences (K64 - K32) in FP32 (corr32: Nnode * dof_n^2).
 * Adding corr32 to the diagonal blocks keeps the rounded operator's row sums
 * equal to K's, so rigid translations stay in its null space; without it the
 * FP32 operator of a cracked body stalls the mixed solve near 1e-6 error. */
int efea_split_f32(int Nnode, int dof_n, const int32_t* node_ptr, const double* values64, float* values32,
                   float* corr32, void* stream);

/* y = K x on the canonical CSR value array indexed through node adjacency
 * (node_ptr, node_cols): rows of `known` DOFs (uint8 per DOF, may be NULL)
 * return 0.  precision selects FP64 or FP32 values; x, y are FP64.  Runs the
 * solver's own step-1 kernel (beta = 0); scratch: efea_spmv_scratch_bytes()
 * device bytes, no initialisation needed.
 * Replaces the scipy products K[dofs] @ u of _Simu.Calc_Reaction (_simu.py:212-260)
 * and Calc_Energy (_simu.py:177-210). */
size_t efea_spmv_scratch_bytes(int Nnode, int dof_n);
int efea_spmv(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
              const void* values, const uint8_t* known, const double* x, double* y, void* scratch, void* stream);
/* The same product launched `reps` times back to back (kernel timing). */
int efea_spmv_repeat(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const void* values, const uint8_t* known, const double* x, double* y,
                     void* scratch, int reps, void* stream);

/* Timing harness of the solver's dominant kernel (no counterpart in the
 * reference; for benchmarks and profiles): `reps` back-to-back launches of the
 * step-1 kernel (p = z + beta p; q = A p, masked; partials of p.q) exactly as a
 * solve launches it - k_pcg2_step1 with the r.z partials of `nagg` aggregate
 * blocks when two_level, k_pcg_step1 with those of the node grid otherwise -
 * in the state of a running solve whose stopping tests never fire.  z: input
 * (Nnode*dof_n doubles); q: output; corr32 as in efea_solve_pcg (mode 32);
 * scratch: efea_step1_bench_scratch_bytes(Nnode, dof_n, nagg) bytes. */
size_t efea_step1_bench_scratch_bytes(int Nnode, int dof_n, int nagg);
int efea_step1_bench(int two_level, int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const void* values, const float* corr32, const uint8_t* known,
                     int nagg, const double* z, double* q, void* scratch, int reps, void* stream);

/* Block-Jacobi setup: inverse of each node's dof_n x dof_n diagonal block of the
 * Dirichlet-reduced matrix, stored in `precision`. */
int efea_bj_setup(int precision, int Nnode, int dof_n, const int32_t* node_ptr, const int32_t* node_cols,
                  const double* values, const uint8_t* known, void* Minv, void* stream);

size_t efea_pcg_workspace_bytes(int Nnode, int dof_n, int group);
/* Lanes per node (`group`) of the solver kernels for a pattern of Nnode nodes
 * with at most `max_neighbours` neighbours (self included): 2, 4, 8 or 16, from
 * timings of the PCG step-1 kernel on MI355X (`python -m easyfea_amd.autotune`).
 * Supported (dof_n, group) pairs: (1, 2|4|8|16), (2, 4|8|16) and (3, 8|16|32). */
int efea_default_group(int dof_n, int max_neighbours, int Nnode);

/* Dirichlet-reduced solve A_ii x_i = b_i - A_ic x_c by block-Jacobi PCG on the
 * full-length vectors (known DOFs masked, x holds the Dirichlet values there and
 * the initial guess elsewhere; x is overwritten with the solution).
 * mode 64: FP64 matrix/preconditioner.  mode 32: FP32 matrix values (with the
 * efea_split_f32 correction, corr32, which may be NULL) and preconditioner in
 * the Krylov products, FP64 vectors, and the recursive residual replaced by
 * the FP64 r = b - K64 x at every poll (replace_ratio = 0), or whenever it
 * falls below replace_ratio times its running maximum (replace_ratio > 0).  Stops when ||r|| <= tol_residual ||r0|| or when the delayed
 * A-norm error estimate <= tol_anorm (relative).  Synchronises `stream`.
 * The host reads the solver state back (and, in mode 32, replaces the
 * residual) after `first_check` iterations (<= 0: check_every), then every
 * `check_every` iterations; kernels launched after convergence return at once.
 * A caller that expects about n iterations (a sequence of similar solves)
 * passes first_check ~ n and a small check_every to save read-backs.
 * info (host, 4 doubles): iterations, final true relative residual, residual
 * replacements, status (1 converged, 2 breakdown or non-finite values,
 * 3 diverged, 0 iteration limit).  max_blocks caps the blocks of each kernel's
 * grid (<= 0: 1024); it fixes the reduction order, so solves agree bitwise
 * only with the same cap.  Small caps suit many systems batched together.
 * Replaces __Solver_1 (Simulations/Solvers.py:502-553) with _Solve_Axb's
 * backends (Solvers.py:225-394; spsolve / pypardiso / cg). */
int efea_solve_pcg(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                   const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                   const void* Minv,
                   const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                   double tol_anorm, double replace_ratio, int maxit, int check_every, int first_check,
                   int max_blocks, void* stream, double* info);

/* efea_solve_pcg without a workspace, for callers that cannot allocate
 * device or pinned memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * solver's workspace lives in a static device array of
 * EFEA_NOALLOC_SCRATCH_BYTES (16 MiB) per GPU, held until the process ends
 * (systems up to about 418,000 DOF; larger ones return EFEA_ERR_SIZE before
 * any device work), and the state read-backs go to pageable host memory on
 * the stack; the call makes no hipMalloc, hipFree, hipHostMalloc or other
 * allocation.  Arguments, iterates, results and info are those of
 * efea_solve_pcg (the same kernels on the same grids: bitwise the same
 * solution).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams or threads) race. */
#define EFEA_NOALLOC_SCRATCH_BYTES (16u << 20)
int efea_solve_pcg_noalloc(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                           const int32_t* node_cols, const double* values64, const float* values32,
                           const float* corr32, const uint8_t* known, const void* Minv, const double* b, double* x,
                           double tol_residual, double tol_anorm, double replace_ratio, int maxit, int check_every,
                           int first_check, int max_blocks, void* stream, double* info);

/* Two-level preconditioner: M^-1 = B^-1 + P Ac^-1 P^T with B the nodal block
 * Jacobi and P the rigid-body modes (2D: 3, 3D: 6, scalar: 1 per aggregate) of
 * node aggregates, zero at known DOFs; rel (Nnode, 3) holds node coordinates
 * relative to their aggregate's centroid.
 * efea_coarse_setup assembles Ac = P^T A P from (node, neighbour position)
 * pairs grouped by aggregate block (blk_I[k], blk_J[k]) = pairs
 * pair_ptr[k]..pair_ptr[k+1]-1, and inverts it in place (blocked Gauss-Jordan,
 * FP64; above order 1024 with 256-column steps whose products are rocBLAS
 * DGEMMs) into Ainv (ld x ld, ld = efea_coarse_ld; scratch:
 * efea_coarse_scratch_bytes(ld) bytes).  Diagonal entries of modes that
 * vanish on the unknown DOFs are set to 1, the others scaled by
 * 1 + diag_shift (0 in general; a small shift keeps the inverse of a nearly
 * singular Ac positive definite, which no-pivot elimination cannot promise).
 * efea_solve_pcg2 is efea_solve_pcg with that preconditioner; agg_ptr /
 * agg_nodes list the nodes of each aggregate.  EasyFEA solves the reduced
 * system directly (Solvers.py:225-394); the preconditioner changes the
 * iteration count, not the solution. */
int efea_coarse_ld(int nagg, int dof_n);
size_t efea_coarse_scratch_bytes(int ld);
int efea_coarse_setup(int dof_n, int nblocks, const int32_t* pair_ptr, const int32_t* pair_node,
                      const int32_t* pair_pos, const int32_t* blk_I, const int32_t* blk_J, const int32_t* node_ptr,
                      const int32_t* node_cols, const double* values64, const uint8_t* known, const double* rel,
                      int ld, double diag_shift, double* Ainv, double* scratch, void* stream);
size_t efea_pcg2_workspace_bytes(int Nnode, int dof_n, int nagg);
int efea_solve_pcg2(int mode, int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                    const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                    const void* Minv, int nagg,
                    const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                    const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                    double tol_anorm, double replace_ratio, int maxit, int check_every, int first_check,
                    int max_blocks, void* stream, double* info);

/* Pipelined block-Jacobi PCG (Ghysels and Vanroose's preconditioned pipelined
 * CG with the nodal block-Jacobi inverse of efea_bj_setup): one kernel per
 * iteration instead of efea_solve_pcg's two, for latency-bound systems.  The
 * preconditioner is node-local, so the launch that updates w = A u also forms
 * the m = M w the next launch multiplies; every block reduces the previous
 * launch's dot-product partials itself.  Same arguments, stopping test and
 * info as efea_solve_pcg (precision 64: FP64 throughout; 32: the iterations'
 * products use values32 + corr32, both required, and Minv is FP32, the
 * start-up and the residual replacements use values64), without
 * replace_ratio and max_blocks:
 * every replace_every iterations (0: 64, or 32 with FP32 values) the
 * recursive vectors are recomputed from x and the direction with the FP64
 * matrix.  The grid covers
 * Nnode * group lanes in one pass: at most efea_pcgp_max_lanes() lanes
 * (EFEA_ERR_ARG beyond).  A solve whose recurrences break down returns
 * EFEA_ERR_NOCONV with info status 2; the caller redoes it with
 * efea_solve_pcg from the same start.  Workspace: efea_pcgp_workspace_bytes.
 * Replaces __Solver_1 (Simulations/Solvers.py:502-553) with _Solve_Axb's
 * backends (Solvers.py:225-394), as efea_solve_pcg does. */
size_t efea_pcgp_workspace_bytes(int Nnode, int dof_n);
int efea_pcgp_max_lanes(void);
int efea_solve_pcgp(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                    const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                    const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                    size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                    int check_every, int first_check, void* stream, double* info);

/* efea_solve_pcgp split in two, for a caller with host work to overlap with
 * the solve (the staggered loop prepares the displacement system while the
 * damage solve runs).  efea_pcgp_start queues the start-up passes, the first
 * first_check (>= 1) iterations, the residual estimate and the read-back of
 * the solver state on `stream` and returns without waiting (no info).
 * efea_pcgp_finish, called with the same arguments on the same stream before
 * the next efea_pcgp_start of the calling thread, waits for that read-back
 * and, if the stopping test had not been met, continues the solve with a
 * read-back every check_every iterations, as efea_solve_pcgp does.  The
 * iterates, x and info are those of efea_solve_pcgp with the same
 * first_check.  *resumed (may be null) is set to 1 when iterations ran after
 * the start's: x then changed after the work the start queued, and what the
 * caller queued in between from x is stale.  Replaces the same reference
 * call as efea_solve_pcgp (__Solver_1, Simulations/Solvers.py:502-553). */
int efea_pcgp_start(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                    const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                    const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                    size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                    int check_every, int first_check, void* stream);
int efea_pcgp_finish(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                     const uint8_t* known, const void* Minv, const double* b, double* x, void* workspace,
                     size_t workspace_bytes, double tol_residual, double tol_anorm, int replace_every, int maxit,
                     int check_every, int first_check, void* stream, double* info, int* resumed);

/* Right-preconditioned BiCGStab with the nodal block-Jacobi inverse of
 * efea_bj_setup (mode 64, FP64 throughout) for non-symmetric systems: the
 * displacement stiffness of the phase-field splits Zhang and *_PM / *_MP.
 * Same matrix, mask, b / x conventions and info as efea_solve_pcg; stops when
 * the recursive residual norm is below tol_residual times the initial one;
 * info[1] is the true relative residual of the returned x.  Breakdown or no
 * convergence within maxit: EFEA_ERR_NOCONV.  Synchronises the stream every
 * check_every iterations.  Workspace: efea_bicgstab_workspace_bytes.
 * Replaces the non-symmetric branch of _Solve_Axb (Simulations/Solvers.py:
 * 225-394), which EasyFEA solves with a sparse direct solver. */
size_t efea_bicgstab_workspace_bytes(int Nnode, int dof_n);
int efea_solve_bicgstab(int Nnode, int dof_n, int group, const int32_t* node_ptr, const int32_t* node_cols,
                        const double* values, const uint8_t* known, const double* Minv, const double* b, double* x,
                        void* workspace, size_t workspace_bytes, double tol_residual, int maxit, int check_every,
                        void* stream, double* info);

/* Pipelined two-level PCG (Ghysels and Vanroose's preconditioned pipelined
 * CG) with the preconditioner of efea_solve_pcg2: two kernels per iteration
 * instead of three (the restriction P^T w and both dot products ride on the
 * product n = A m), for latency-bound systems.  precision 64: FP64
 * throughout; 32: the iterations' products use values32 + corr32 (both
 * required; efea_split_f32 makes them) and Minv is FP32, the start-up and
 * replacement passes use values64.  Same arguments, stopping test and info as efea_solve_pcg2 (mode
 * 64 or 32 with FP64 vectors), plus the node chunks
 * the product pass runs over, one block each (at most 1024): runs of at most
 * W = 256 / group nodes of one aggregate in agg_nodes order.  Chunk c owns the
 * W positions c * W + t of chunk_rows (3 * nchunk * W ints): position j holds
 * its node (-1: none), then node_ptr[node] and node_ptr[node + 1]; chunk c
 * writes its share of P^T w to slot chunk_slot[c] = a * chunks_per_agg + (its
 * index among its aggregate a's chunks).  agg_first (256 * nagg ints): the
 * first 256 nodes of each aggregate in agg_nodes order, -1 past its end.
 * Workspace: efea_pcg2p_workspace_bytes(Nnode, dof_n,
 * nagg * chunks_per_agg).  Every 64 iterations (32 with FP32 values) the
 * recursive vectors are recomputed from x and the direction with the FP64
 * matrix (residual replacement). */
/* Replaces __Solver_1 (Simulations/Solvers.py:502-553) with _Solve_Axb's
 * backends (Solvers.py:225-394), as efea_solve_pcg does. */
size_t efea_pcg2p_workspace_bytes(int Nnode, int dof_n, int nslot);
int efea_solve_pcg2p(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                     const uint8_t* known, const void* Minv, int nagg,
                     const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                     int nchunk, const int32_t* chunk_rows, const int32_t* chunk_slot, int chunks_per_agg,
                     const int32_t* agg_first, const double* b, double* x, void* workspace, size_t workspace_bytes,
                     double tol_residual, double tol_anorm, int maxit, int check_every, int first_check, void* stream,
                     double* info);

/* efea_solve_pcg2p split in two, as efea_pcgp_start / efea_pcgp_finish split
 * efea_solve_pcgp: efea_pcg2p_start queues the start-up, the first first_check
 * (>= 1) iterations, the residual estimate and the state read-back on
 * `stream` and returns without waiting; efea_pcg2p_finish, called with the
 * same arguments on the same stream before the next deferred start of the
 * calling thread, waits and, if the stopping test had not been met, continues
 * with a read-back every check_every iterations.  Same iterates, x and info
 * as efea_solve_pcg2p with the same first_check (the residual replacements
 * follow the iteration count); *resumed (may be null) = 1 when iterations ran
 * after the start's.  Replaces the same reference call as efea_solve_pcg2p. */
int efea_pcg2p_start(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                     const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                     const uint8_t* known, const void* Minv, int nagg,
                     const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                     int nchunk, const int32_t* chunk_rows, const int32_t* chunk_slot, int chunks_per_agg,
                     const int32_t* agg_first, const double* b, double* x, void* workspace, size_t workspace_bytes,
                     double tol_residual, double tol_anorm, int maxit, int check_every, int first_check, void* stream);
int efea_pcg2p_finish(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                      const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                      const uint8_t* known, const void* Minv, int nagg,
                      const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                      int nchunk, const int32_t* chunk_rows, const int32_t* chunk_slot, int chunks_per_agg,
                      const int32_t* agg_first, const double* b, double* x, void* workspace, size_t workspace_bytes,
                      double tol_residual, double tol_anorm, int maxit, int check_every, int first_check,
                      void* stream, double* info, int* resumed);

/* Timing harness of efea_solve_pcg2p's dominant kernel (the product pass:
 * n = A m, the eight vector updates, the dot-product and restriction
 * partials): reps launches back to back on `stream`, on the solve's chunks,
 * matrix (precision 64: values64; 32: values32 + corr32), mask and aggregate
 * coordinates, in the state of an iteration past the first, after four fill
 * and two state kernels; x (Ndof doubles) and the workspace
 * (efea_pcg2p_workspace_bytes) are overwritten.  Benchmark support, no
 * counterpart in the reference. */
int efea_pp_iter_bench(int precision, int Nnode, int dof_n, int group, const int32_t* node_ptr,
                       const int32_t* node_cols, const double* values64, const float* values32, const float* corr32,
                       const uint8_t* known, int nagg, const double* rel, int nchunk, const int32_t* chunk_rows,
                       const int32_t* chunk_slot, int chunks_per_agg, double* x, void* workspace,
                       size_t workspace_bytes, int reps, void* stream);

/* Continuous batching of Dirichlet-reduced solves (the serving loop).
 * A batch holds up to `capacity` systems of one (mode, dof_n, group) class and
 * advances all of them with one launch per PCG step; systems join
 * (efea_batch_add) and leave (efea_batch_pop) between calls of efea_batch_run.
 * Every system runs on the grid it would get alone, so its iterates are
 * bitwise those of efea_solve_pcg (efea_solve_pcg2 in a two-level batch) with
 * check_every = the `iterations` passed to
 * efea_batch_run.  All calls on one batch use one stream; the device buffers of
 * a system (matrix, Minv, b, x, workspace of efea_pcg_workspace_bytes bytes,
 * or efea_pcg2_workspace_bytes in a two-level batch, coarse arrays) stay alive
 * until it is popped.  two_level = 1: every system of the batch uses the
 * two-level preconditioner (efea_solve_pcg2's iterates; the coarse arguments
 * of efea_batch_add are used, at most 4096 aggregates per system), 0: block
 * Jacobi (efea_solve_pcg's; coarse arguments ignored).  The handle is a host
 * pointer.
 * Replaces the per-simulation solves of __Solver_1 (Simulations/Solvers.py:502-553)
 * for many simulations at once; the reference has no batched path. */
int efea_batch_create(int mode, int dof_n, int group, int two_level, int capacity, void** handle);
int efea_batch_destroy(void* handle);
/* Starts one solve (asynchronous); returns its id (>= 0) or an error code. */
int efea_batch_add(void* handle, int Nnode, const int32_t* node_ptr, const int32_t* node_cols,
                   const double* values64, const float* values32, const float* corr32, const uint8_t* known,
                   const void* Minv, int nagg,
                   const int32_t* agg_ptr, const int32_t* agg_nodes, const double* rel, const double* Ainv, int ld,
                   const double* b, double* x, void* workspace, size_t workspace_bytes, double tol_residual,
                   double tol_anorm, double replace_ratio, int maxit, int max_blocks, void* stream);
/* Runs `iterations` PCG iterations of every active system, then polls their
 * states (synchronises `stream`); returns how many systems finished.
 * efea_batch_run = efea_batch_launch + efea_batch_wait. */
int efea_batch_run(void* handle, int iterations, void* stream);
/* The two halves of efea_batch_run, so that the host can work while the GPU
 * runs the chunk: efea_batch_launch enqueues the iterations and the state
 * read-back and returns at once (one chunk in flight per batch; a second
 * launch before the wait is an error); efea_batch_wait synchronises `stream`,
 * reads the states and returns how many systems finished (0 with nothing in
 * flight).  Systems added between the two join at the next launch; the
 * caller orders work from other streams before efea_batch_add (an event the
 * batch stream waits on). */
int efea_batch_launch(void* handle, int iterations, void* stream);
int efea_batch_wait(void* handle, void* stream);
/* Pops one finished system: *id and info (host, 4 doubles, as efea_solve_pcg).
 * Returns 0, or 1 with *id = -1 when none is waiting. */
int efea_batch_pop(void* handle, int* id, double* info);
/* Number of systems still iterating. */
int efea_batch_active(void* handle);

/* Deterministic dot product (b may be NULL: sum of a); scratch holds 1024 doubles. */
int efea_dot(int64_t n, const double* a, const double* b, double* scratch, double* out, void* stream);

/* out = sum_k y[idx[k]] in index order (reaction resultants). */
int efea_gather_sum(int64_t n, const int32_t* idx, const double* y, double* out, void* stream);

/* Energy splits of the phase-field kernels (PhaseField.SplitType,
 * Models/_phasefield.py:55-78).  Codes 0-8 give a symmetric stiffness
 * g c+ + c- (PCG); Zhang and the *_PM / *_MP variants (9-13) give a
 * non-symmetric one, solved by efea_solve_bicgstab. */
#define EFEA_SPLIT_AMOR 0
#define EFEA_SPLIT_MIEHE 1
#define EFEA_SPLIT_BOURDIN 2
#define EFEA_SPLIT_STRESS 3
#define EFEA_SPLIT_HE 4
#define EFEA_SPLIT_ANISOT_STRAIN 5
#define EFEA_SPLIT_ANISOT_STRAIN_NOCROSS 6
#define EFEA_SPLIT_ANISOT_STRESS 7
#define EFEA_SPLIT_ANISOT_STRESS_NOCROSS 8
/* non-symmetric c+ / c- (the displacement system needs a non-symmetric solver) */
#define EFEA_SPLIT_ZHANG 9
#define EFEA_SPLIT_ANISOT_STRAIN_PM 10
#define EFEA_SPLIT_ANISOT_STRAIN_MP 11
#define EFEA_SPLIT_ANISOT_STRESS_PM 12
#define EFEA_SPLIT_ANISOT_STRESS_MP 13

/* Phase field, damage problem: psi+ at the mass points with the History
 * irreversibility (psi+ <- H_old where smaller; H_old may be NULL), then
 * Kd_e (Ne, nPe, nPe), Fd_e (Ne, nPe).  split EFEA_SPLIT_*, regu 0 AT1 / 1 AT2.
 * Replaces PhaseField.__Calc_psiPlus_e_pg and __Construct_Damage_Matrix
 * (Simulations/_phasefield.py:494-573) with Models/_phasefield.py:236-571, 751-1128. */
int efea_pf_damage_ke(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                      const double* tables_mass, int nPg_mass, const double* tables_rigi, int nPg_rigi,
                      const double* u, const double* H_old, const double* E, int E_stride, double nu,
                      int plane_stress, const double* Gc, int Gc_stride, double l0, int split, int regu,
                      double thickness, double* psiP, double* Kd, double* Fd, void* stream);

/* Phase field, displacement problem: K_e with c = g(d) c+ + c- at the rigi points.
 * Replaces PhaseField.__Construct_Elastic_Matrix (Simulations/_phasefield.py:444-482). */
int efea_pf_elastic_ke(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                       const double* tables_rigi, int nPg_rigi, const double* u, const double* d,
                       const double* E, int E_stride, double nu, int plane_stress, int split, double thickness,
                       double* Ke, void* stream);

/* Phase-field results per element (PhaseField.Result strain / stress options,
 * Simulations/_phasefield.py:682-760 with _Calc_Sigma_e_pg :868-916): the
 * damaged stress g(d) c+ eps + c- eps at the rigi points; out (Ne, 9) as
 * efea_elastic_strain_stress in 2D. */
int efea_pf_strain_stress(int elem_type, int64_t Ne, const int32_t* connect, const double* coord,
                          const double* tables_rigi, int nPg_rigi, const double* u, const double* d, const double* E,
                          int E_stride, double nu, int plane_stress, int split, double thickness, double* out,
                          void* stream);

/* Pointwise 2D split: eps (n, 3) -> c+, c- (n, 3, 3), psi+, psi- (n).
 * Replaces PhaseField.Calc_C / Calc_psi_e_pg (Models/_phasefield.py:335-431). */
int efea_pf_split_points(int64_t n, int split, const double* eps, const double* E, int E_stride, double nu,
                         int plane_stress, double* cP, double* cM, double* psiP, double* psiM, void* stream);

/* ---- time integration (parabolic and hyperbolic schemes) ---------------- */

/* Time schemes, the reference's AlgoType (Simulations/Solvers.py:64-190). */
#define EFEA_SCHEME_ELLIPTIC 0
#define EFEA_SCHEME_PARABOLIC 1
#define EFEA_SCHEME_NEWMARK 2
#define EFEA_SCHEME_MIDPOINT 3
#define EFEA_SCHEME_HHT 4
#define EFEA_SCHEME_HHT_NEWMARK 5
#define EFEA_SCHEME_EULER_IMPLICIT 6
#define EFEA_SCHEME_EULER_EXPLICIT 7

/* Mass / capacity element matrices over the mass rule:
 * Ke[e, (a,i), (b,j)] = delta_ij sum_p coef wJ N_a N_b x scale, (Ne, nPe dof_n, nPe dof_n),
 * dof_n 1 or the element dimension.  coef: per point (coef_stride = nPg), per
 * element (coef_stride = 1) or NULL (then coef_const).  precision 64.
 * Replaces Operators.Bilinear.UV with dof_n (FEM/Operators/Bilinear.py:42-59),
 * the M_e of Elastic.Construct_local_matrix_system (Simulations/_elastic.py:135-148)
 * and the C_e of Thermal.Construct_local_matrix_system (Simulations/_thermal.py:114-130). */
int efea_uv_ke(int elem_type, int precision, int64_t Ne, const int32_t* connect, const double* coord,
               const double* tables_mass, int nPg_mass, const double* coef, int coef_stride, double coef_const,
               int dof_n, double scale, void* Ke, void* stream);

/* out = c0 x0 + c1 x1 + c2 x2 + c3 x3 over n entries (x1..x3 may be NULL: term absent;
 * out may be one of the x).  Forms A = coefK K + coefC C + coefM M over one pattern's
 * values (Simulations/_simu.py:1895-1899) and the history vectors of the
 * time-scheme right-hand sides (_Solver_Apply_Neumann, _simu.py:1735-1860). */
int efea_lincomb(int64_t n, double c0, const double* x0, double c1, const double* x1, double c2, const double* x2,
                 double c3, const double* x3, double* out, void* stream);

/* The state after a time-step solve: (u, v, a) from (u_n, v_n, a_n) and the
 * solve's unknown x (u^{n+1}, or a^n for EFEA_SCHEME_EULER_EXPLICIT), the
 * reference's expressions evaluated in its operation order (bit-identical to
 * them).  v_out / a_out may be NULL; outputs may alias inputs.
 * Replaces _Simu._Solver_Update_solutions (Simulations/_simu.py:1552-1640). */
int efea_time_update(int scheme, int64_t n, double dt, double beta, double gamma, double alpha, const double* u_n,
                     const double* v_n, const double* a_n, const double* x, double* u_out, double* v_out,
                     double* a_out, void* stream);

/* ZZ1 smoothed deformation energy per element (Simulations/_elastic.py:323-418, the
 * smoothedStress branch of _Calc_Psi_Elas used by _Calc_ZZ1): out[e] = 1/2 sum_p wJ sig_s . eps
 * (x thickness in 2D) at the stiffness points, eps = B u_e and sig_s = sum_a N_a sig_n[a], with
 * sig_n (Nn, ns) the nodal average of the element-mean stresses (tensor components, shear not
 * scaled by sqrt 2, as efea_elastic_strain_stress reports them, then efea_node_average). */
int efea_zz1_energy(int elem_type, int64_t Ne, const int32_t* connect, const double* coord, const double* tables,
                    int nPg, const double* u, const double* sig_n, double thickness, double* out, void* stream);

/* Workspace of efea_solve_bounded_lsq for N unknowns (bytes). */
size_t efea_bounded_lsq_workspace_bytes(int N);

/* Bound-constrained least squares on the device: x = argmin |A x - b|^2 subject to
 * lb <= x <= ub, for a symmetric scalar CSR matrix A (row_ptr N+1, cols, values, int32
 * indices), started from x clamped to the bounds (x in/out).  Primal-dual active-set updates
 * with CGLS on the free set: each pass reduces the column-scaled free-set gradient norm by
 * tol, or stops after maxit iterations (rounded up to a multiple of 16: the host checks the
 * norm every 16 iterations); done when no variable violates the first-order conditions by
 * more than kkt_tol, after which x is clamped to [lb, ub] (a move of at most kkt_tol);
 * EFEA_ERR_NOCONV after max_outer updates.  info (4 doubles, may be NULL):
 * active-set updates, CGLS iterations, violations left, first scaled gradient norm^2.
 * Synchronises the stream.  Replaces scipy.optimize.lsq_linear(A, b, bounds=(lb, ub),
 * tol=1e-10, method="trf") of the BoundConstrain damage solve (Simulations/Solvers.py:376-383,
 * bounds from Simulations/_phasefield.py:148-163). */
int efea_solve_bounded_lsq(int N, const int32_t* row_ptr, const int32_t* cols, const double* values,
                           const double* b, const double* lb, const double* ub, double* x, void* workspace,
                           size_t workspace_bytes, double tol, double kkt_tol, int maxit, int max_outer,
                           double* info, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EFEA_H */
