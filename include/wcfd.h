/*
 * wcfd.h -- C ABI of libwcfd.so, the gfx950 HIP implementation of the
 * SIMPLE/PISO inner loop of JohnCSu/WarpCFD.
 *
 * Conventions
 *  - Every pointer is a device pointer unless marked [host].  The caller owns
 *    all inputs, outputs and workspaces, and the library never allocates or
 *    frees device memory: scratch space comes from a caller-provided
 *    workspace whose size wcfd_*_workspace_bytes() returns (the one
 *    exception, wcfd_batch_convergence_report_noalloc, uses a static device
 *    array instead).  Host-side, the Krylov solves that poll for their stop
 *    (wcfd_cg*, wcfd_bicgstab*, wcfd_simple_iteration and the batched forms
 *    with graph_groups 0) keep one small pinned host buffer (hipHostMalloc,
 *    4 x max(64, K) flag records) and five hipEvents per host thread,
 *    created on that thread's first polled solve, grown for a larger K, and
 *    kept until the process ends; the fixed-budget form (graph_groups > 0)
 *    uses neither.
 *  - Arrays are structure-of-arrays in the device layout: cells renumbered
 *    (reverse Cuthill-McKee), internal faces first then boundary faces.
 *    "[4][C]" means 4 contiguous rows of C floats (u, v, w, p).
 *  - Every function returns WCFD_OK or an error code; on WCFD_EHIP the HIP
 *    error is kept per thread and readable with wcfd_last_hip_error().
 *  - `stream` is a hipStream_t passed as void* (0 = the null stream).
 *  - Functions keep no global mutable state, except the opt-in profiler and
 *    the Krylov launch shape, fold limit and stop-test placement
 *    (wcfd_set_krylov_launch / _fold / _inline_check),
 *    which are process-wide: two callers that tune for different meshes must
 *    not interleave their set and solve calls.  The fold limit never changes
 *    results; the launch shape sets the order of the float64 partial sums.
 *    Calls on different streams with different workspaces may run
 *    concurrently.
 *
 * Each entry point cites the reference function it replaces
 * (paths relative to the JohnCSu/WarpCFD repository root).
 */
#ifndef WCFD_H
#define WCFD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Field type of the library build: libwcfd.so computes in float (the
 * reference's default float_dtype=wp.float32, FV/model.py:55); libwcfd64.so is
 * the same source built with -DWCFD_REAL=double (float_dtype=wp.float64).
 * Every field, geometry and matrix array is wcfd_real; scalar parameters
 * (viscosity, relaxation, dt, scales) are float in both builds; reductions
 * are double in both.  wcfd_real_bytes() tells a caller which build it
 * loaded. */
#ifndef WCFD_REAL
#define WCFD_REAL float
#endif
typedef WCFD_REAL wcfd_real;

#define WCFD_OK 0
#define WCFD_EINVAL 1
#define WCFD_EHIP 2
#define WCFD_ENOTCONV 3
#define WCFD_ENAN 4
#define WCFD_ESCRATCH 5 /* input larger than a no-allocation entry point's static scratch */
#define WCFD_EMESH 6     /* a face shared by more than two cells (wcfd_match_faces) */
#define WCFD_ECAPACITY 7 /* an output array too small for the result (wcfd_match_faces) */

#define WCFD_DIRICHLET 1
#define WCFD_NEUMANN 2

/* CG variants (same stopping rule and, in exact arithmetic, the same
 * iterates) */
#define WCFD_CG_STANDARD 0          /* two reductions, two kernels per iteration */
#define WCFD_CG_SINGLE_REDUCTION 1  /* Chronopoulos-Gear: one fused kernel per iteration */

/* scheme flags */
#define WCFD_SCHEME_CORRECTED 1        /* over-relaxed non-orthogonal Laplacian
                                          (laplacianScheme.py:39-66) */
#define WCFD_SCHEME_IMPLICIT_OUTFLOW 2 /* not in the reference: implicit upwind
                                          convective outflow on Neumann faces */

/* Read-only mesh description (replaces the Cell/Face struct arrays of
 * warp_cfd/FV/mesh_structs.py:5-64 built by init_structs 141-251). */
typedef struct wcfd_mesh {
  int32_t n_cells;      /* C */
  int32_t n_faces;      /* F = Fi + Fb */
  int32_t n_internal;   /* Fi, faces 0..Fi-1 */
  int32_t n_boundary;   /* Fb, faces Fi..F-1 */
  int32_t nnz;          /* nnz of the C x C CSR pattern */
  int32_t n_slots;      /* slots per cell in use: 6 (hexahedra; 0 also means 6), 5 (wedges),
                           4 (tetrahedra).  Slots k >= n_slots hold the null face and zero
                           matrix coefficients; the Krylov kernels skip them.  Any other
                           value makes every entry point return WCFD_EINVAL. */
  const int32_t* owner;     /* [F]    owner cell (the reference's owner) */
  const int32_t* neigh;     /* [Fi]   neighbour cell */
  const wcfd_real* psi;         /* [Fi]   |x_f - x_O| / |x_O - x_N| (mesh_structs.py:206) */
  const wcfd_real* skew;        /* [3][Fi] x_f - (psi x_O + (1-psi) x_N) */
  const wcfd_real* area;        /* [F]    owner-side face area (mesh_structs.py:200) */
  const wcfd_real* normal;      /* [3][F] owner-side normal */
  const wcfd_real* dist;        /* [3][Fi] x_N - x_O */
  const wcfd_real* dmag;        /* [F]    |x_N - x_O| (internal), |x_f - x_P| (boundary) */
  const wcfd_real* volume;      /* [C] */
  const int32_t* slot_face; /* [6][C] face of local face k of cell c */
  const int8_t* slot_side;  /* [6][C] 0 if c owns that face, 1 otherwise */
  const wcfd_real* slot_sn;     /* [3][6][C] A_f * n_{c,k} (the cell's own normal) */
  const int32_t* slot_pos;  /* [6][C] CSR index of (c, neighbour) or -1 */
  const int32_t* diag_pos;  /* [C]    CSR index of (c, c) */
  const int32_t* row_ptr;   /* [C+1] */
  const int32_t* col;       /* [nnz]  columns sorted within a row */
  const int32_t* slot_nbr;  /* [6][C] neighbour in slot k (c itself at boundary slots) */
  const int16_t* slot_dnb;  /* [6][C] slot_nbr - c, or NULL when the band exceeds int16 */
} wcfd_mesh;

/* Matrices of the device path (the assembly outputs and the *_ell solvers)
 * are stored slot-major, "ELL": values[c] is the diagonal of row c and
 * values[(1+k)*C + c] the coefficient of (c, slot_nbr[k][c]), 0 at boundary
 * slots -- 7*C floats.  wcfd_ell_to_csr() writes the same matrix on the CSR
 * pattern (row_ptr, col), which is the reference's pattern (bit-exact). */

/* Boundary conditions per boundary face (warp_cfd/FV/boundary/boundary.py). */
typedef struct wcfd_bc {
  const uint8_t* type;  /* [4][Fb] 1 Dirichlet, 2 Neumann */
  const wcfd_real* value;   /* [4][Fb] value or normal gradient */
} wcfd_bc;

/* Batch of independent cases on one mesh descriptor (no counterpart in the
 * reference, which runs one case per FVM object).  The cases' meshes are
 * placed side by side in one wcfd_mesh: case k owns the cells
 * [cell_off[k], cell_end[k]) and a range of internal and of boundary
 * faces; no face or matrix coefficient couples two cases.  Cells and faces
 * that belong to no case (free space of a serving arena) have case -1.
 * Row-parallel kernels (Krylov, continuity) give every case the workgroups
 * [blk_off[k], blk_off[k+1]) of `block` threads, each owning `chunk`
 * consecutive rows, so a case's arithmetic -- including every float64 sum --
 * is bitwise the same alone or beside any other cases.  Kernels skip cases
 * whose active flag is 0; the SIMPLE step clears the flag of a case once its
 * convergence monitor (wcfd_monitor_update, below) reports it converged. */
typedef struct wcfd_batch {
  int32_t n_cases;          /* K */
  int32_t n_blocks;         /* G = blk_off[K] */
  int32_t chunk;            /* rows per workgroup */
  int32_t block;            /* threads per workgroup (64..512, multiple of 64) */
  const int32_t* cell_off;  /* [K+1] first cell of case k */
  const int32_t* cell_end;  /* [K] one past the last cell of case k; NULL: cell_off[k+1] */
  const int32_t* blk_off;   /* [K+1] */
  const int32_t* blk_case;  /* [G] case of each workgroup */
  const int32_t* cell_case; /* [C] case of each cell, -1 if none */
  const int32_t* face_case; /* [F] case of each face, -1 if none */
  uint8_t* active;          /* [K] 1 = iterate the case */
  const wcfd_real* nu;          /* [K] kinematic viscosity per case */
  const int32_t* ref_cell;  /* [K] reference-pressure row per case, -1 none */
  const wcfd_real* ref_value;   /* [K] */
  const double* stop;       /* [K][3] stop_c, stop_u, floor_u of wcfd_monitor_update;
                               stop_c = 0: never stop */
  double* monitor;          /* [K][WCFD_MONITOR_LEN] monitor records, zeroed at the start */
  int32_t* outer_iters;     /* [K] SIMPLE iterations run per case */
  int32_t max_case_blocks;  /* largest blk_off[k+1] - blk_off[k]; 0: taken as n_blocks */
  /* Compacted face/cell launches: [2][n] workgroup item ranges (starts, then
   * ends; at most 256 items each) covering only the cases to iterate --
   * cells, all faces (internal then boundary numbering), boundary faces
   * (0-based).  NULL: the face and cell kernels cover the whole arena. */
  int32_t n_cell_blk, n_face_blk, n_bface_blk;
  const int32_t* cell_blk;
  const int32_t* face_blk;
  const int32_t* bface_blk;
} wcfd_batch;

/* Convergence monitor of a steady run (no counterpart in the reference,
 * whose check_convergence, model.py:278-315 / Convergence.py:7-106, prints
 * norms on the host).  Called once per outer iteration with the continuity
 * residual c and the relative velocity change d = ||u^n - u^(n-1)|| /
 * ||u^n|| (float64 sums).  SIMPLE converges linearly, so the distance to
 * the fixed point is estimated as E = dbar rho / (1 - rho), dbar the mean of
 * d over the last 20 iterations and rho = (D_now / D_prev)^(1/20) from the
 * sums of d over the last two windows of 20 (E = inf while rho >= 1 or
 * fewer than 40 iterations are known).  Returns 1 when
 *   c <= stop_c  and  (E <= stop_u  or  dbar <= floor_u),
 * else 0.  record: [host] WCFD_MONITOR_LEN doubles, zeroed before the first
 * call; record[1..5] = c, d, dbar, rho, E.  The batched SIMPLE step runs the
 * same function on the device. */
#define WCFD_MONITOR_LEN 48
int wcfd_monitor_update(double* record, double continuity, double velocity_change,
                        double stop_c, double stop_u, double floor_u);

/* ---------------------------------------------------------------- errors */
int wcfd_last_hip_error(void);
const char* wcfd_error_string(int code); /* [host] static string */
int wcfd_abi_version(void);
int wcfd_real_bytes(void);  /* sizeof(wcfd_real): 4 (libwcfd.so) or 8 (libwcfd64.so) */

/* ------------------------------------------------------------ face/cell ops */

/* FVM.set_boundary_conditions (warp_cfd/FV/model.py:206-208,
 * FV/boundary/conditions.py:6-19): Dirichlet face values for the fields in
 * field_mask (bit q = field q). */
int wcfd_set_boundary_conditions(const wcfd_mesh* m, wcfd_bc bc, wcfd_real* face_values,
                                 int field_mask, void* stream);

/* FVM.face_interpolation (model.py:210-229; interpolation_kernels.py:5-54;
 * skewness_functions.py:7-20): internal faces linear in psi plus skew
 * correction with cell_grad, Neumann boundary faces |x_f-x_P| g + phi_P.
 * skew_mask selects the fields whose skew term is evaluated (fields whose
 * gradients are known to be zero may be skipped).  If mass_flux != NULL the
 * face flux (u_f . n_O) A_f is written too (model_kernels.py:46-60); if
 * rua_cell != NULL the face value of rUA (solver_kernels.py:22-36) is
 * written to rua_face. */
int wcfd_face_interpolation(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* cell_values,
                            const wcfd_real* cell_grad, wcfd_real* face_values, int field_mask,
                            int skew_mask, wcfd_real* mass_flux, const wcfd_real* rua_cell,
                            wcfd_real* rua_face, void* stream);

/* FVM.face_interpolation(upwind=True) (model.py:210-229 with
 * interpolation_functions.py:18-28): an internal face takes its owner's value
 * where mass_flux (per face, read) is > 0, else its neighbour's, plus the
 * same skew correction as the linear scheme for the fields in skew_mask;
 * Neumann boundary faces as wcfd_face_interpolation. */
int wcfd_face_interpolation_upwind(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* cell_values,
                                   const wcfd_real* cell_grad, const wcfd_real* mass_flux,
                                   wcfd_real* face_values, int field_mask, int skew_mask,
                                   void* stream);

/* FVM.calculate_gradients (model.py:232-239; model_kernels.py:72-90):
 * Green-Gauss gradient of the fields in field_mask; when zero_others != 0 the
 * gradients of the other fields are zeroed, as the reference's zero_() does. */
int wcfd_green_gauss(const wcfd_mesh* m, const wcfd_real* face_values, wcfd_real* cell_grad,
                     int field_mask, int zero_others, void* stream);

/* FVM.calculate_mass_flux (model.py:264-266; model_kernels.py:46-60). */
int wcfd_mass_flux(const wcfd_mesh* m, const wcfd_real* face_values, wcfd_real* mass_flux,
                   void* stream);

/* FVM.divFlux (model.py:269-275; model_kernels.py:27-34). */
int wcfd_divergence(const wcfd_mesh* m, const wcfd_real* mass_flux, wcfd_real* div, void* stream);

/* Momentum system, fused: ConvectionTerm upwind (convection.py:44-88,
 * convectionScheme.py:28-52) + DiffusionTerm (diffusion.py:113-197,
 * laplacianScheme.py:14-66) + GradTerm (grad.py:22-27) [+ DdtTerm
 * (ddt.py:25-33) when dt > 0] -> Equation.form_system (equation.py:99-124,
 * equation_kernels.py:52-122) -> Equation.relax (equation_kernels.py:128-152)
 * when alpha > 0.  One C x C value array shared by u, v, w (requires the same
 * BC type for u, v, w on every face); three right-hand sides rhs[3][C].
 * If compute_gradients != 0 the Green-Gauss gradients of all 4 fields are
 * computed first (model_kernels.py:72-90) into cell_grad and used for the p
 * term.  scheme_flags: WCFD_SCHEME_* bits.  values: [7][C] ELL. */
int wcfd_assemble_momentum(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* cell_values,
                           const wcfd_real* face_values, const wcfd_real* mass_flux,
                           wcfd_real* cell_grad, int compute_gradients, float nu,
                           float inv_density, float alpha_u, float dt, int scheme_flags,
                           wcfd_real* values, wcfd_real* rhs, void* stream);

/* Pressure equation, fused: divFlux (model_kernels.py:27-34) + DiffusionTerm
 * with per-face viscosity rUA_f (diffusion.py:40-49) -> form_system + add_RHS
 * (equation.py:99-124,169-170) -> replace_row (equation_kernels.py:155-175)
 * when ref_cell >= 0.  If compute_grad_p != 0 the Green-Gauss gradient of p
 * is computed into cell_grad (other gradients untouched).  div_out may be
 * NULL.  values: [7][C] ELL. */
int wcfd_assemble_pressure(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* face_values,
                           const wcfd_real* mass_flux, const wcfd_real* rua_face, wcfd_real* cell_grad,
                           int compute_grad_p, int correction, int ref_cell, float ref_value,
                           wcfd_real* values, wcfd_real* rhs, wcfd_real* div_out, void* stream);

/* ------------------------------------------------- composable terms
 * The reference's term-by-term API (the FV/terms package), for callers that build
 * their own equations; the SIMPLE step uses the fused assembly above.
 *
 * Face-based term weights are the reference's Term.weights (C, 6, O, 3) in
 * the layout w[O][3][6][C]: output o, part (0 owner-cell coefficient, 1
 * neighbour coefficient, 2 explicit), local face k, cell c (device order).
 * Cell-based weights: w[O][2][C] (implicit diagonal, explicit).  `outputs`:
 * [host] n_outputs (1..4) global field indices (0..3). */
#define WCFD_CONV_UPWIND 0        /* convectionScheme.py:28-52 (tie -> neighbour) */
#define WCFD_CONV_CENTRAL 1       /* convectionScheme.py:3-26 */
#define WCFD_CONV_UPWIND_LINEAR 2 /* convectionScheme.py:54-88 */

/* ConvectionTerm.calculate_weights (convection.py:33-88): scheme weights
 * times the face's mass flux, plus the skew correction of internal faces.
 * face_c2f: [2][3][Fi] per internal face, x_f - x_owner and the owner's
 * centroid-to-face vector of the neighbour's local face index (the vector
 * upwindLinear reads when the neighbour is upwind, convectionScheme.py:82);
 * required for WCFD_CONV_UPWIND_LINEAR only.  flags: WCFD_SCHEME_IMPLICIT_
 * OUTFLOW. */
int wcfd_convection_weights(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* cell_grad,
                            const wcfd_real* face_values, const wcfd_real* mass_flux,
                            const wcfd_real* face_c2f, int scheme, int flags, const int32_t* outputs,
                            int n_outputs, wcfd_real* weights, void* stream);

/* DiffusionTerm.calculate_weights (diffusion.py:40-49, 52-197; laplacianScheme.py
 * central_difference / central_difference_corrected).  viscosity_faces: [F]
 * device face order, or NULL for the scalar `viscosity`.  bc_mode 0: the
 * boundary conditions; 1: the von_neumann= (and dirchlet=, which the
 * reference builds with the same kernel) option, explicit nu * bc_value * A
 * on every boundary face, one output only. */
int wcfd_diffusion_weights(const wcfd_mesh* m, wcfd_bc bc, const wcfd_real* cell_grad,
                           const wcfd_real* face_values, const wcfd_real* viscosity_faces,
                           float viscosity, int correction, int bc_mode, float bc_value,
                           const int32_t* outputs, int n_outputs, wcfd_real* weights, void* stream);

/* DdtTerm.calculate_weights, backwards_Euler (ddt.py:18-33): (V/dt, -phi V/dt)
 * per output, indexed by local output. */
int wcfd_ddt_weights(const wcfd_mesh* m, const wcfd_real* cell_values, float dt,
                     const int32_t* outputs, int n_outputs, wcfd_real* weights, void* stream);

/* GradTerm.calculate_weights (grad.py:17-28): w[d][c] = (grad phi_field)_c[d] V_c,
 * an explicit term (component-major [3][C]). */
int wcfd_grad_weights(const wcfd_mesh* m, const wcfd_real* cell_grad, int field, wcfd_real* weights,
                      void* stream);

/* Equation of n_outputs outputs: values[O][7][C] (one ELL plane per output:
 * the reference's BSR rows O*c + o), rhs[O][C].  Equation.form_system
 * (equation.py:99-124) with calculate_BSR_values_kernel and
 * calculate_RHS_values_kernel (equation_kernels.py:52-101) for a face-based
 * term, add_Implicit_cell_based_term_kernel (104-122) for a cell-based one,
 * rhs += scale arr for an explicit term or add_RHS (equation.py:169-170). */
int wcfd_equation_add_face_term(const wcfd_mesh* m, const wcfd_real* weights, int n_outputs,
                                float scale, wcfd_real* values, wcfd_real* rhs, void* stream);
int wcfd_equation_add_cell_term(const wcfd_mesh* m, const wcfd_real* weights, int n_outputs,
                                float scale, wcfd_real* values, wcfd_real* rhs, void* stream);
int wcfd_equation_add_rhs(int32_t n, const wcfd_real* arr, float scale, wcfd_real* rhs, void* stream);
/* Equation.relax (equation.py:173-181; implicit_relaxation_kernel,
 * equation_kernels.py:128-152) with the outputs' current cell values. */
int wcfd_equation_relax(const wcfd_mesh* m, const wcfd_real* cell_values, const int32_t* outputs,
                        int n_outputs, float alpha, wcfd_real* values, wcfd_real* rhs, void* stream);
/* Equation.replace_row (equation.py:184-193): identity row of (output, cell)
 * with right-hand side `value` (the reference reads it from a length-1 array
 * at index row_id, T7). */
int wcfd_equation_replace_row(const wcfd_mesh* m, int n_outputs, int output, int32_t cell,
                              float value, wcfd_real* values, wcfd_real* rhs, void* stream);

/* ELL values [7][C] -> CSR values [nnz] on (row_ptr, col). */
int wcfd_ell_to_csr(const wcfd_mesh* m, const wcfd_real* ell_values, wcfd_real* csr_values, void* stream);

/* calculate_rUA + get_HbyA + replace_cell_values (solver_kernels.py:7-16,
 * 39-43; model_kernels.py:104-109): rUA = V / (a_PP rho), HbyA += rUA grad p,
 * cell_values[u,v,w] = HbyA.  HbyA is [3][C]; values is the momentum ELL
 * matrix (its diagonal a_PP). */
int wcfd_rua_hbya(const wcfd_mesh* m, const wcfd_real* values, float density,
                  const wcfd_real* cell_grad, wcfd_real* hbya, wcfd_real* cell_values, wcfd_real* rua,
                  void* stream);

/* FVM.relax (model.py:256-262,381-384): p = alpha p_new + (1-alpha) p;
 * alpha <= 0 means replace (model.py:147). */
int wcfd_relax_pressure(const wcfd_mesh* m, const wcfd_real* p_new, float alpha,
                        wcfd_real* cell_values, void* stream);

/* Corrector tail (incompressible.py:150-156): face_interpolation('p'),
 * calculate_gradients('p') (zeroes u,v,w gradients, T3), HbyA -= rUA grad p,
 * cell_values[u,v,w] = HbyA.  Also the float64 continuity residual
 * sum_c|sum_f +-F_c| / sum_c sum_f |F_c| of the corrected flux
 * F_c = F - rUA_f A_f (p_N - p_O)/|d|, written to resid[0..1]
 * (numerator, denominator) if resid != NULL; resid[2..3] are set to 0
 * (the SIMPLE step writes ||u^n - u^(n-1)||^2, ||u^n||^2 there).  Needs a
 * workspace of wcfd_reduce_workspace_bytes(F) bytes. */
int wcfd_correct_velocity(const wcfd_mesh* m, wcfd_bc bc, wcfd_real* cell_values,
                          wcfd_real* cell_grad, wcfd_real* face_values, const wcfd_real* mass_flux,
                          const wcfd_real* rua, const wcfd_real* rua_face, wcfd_real* hbya,
                          double* resid, void* workspace, void* stream);

size_t wcfd_reduce_workspace_bytes(int32_t n_faces);

/* Convergence report of an outer iteration: the norms FVM.check_convergence
 * (model.py:278-315) hands to Convergence/Criteria (Convergence.py:7-106),
 * computed on the device in float64 and written to out[0..8] (device):
 *   out[0..2] ||A u_k - b_k||_2, out[3..5] ||b_k||_2, k = x, y, z, for the
 *             momentum matrix `values` [7][C] (ELL, shared by the three
 *             components), u and b [3][C] (momentum_x/y/z);
 *   out[6]    max_k,c |c_k| with c_k = rua * corr_k (float32, as
 *             get_gradient('p', coeff = rUA)), or corr_k when rua == NULL;
 *             corr [3][C] (velocity_correction);
 *   out[7]    max |p|, p [C] (pressure_correction: the pressure solve result);
 *   out[8]    sum_c |div F|_c of the face fluxes `flux` [F] (continuity and
 *             local_continuity).
 * A NULL group (values, corr, p, flux) leaves its outputs 0.  Deterministic
 * (fixed reduction order).  workspace: wcfd_convergence_workspace_bytes(). */
int wcfd_convergence_report(const wcfd_mesh* m, const wcfd_real* values, const wcfd_real* u,
                            const wcfd_real* b, const wcfd_real* rua, const wcfd_real* corr, const wcfd_real* p,
                            const wcfd_real* flux, void* workspace, double* out, void* stream);
size_t wcfd_convergence_workspace_bytes(void);

/* The same report for every case of a batch arena (the per-case form of
 * model.py:278-315): out[9k .. 9k+8] (device) for case k, from the case's
 * own workgroups (bt's cell and workgroup ranges; finished cases, which have
 * none, report 0), so a case's report does not depend on the other cases.
 * workspace: wcfd_batch_convergence_workspace_bytes(bt->n_blocks). */
int wcfd_batch_convergence_report(const wcfd_mesh* m, const wcfd_batch* bt, const wcfd_real* values,
                                  const wcfd_real* u, const wcfd_real* b, const wcfd_real* rua,
                                  const wcfd_real* corr, const wcfd_real* p, const wcfd_real* flux,
                                  void* workspace, double* out, void* stream);
size_t wcfd_batch_convergence_workspace_bytes(int32_t n_blocks);

/* wcfd_batch_convergence_report with no workspace argument and no allocation
 * of any kind.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc.  The float64 partials live in a static
 * device array of WCFD_NOALLOC_MAX_BLOCKS x 9 doubles (1.18 MB) that each GPU
 * the library runs on holds until the process ends; its address is looked up
 * on the first call on each GPU (the caller's current device).  Returns
 * WCFD_ESCRATCH when bt->n_blocks > WCFD_NOALLOC_MAX_BLOCKS.  Results are
 * bitwise those of wcfd_batch_convergence_report (same kernels, same order).
 * Every call on one GPU shares the array: calls that overlap on one GPU (other
 * streams or host threads) race and must be serialised by the caller. */
#define WCFD_NOALLOC_MAX_BLOCKS 16384
int wcfd_batch_convergence_report_noalloc(const wcfd_mesh* m, const wcfd_batch* bt,
                                          const wcfd_real* values, const wcfd_real* u, const wcfd_real* b,
                                          const wcfd_real* rua, const wcfd_real* corr, const wcfd_real* p,
                                          const wcfd_real* flux, double* out, void* stream);

/* ------------------------------------------------------------------ Krylov */

/* y[q] = A x[q] for q < nvec, CSR (row_ptr, col, values) of size n. */
int wcfd_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const wcfd_real* values,
              int nvec, const wcfd_real* x, wcfd_real* y, void* stream);

/* float64 dot product of two float32 vectors, result to out[0] (device).
 * Workspace: wcfd_reduce_workspace_bytes(0) bytes. */
int wcfd_dot_f64(int32_t n, const wcfd_real* a, const wcfd_real* b, double* out, void* workspace,
                 void* stream);

/* y[q] = A x[q] for the ELL matrix `values` [7][C] of mesh m. */
int wcfd_spmv_ell(const wcfd_mesh* m, const wcfd_real* values, int nvec, const wcfd_real* x, wcfd_real* y,
                  void* stream);

size_t wcfd_krylov_workspace_bytes(int32_t n, int nvec);

/* Result of a Krylov solve (host struct). */
typedef struct wcfd_solve_info {
  int32_t iterations;
  int32_t converged;
  double residual;   /* ||r||_2 over all components */
  double threshold;  /* max(tol ||b||, atol) */
} wcfd_solve_info;

/* warp.optim.linear.cg with the Jacobi preconditioner, as called by
 * Equation.solve_Axb (equation.py:195-200; incompressible.py:32,142).
 * x is updated in place (warm start).  Stops when ||r|| <= max(tol ||b||,
 * atol), checked before the first iteration and every check_every
 * iterations, or at maxiter.  variant: WCFD_CG_*.  info is [host]. */
int wcfd_cg(int32_t n, const int32_t* row_ptr, const int32_t* col, const wcfd_real* values,
            const int32_t* diag_pos, int variant, const wcfd_real* b, wcfd_real* x, double atol,
            double tol, int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
            void* stream);

/* warp.optim.linear.bicgstab, right Jacobi preconditioning, as called by
 * Equation.solve_Axb for the momentum system (equation.py:195-200;
 * incompressible.py:116).  nvec right-hand sides b[nvec][n] share the matrix
 * and all scalars (the reference's single interleaved 3C solve). */
int wcfd_bicgstab(int32_t n, const int32_t* row_ptr, const int32_t* col, const wcfd_real* values,
                  const int32_t* diag_pos, int nvec, const wcfd_real* b, wcfd_real* x, double atol,
                  double tol, int maxiter, int check_every, void* workspace,
                  wcfd_solve_info* info, void* stream);

/* Launch shape of the Krylov kernels: `grid` workgroups (grid-stride,
 * 1..65536; 0 = one row per thread, i.e. ceil(n / block) workgroups) of
 * `block` threads (64..512, multiple of 64).  Every workgroup writes one
 * float64 partial per dot product.  When a case has at most `max_inline`
 * workgroups (wcfd_set_krylov_fold; default 2048), every workgroup of the
 * next kernel folds all of them in its prologue; above that, a fold kernel
 * (one workgroup per case) reduces them to one value first, so the fold
 * work grows linearly rather than quadratically with the grid.  Both forms
 * sum in the same order and give bitwise the same results.  Process-wide;
 * chosen by the autotuner (warpcfd_amd/tune.py).  No counterpart in the
 * reference (Warp picks its own launch shapes). */
int wcfd_set_krylov_launch(int grid, int block);
int wcfd_get_krylov_launch(int* grid, int* block); /* [host] outputs */
int wcfd_set_krylov_fold(int max_inline);           /* max_inline >= 0 */
int wcfd_get_krylov_fold(void);

/* Where the stop test of a single-case, single-reduction CG solve is made:
 * on (1, the default) the first iteration kernel of every check group makes
 * it in its prologue, from the partials it folds anyway; off (0) a check
 * kernel follows every group.  The iterations run, the solution and the
 * reported residual are bitwise the same either way; on saves one launch
 * per group.  Batched solves and fixed launch budgets always use the check
 * kernel.  Process-wide.  No counterpart in the reference (warp.optim.linear
 * tests the residual on the host every check_every iterations). */
int wcfd_set_krylov_inline_check(int on);           /* 0 or 1 */
int wcfd_get_krylov_inline_check(void);

/* Persistent single-reduction CG for solves whose widest case has at most
 * max_rows rows (its virtual workgroups times the chunk); 0, the default:
 * one launch per iteration always.  One cooperative launch runs the whole
 * solve.  Every case of a batch, and a single case of at most 8 192 rows,
 * sits on the CUs of ONE XCD (its workgroups share that XCD's L2 and meet at
 * a barrier in memory once per iteration), one row per thread (chunk <=
 * block), one or two virtual workgroups per workgroup.  A larger single case
 * spreads over every XCD in contiguous bands (a two-level barrier, agent-scope
 * stores), one to four rows per thread.  Fixed launch budgets and other
 * shapes take the per-iteration kernels.  The iterations, the solution and
 * the reported residual are bitwise those of the per-iteration kernels at
 * the same launch shape.  Measured on MI355X: on one XCD an iteration costs
 * about 4 us plus 0.2 us per thousand rows (one XCD's gather rate), against
 * 5.5-6 us per launch, so it pays below about 8 000 rows per case; the
 * 320 k-row channel over every XCD takes 8.06 us per iteration, as many as a
 * launch, and saves the kernel boundaries.  A launch whose
 * workgroups do not land as planned, or whose barrier waits longer than
 * 100 ms, aborts before x is written; the solve then runs on the
 * per-iteration kernels and the setting turns itself off.  Process-wide.  No
 * counterpart in the reference (warp.optim.linear launches per iteration). */
int wcfd_set_krylov_persistent(int max_rows);       /* max_rows >= 0 */
int wcfd_get_krylov_persistent(void);
/* Solves run by the persistent kernel and solves that fell back after an
 * abort, since the library was loaded. [host] outputs */
int wcfd_krylov_persistent_stats(int64_t* solves, int64_t* fallbacks);
/* Test hook for the abort path (no counterpart in the reference): launch the
 * persistent kernel with `slots` fewer workgroup slots per XCD than its plan
 * needs, so that every workgroup finds the plan does not fit, sets the abort
 * flag and returns at entry; the solve then falls back.  0 (the default): off. */
int wcfd_debug_krylov_persistent_shrink(int slots);

/* wcfd_cg / wcfd_bicgstab on the ELL matrix `values` [7][C] of mesh m. */
int wcfd_cg_ell(const wcfd_mesh* m, const wcfd_real* values, int variant, const wcfd_real* b, wcfd_real* x,
                double atol, double tol, int maxiter, int check_every, void* workspace,
                wcfd_solve_info* info, void* stream);
int wcfd_bicgstab_ell(const wcfd_mesh* m, const wcfd_real* values, int nvec, const wcfd_real* b, wcfd_real* x,
                      double atol, double tol, int maxiter, int check_every, void* workspace,
                      wcfd_solve_info* info, void* stream);

/* ------------------------------------------------------ whole SIMPLE step */

typedef struct wcfd_simple_params {
  float nu;            /* kinematic viscosity mu/rho (incompressible.py:89) */
  float density;
  float alpha_u;       /* <= 0: no momentum relaxation */
  float alpha_p;       /* <= 0: replace p */
  float dt;            /* > 0: transient (DdtTerm); 0: steady */
  int32_t scheme_flags; /* WCFD_SCHEME_* bits */
  int32_t ref_cell;    /* device cell id of the reference-pressure row, -1 none */
  float ref_value;
  int32_t n_inner;     /* NUM_INNER_LOOPS */
  int32_t n_ortho;     /* NUM_ORTHOGONAL_CORRECTORS */
  double atol, tol;    /* Krylov tolerances: stop at ||r|| <= max(tol ||b||, rtol0 ||r_0||, atol) */
  double rtol0;        /* relative to the warm-start residual r_0; 0 = off (the reference) */
  int32_t maxiter, check_every;
  int32_t first_iteration; /* 1: u,v,w gradients may be non-zero on entry */
  int32_t cg_variant;      /* WCFD_CG_* for the pressure equation */
  int32_t check_every_u;   /* check interval of the momentum solve; 0: check_every */
  int32_t p_refine;        /* 0: the reference's pressure solve; n (<= 8): mixed-precision
                              iterative refinement, 1 + n warm-started CG sweeps, each from
                              a residual b - A p formed in float64 */
  int32_t graph_groups_u;  /* 0: the Krylov solves poll the host (adaptive).  > 0 (both):
                              fixed launch budget for stream capture -- the momentum and */
  int32_t graph_groups_p;  /* pressure solves launch exactly this many check groups, with no
                              event or host read (maxiter clipped to groups * check_every;
                              cases done earlier exit at entry).  Pass info = NULL. */
  int32_t maxiter_u;       /* momentum-solve iteration cap; 0: maxiter (the reference) */
} wcfd_simple_params;

/* Device state of one case. */
typedef struct wcfd_simple_state {
  wcfd_real* cell_values;  /* [4][C] */
  wcfd_real* cell_grad;    /* [4][3][C] */
  wcfd_real* face_values;  /* [4][F] */
  wcfd_real* mass_flux;    /* [F] */
  wcfd_real* hbya;         /* [3][C] momentum solution / HbyA */
  wcfd_real* p_x;          /* [C] pressure solution (warm start) */
  wcfd_real* rua;          /* [C] */
  wcfd_real* rua_face;     /* [F] */
  wcfd_real* mom_values;   /* [7][C] ELL */
  wcfd_real* mom_rhs;      /* [3][C] */
  wcfd_real* p_values;     /* [7][C] ELL */
  wcfd_real* p_rhs;        /* [C] */
  double* resid;       /* [4] (batch: [K][4]) continuity numerator, denominator,
                          ||u^n - u^(n-1)||^2, ||u^n||^2 */
} wcfd_simple_state;

typedef struct wcfd_simple_info {
  wcfd_solve_info momentum;
  wcfd_solve_info pressure; /* last corrector */
  double continuity;        /* numerator / denominator */
  double velocity_change;   /* ||u^n - u^(n-1)|| / ||u^n|| */
} wcfd_simple_info;

size_t wcfd_simple_workspace_bytes(int32_t n_cells, int32_t n_faces);

/* One outer iteration of IncompressibleSolver.run (incompressible.py:95-156)
 * on the device.  info is [host] and may be NULL (then no host sync at the
 * end; Krylov checks still sync every check_every iterations). */
int wcfd_simple_iteration(const wcfd_mesh* m, wcfd_bc bc, const wcfd_simple_params* prm,
                          const wcfd_simple_state* st, void* workspace,
                          wcfd_simple_info* info, void* stream);

/* ------------------------------------------------------------ batched path
 * The same operations on a wcfd_batch (declared above).  Every case has its
 * own Krylov scalars, stopping test and iteration count; `info` arrays are
 * [host] [K].  Workspaces are sized for (C, K, G). */
size_t wcfd_batch_krylov_workspace_bytes(int32_t n, int32_t n_cases, int32_t n_blocks);
int wcfd_batch_cg_ell(const wcfd_mesh* m, const wcfd_batch* b, const wcfd_real* values, int variant,
                      const wcfd_real* rhs, wcfd_real* x, double atol, double tol, int maxiter,
                      int check_every, void* workspace, wcfd_solve_info* info, void* stream);
int wcfd_batch_bicgstab_ell(const wcfd_mesh* m, const wcfd_batch* b, const wcfd_real* values,
                            int nvec, const wcfd_real* rhs, wcfd_real* x, double atol, double tol,
                            int maxiter, int check_every, void* workspace,
                            wcfd_solve_info* info, void* stream);

size_t wcfd_batch_simple_workspace_bytes(int32_t n_cells, int32_t n_faces, int32_t n_cases,
                                         int32_t n_blocks);

/* One SIMPLE outer iteration of every active case of the batch
 * (IncompressibleSolver.run, incompressible.py:95-156, per case).  prm's
 * nu, ref_cell and ref_value are ignored (per case in the batch).
 * st->resid is [K][4].  After the corrector each active case's monitor
 * record is updated (wcfd_monitor_update with b->stop[k]); converged cases
 * get active = 0 on the device, with no host round trip, and
 * b->outer_iters[k] counts the iterations each case ran.  info is [host]
 * [K] or NULL (no sync at the end). */
int wcfd_batch_simple_iteration(const wcfd_mesh* m, const wcfd_batch* b, wcfd_bc bc,
                                const wcfd_simple_params* prm, const wcfd_simple_state* st,
                                void* workspace, wcfd_simple_info* info, void* stream);

/* ---------------------------------------------------------------- profiler
 * No counterpart in the reference (it has no timers).  While active, every
 * kernel launch of the library records a start/stop hipEvent pair on its
 * stream (up to `capacity` launches); wcfd_profile_end() waits for them and
 * returns per-kernel total milliseconds and launch counts, indexed by kernel
 * id (names from wcfd_kernel_name). */
int wcfd_kernel_count(void);
const char* wcfd_kernel_name(int id); /* [host] static string */
int wcfd_profile_begin(int capacity);
int wcfd_profile_end(double* total_ms, int32_t* counts, int n_ids); /* [host] arrays */
/* Krylov solves keep launching while the host has not yet seen their stop;
 * those launches exit at entry.  wcfd_profile_end counts only launches that
 * did an iteration's work (a polled solve retires, when it returns, its
 * launches at or beyond the iteration count it ran); this returns, per
 * kernel id, how many launches of the last profiled region were left out. */
int wcfd_profile_skipped(int32_t* counts, int n_ids); /* [host] array */
/* The launches recorded so far (before wcfd_profile_end), in launch order:
 * kernel id (-1 - id for a launch that exited at entry) and the begin and end
 * of each, in ms from the first launch's begin.  Returns the number written
 * (at most capacity), or minus an error code. */
int wcfd_profile_timeline(int32_t* ids, double* start_ms, double* end_ms,
                          int capacity); /* [host] arrays */

/* Face matching of the mesh preprocessing, on the host (no GPU work; host
 * scratch of about 20 bytes per slot plus 4 per node, freed on return).
 * Replaces the unique-face numbering and the pairing of shared faces in
 * Mesh.get_mesh_properties (warp_cfd/preprocess/mesh.py:108-148 and
 * 220-302, np.unique over the sorted node ids and the owner/neighbour pair
 * search).  face_nodes [host] (n_cells, faces_per_cell, nodes_per_face)
 * node ids, nodes_per_face <= 4, one face set (every face of the set has
 * that many nodes).  Writes face_of_slot [host] (n_cells * faces_per_cell):
 * the face id of each (cell, local face) slot; unique_nodes [host]
 * (capacity, nodes_per_face): each face's node ids in ascending order,
 * faces numbered in the lexicographic order of those rows; face_slots
 * [host] (capacity, 2): the face's slots c * faces_per_cell + k in
 * ascending order, -1 in the second for a boundary face.  Returns the
 * number of faces, or minus WCFD_EINVAL (bad sizes, a negative node id),
 * WCFD_EMESH (a face held by three or more slots) or WCFD_ECAPACITY (more
 * than `capacity` faces). */
int64_t wcfd_match_faces(const int32_t* face_nodes, int64_t n_cells, int32_t faces_per_cell,
                         int32_t nodes_per_face, int32_t* face_of_slot, int32_t* unique_nodes,
                         int32_t* face_slots, int64_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* WCFD_H */
