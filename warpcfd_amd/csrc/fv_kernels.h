// Internal launchers shared by the wcfd translation units.
#pragma once
#include "common.h"

namespace wcfd {

// Laplacian factor of internal face f: 1/|d| (laplacianScheme.py:14-37), or
// with `correction` the over-relaxed |n1|/|d|, n1 = d / (n_f . d)
// (laplacianScheme.py:39-66).
__device__ inline real lap_factor(const wcfd_mesh& m, int f, int correction) {
  const real dm = m.dmag[f];
  if (!correction) return 1.0f / dm;
  const int Fi = m.n_internal, F = m.n_faces;
  const real d0 = m.dist[f], d1 = m.dist[Fi + f], d2 = m.dist[2 * Fi + f];
  const real n0 = m.normal[f], n1 = m.normal[F + f], n2 = m.normal[2 * F + f];
  const real inv_nd = 1.0f / (n0 * d0 + n1 * d1 + n2 * d2);
  const real e0 = d0 * inv_nd, e1 = d1 * inv_nd, e2 = d2 * inv_nd;
  const real n1m = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  return 1.0f / dm * n1m;
}

// The same factor with the face arrays read through buffer descriptors at
// byte offset f4 = 4 f (dmag, and dist / normal planes when corrected).
__device__ inline real lap_factor_b(const wcfd_mesh& m, const FBuf& dmag, uint32_t f4,
                                     int correction) {
  const real dm = dmag(f4);
  if (!correction) return 1.0f / dm;
  const uint32_t Fi4 = (uint32_t)m.n_internal * 4u, F4 = (uint32_t)m.n_faces * 4u;
  const FBuf dist(m.dist, 3u * Fi4), nrm(m.normal, 3u * F4);
  const real d0 = dist(f4), d1 = dist(Fi4 + f4), d2 = dist(2u * Fi4 + f4);
  const real n0 = nrm(f4), n1 = nrm(F4 + f4), n2 = nrm(2u * F4 + f4);
  const real inv_nd = 1.0f / (n0 * d0 + n1 * d1 + n2 * d2);
  const real e0 = d0 * inv_nd, e1 = d1 * inv_nd, e2 = d2 * inv_nd;
  const real n1m = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  return 1.0f / dm * n1m;
}

// a mesh descriptor with every array the face and cell kernels read
bool mesh_ok(const wcfd_mesh* m);

// Face and cell operators (fv_kernels.hip).  `bd` carries a batch's per-cell
// and per-face case lookups (all NULL for a single case).
int launch_set_bc(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, real* fv, int mask,
                  hipStream_t s);
int launch_face_interp(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* cv,
                       const real* cg, real* fv, int mask, int skew_mask, real* flux,
                       const real* rua, real* rua_f, int apply_dirichlet, hipStream_t s);
int launch_green_gauss(const wcfd_mesh& m, const BatchDev& bd, const real* fv, real* cg,
                       int mask, int zero_others, hipStream_t s);
int launch_assemble_momentum(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* cv,
                             const real* fv, const real* flux, real* cg, int compute_grad,
                             real nu, real inv_density, real alpha, real dt, int flags,
                             real* values, real* rhs, hipStream_t s);
int launch_assemble_pressure(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* fv,
                             const real* flux, const real* rua_f, real* cg, int compute_grad_p,
                             int correction, int ref_cell, real ref_value, real* values,
                             real* rhs, real* div_out, hipStream_t s);
// u_prev (may be NULL): receives u of the previous outer iteration
int launch_rua_hbya(const wcfd_mesh& m, const BatchDev& bd, const real* values, real density,
                    const real* cg, real* hbya, real* cv, real* rua, real* u_prev,
                    hipStream_t s);
int launch_relax_p(const wcfd_mesh& m, const BatchDev& bd, const real* pn, real alpha, real* cv,
                   hipStream_t s);
// Corrector tail.  sg segments the cells for the float64 partials (one case,
// or the batch); resid gets [K][4] (continuity numerator and denominator,
// ||u - u_prev||^2, ||u||^2; the last two 0 without u_prev).  With a batch
// (`bt`), every active case's iteration count and convergence monitor are
// updated and converged cases are deactivated.
int launch_correct_velocity(const wcfd_mesh& m, const BatchDev& bd, const Seg& sg, int block,
                            const wcfd_batch* bt, wcfd_bc bc, real* cv, real* cg, real* fv,
                            const real* flux, const real* rua, const real* rua_f, real* hbya,
                            const real* u_prev, double* resid, real* fc_scratch,
                            double* partials, hipStream_t s);

// Krylov solvers (krylov.hip).  Stopping threshold
// max(tol ||b||, rtol0 ||r_0||, atol); the reference's call has tol = atol
// and rtol0 = 0.  refine > 0 (CG only): that many extra warm-started sweeps,
// every sweep starting from a float64 residual.  fixed_groups > 0: launch
// exactly that many check groups without polling (stream capture), maxiter
// clipped to fixed_groups * check_every.
struct Tol {
  double atol, tol, rtol0;
  int refine;
  int fixed_groups;
};

struct CsrView {
  int32_t n;
  const int32_t* row_ptr;
  const int32_t* col;
  const real* values;
  const int32_t* diag_pos;
};

int solve_cg(const CsrView& A, int variant, const real* b, real* x, const Tol& tl, int maxiter,
             int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s);
int solve_bicgstab(const CsrView& A, int nvec, const real* b, real* x, const Tol& tl,
                   int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                   hipStream_t s);
// `bt` NULL: one case, launch shape from wcfd_set_krylov_launch.  info: [K].
int solve_cg_ell(const wcfd_mesh& m, const wcfd_batch* bt, const real* values, int variant,
                 const real* b, real* x, const Tol& tl, int maxiter, int check_every,
                 void* workspace, wcfd_solve_info* info, hipStream_t s);
int solve_bicgstab_ell(const wcfd_mesh& m, const wcfd_batch* bt, const real* values, int nvec,
                       const real* b, real* x, const Tol& tl, int maxiter, int check_every,
                       void* workspace, wcfd_solve_info* info, hipStream_t s);
size_t krylov_workspace_bytes(int32_t n, int nvec);
size_t krylov_workspace_bytes_batch(int32_t n, int nvec, int n_cases, int max_blocks);
bool batch_ok(const wcfd_mesh* m, const wcfd_batch* bt);

}  // namespace wcfd
