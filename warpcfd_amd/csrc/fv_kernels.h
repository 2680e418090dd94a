// Internal launchers shared by the wcfd translation units.
#pragma once
#include "common.h"

namespace wcfd {

int launch_set_bc(const wcfd_mesh& m, wcfd_bc bc, float* fv, int mask, hipStream_t s);
int launch_face_interp(const wcfd_mesh& m, wcfd_bc bc, const float* cv, const float* cg, float* fv,
                       int mask, int skew_mask, float* flux, const float* rua, float* rua_f,
                       int apply_dirichlet, hipStream_t s);
int launch_green_gauss(const wcfd_mesh& m, const float* fv, float* cg, int mask, int zero_others,
                       hipStream_t s);
int launch_assemble_momentum(const wcfd_mesh& m, wcfd_bc bc, const float* cv, const float* fv,
                             const float* flux, float* cg, int compute_grad, float nu,
                             float inv_density, float alpha, float dt, int correction,
                             float* values, float* rhs, hipStream_t s);
int launch_assemble_pressure(const wcfd_mesh& m, wcfd_bc bc, const float* fv, const float* flux,
                             const float* rua_f, float* cg, int compute_grad_p, int correction,
                             int ref_cell, float ref_value, float* values, float* rhs,
                             float* div_out, hipStream_t s);
int launch_rua_hbya(const wcfd_mesh& m, const float* values, float density, const float* cg,
                    float* hbya, float* cv, float* rua, hipStream_t s);
int launch_relax_p(const wcfd_mesh& m, const float* pn, float alpha, float* cv, hipStream_t s);
int launch_correct_velocity(const wcfd_mesh& m, wcfd_bc bc, float* cv, float* cg, float* fv,
                            const float* flux, const float* rua, const float* rua_f, float* hbya,
                            double* resid, float* fc_scratch, double* partials, hipStream_t s);

// Krylov solvers (krylov.hip)
struct CsrView {
  int32_t n;
  const int32_t* row_ptr;
  const int32_t* col;
  const float* values;
  const int32_t* diag_pos;
};

int solve_cg(const CsrView& A, int variant, const float* b, float* x, double atol, double tol,
             int maxiter, int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s);
int solve_bicgstab(const CsrView& A, int nvec, const float* b, float* x, double atol, double tol,
                   int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                   hipStream_t s);
int solve_cg_ell(const wcfd_mesh& m, const float* values, int variant, const float* b, float* x,
                 double atol, double tol, int maxiter, int check_every, void* workspace,
                 wcfd_solve_info* info, hipStream_t s);
int solve_bicgstab_ell(const wcfd_mesh& m, const float* values, int nvec, const float* b, float* x,
                       double atol, double tol, int maxiter, int check_every, void* workspace,
                       wcfd_solve_info* info, hipStream_t s);
size_t krylov_workspace_bytes(int32_t n, int nvec);

}  // namespace wcfd
