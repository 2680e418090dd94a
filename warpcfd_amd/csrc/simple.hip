// One SIMPLE / PISO outer iteration on the device (incompressible.py:95-156).
//
// Kernel sequence (each line one launch unless noted):
//   face values of u,v,w,p (+ Dirichlet BCs) and face fluxes
//   Green-Gauss gradients of all fields + momentum matrix / 3 rhs / relaxation
//   BiCGSTAB on the shared-matrix 3-vector momentum system     (3 launches / it.)
//   rUA, HbyA = u* + rUA grad p, cell values <- HbyA
//   per corrector:
//     face values + fluxes of HbyA, rUA_f
//     grad p + divergence + pressure matrix / rhs (+ reference row)
//     CG on the pressure equation                                (2 launches / it.)
//     p relaxation
//     face p + corrected flux F_c ; grad p, HbyA -= rUA grad p, continuity residual
#include "fv_kernels.h"

namespace wcfd {

struct SimpleWork {
  void* krylov;
  double* partials;
  float* fc;
};

static SimpleWork carve_simple(void* ws, int32_t C, int32_t F) {
  Carver c(ws);
  SimpleWork w;
  w.krylov = c.take<char>(krylov_workspace_bytes(C, 3));
  w.partials = c.take<double>(2 * kMaxReduceBlocks);
  w.fc = c.take<float>(F);
  return w;
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

size_t wcfd_simple_workspace_bytes(int32_t n_cells, int32_t n_faces) {
  if (n_cells <= 0 || n_faces <= 0) return 0;
  Carver c(nullptr);
  c.take<char>(krylov_workspace_bytes(n_cells, 3));
  c.take<double>(2 * kMaxReduceBlocks);
  c.take<float>(n_faces);
  return c.off;
}

int wcfd_simple_iteration(const wcfd_mesh* mp, wcfd_bc bc, const wcfd_simple_params* prm,
                          const wcfd_simple_state* st, void* workspace, wcfd_simple_info* info,
                          void* stream) {
  if (!mp || !prm || !st || !workspace || !bc.type || !bc.value || !mp->slot_nbr) return WCFD_EINVAL;
  const wcfd_mesh& m = *mp;
  if (m.n_cells <= 0 || m.n_faces < m.n_internal || prm->density == 0.f) return WCFD_EINVAL;
  if (prm->ref_cell >= m.n_cells) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  SimpleWork w = carve_simple(workspace, m.n_cells, m.n_faces);
  int e;
  wcfd_simple_info local{};
  wcfd_simple_info* inf = info ? info : &local;

  // u,v,w gradients are zero on entry unless this is the first iteration
  // (calculate_gradients('p') zeroes them, model.py:237, T3)
  const int skew_entry = prm->first_iteration ? 15 : 8;
  if ((e = launch_face_interp(m, bc, st->cell_values, st->cell_grad, st->face_values, 15,
                              skew_entry, st->mass_flux, nullptr, nullptr, 1, s)))
    return e;
  if ((e = launch_assemble_momentum(m, bc, st->cell_values, st->face_values, st->mass_flux,
                                    st->cell_grad, 1, prm->nu, 1.0f / prm->density, prm->alpha_u,
                                    prm->dt, prm->scheme_flags, st->mom_values, st->mom_rhs, s)))
    return e;
  e = solve_bicgstab_ell(m, st->mom_values, 3, st->mom_rhs, st->hbya, prm->atol, prm->tol,
                         prm->maxiter, prm->check_every, w.krylov, &inf->momentum, s);
  if (e) return e;
  if ((e = launch_rua_hbya(m, st->mom_values, prm->density, st->cell_grad, st->hbya,
                           st->cell_values, st->rua, s)))
    return e;
  const int n_inner = prm->n_inner > 0 ? prm->n_inner : 1;
  const int n_ortho = prm->n_ortho > 0 ? prm->n_ortho : 1;
  for (int ii = 0; ii < n_inner; ++ii) {
    for (int io = 0; io < n_ortho; ++io) {
      const int skew = (ii == 0 && io == 0) ? 15 : 8;
      if ((e = launch_face_interp(m, bc, st->cell_values, st->cell_grad, st->face_values, 15, skew,
                                  st->mass_flux, st->rua, st->rua_face, 1, s)))
        return e;
      if ((e = launch_assemble_pressure(m, bc, st->face_values, st->mass_flux, st->rua_face,
                                        st->cell_grad, 1, prm->scheme_flags & WCFD_SCHEME_CORRECTED,
                                        prm->ref_cell,
                                        prm->ref_value, st->p_values, st->p_rhs, nullptr, s)))
        return e;
      e = solve_cg_ell(m, st->p_values, prm->cg_variant, st->p_rhs, st->p_x, prm->atol, prm->tol,
                       prm->maxiter, prm->check_every, w.krylov, &inf->pressure, s);
      if (e) return e;
      if ((e = launch_relax_p(m, st->p_x, prm->alpha_p, st->cell_values, s))) return e;
      if ((e = launch_correct_velocity(m, bc, st->cell_values, st->cell_grad, st->face_values,
                                       st->mass_flux, st->rua, st->rua_face, st->hbya, st->resid,
                                       w.fc, w.partials, s)))
        return e;
    }
  }
  if (info && st->resid) {
    double r[2];
    if ((e = check_hip(hipMemcpyAsync(r, st->resid, sizeof(r), hipMemcpyDeviceToHost, s)))) return e;
    if ((e = check_hip(hipStreamSynchronize(s)))) return e;
    info->continuity = r[1] > 0.0 ? r[0] / r[1] : 0.0;
  }
  return WCFD_OK;
}

}  // extern "C"
