// One SIMPLE / PISO outer iteration on the device (incompressible.py:95-156),
// for one case or for every active case of a batch.
//
// Kernel sequence (each line one launch unless noted):
//   face values of u,v,w,p (+ Dirichlet BCs) and face fluxes
//   Green-Gauss gradients of all fields + momentum matrix / 3 rhs / relaxation
//   BiCGSTAB on the shared-matrix 3-vector momentum system     (3 launches / it.)
//   rUA, HbyA = u* + rUA grad p, cell values <- HbyA
//   per corrector:
//     face values + fluxes of HbyA, rUA_f
//     grad p + divergence + pressure matrix / rhs (+ reference row)
//     CG on the pressure equation                       (1 or 2 launches / it.)
//     p relaxation
//     face p + corrected flux F_c ; grad p, HbyA -= rUA grad p, continuity
//     residual per case (+ the batch's stop test)
#include <math.h>

#include <vector>

#include "fv_kernels.h"

namespace wcfd {

struct SimpleWork {
  void* krylov;
  double* partials;
  real* fc;
  real* u_prev;
};

static SimpleWork carve_simple(void* ws, int32_t C, int32_t F, int K, int G, bool batch) {
  Carver c(ws);
  SimpleWork w;
  w.krylov = c.take<char>(batch ? krylov_workspace_bytes_batch(C, 3, K, G)
                                : krylov_workspace_bytes(C, 3));
  w.partials = c.take<double>(4 * (size_t)(batch ? G : cell_reduce_blocks(C)));
  w.fc = c.take<real>(F);
  w.u_prev = c.take<real>(3 * (size_t)C);
  return w;
}

static size_t simple_bytes(int32_t C, int32_t F, int K, int G, bool batch) {
  Carver c(nullptr);
  c.take<char>(batch ? krylov_workspace_bytes_batch(C, 3, K, G) : krylov_workspace_bytes(C, 3));
  c.take<double>(4 * (size_t)(batch ? G : cell_reduce_blocks(C)));
  c.take<real>(F);
  c.take<real>(3 * (size_t)C);
  return c.off;
}

static int simple_iteration(const wcfd_mesh& m, const wcfd_batch* bt, wcfd_bc bc,
                            const wcfd_simple_params* prm, const wcfd_simple_state* st,
                            void* workspace, wcfd_simple_info* info, hipStream_t s) {
  const int K = bt ? bt->n_cases : 1;
  const int Gb = bt ? bt->n_blocks : 0;
  SimpleWork w = carve_simple(workspace, m.n_cells, m.n_faces, K, Gb, bt != nullptr);
  const BatchDev bd = batch_dev(bt);
  // continuity partials: the batch's segmentation, or one case over kBlock
  const Seg sg = bt ? batch_seg(*bt, m.n_cells)
                    : single_seg(m.n_cells, kBlock, (int)cell_reduce_blocks(m.n_cells));
  const int sg_block = bt ? bt->block : kBlock;
  std::vector<wcfd_solve_info> mom(K), pres(K);
  int e;

  // u,v,w gradients are zero on entry unless this is the first iteration
  // (calculate_gradients('p') zeroes them, model.py:237, T3)
  const int skew_entry = prm->first_iteration ? 15 : 8;
  if ((e = launch_face_interp(m, bd, bc, st->cell_values, st->cell_grad, st->face_values, 15,
                              skew_entry, st->mass_flux, nullptr, nullptr, 1, s)))
    return e;
  if ((e = launch_assemble_momentum(m, bd, bc, st->cell_values, st->face_values, st->mass_flux,
                                    st->cell_grad, 1, prm->nu, 1.0f / prm->density, prm->alpha_u,
                                    prm->dt, prm->scheme_flags, st->mom_values, st->mom_rhs, s)))
    return e;
  const Tol tl{prm->atol, prm->tol, prm->rtol0, 0, prm->graph_groups_u};
  const Tol tp{prm->atol, prm->tol, prm->rtol0, prm->p_refine, prm->graph_groups_p};
  const int ce_u = prm->check_every_u > 0 ? prm->check_every_u : prm->check_every;
  const int mi_u = prm->maxiter_u > 0 ? prm->maxiter_u : prm->maxiter;
  e = solve_bicgstab_ell(m, bt, st->mom_values, 3, st->mom_rhs, st->hbya, tl, mi_u, ce_u,
                         w.krylov, mom.data(), s);
  if (e) return e;
  if ((e = launch_rua_hbya(m, bd, st->mom_values, prm->density, st->cell_grad, st->hbya,
                           st->cell_values, st->rua, w.u_prev, s)))
    return e;
  const int n_inner = prm->n_inner > 0 ? prm->n_inner : 1;
  const int n_ortho = prm->n_ortho > 0 ? prm->n_ortho : 1;
  for (int ii = 0; ii < n_inner; ++ii) {
    for (int io = 0; io < n_ortho; ++io) {
      const bool last = (ii == n_inner - 1) && (io == n_ortho - 1);
      const int skew = (ii == 0 && io == 0) ? 15 : 8;
      if ((e = launch_face_interp(m, bd, bc, st->cell_values, st->cell_grad, st->face_values, 15,
                                  skew, st->mass_flux, st->rua, st->rua_face, 1, s)))
        return e;
      if ((e = launch_assemble_pressure(m, bd, bc, st->face_values, st->mass_flux, st->rua_face,
                                        st->cell_grad, 1, prm->scheme_flags & WCFD_SCHEME_CORRECTED,
                                        prm->ref_cell, prm->ref_value, st->p_values, st->p_rhs,
                                        nullptr, s)))
        return e;
      e = solve_cg_ell(m, bt, st->p_values, prm->cg_variant, st->p_rhs, st->p_x, tp,
                       prm->maxiter, prm->check_every, w.krylov, pres.data(), s);
      if (e) return e;
      if ((e = launch_relax_p(m, bd, st->p_x, prm->alpha_p, st->cell_values, s))) return e;
      // the batch's stop test and iteration count run after the last corrector only
      if ((e = launch_correct_velocity(m, bd, sg, sg_block, last ? bt : nullptr, bc,
                                       st->cell_values, st->cell_grad, st->face_values,
                                       st->mass_flux, st->rua, st->rua_face, st->hbya, w.u_prev,
                                       st->resid, w.fc, w.partials, s)))
        return e;
    }
  }
  if (info) {
    for (int k = 0; k < K; ++k) {
      info[k].momentum = mom[k];
      info[k].pressure = pres[k];
      info[k].continuity = 0.0;
      info[k].velocity_change = 0.0;
    }
    if (st->resid) {
      std::vector<double> r(4 * (size_t)K);
      if ((e = check_hip(hipMemcpyAsync(r.data(), st->resid, sizeof(double) * 4 * K,
                                        hipMemcpyDeviceToHost, s))))
        return e;
      if ((e = check_hip(hipStreamSynchronize(s)))) return e;
      for (int k = 0; k < K; ++k) {
        const double* q = r.data() + 4 * k;
        info[k].continuity = q[1] > 0.0 ? q[0] / q[1] : 0.0;
        info[k].velocity_change = q[3] > 0.0 ? sqrt(q[2] / q[3]) : 0.0;
      }
    }
  }
  return WCFD_OK;
}

static bool params_ok(const wcfd_simple_params* prm) {
  return prm && prm->density != 0.f && prm->check_every >= 0 && prm->check_every_u >= 0 &&
         prm->rtol0 >= 0.0 && prm->p_refine >= 0 && prm->p_refine <= 8 &&
         prm->graph_groups_u >= 0 && prm->graph_groups_p >= 0 && prm->maxiter_u >= 0 &&
         (prm->graph_groups_u > 0) == (prm->graph_groups_p > 0) &&
         (prm->cg_variant == WCFD_CG_STANDARD || prm->cg_variant == WCFD_CG_SINGLE_REDUCTION);
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

size_t wcfd_simple_workspace_bytes(int32_t n_cells, int32_t n_faces) {
  if (n_cells <= 0 || n_faces <= 0) return 0;
  return simple_bytes(n_cells, n_faces, 1, 0, false);
}

size_t wcfd_batch_simple_workspace_bytes(int32_t n_cells, int32_t n_faces, int32_t n_cases,
                                         int32_t n_blocks) {
  if (n_cells <= 0 || n_faces <= 0 || n_cases <= 0 || n_blocks <= 0) return 0;
  return simple_bytes(n_cells, n_faces, n_cases, n_blocks, true);
}

int wcfd_simple_iteration(const wcfd_mesh* mp, wcfd_bc bc, const wcfd_simple_params* prm,
                          const wcfd_simple_state* st, void* workspace, wcfd_simple_info* info,
                          void* stream) {
  if (!mp || !params_ok(prm) || !st || !workspace || !bc.type || !bc.value || !mp->slot_nbr ||
      !slots_ok(mp))
    return WCFD_EINVAL;
  const wcfd_mesh& m = *mp;
  if (m.n_cells <= 0 || m.n_faces < m.n_internal) return WCFD_EINVAL;
  if (prm->ref_cell >= m.n_cells) return WCFD_EINVAL;
  return simple_iteration(m, nullptr, bc, prm, st, workspace, info, (hipStream_t)stream);
}

int wcfd_batch_simple_iteration(const wcfd_mesh* mp, const wcfd_batch* bt, wcfd_bc bc,
                                const wcfd_simple_params* prm, const wcfd_simple_state* st,
                                void* workspace, wcfd_simple_info* info, void* stream) {
  if (!batch_ok(mp, bt) || !params_ok(prm) || !st || !workspace || !bc.type || !bc.value ||
      !mp->slot_nbr || !slots_ok(mp) || !st->resid)
    return WCFD_EINVAL;
  if (!bt->cell_case || !bt->face_case || !bt->nu || !bt->ref_cell || !bt->ref_value ||
      !bt->stop || !bt->monitor || !bt->outer_iters)
    return WCFD_EINVAL;
  if (mp->n_faces < mp->n_internal) return WCFD_EINVAL;
  return simple_iteration(*mp, bt, bc, prm, st, workspace, info, (hipStream_t)stream);
}

}  // extern "C"
