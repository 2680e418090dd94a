// Term weights and equation assembly of the composable FV API (gfx950):
// the reference's ConvectionTerm / DiffusionTerm / DdtTerm / GradTerm
// (warp_cfd/FV/terms/*.py with Implicit_Schemes/*.py) and Equation
// (terms/equation.py, kernels/equation_kernels.py).  The SIMPLE step uses the
// fused assembly kernels of fv_kernels.hip instead; these kernels serve
// callers that build their own equations term by term.
//
// Device layouts:
//   face-based weights  w[O][3][6][C]: output o, part j (0 owner-cell
//                        coefficient, 1 neighbour coefficient, 2 explicit),
//                        local face k of cell c -- the reference's
//                        Term.weights (C, 6, O, 3);
//   cell-based weights  w[O][2][C]: (implicit diagonal, explicit);
//   equation            values[O][7][C] (one ELL plane per output: the
//                        reference's BSR rows O*c + o), rhs[O][C].
// Every kernel is cell-centric (one thread per cell, its six slots in local
// face order), so no atomics are needed and sums run in the reference's
// per-row order.
#include "fv_kernels.h"

namespace wcfd {

// local output o -> global field g[o]
struct Outs {
  int n;
  int g[4];
};

__device__ inline size_t widx(int o, int j, int k, int c, int C) {
  return (((size_t)o * 3 + j) * 6 + k) * (size_t)C + c;
}

// --------------------------------------------------------- convection
// convection.py:44-88 with convectionScheme.py: central_difference (3-26),
// upwind (28-52, tie -> neighbour, T6), upwindLinear (54-88).  Internal faces
// also get the skew correction of the LOCAL output index o (convection.py:
// 72-74 passes `output`), as explicit part.  upwindLinear's explicit part is
// grad(phi)_upwind . r, r from the face's precomputed vectors fc2f[2][3][Fi]:
// [0] = x_f - x_owner, [1] = the owner's centroid-to-face vector of the
// neighbour's local face index (the reference reads
// owner_cell.cell_centroid_to_face_centroid[j[1]] when the neighbour is
// upwind, convectionScheme.py:82).  Boundary faces: F_b phi_b explicit, or,
// with implicit_outflow on a Neumann outflow face, F_b implicit plus
// F_b |x_f - x_P| g explicit (not in the reference).
__global__ void k_convection_weights(wcfd_mesh m, wcfd_bc bc, const real* __restrict__ cg,
                                     const real* __restrict__ fv,
                                     const real* __restrict__ flux,
                                     const real* __restrict__ fc2f, int scheme,
                                     int implicit_outflow, Outs outs, real* __restrict__ w) {
  const int c = xcd_item();
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal, Fb = m.n_boundary;
  if (c >= C) return;
  for (int k = 0; k < 6; ++k) {
    const int f = m.slot_face[k * C + c];
    const real Fl = flux[f];
    if (f >= Fi) {
      const int j = f - Fi;
      for (int o = 0; o < outs.n; ++o) {
        const int go = outs.g[o];
        real w0 = 0.f, w2;
        if (implicit_outflow && Fl > 0.f && bc.type[go * Fb + j] == WCFD_NEUMANN) {
          w0 = Fl;
          w2 = Fl * (m.dmag[f] * bc.value[go * Fb + j]);
        } else {
          w2 = Fl * fv[(size_t)go * F + f];
        }
        w[widx(o, 0, k, c, C)] = w0;
        w[widx(o, 1, k, c, C)] = 0.f;
        w[widx(o, 2, k, c, C)] = w2;
      }
      continue;
    }
    const int on = m.owner[f], nn = m.neigh[f];
    const real psi = m.psi[f];
    const real psi2 = 1.0f - psi;
    const bool up = Fl > 0.f;
    real wo, wn;
    if (scheme == WCFD_CONV_CENTRAL) {
      wo = psi;
      wn = psi2;
    } else {
      wo = up ? 1.f : 0.f;
      wn = up ? 0.f : 1.f;
    }
    const bool own = m.slot_side[k * C + c] == 0;
    const real r0 = m.skew[f], r1 = m.skew[Fi + f], r2 = m.skew[2 * Fi + f];
    for (int o = 0; o < outs.n; ++o) {
      real wx = 0.f;
      if (scheme == WCFD_CONV_UPWIND_LINEAR) {
        const int go = outs.g[o];
        const int u = up ? on : nn;
        const real* d = fc2f + (size_t)(up ? 0 : 3) * Fi;
        const real* gu = cg + (size_t)(3 * go) * C;
        wx = gu[u] * d[f] + gu[C + u] * d[Fi + f] + gu[2 * C + u] * d[2 * Fi + f];
      }
      const real* g = cg + (size_t)(3 * o) * C;
      const real g0 = g[on] * psi + psi2 * g[nn];
      const real g1 = g[C + on] * psi + psi2 * g[C + nn];
      const real g2 = g[2 * C + on] * psi + psi2 * g[2 * C + nn];
      wx = wx + (g0 * r0 + g1 * r1 + g2 * r2);
      real w0, w1, w2;
      if (own) {
        w0 = wo * Fl;
        w1 = wn * Fl;
        w2 = wx * Fl;
      } else {
        const real mf = -Fl;
        w0 = wn * mf;
        w1 = wo * mf;
        w2 = wx * mf;
      }
      w[widx(o, 0, k, c, C)] = w0;
      w[widx(o, 1, k, c, C)] = w1;
      w[widx(o, 2, k, c, C)] = w2;
    }
  }
}

// ---------------------------------------------------------- diffusion
// diffusion.py:40-49, boundary kernels 52-152, internal 155-197, with
// laplacianScheme.py central_difference (14-37) or central_difference_
// corrected (39-66).  nu: per face (device face order) or the scalar.
// bc_mode 0: the faces' boundary conditions; bc_mode 1: every boundary face
// gets the explicit nu * bc_value * A_f on output 0 (the von_neumann option;
// the reference's dirchlet option builds the same kernel, diffusion.py:33).
__global__ void k_diffusion_weights(wcfd_mesh m, wcfd_bc bc, const real* __restrict__ cg,
                                    const real* __restrict__ fv,
                                    const real* __restrict__ nu_face, real nu_scalar,
                                    int correction, int bc_mode, real bc_value, Outs outs,
                                    real* __restrict__ w) {
  const int c = xcd_item();
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal, Fb = m.n_boundary;
  if (c >= C) return;
  for (int k = 0; k < 6; ++k) {
    const int f = m.slot_face[k * C + c];
    const real nu = nu_face ? nu_face[f] : nu_scalar;
    const real A = m.area[f];
    if (f >= Fi) {
      const int j = f - Fi;
      if (bc_mode == 1) {
        w[widx(0, 0, k, c, C)] = 0.f;
        w[widx(0, 1, k, c, C)] = 0.f;
        w[widx(0, 2, k, c, C)] = nu * bc_value * A;
        continue;
      }
      for (int o = 0; o < outs.n; ++o) {
        const int go = outs.g[o];
        real w0, w2;
        if (bc.type[go * Fb + j] == WCFD_NEUMANN) {
          w0 = 0.f;
          w2 = nu * bc.value[go * Fb + j] * A;
        } else {
          const real wt = nu * A / m.dmag[f];
          w0 = -wt;
          w2 = wt * fv[(size_t)go * F + f];
        }
        w[widx(o, 0, k, c, C)] = w0;
        w[widx(o, 1, k, c, C)] = 0.f;
        w[widx(o, 2, k, c, C)] = w2;
      }
      continue;
    }
    const real wt = nu * A;
    const real f1 = lap_factor(m, f, correction);
    const real f0 = -f1;
    real n2[3] = {0.f, 0.f, 0.f};
    real psi = 0.f, psi2 = 0.f;
    int on = 0, nn = 0;
    if (correction) {
      const real d0 = m.dist[f], d1 = m.dist[Fi + f], d2 = m.dist[2 * Fi + f];
      const real nf0 = m.normal[f], nf1 = m.normal[F + f], nf2 = m.normal[2 * F + f];
      const real inv_nd = 1.0f / (nf0 * d0 + nf1 * d1 + nf2 * d2);
      n2[0] = nf0 - d0 * inv_nd;
      n2[1] = nf1 - d1 * inv_nd;
      n2[2] = nf2 - d2 * inv_nd;
      psi = m.psi[f];
      psi2 = 1.0f - psi;
      on = m.owner[f];
      nn = m.neigh[f];
    }
    const bool own = m.slot_side[k * C + c] == 0;
    for (int o = 0; o < outs.n; ++o) {
      real fx = 0.f;
      if (correction) {
        const real* g = cg + (size_t)(3 * outs.g[o]) * C;
        const real g0 = psi * g[on] + psi2 * g[nn];
        const real g1 = psi * g[C + on] + psi2 * g[C + nn];
        const real g2 = psi * g[2 * C + on] + psi2 * g[2 * C + nn];
        fx = g0 * n2[0] + g1 * n2[1] + g2 * n2[2];
      }
      real w0, w1, w2;
      if (own) {
        w0 = f0 * wt;
        w1 = f1 * wt;
        w2 = fx * wt;
      } else {
        w0 = -f1 * wt;
        w1 = -f0 * wt;
        w2 = -fx * wt;
      }
      w[widx(o, 0, k, c, C)] = w0;
      w[widx(o, 1, k, c, C)] = w1;
      w[widx(o, 2, k, c, C)] = w2;
    }
  }
}

// ---------------------------------------------------- ddt and grad
// ddt.py:20-33 (backwards Euler): w[o] = (V/dt, -phi V/dt).  The reference
// indexes its (C, O, 2) weights by the global field (ddt.py:31-32); here by
// the local output, which is the same for u, v, w.
__global__ void k_ddt_weights(wcfd_mesh m, const real* __restrict__ cv, real dt, Outs outs,
                              real* __restrict__ w) {
  const int c = xcd_item();
  const int C = m.n_cells;
  if (c >= C) return;
  const real vdt = m.volume[c] / dt;
  for (int o = 0; o < outs.n; ++o) {
    w[((size_t)o * 2 + 0) * C + c] = vdt;
    w[((size_t)o * 2 + 1) * C + c] = -cv[(size_t)outs.g[o] * C + c] * vdt;
  }
}

// grad.py:22-27: w[d][c] = (grad phi)_c[d] V_c (explicit, component-major)
__global__ void k_grad_weights(wcfd_mesh m, const real* __restrict__ cg, int field,
                               real* __restrict__ w) {
  const int c = xcd_item();
  const int C = m.n_cells;
  if (c >= C) return;
  const real V = m.volume[c];
#pragma unroll
  for (int d = 0; d < 3; ++d) w[(size_t)d * C + c] = cg[(size_t)(3 * field + d) * C + c] * V;
}

// ------------------------------------------------------------ equation
// calculate_BSR_values_kernel (equation_kernels.py:52-79): diagonal +=
// scale w0 over all six faces, off-diagonal += scale w1 on internal faces;
// calculate_RHS_values_kernel (84-101): b -= scale w2 on boundary faces only
// (T2).
__global__ void k_add_face_term(wcfd_mesh m, const real* __restrict__ w, int O, real scale,
                                real* __restrict__ values, real* __restrict__ rhs) {
  const int c = xcd_item();
  const int C = m.n_cells, Fi = m.n_internal;
  if (c >= C) return;
  int fk[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) fk[k] = m.slot_face[k * C + c];
  for (int o = 0; o < O; ++o) {
    real* v = values + (size_t)o * 7 * C;
    real d = v[c];
#pragma unroll
    for (int k = 0; k < 6; ++k) d = d + scale * w[widx(o, 0, k, c, C)];
    v[c] = d;
    real b = rhs[(size_t)o * C + c];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (fk[k] < Fi)
        v[(size_t)(k + 1) * C + c] = v[(size_t)(k + 1) * C + c] + scale * w[widx(o, 1, k, c, C)];
      else
        b = b + (-scale * w[widx(o, 2, k, c, C)]);
    }
    rhs[(size_t)o * C + c] = b;
  }
}

// add_Implicit_cell_based_term_kernel (equation_kernels.py:104-122)
__global__ void k_add_cell_term(int C, const real* __restrict__ w, int O, real scale,
                                real* __restrict__ values, real* __restrict__ rhs) {
  const int c = xcd_item();
  if (c >= C) return;
  for (int o = 0; o < O; ++o) {
    real* v = values + (size_t)o * 7 * C;
    v[c] = v[c] + scale * w[((size_t)o * 2 + 0) * C + c];
    rhs[(size_t)o * C + c] = rhs[(size_t)o * C + c] + (-scale * w[((size_t)o * 2 + 1) * C + c]);
  }
}

// rhs += scale * arr (equation.py:120-121, 169-170)
__global__ void k_add_rhs(int n, const real* __restrict__ arr, real scale,
                          real* __restrict__ rhs) {
  const int i = xcd_item();
  if (i < n) rhs[i] = rhs[i] + arr[i] * scale;
}

// implicit_relaxation_kernel (equation_kernels.py:128-152)
__global__ void k_relax_equation(int C, const real* __restrict__ cv, Outs outs, real alpha,
                                 real* __restrict__ values, real* __restrict__ rhs) {
  const int c = xcd_item();
  if (c >= C) return;
  for (int o = 0; o < outs.n; ++o) {
    real* v = values + (size_t)o * 7 * C;
    const real a = v[c];
    rhs[(size_t)o * C + c] =
        rhs[(size_t)o * C + c] + (1.0f - alpha) / alpha * a * cv[(size_t)outs.g[o] * C + c];
    v[c] = a / alpha;
  }
}

// replace_row_kernel (equation_kernels.py:155-175): an identity row.  The
// reference reads the right-hand side from a length-1 array at index row_id
// (T7); the given value is used here for any row.
__global__ void k_replace_row(int C, int o, int c, real value, real* __restrict__ values,
                              real* __restrict__ rhs) {
  const int k = threadIdx.x;
  if (k < 7) values[((size_t)o * 7 + k) * C + c] = k == 0 ? 1.0f : 0.0f;
  if (k == 0) rhs[(size_t)o * C + c] = value;
}

static bool outs_of(const int32_t* outputs, int n, Outs& o) {
  if (!outputs || n < 1 || n > 4) return false;
  o.n = n;
  for (int i = 0; i < 4; ++i) o.g[i] = 0;
  for (int i = 0; i < n; ++i) {
    if (outputs[i] < 0 || outputs[i] > 3) return false;
    o.g[i] = outputs[i];
  }
  return true;
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

int wcfd_convection_weights(const wcfd_mesh* m, wcfd_bc bc, const real* cell_grad,
                            const real* face_values, const real* mass_flux,
                            const real* face_c2f, int scheme, int flags, const int32_t* outputs,
                            int n_outputs, real* weights, void* stream) {
  Outs o;
  if (!mesh_ok(m) || !cell_grad || !face_values || !mass_flux || !weights || !bc.type ||
      !bc.value || !outs_of(outputs, n_outputs, o))
    return WCFD_EINVAL;
  if (scheme != WCFD_CONV_UPWIND && scheme != WCFD_CONV_CENTRAL &&
      scheme != WCFD_CONV_UPWIND_LINEAR)
    return WCFD_EINVAL;
  if (scheme == WCFD_CONV_UPWIND_LINEAR && !face_c2f) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_TERM_WEIGHTS, s, k_convection_weights, grid_for(m->n_cells), *m, bc, cell_grad,
              face_values, mass_flux, face_c2f, scheme,
              (flags & WCFD_SCHEME_IMPLICIT_OUTFLOW) ? 1 : 0, o, weights);
  return check_launch();
}

int wcfd_diffusion_weights(const wcfd_mesh* m, wcfd_bc bc, const real* cell_grad,
                           const real* face_values, const real* viscosity_faces,
                           float viscosity, int correction, int bc_mode, float bc_value,
                           const int32_t* outputs, int n_outputs, real* weights, void* stream) {
  Outs o;
  if (!mesh_ok(m) || !face_values || !weights || !outs_of(outputs, n_outputs, o))
    return WCFD_EINVAL;
  if (bc_mode != 0 && bc_mode != 1) return WCFD_EINVAL;
  if (bc_mode == 0 && (!bc.type || !bc.value)) return WCFD_EINVAL;
  if (bc_mode == 1 && n_outputs != 1) return WCFD_EINVAL;
  if (correction && !cell_grad) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_TERM_WEIGHTS, s, k_diffusion_weights, grid_for(m->n_cells), *m, bc, cell_grad,
              face_values, viscosity_faces, viscosity, correction ? 1 : 0, bc_mode, bc_value, o,
              weights);
  return check_launch();
}

int wcfd_ddt_weights(const wcfd_mesh* m, const real* cell_values, float dt,
                     const int32_t* outputs, int n_outputs, real* weights, void* stream) {
  Outs o;
  if (!mesh_ok(m) || !cell_values || !weights || !(dt > 0.f) ||
      !outs_of(outputs, n_outputs, o))
    return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_TERM_WEIGHTS, s, k_ddt_weights, grid_for(m->n_cells), *m, cell_values, dt, o,
              weights);
  return check_launch();
}

int wcfd_grad_weights(const wcfd_mesh* m, const real* cell_grad, int field, real* weights,
                      void* stream) {
  if (!mesh_ok(m) || !cell_grad || !weights || field < 0 || field > 3) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_TERM_WEIGHTS, s, k_grad_weights, grid_for(m->n_cells), *m, cell_grad, field,
              weights);
  return check_launch();
}

int wcfd_equation_add_face_term(const wcfd_mesh* m, const real* weights, int n_outputs,
                                float scale, real* values, real* rhs, void* stream) {
  if (!mesh_ok(m) || !weights || !values || !rhs || n_outputs < 1 || n_outputs > 4)
    return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_EQUATION, s, k_add_face_term, grid_for(m->n_cells), *m, weights, n_outputs, scale,
              values, rhs);
  return check_launch();
}

int wcfd_equation_add_cell_term(const wcfd_mesh* m, const real* weights, int n_outputs,
                                float scale, real* values, real* rhs, void* stream) {
  if (!mesh_ok(m) || !weights || !values || !rhs || n_outputs < 1 || n_outputs > 4)
    return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_EQUATION, s, k_add_cell_term, grid_for(m->n_cells), m->n_cells, weights, n_outputs,
              scale, values, rhs);
  return check_launch();
}

int wcfd_equation_add_rhs(int32_t n, const real* arr, float scale, real* rhs, void* stream) {
  if (n <= 0 || !arr || !rhs) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_EQUATION, s, k_add_rhs, grid_for(n), n, arr, scale, rhs);
  return check_launch();
}

int wcfd_equation_relax(const wcfd_mesh* m, const real* cell_values, const int32_t* outputs,
                        int n_outputs, float alpha, real* values, real* rhs, void* stream) {
  Outs o;
  if (!mesh_ok(m) || !cell_values || !values || !rhs || !(alpha > 0.f) ||
      !outs_of(outputs, n_outputs, o))
    return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_EQUATION, s, k_relax_equation, grid_for(m->n_cells), m->n_cells, cell_values, o,
              alpha, values, rhs);
  return check_launch();
}

int wcfd_equation_replace_row(const wcfd_mesh* m, int n_outputs, int output, int32_t cell,
                              float value, real* values, real* rhs, void* stream) {
  if (!mesh_ok(m) || !values || !rhs || n_outputs < 1 || n_outputs > 4 || output < 0 ||
      output >= n_outputs || cell < 0 || cell >= m->n_cells)
    return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_replace_row, dim3(1), dim3(kWave), 0, s, m->n_cells, output, cell, value,
                     values, rhs);
  return check_launch();
}

}  // extern "C"
