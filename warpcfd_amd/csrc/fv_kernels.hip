// Face and cell kernels of the finite-volume operators (gfx950).
//
// Every face->cell reduction is a cell-centric gather over the cell's six
// "slots" (its local faces in the reference's order), so no atomics are used
// and results are bitwise reproducible run to run.  The reference
// accumulates with wp.atomic_add in launch order (model_kernels.py:32-34,90;
// equation_kernels.py:75,79,101); the sums here run over the same terms in
// local-face order.
#include "fv_kernels.h"
#include "monitor.h"

#include <mutex>

static_assert(wcfd::kMonHead + 2 * wcfd::kMonWindow <= WCFD_MONITOR_LEN,
              "monitor record larger than WCFD_MONITOR_LEN");

namespace wcfd {

thread_local int g_last_hip_error = 0;

// ------------------------------------------------------------------ BCs
__global__ void k_set_bc(BatchDev bd, int F, int Fi, int Fb, const uint8_t* __restrict__ type,
                         const real* __restrict__ value, real* __restrict__ fv, int mask) {
  const int j = item_index(bd.bfaces);
  if (j >= Fb) return;
  int f = Fi + j;
  if (bd.face_skip(f)) return;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if ((mask >> q) & 1)
      if (type[q * Fb + j] == WCFD_DIRICHLET) fv[q * F + f] = value[q * Fb + j];
}

// --------------------------------------------------- face interpolation
// interpolation_kernels.py:5-54, interpolation_functions.py:6-14,
// skewness_functions.py:7-20, model_kernels.py:46-60, solver_kernels.py:22-36.
// k_face_p_fc gathers its two cells' values through buffer descriptors at
// the cells' byte offsets (one offset register per cell, shared by every
// cell array) and streams the face arrays at the face's byte offset; for
// k_face_interp that form measured 3% slower at 8.6 M cells, so it keeps
// plain indexing.
__device__ inline FBuf cell_plane(const wcfd_mesh& m, const real* a, int q) {
  return FBuf(a + (size_t)q * m.n_cells, (uint32_t)m.n_cells * 4u);
}

// An internal face in three phases: every load (indices, then the cell
// values and gradients of every field, the face normal and area), then the
// arithmetic, then every store.  With no store between two fields, all of
// the face's gathers are in flight together rather than one field's at a
// time (2-3 % faster at 4.4-8.6 M cells; two faces per thread this way
// needs 104 VGPRs and lost 25 %).
__device__ inline void face_interp_internal(const wcfd_mesh& m, int f,
                                            const real* __restrict__ cv,
                                            const real* __restrict__ cg, real* __restrict__ fv,
                                            int mask, int skew_mask, real* __restrict__ flux,
                                            const real* __restrict__ rua,
                                            real* __restrict__ rua_f) {
  constexpr bool NT = WCFD_NT_FACE;
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal;
  const int o = ld_once<NT>(m.owner + f), n = ld_once<NT>(m.neigh + f);
  const real psi = ld_once<NT>(m.psi + f);
  const real s0 = ld_once<NT>(m.skew + f), s1 = ld_once<NT>(m.skew + Fi + f),
             s2 = ld_once<NT>(m.skew + 2 * Fi + f);
  // loads of every field
  real co[4], cn[4], go[4][3], gn[4][3], old[3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((mask >> q) & 1) {
      const real* gq = cg + (size_t)(3 * q) * C;
      co[q] = cv[q * C + o];
      cn[q] = cv[q * C + n];
      if ((skew_mask >> q) & 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          go[q][d] = gq[d * C + o];
          gn[q][d] = gq[d * C + n];
        }
      }
    } else if (q < 3 && flux) {
      old[q] = ld_once<NT>(fv + q * F + f);
    }
  }
  real nr[3], ar = 0.f, ro = 0.f, rn = 0.f;
  if (flux) {
    nr[0] = ld_once<NT>(m.normal + f);
    nr[1] = ld_once<NT>(m.normal + F + f);
    nr[2] = ld_once<NT>(m.normal + 2 * F + f);
    ar = ld_once<NT>(m.area + f);
  }
  if (rua_f) {
    ro = rua[o];
    rn = rua[n];
  }
  // arithmetic
  const real psi2 = 1.0f - psi;
  real val[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((mask >> q) & 1) {
      real v = co[q] * psi + psi2 * cn[q];
      if ((skew_mask >> q) & 1) {
        const real g0 = go[q][0] * psi + psi2 * gn[q][0];
        const real g1 = go[q][1] * psi + psi2 * gn[q][1];
        const real g2 = go[q][2] * psi + psi2 * gn[q][2];
        v = v + (g0 * s0 + g1 * s1 + g2 * s2);
      }
      val[q] = v;
    } else {
      val[q] = (q < 3 && flux) ? old[q] : 0.0f;
    }
  }
  // stores
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if ((mask >> q) & 1) st_once<NT>(fv + q * F + f, val[q]);
  if (rua_f) st_once<NT>(rua_f + f, psi * ro + psi2 * rn);
  if (flux) st_once<NT>(flux + f, (val[0] * nr[0] + val[1] * nr[1] + val[2] * nr[2]) * ar);
}

__device__ inline void face_interp_boundary(const wcfd_mesh& m, wcfd_bc bc, int f,
                                            const real* __restrict__ cv, real* __restrict__ fv,
                                            int mask, real* __restrict__ flux,
                                            const real* __restrict__ rua,
                                            real* __restrict__ rua_f, int apply_dirichlet) {
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal, Fb = m.n_boundary;
  real val[4];
  const int j = f - Fi;
  const int P = m.owner[f];
  const real d = m.dmag[f];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool want = ((mask >> q) & 1) || (q < 3 && flux);
    if (!want) continue;
    real v;
    const uint8_t t = bc.type[q * Fb + j];
    if (((mask >> q) & 1) && t == WCFD_NEUMANN) {
      v = d * bc.value[q * Fb + j] + cv[q * C + P];
      fv[q * F + f] = v;
    } else if (((mask >> q) & 1) && apply_dirichlet && t == WCFD_DIRICHLET) {
      v = bc.value[q * Fb + j];
      fv[q * F + f] = v;
    } else {
      v = fv[q * F + f];
    }
    val[q] = v;
  }
  if (rua_f) rua_f[f] = rua[P];
  if (flux) {
    const real n0 = m.normal[f], n1 = m.normal[F + f], n2 = m.normal[2 * F + f];
    flux[f] = (val[0] * n0 + val[1] * n1 + val[2] * n2) * m.area[f];
  }
}

__global__ void k_face_interp(wcfd_mesh m, BatchDev bd, wcfd_bc bc, const real* __restrict__ cv,
                              const real* __restrict__ cg, real* __restrict__ fv,
                              int mask, int skew_mask, real* __restrict__ flux,
                              const real* __restrict__ rua, real* __restrict__ rua_f,
                              int apply_dirichlet) {
  const int f = item_index(bd.faces);
  if (f >= m.n_faces || bd.face_skip(f)) return;
  if (f < m.n_internal) {
    face_interp_internal(m, f, cv, cg, fv, mask, skew_mask, flux, rua, rua_f);
  } else {
    face_interp_boundary(m, bc, f, cv, fv, mask, flux, rua, rua_f, apply_dirichlet);
  }
}

// face_interpolation(upwind=True) (model.py:210-229 with
// interpolation_functions.py:18-28): an internal face takes its owner's
// value where its mass flux is > 0 and its neighbour's otherwise, plus the
// linear scheme's skew correction (the same expression as
// face_interp_internal); boundary faces as in the linear scheme (Neumann
// faces only).  Off the solver's path, so one plain face per thread.
__global__ void k_face_interp_upwind(wcfd_mesh m, wcfd_bc bc, const real* __restrict__ cv,
                                     const real* __restrict__ cg,
                                     const real* __restrict__ flux, real* __restrict__ fv,
                                     int mask, int skew_mask) {
  const int f = xcd_item();
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal;
  if (f >= F) return;
  if (f >= Fi) {
    face_interp_boundary(m, bc, f, cv, fv, mask, nullptr, nullptr, nullptr, 0);
    return;
  }
  const int o = m.owner[f], n = m.neigh[f];
  const real psi = m.psi[f];
  const real psi2 = 1.0f - psi;
  const real s0 = m.skew[f], s1 = m.skew[Fi + f], s2 = m.skew[2 * Fi + f];
  const bool up = flux[f] > 0.f;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((mask >> q) & 1)) continue;
    real v = up ? cv[q * C + o] : cv[q * C + n];
    if ((skew_mask >> q) & 1) {
      const real* gq = cg + (size_t)(3 * q) * C;
      const real g0 = gq[o] * psi + psi2 * gq[n];
      const real g1 = gq[C + o] * psi + psi2 * gq[C + n];
      const real g2 = gq[2 * C + o] * psi + psi2 * gq[2 * C + n];
      v = v + (g0 * s0 + g1 * s1 + g2 * s2);
    }
    fv[q * F + f] = v;
  }
}

// Owner data of a workgroup's internal faces staged in LDS.  Internal faces
// are numbered by owner (per case in a batch, and the compacted item ranges
// never span two cases), so the owners of one workgroup's faces are one cell
// range [o_lo, o_hi].  The workgroup loads those cells' values, skewed
// gradients and rUA once, coalesced, into LDS while each thread's neighbour
// gathers are in flight; after the barrier a face reads its owner's data
// from LDS.  A range longer than kStage cells, or an owner outside it (an
// uncompacted launch over several cases), falls back to global loads, so the
// kernel is correct for any face order.  The arithmetic is face_interp_
// internal's, so the two forms agree bit for bit.
constexpr int kStage = kBlock;

__device__ inline void item_range(const Items& it, int n, int& a, int& b) {
  const int vb = (int)xcd_remap(blockIdx.x, gridDim.x);
  if (it.blk) {
    a = it.blk[vb];
    b = it.blk[it.G + vb];
  } else {
    a = vb * (int)blockDim.x;
    b = min(a + (int)blockDim.x, n);
  }
}

__global__ __launch_bounds__(kBlock) void k_face_interp_lds(
    wcfd_mesh m, BatchDev bd, wcfd_bc bc, const real* __restrict__ cv,
    const real* __restrict__ cg, real* __restrict__ fv, int mask, int skew_mask,
    real* __restrict__ flux, const real* __restrict__ rua, real* __restrict__ rua_f,
    int apply_dirichlet) {
  __shared__ real s_cv[4][kStage];
  __shared__ real s_cg[4][3][kStage];
  __shared__ real s_rua[kStage];
  const int C = m.n_cells, F = m.n_faces, Fi = m.n_internal;
  int a, b;
  item_range(bd.faces, F, a, b);
  const int f = a + (int)threadIdx.x;
  const bool live = f < b && f < F && !bd.face_skip(f);
  const int ie = min(b, Fi);              // internal faces of the workgroup: [a, ie)
  // uniform: the staged owner range
  int o_lo = 0, R = 0;
  if (a < ie) {
    o_lo = m.owner[a];
    R = m.owner[ie - 1] - o_lo + 1;
    if (R < 1 || R > kStage) R = 0;       // not staged: global loads
  }
  const bool in = live && f < Fi;
  // this face's indices and its neighbour's data
  int o = 0, n = 0;
  real psi = 0.f, s0 = 0.f, s1 = 0.f, s2 = 0.f;
  real cn[4], gn[4][3], old[3], nr[3], ar = 0.f, rn = 0.f;
  if (in) {
    o = m.owner[f];
    n = m.neigh[f];
    psi = m.psi[f];
    s0 = m.skew[f];
    s1 = m.skew[Fi + f];
    s2 = m.skew[2 * Fi + f];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if ((mask >> q) & 1) {
        cn[q] = cv[q * C + n];
        if ((skew_mask >> q) & 1) {
          const real* gq = cg + (size_t)(3 * q) * C;
#pragma unroll
          for (int d = 0; d < 3; ++d) gn[q][d] = gq[d * C + n];
        }
      } else if (q < 3 && flux) {
        old[q] = fv[q * F + f];
      }
    }
    if (flux) {
      nr[0] = m.normal[f];
      nr[1] = m.normal[F + f];
      nr[2] = m.normal[2 * F + f];
      ar = m.area[f];
    }
    if (rua_f) rn = rua[n];
  }
  // stage the owners' data
  for (int i = (int)threadIdx.x; i < R; i += (int)blockDim.x) {
    const int c = o_lo + i;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if ((mask >> q) & 1) {
        s_cv[q][i] = cv[q * C + c];
        if ((skew_mask >> q) & 1) {
          const real* gq = cg + (size_t)(3 * q) * C;
#pragma unroll
          for (int d = 0; d < 3; ++d) s_cg[q][d][i] = gq[d * C + c];
        }
      }
    }
    if (rua_f) s_rua[i] = rua[c];
  }
  __syncthreads();
  if (in) {
    const int j = o - o_lo;
    const bool st = j >= 0 && j < R;
    real co[4], go[4][3], ro = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if ((mask >> q) & 1) {
        co[q] = st ? s_cv[q][j] : cv[q * C + o];
        if ((skew_mask >> q) & 1) {
          const real* gq = cg + (size_t)(3 * q) * C;
#pragma unroll
          for (int d = 0; d < 3; ++d) go[q][d] = st ? s_cg[q][d][j] : gq[d * C + o];
        }
      }
    }
    if (rua_f) ro = st ? s_rua[j] : rua[o];
    const real psi2 = 1.0f - psi;
    real val[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if ((mask >> q) & 1) {
        real v = co[q] * psi + psi2 * cn[q];
        if ((skew_mask >> q) & 1) {
          const real g0 = go[q][0] * psi + psi2 * gn[q][0];
          const real g1 = go[q][1] * psi + psi2 * gn[q][1];
          const real g2 = go[q][2] * psi + psi2 * gn[q][2];
          v = v + (g0 * s0 + g1 * s1 + g2 * s2);
        }
        val[q] = v;
      } else {
        val[q] = (q < 3 && flux) ? old[q] : 0.0f;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if ((mask >> q) & 1) fv[q * F + f] = val[q];
    if (rua_f) rua_f[f] = psi * ro + psi2 * rn;
    if (flux) flux[f] = (val[0] * nr[0] + val[1] * nr[1] + val[2] * nr[2]) * ar;
  } else if (live) {
    face_interp_boundary(m, bc, f, cv, fv, mask, flux, rua, rua_f, apply_dirichlet);
  }
}

// ------------------------------------------------------------ Green-Gauss
// Cell kernels read their cell's arrays at a uniform base plus the cell's
// byte offset c4 = 4 c, and the faces' arrays through buffer descriptors at
// f4 = 4 f, so the six faces of a cell need six offset registers however
// many face arrays are read (rather than a 64-bit address per array and
// face).

// the six faces of cell c as byte offsets (NT: nontemporal table loads)
template <bool NT = false>
__device__ inline void slot_faces(const wcfd_mesh& m, uint32_t c4, uint32_t (&f4)[6]) {
  const size_t C = m.n_cells;
#pragma unroll
  for (int k = 0; k < 6; ++k) f4[k] = (uint32_t)ld_cell<NT>(m.slot_face + k * C, c4) * 4u;
}

// model_kernels.py:72-90: grad_c = sum_k phi_f A_f n_{c,k} / V_c.
template <bool NT = false>
__device__ inline void gg_field(const wcfd_mesh& m, uint32_t c4, const uint32_t (&f4)[6],
                                const FBuf& phi_f, real V, real (&g)[3]) {
  const size_t C = m.n_cells;
  real a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const real phi = phi_f(f4[k]);
    a0 = a0 + (phi * ld_cell<NT>(m.slot_sn + (0 * 6 + k) * C, c4)) / V;
    a1 = a1 + (phi * ld_cell<NT>(m.slot_sn + (1 * 6 + k) * C, c4)) / V;
    a2 = a2 + (phi * ld_cell<NT>(m.slot_sn + (2 * 6 + k) * C, c4)) / V;
  }
  g[0] = a0; g[1] = a1; g[2] = a2;
}

__device__ inline FBuf face_plane(const wcfd_mesh& m, const real* a, int q) {
  return FBuf(a + (size_t)q * m.n_faces, (uint32_t)m.n_faces * 4u);
}

__global__ void k_green_gauss(wcfd_mesh m, Items items, const real* __restrict__ fv,
                              real* __restrict__ cg, int mask, int zero_others) {
  const int c = item_index(items);
  const int C = m.n_cells;
  if (c >= C) return;
  const uint32_t c4 = (uint32_t)c * 4u;
  uint32_t f4[6];
  slot_faces(m, c4, f4);
  const real V = ld_off(m.volume, c4);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    real g[3] = {0.f, 0.f, 0.f};
    if ((mask >> q) & 1) {
      gg_field(m, c4, f4, face_plane(m, fv, q), V, g);
    } else if (!zero_others) {
      continue;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) st_plain(cg + (size_t)(3 * q + d) * C, c4, g[d]);
  }
}

__global__ void k_mass_flux(wcfd_mesh m, const real* __restrict__ fv, real* __restrict__ flux) {
  const int f = xcd_item();
  const int F = m.n_faces;
  if (f >= F) return;
  flux[f] = (fv[f] * m.normal[f] + fv[F + f] * m.normal[F + f] + fv[2 * F + f] * m.normal[2 * F + f]) *
            m.area[f];
}

template <bool NT = false>
__device__ inline real cell_div(const wcfd_mesh& m, uint32_t c4, const uint32_t (&f4)[6],
                                 const FBuf& flux) {
  const size_t C = m.n_cells;
  real acc = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const real fl = flux(f4[k]);
    acc = acc + (ld_cell<NT>(m.slot_side + k * C, c4 >> 2) == 0 ? fl : -fl);
  }
  return acc;
}

__global__ void k_divergence(wcfd_mesh m, const real* __restrict__ flux, real* __restrict__ div) {
  const int c = xcd_item();
  const int C = m.n_cells;
  if (c >= C) return;
  const uint32_t c4 = (uint32_t)c * 4u;
  uint32_t f4[6];
  slot_faces(m, c4, f4);
  st_plain(div, c4, cell_div(m, c4, f4, FBuf(flux, (uint32_t)m.n_faces * 4u)));
}

// ------------------------------------------------------ momentum assembly
__global__ void k_assemble_momentum(wcfd_mesh m, BatchDev bd, wcfd_bc bc,
                                    const real* __restrict__ cv, const real* __restrict__ fv,
                                    const real* __restrict__ flux, real* __restrict__ cg,
                                    int compute_grad, real nu_scalar,
                                    real inv_density, real alpha, real dt, int flags,
                                    real* __restrict__ values, real* __restrict__ rhs) {
  constexpr bool NT = WCFD_NT_CELL;
  const int correction = flags & WCFD_SCHEME_CORRECTED;
  const bool implicit_outflow = (flags & WCFD_SCHEME_IMPLICIT_OUTFLOW) != 0;
  const int c = item_index(bd.cells);
  const int C = m.n_cells, Fi = m.n_internal, Fb = m.n_boundary;
  if (c >= C || bd.cell_skip(c)) return;
  const real nu = bd.nu ? bd.nu[bd.cell_case[c]] : nu_scalar;
  const uint32_t c4 = (uint32_t)c * 4u, Fi4 = (uint32_t)Fi * 4u;
  const size_t Cz = C;
  uint32_t f4[6];
  slot_faces<NT>(m, c4, f4);
  int side[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) side[k] = ld_cell<NT>(m.slot_side + k * Cz, (uint32_t)c);
  const FBuf Fl_(flux, (uint32_t)m.n_faces * 4u);
  const FBuf area(m.area, (uint32_t)m.n_faces * 4u), dmag(m.dmag, (uint32_t)m.n_faces * 4u);
  const real V = ld_cell<NT>(m.volume, c4);
  // (storing the gradients at the end, after every load, measured 1.5 %
  // slower here, unlike the face and corrector kernels)
  real gp[3];
  if (compute_grad) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      real g[3];
      gg_field<NT>(m, c4, f4, face_plane(m, fv, q), V, g);
      st_cell<NT>(cg + (3 * q + 0) * Cz, c4, g[0]);
      st_cell<NT>(cg + (3 * q + 1) * Cz, c4, g[1]);
      st_cell<NT>(cg + (3 * q + 2) * Cz, c4, g[2]);
      if (q == 3) { gp[0] = g[0]; gp[1] = g[1]; gp[2] = g[2]; }
    }
  } else {
    gp[0] = ld_cell<NT>(cg + 9 * Cz, c4);
    gp[1] = ld_cell<NT>(cg + 10 * Cz, c4);
    gp[2] = ld_cell<NT>(cg + 11 * Cz, c4);
  }
  real u[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) u[q] = ld_cell<NT>(cv + q * Cz, c4);
  real diag = 0.f;
  real b[3] = {0.f, 0.f, 0.f};
  real off[6];
  // DdtTerm first (incompressible.py:110): diag += V/dt, b += phi V/dt
  if (dt > 0.f) {
    const real vdt = V / dt;
    diag = diag + vdt;
#pragma unroll
    for (int q = 0; q < 3; ++q) b[q] = b[q] + (-(-u[q] * vdt));
  }
  // ConvectionTerm (scale +1), upwind, tie -> neighbour (T6).  Boundary
  // faces go to b with the lagged face value (convection.py:63-67); with
  // WCFD_SCHEME_IMPLICIT_OUTFLOW a Neumann outflow face puts F_b on the
  // diagonal and keeps only F_b |x_f - x_P| g explicit.
  unsigned outflow = 0u;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const real Fl = Fl_(f4[k]);
    off[k] = 0.f;
    if (f4[k] < Fi4) {
      const bool up = Fl > 0.f;
      if (side[k] == 0) {
        diag = diag + (up ? Fl : 0.f);
        off[k] = up ? 0.f : Fl;
      } else {
        diag = diag + (up ? 0.f : -Fl);
        off[k] = up ? -Fl : 0.f;
      }
    } else if (implicit_outflow && Fl > 0.f && bc.type[(f4[k] - Fi4) >> 2] == WCFD_NEUMANN) {
      diag = diag + Fl;
      outflow |= 1u << k;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (f4[k] >= Fi4) {
      const real Fl = Fl_(f4[k]);
      if ((outflow >> k) & 1u) {
        const int j = (int)((f4[k] - Fi4) >> 2);
        const real dm = dmag(f4[k]);
#pragma unroll
        for (int q = 0; q < 3; ++q) b[q] = b[q] + (-(Fl * (dm * bc.value[q * Fb + j])));
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) b[q] = b[q] + (-(Fl * face_plane(m, fv, q)(f4[k])));
      }
    }
  }
  // DiffusionTerm (scale -1)
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (f4[k] < Fi4) {
      const real w = lap_factor_b(m, dmag, f4[k], correction) * (nu * area(f4[k]));
      diag = diag + w;
      off[k] = off[k] + (-w);
    } else {
      const int j = (int)((f4[k] - Fi4) >> 2);
      if (bc.type[j] == WCFD_NEUMANN) {
        const real A = area(f4[k]);
#pragma unroll
        for (int q = 0; q < 3; ++q) b[q] = b[q] + (nu * bc.value[q * Fb + j]) * A;
      } else {
        const real w = (nu * area(f4[k])) / dmag(f4[k]);
        diag = diag + w;
#pragma unroll
        for (int q = 0; q < 3; ++q) b[q] = b[q] + w * face_plane(m, fv, q)(f4[k]);
      }
    }
  }
  // explicit -(1/rho) grad(p) V (incompressible.py:107, grad.py:22-27)
#pragma unroll
  for (int q = 0; q < 3; ++q) b[q] = b[q] + (gp[q] * V) * (-inv_density);
  // implicit under-relaxation (equation_kernels.py:128-152)
  if (alpha > 0.f) {
    const real r = (1.0f - alpha) / alpha;
#pragma unroll
    for (int q = 0; q < 3; ++q) b[q] = b[q] + r * diag * u[q];
    diag = diag / alpha;
  }
  // slot-major (ELL) matrix: diagonal, then the 6 slot coefficients
  st_cell<NT>(values, c4, diag);
#pragma unroll
  for (int k = 0; k < 6; ++k) st_cell<NT>(values + (k + 1) * Cz, c4, off[k]);
#pragma unroll
  for (int q = 0; q < 3; ++q) st_cell<NT>(rhs + q * Cz, c4, b[q]);
}

// ------------------------------------------------------ pressure assembly
__global__ void k_assemble_pressure(wcfd_mesh m, BatchDev bd, wcfd_bc bc,
                                    const real* __restrict__ fv,
                                    const real* __restrict__ flux,
                                    const real* __restrict__ rua_f, real* __restrict__ cg,
                                    int compute_grad_p, int correction, int ref_cell,
                                    real ref_value, real* __restrict__ values,
                                    real* __restrict__ rhs, real* __restrict__ div_out) {
  const int c = item_index(bd.cells);
  const int C = m.n_cells, Fi = m.n_internal, Fb = m.n_boundary;
  if (c >= C || bd.cell_skip(c)) return;
  if (bd.ref_cell) {
    const int k = bd.cell_case[c];
    ref_cell = bd.ref_cell[k];
    ref_value = bd.ref_value[k];
  }
  constexpr bool NT = WCFD_NT_CELL;
  const uint32_t c4 = (uint32_t)c * 4u, Fi4 = (uint32_t)Fi * 4u, F4 = (uint32_t)m.n_faces * 4u;
  const size_t Cz = C;
  uint32_t f4[6];
  slot_faces<NT>(m, c4, f4);
  const FBuf pf = face_plane(m, fv, 3);
  if (compute_grad_p) {
    real g[3];
    gg_field<NT>(m, c4, f4, pf, ld_cell<NT>(m.volume, c4), g);
    st_cell<NT>(cg + 9 * Cz, c4, g[0]);
    st_cell<NT>(cg + 10 * Cz, c4, g[1]);
    st_cell<NT>(cg + 11 * Cz, c4, g[2]);
  }
  const real div = cell_div<NT>(m, c4, f4, FBuf(flux, F4));
  if (div_out) st_cell<NT>(div_out, c4, div);
  const FBuf ruaf(rua_f, F4), area(m.area, F4), dmag(m.dmag, F4);
  real diag = 0.f, b = 0.f;
  real off[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    off[k] = 0.f;
    if (f4[k] < Fi4) {
      const real w = lap_factor_b(m, dmag, f4[k], correction) * (ruaf(f4[k]) * area(f4[k]));
      diag = diag + (-w);
      off[k] = w;
    } else {
      const int j = (int)((f4[k] - Fi4) >> 2);
      if (bc.type[3 * Fb + j] == WCFD_NEUMANN) {
        b = b + (-((ruaf(f4[k]) * bc.value[3 * Fb + j]) * area(f4[k])));
      } else {
        const real w = (ruaf(f4[k]) * area(f4[k])) / dmag(f4[k]);
        diag = diag + (-w);
        b = b + (-(w * pf(f4[k])));
      }
    }
  }
  b = b + div;
  if (c == ref_cell) {
    diag = 1.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) off[k] = 0.f;
    b = ref_value;
  }
  st_cell<NT>(values, c4, diag);
#pragma unroll
  for (int k = 0; k < 6; ++k) st_cell<NT>(values + (k + 1) * Cz, c4, off[k]);
  st_cell<NT>(rhs, c4, b);
}

// ------------------------------------------------------------ rUA, HbyA
__global__ void k_rua_hbya(wcfd_mesh m, BatchDev bd, const real* __restrict__ values,
                           real density, const real* __restrict__ cg, real* __restrict__ hbya,
                           real* __restrict__ cv, real* __restrict__ rua,
                           real* __restrict__ u_prev) {
  const int c = item_index(bd.cells);
  const int C = m.n_cells;
  if (c >= C || bd.cell_skip(c)) return;
  const real r = m.volume[c] / (values[c] * density);
  rua[c] = r;
  if (u_prev) {   // u of the previous outer iteration, for the velocity change
#pragma unroll
    for (int q = 0; q < 3; ++q) u_prev[q * C + c] = cv[q * C + c];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const real h = hbya[q * C + c] + cg[(9 + q) * C + c] * r;
    hbya[q * C + c] = h;
    cv[q * C + c] = h;
  }
}

__global__ void k_relax_p(BatchDev bd, int C, const real* __restrict__ pn, real alpha,
                          real* __restrict__ cv) {
  const int c = item_index(bd.cells);
  if (c >= C || bd.cell_skip(c)) return;
  real* p = cv + (size_t)3 * C;
  p[c] = alpha > 0.f ? alpha * pn[c] + (1.0f - alpha) * p[c] : pn[c];
}

// face_interpolation('p') + the corrected conservative flux F_c per face.
__global__ void k_face_p_fc(wcfd_mesh m, BatchDev bd, wcfd_bc bc, const real* __restrict__ cv,
                            const real* __restrict__ cg, real* __restrict__ fv,
                            const real* __restrict__ flux, const real* __restrict__ rua_f,
                            real* __restrict__ fc) {
  const int f = item_index(bd.faces);
  const int F = m.n_faces, Fi = m.n_internal, Fb = m.n_boundary;
  if (f >= F || bd.face_skip(f)) return;
  constexpr bool NT = WCFD_NT_FACE;
  const uint32_t f4 = (uint32_t)f * 4u;
  const size_t Fz = F, Fiz = Fi;
  const FBuf p = cell_plane(m, cv, 3);
  if (f < Fi) {
    const uint32_t o4 = (uint32_t)ld_cell<NT>(m.owner, f4) * 4u,
                   n4 = (uint32_t)ld_cell<NT>(m.neigh, f4) * 4u;
    const real psi = ld_cell<NT>(m.psi, f4);
    const real psi2 = 1.0f - psi;
    const FBuf g0b = cell_plane(m, cg, 9), g1b = cell_plane(m, cg, 10), g2b = cell_plane(m, cg, 11);
    const real po = p(o4), pn = p(n4);
    real v = po * psi + psi2 * pn;
    real g0 = g0b(o4) * psi + psi2 * g0b(n4);
    real g1 = g1b(o4) * psi + psi2 * g1b(n4);
    real g2 = g2b(o4) * psi + psi2 * g2b(n4);
    v = v + (g0 * ld_cell<NT>(m.skew, f4) + g1 * ld_cell<NT>(m.skew + Fiz, f4) +
             g2 * ld_cell<NT>(m.skew + 2 * Fiz, f4));
    st_cell<NT>(fv + 3 * Fz, f4, v);
    if (fc)
      st_cell<NT>(fc, f4, ld_cell<NT>(flux, f4) - (ld_cell<NT>(rua_f, f4) * ld_cell<NT>(m.area, f4) /
                                                   ld_cell<NT>(m.dmag, f4)) * (pn - po));
  } else {
    const int j = f - Fi;
    const uint32_t p4 = (uint32_t)ld_off(m.owner, f4) * 4u;
    const uint8_t t = bc.type[3 * Fb + j];
    if (t == WCFD_NEUMANN)
      st_plain(fv + 3 * Fz, f4, ld_off(m.dmag, f4) * bc.value[3 * Fb + j] + p(p4));
    if (fc) {
      real F_ = ld_off(flux, f4);
      if (t == WCFD_DIRICHLET)
        F_ = F_ - (ld_off(rua_f, f4) * ld_off(m.area, f4) / ld_off(m.dmag, f4)) *
                      (bc.value[3 * Fb + j] - p(p4));
      st_plain(fc, f4, F_);
    }
  }
}

// calculate_gradients('p') + HbyA -= rUA grad p + replace_cell_values, plus
// the float64 continuity partial sums of the corrected flux: one pair per
// workgroup of the segmentation (one case, or a batch), folded per case by
// k_fold_resid.
__global__ void k_grad_p_correct(wcfd_mesh m, Seg sg, const real* __restrict__ fv,
                                 real* __restrict__ cg, real* __restrict__ cv,
                                 real* __restrict__ hbya, const real* __restrict__ rua,
                                 const real* __restrict__ fc, const real* __restrict__ u_prev,
                                 double* __restrict__ partials) {
  const size_t Cz = m.n_cells;
  const BlockRows br = block_rows(sg);
  if (sg.active && !sg.active[br.k]) return;   // whole workgroup: one case
  const FBuf pf = face_plane(m, fv, 3);
  const FBuf fcb(fc, (uint32_t)m.n_faces * 4u);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  constexpr bool NT = WCFD_NT_CELL;
  for (int c = br.begin + (int)threadIdx.x; c < br.end; c += blockDim.x) {
    const uint32_t c4 = (uint32_t)c * 4u;
    uint32_t f4[6];
    slot_faces<NT>(m, c4, f4);
    // every load (the cell's gradient sums, rUA, HbyA, the previous
    // velocity, the corrected face fluxes) before the first store
    real g[3];
    gg_field<NT>(m, c4, f4, pf, ld_cell<NT>(m.volume, c4), g);
    const real r = ld_cell<NT>(rua, c4);
    real hb[3], up[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      hb[q] = ld_cell<NT>(hbya + q * Cz, c4);
      up[q] = (fc && u_prev) ? ld_cell<NT>(u_prev + q * Cz, c4) : 0.f;
    }
    real fk[6];
    int sd[6];
    if (fc) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        fk[k] = fcb(f4[k]);
        sd[k] = ld_cell<NT>(m.slot_side + k * Cz, (uint32_t)c);
      }
    }
    real h[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      h[q] = hb[q] - r * g[q];
      if (fc && u_prev) {
        const double dq = (double)h[q] - (double)up[q];
        acc[2] += dq * dq;
        acc[3] += (double)h[q] * (double)h[q];
      }
    }
    if (fc) {
      double num = 0.0, den = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const double v = (double)fk[k];
        num += sd[k] == 0 ? v : -v;
        den += fabs(v);
      }
      acc[0] += fabs(num);
      acc[1] += den;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) st_cell<NT>(cg + i * Cz, c4, (real)0);
    st_cell<NT>(cg + 9 * Cz, c4, g[0]);
    st_cell<NT>(cg + 10 * Cz, c4, g[1]);
    st_cell<NT>(cg + 11 * Cz, c4, g[2]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      st_cell<NT>(hbya + q * Cz, c4, h[q]);
      st_cell<NT>(cv + q * Cz, c4, h[q]);
    }
  }
  if (fc) store_partials<4>(acc, partials, sg.G, br.vb);
}

// One workgroup per case: fold the case's partials to resid[4k .. 4k+3]
// (continuity numerator, denominator, ||du||^2, ||u||^2).  With a batch `bt`,
// count the iteration and run the convergence monitor; a converged case is
// deactivated.
__global__ void k_fold_resid(Seg sg, const double* __restrict__ partials,
                             double* __restrict__ resid, const double* __restrict__ stop,
                             double* __restrict__ monitor, int32_t* __restrict__ outer_iters,
                             uint8_t* __restrict__ active) {
  const int k = blockIdx.x;
  if (sg.active && !sg.active[k]) return;
  int b0, nb;
  case_blocks(sg, k, b0, nb);
  const double* pp[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) pp[i] = partials + (size_t)i * sg.G + b0;
  double v[4];
  fold_n<4, 8>(pp, nb, v);
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) resid[4 * k + i] = v[i];
  if (outer_iters) outer_iters[k] += 1;
  if (stop && monitor && active) {
    const double c = v[1] > 0.0 ? v[0] / v[1] : 0.0;
    const double d = v[3] > 0.0 ? sqrt(v[2] / v[3]) : 0.0;
    const double* st = stop + 3 * k;
    const bool conv = monitor_update(monitor + (size_t)k * WCFD_MONITOR_LEN, c, d, st[0], st[1],
                                     st[2]);
    if (st[0] > 0.0 && conv) active[k] = 0;
  }
}

// ELL (slot-major) values -> values on the CSR pattern (row_ptr, col).
__global__ void k_ell_to_csr(wcfd_mesh m, const real* __restrict__ ell, real* __restrict__ csr) {
  const int c = xcd_item();
  const int C = m.n_cells;
  if (c >= C) return;
  csr[m.diag_pos[c]] = ell[c];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int p = m.slot_pos[k * C + c];
    if (p >= 0) csr[p] = ell[(k + 1) * C + c];
  }
}

// ========================================================== host launchers
int launch_set_bc(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, real* fv, int mask,
                  hipStream_t s) {
  const unsigned g = items_grid(bd.bfaces, m.n_boundary);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_SET_BC, s, k_set_bc, g, bd, m.n_faces, m.n_internal, m.n_boundary, bc.type,
              bc.value, fv, mask);
  return check_launch();
}

int launch_face_interp(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* cv,
                       const real* cg, real* fv, int mask, int skew_mask, real* flux,
                       const real* rua, real* rua_f, int apply_dirichlet, hipStream_t s) {
  const unsigned g = items_grid(bd.faces, m.n_faces);
  if (g == 0) return WCFD_OK;
#if WCFD_FACE_LDS
  WCFD_LAUNCH(K_FACE_INTERP, s, k_face_interp_lds, g, m, bd, bc, cv, cg, fv, mask, skew_mask,
              flux, rua, rua_f, apply_dirichlet);
#else
  WCFD_LAUNCH(K_FACE_INTERP, s, k_face_interp, g, m, bd, bc, cv, cg, fv, mask, skew_mask, flux,
              rua, rua_f, apply_dirichlet);
#endif
  return check_launch();
}

int launch_green_gauss(const wcfd_mesh& m, const BatchDev& bd, const real* fv, real* cg,
                       int mask, int zero_others, hipStream_t s) {
  const unsigned g = items_grid(bd.cells, m.n_cells);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_GREEN_GAUSS, s, k_green_gauss, g, m, bd.cells, fv, cg, mask, zero_others);
  return check_launch();
}

int launch_assemble_momentum(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* cv,
                             const real* fv, const real* flux, real* cg, int compute_grad,
                             real nu, real inv_density, real alpha, real dt, int flags,
                             real* values, real* rhs, hipStream_t s) {
  const unsigned g = items_grid(bd.cells, m.n_cells);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_ASM_MOM, s, k_assemble_momentum, g, m, bd, bc, cv, fv, flux, cg, compute_grad, nu,
              inv_density, alpha, dt, flags, values, rhs);
  return check_launch();
}

int launch_assemble_pressure(const wcfd_mesh& m, const BatchDev& bd, wcfd_bc bc, const real* fv,
                             const real* flux, const real* rua_f, real* cg, int compute_grad_p,
                             int correction, int ref_cell, real ref_value, real* values,
                             real* rhs, real* div_out, hipStream_t s) {
  const unsigned g = items_grid(bd.cells, m.n_cells);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_ASM_P, s, k_assemble_pressure, g, m, bd, bc, fv, flux, rua_f, cg, compute_grad_p,
              correction, ref_cell, ref_value, values, rhs, div_out);
  return check_launch();
}

int launch_rua_hbya(const wcfd_mesh& m, const BatchDev& bd, const real* values, real density,
                    const real* cg, real* hbya, real* cv, real* rua, real* u_prev,
                    hipStream_t s) {
  const unsigned g = items_grid(bd.cells, m.n_cells);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_RUA_HBYA, s, k_rua_hbya, g, m, bd, values, density, cg, hbya, cv, rua, u_prev);
  return check_launch();
}

int launch_relax_p(const wcfd_mesh& m, const BatchDev& bd, const real* pn, real alpha, real* cv,
                   hipStream_t s) {
  const unsigned g = items_grid(bd.cells, m.n_cells);
  if (g == 0) return WCFD_OK;
  WCFD_LAUNCH(K_RELAX_P, s, k_relax_p, g, bd, m.n_cells, pn, alpha, cv);
  return check_launch();
}

int launch_correct_velocity(const wcfd_mesh& m, const BatchDev& bd, const Seg& sg, int block,
                            const wcfd_batch* bt, wcfd_bc bc, real* cv, real* cg, real* fv,
                            const real* flux, const real* rua, const real* rua_f, real* hbya,
                            const real* u_prev, double* resid, real* fc_scratch,
                            double* partials, hipStream_t s) {
  real* fc = resid ? fc_scratch : nullptr;
  const unsigned g = items_grid(bd.faces, m.n_faces);
  if (g > 0)
    WCFD_LAUNCH(K_FACE_P_FC, s, k_face_p_fc, g, m, bd, bc, cv, cg, fv, flux, rua_f, fc);
  int e = check_launch();
  if (e) return e;
  WCFD_LAUNCH_B(K_GRAD_P_CORR, s, k_grad_p_correct, sg.G, block, m, sg, fv, cg, cv, hbya, rua, fc,
                u_prev, partials);
  if (resid)
    WCFD_LAUNCH(K_FOLD, s, k_fold_resid, sg.n_cases, sg, partials, resid,
                bt ? bt->stop : nullptr, bt ? bt->monitor : nullptr,
                bt ? bt->outer_iters : nullptr, bt ? bt->active : nullptr);
  return check_launch();
}

}  // namespace wcfd

// ============================================================== C ABI
using namespace wcfd;

// Kernels address field arrays with 32-bit byte offsets per lane, and the
// widest buffer view spans three components of a face array: meshes whose
// 3 max(C, F) reals reach 2^32 bytes are refused.
bool wcfd::mesh_ok(const wcfd_mesh* m) {
  if (!m || m->n_cells <= 0) return false;
  const uint64_t widest = (uint64_t)(m->n_faces > m->n_cells ? m->n_faces : m->n_cells);
  return m->n_faces >= m->n_internal && m->n_internal >= 0 &&
         m->n_boundary == m->n_faces - m->n_internal && m->owner && m->slot_face && m->volume &&
         slots_ok(m) && 3u * widest * sizeof(real) < (1ull << 32);
}

extern "C" {

int wcfd_last_hip_error(void) { return g_last_hip_error; }

const char* wcfd_error_string(int code) {
  switch (code) {
    case WCFD_OK: return "ok";
    case WCFD_EINVAL: return "invalid argument";
    case WCFD_EHIP: return "HIP runtime error";
    case WCFD_ENOTCONV: return "solver did not converge";
    case WCFD_ENAN: return "non-finite residual";
    case WCFD_ESCRATCH:
      return "input exceeds the static scratch of the no-allocation entry point "
             "(WCFD_NOALLOC_MAX_BLOCKS workgroups)";
    case WCFD_EMESH: return "a face is shared by more than two cells";
    case WCFD_ECAPACITY: return "output array too small";
    default: return "unknown error";
  }
}

int wcfd_abi_version(void) { return 3; }

int wcfd_real_bytes(void) { return (int)sizeof(real); }

int wcfd_monitor_update(double* record, double continuity, double velocity_change,
                        double stop_c, double stop_u, double floor_u) {
  if (!record) return -WCFD_EINVAL;
  return monitor_update(record, continuity, velocity_change, stop_c, stop_u, floor_u) ? 1 : 0;
}

int wcfd_set_boundary_conditions(const wcfd_mesh* m, wcfd_bc bc, real* face_values,
                                 int field_mask, void* stream) {
  if (!mesh_ok(m) || !face_values || !bc.type || !bc.value) return WCFD_EINVAL;
  return launch_set_bc(*m, BatchDev{}, bc, face_values, field_mask, (hipStream_t)stream);
}

int wcfd_face_interpolation(const wcfd_mesh* m, wcfd_bc bc, const real* cell_values,
                            const real* cell_grad, real* face_values, int field_mask,
                            int skew_mask, real* mass_flux, const real* rua_cell,
                            real* rua_face, void* stream) {
  if (!mesh_ok(m) || !cell_values || !cell_grad || !face_values) return WCFD_EINVAL;
  if ((rua_cell == nullptr) != (rua_face == nullptr)) return WCFD_EINVAL;
  return launch_face_interp(*m, BatchDev{}, bc, cell_values, cell_grad, face_values, field_mask & 15,
                            skew_mask & 15, mass_flux, rua_cell, rua_face, 0, (hipStream_t)stream);
}

int wcfd_face_interpolation_upwind(const wcfd_mesh* m, wcfd_bc bc, const real* cell_values,
                                   const real* cell_grad, const real* mass_flux,
                                   real* face_values, int field_mask, int skew_mask,
                                   void* stream) {
  if (!mesh_ok(m) || !cell_values || !cell_grad || !mass_flux || !face_values) return WCFD_EINVAL;
  if (m->n_faces <= 0) return WCFD_OK;
  WCFD_LAUNCH(K_FACE_INTERP, (hipStream_t)stream, k_face_interp_upwind, grid_for(m->n_faces), *m,
              bc, cell_values, cell_grad, mass_flux, face_values, field_mask & 15, skew_mask & 15);
  return check_launch();
}

int wcfd_green_gauss(const wcfd_mesh* m, const real* face_values, real* cell_grad, int field_mask,
                     int zero_others, void* stream) {
  if (!mesh_ok(m) || !face_values || !cell_grad) return WCFD_EINVAL;
  return launch_green_gauss(*m, BatchDev{}, face_values, cell_grad, field_mask & 15, zero_others,
                            (hipStream_t)stream);
}

int wcfd_mass_flux(const wcfd_mesh* m, const real* face_values, real* mass_flux, void* stream) {
  if (!mesh_ok(m) || !face_values || !mass_flux) return WCFD_EINVAL;
  WCFD_LAUNCH(K_MASS_FLUX, (hipStream_t)stream, k_mass_flux, grid_for(m->n_faces), *m, face_values, mass_flux);
  return check_launch();
}

int wcfd_divergence(const wcfd_mesh* m, const real* mass_flux, real* div, void* stream) {
  if (!mesh_ok(m) || !mass_flux || !div) return WCFD_EINVAL;
  WCFD_LAUNCH(K_DIVERGENCE, (hipStream_t)stream, k_divergence, grid_for(m->n_cells), *m, mass_flux, div);
  return check_launch();
}

int wcfd_assemble_momentum(const wcfd_mesh* m, wcfd_bc bc, const real* cell_values,
                           const real* face_values, const real* mass_flux, real* cell_grad,
                           int compute_gradients, float nu, float inv_density, float alpha_u,
                           float dt, int scheme_flags, real* values, real* rhs, void* stream) {
  if (!mesh_ok(m) || !cell_values || !face_values || !mass_flux || !cell_grad || !values || !rhs)
    return WCFD_EINVAL;
  return launch_assemble_momentum(*m, BatchDev{}, bc, cell_values, face_values, mass_flux, cell_grad,
                                  compute_gradients, nu, inv_density, alpha_u, dt, scheme_flags,
                                  values, rhs, (hipStream_t)stream);
}

int wcfd_assemble_pressure(const wcfd_mesh* m, wcfd_bc bc, const real* face_values,
                           const real* mass_flux, const real* rua_face, real* cell_grad,
                           int compute_grad_p, int correction, int ref_cell, float ref_value,
                           real* values, real* rhs, real* div_out, void* stream) {
  if (!mesh_ok(m) || !face_values || !mass_flux || !rua_face || !values || !rhs) return WCFD_EINVAL;
  if (compute_grad_p && !cell_grad) return WCFD_EINVAL;
  if (ref_cell >= m->n_cells) return WCFD_EINVAL;
  return launch_assemble_pressure(*m, BatchDev{}, bc, face_values, mass_flux, rua_face, cell_grad,
                                  compute_grad_p, correction, ref_cell, ref_value, values, rhs,
                                  div_out, (hipStream_t)stream);
}

int wcfd_rua_hbya(const wcfd_mesh* m, const real* values, float density, const real* cell_grad,
                  real* hbya, real* cell_values, real* rua, void* stream) {
  if (!mesh_ok(m) || !values || !cell_grad || !hbya || !cell_values || !rua) return WCFD_EINVAL;
  return launch_rua_hbya(*m, BatchDev{}, values, density, cell_grad, hbya, cell_values, rua, nullptr,
                         (hipStream_t)stream);
}

int wcfd_relax_pressure(const wcfd_mesh* m, const real* p_new, float alpha, real* cell_values,
                        void* stream) {
  if (!mesh_ok(m) || !p_new || !cell_values) return WCFD_EINVAL;
  return launch_relax_p(*m, BatchDev{}, p_new, alpha, cell_values, (hipStream_t)stream);
}

int wcfd_ell_to_csr(const wcfd_mesh* m, const real* ell_values, real* csr_values, void* stream) {
  if (!mesh_ok(m) || !m->slot_pos || !m->diag_pos || !ell_values || !csr_values) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  WCFD_LAUNCH(K_ELL_TO_CSR, s, k_ell_to_csr, grid_for(m->n_cells), *m, ell_values, csr_values);
  return check_launch();
}

// Partial rows for n_faces: a hex mesh has at least 3 faces per cell, so
// cell_reduce_blocks(n_faces) covers one row per thread over its cells.
size_t wcfd_reduce_workspace_bytes(int32_t n_faces) {
  Carver c(nullptr);
  c.take<double>(4 * (size_t)cell_reduce_blocks(n_faces));
  c.take<real>(n_faces > 0 ? n_faces : 1);
  return c.off;
}

int wcfd_correct_velocity(const wcfd_mesh* m, wcfd_bc bc, real* cell_values, real* cell_grad,
                          real* face_values, const real* mass_flux, const real* rua,
                          const real* rua_face, real* hbya, double* resid, void* workspace,
                          void* stream) {
  if (!mesh_ok(m) || !cell_values || !cell_grad || !face_values || !rua || !hbya || !rua_face)
    return WCFD_EINVAL;
  if (resid && (!workspace || !mass_flux)) return WCFD_EINVAL;
  if (cell_reduce_blocks(m->n_cells) > cell_reduce_blocks(m->n_faces)) return WCFD_EINVAL;
  Carver w(workspace);
  double* partials = w.take<double>(4 * (size_t)cell_reduce_blocks(m->n_faces));
  real* fc = w.take<real>(m->n_faces);
  const Seg sg = single_seg(m->n_cells, kBlock, (int)cell_reduce_blocks(m->n_cells));
  return launch_correct_velocity(*m, BatchDev{}, sg, kBlock, nullptr, bc, cell_values, cell_grad,
                                 face_values, mass_flux, rua, rua_face, hbya, nullptr, resid, fc,
                                 partials, (hipStream_t)stream);
}

}  // extern "C"

// ======================================================= convergence report
// model.py:278-315 with Convergence.py:7-106: the norms check_convergence
// reports after an outer iteration, on the device and in float64.  One pass
// over the cells writes per-workgroup partials; one workgroup folds them in a
// fixed order, so the report is deterministic.
namespace wcfd {
namespace {

constexpr int kConvN = 9;           // partial rows, see k_conv_cells
constexpr int kConvMaxBlocks = 1024;

// maximum that keeps a NaN (as numpy's norm(x, inf) does)
__device__ inline double max_nan(double a, double b) { return (a > b || a != a) ? a : b; }

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = max_nan(v, __shfl_xor(v, off, kWave));
  return v;
}

// Block reduction of the report rows: 0-5 and 8 are sums, 6 and 7 maxima.
__device__ inline bool conv_is_max(int j) { return j == 6 || j == 7; }

__device__ inline void conv_block(double (&v)[kConvN], double* __restrict__ out, int stride,
                                  int idx) {
  __shared__ double lds[kConvN * kMaxWaves];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < kConvN; ++j) {
    const double s = conv_is_max(j) ? wave_max(v[j]) : wave_sum(v[j]);
    if (lane == 0) lds[j * kMaxWaves + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < kConvN; ++j) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w)
        t = conv_is_max(j) ? max_nan(t, lds[j * kMaxWaves + w]) : t + lds[j * kMaxWaves + w];
      out[j * stride + idx] = t;
    }
  }
}

// Rows of the partials, per workgroup over a grid-stride range of cells:
// 0-2 sum (A u_k - b_k)^2 and 3-5 sum b_k^2 (k = x, y, z; A u in float64 from
// the float32 coefficients and values), 6 max |c_k| with c_k = rua * corr_k
// in float32 as get_gradient('p', coeff = rUA) forms it (or corr_k when rua
// is NULL), 7 max |p|, 8 sum |div F| with div F summed over the cell's faces
// in float32 in local-face order, as divFlux_kernel (model_kernels.py:27-34).
// Rows are segmented per case as in the Krylov kernels (Seg): workgroup vb
// covers its chunk of one case's cells, so each case gets its own partials.
__global__ __launch_bounds__(kBlock) void k_conv_cells(
    wcfd_mesh m, Seg sg, const real* __restrict__ val, const real* __restrict__ u,
    const real* __restrict__ b, const real* __restrict__ rua, const real* __restrict__ corr,
    const real* __restrict__ p, const real* __restrict__ flux, double* __restrict__ part) {
  const int C = m.n_cells;
  const BlockRows br = block_rows(sg);
  double acc[kConvN];
#pragma unroll
  for (int j = 0; j < kConvN; ++j) acc[j] = 0.0;
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    if (val) {
      double r[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) r[k] = (double)val[i] * (double)u[(size_t)k * C + i];
      for (int s = 0; s < 6; ++s) {
        const double a = (double)val[(size_t)(s + 1) * C + i];
        const int c = m.slot_nbr[(size_t)s * C + i];
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] += a * (double)u[(size_t)k * C + c];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double bk = (double)b[(size_t)k * C + i];
        r[k] -= bk;
        acc[k] += r[k] * r[k];
        acc[3 + k] += bk * bk;
      }
    }
    if (corr) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const real c = rua ? rua[i] * corr[(size_t)k * C + i] : corr[(size_t)k * C + i];
        acc[6] = max_nan(acc[6], fabs((double)c));
      }
    }
    if (p) acc[7] = max_nan(acc[7], fabs((double)p[i]));
    if (flux) {
      const int ns = m.n_slots ? m.n_slots : 6;
      real d = 0.f;
      for (int s = 0; s < ns; ++s) {
        const real fl = flux[m.slot_face[(size_t)s * C + i]];
        d = m.slot_side[(size_t)s * C + i] == 0 ? d + fl : d - fl;
      }
      acc[8] += fabs((double)d);
    }
  }
  conv_block(acc, part, sg.G, br.vb);
}

// One workgroup per case k: out[9k + 0-2] ||A u_k - b_k||_2, out[9k + 3-5]
// ||b_k||_2, out[9k + 6] max |c|, out[9k + 7] max |p|, out[9k + 8] sum |div F|
// (all 0 for a case without workgroups).
__global__ __launch_bounds__(kBlock) void k_conv_fold(Seg sg, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  const int k = blockIdx.x;
  int b0, nb;
  case_blocks(sg, k, b0, nb);
  const int G = sg.G;
  double acc[kConvN];
#pragma unroll
  for (int j = 0; j < kConvN; ++j) {
    double t = 0.0;
    for (int g = (int)threadIdx.x; g < nb; g += blockDim.x) {
      const double v = part[(size_t)j * G + b0 + g];
      t = conv_is_max(j) ? max_nan(t, v) : t + v;
    }
    acc[j] = t;
  }
  __shared__ double res[kConvN];
  conv_block(acc, res, 1, 0);
  __syncthreads();
  if (threadIdx.x < kConvN) {
    const int j = (int)threadIdx.x;
    out[(size_t)kConvN * k + j] = j < 6 ? sqrt(res[j]) : res[j];
  }
}

int launch_conv(const wcfd_mesh& m, const Seg& sg, const real* values, const real* u,
                const real* b, const real* rua, const real* corr, const real* p,
                const real* flux, double* part, double* out, hipStream_t s) {
  if (sg.G > 0)
    WCFD_LAUNCH(K_CONVERGENCE, s, k_conv_cells, sg.G, m, sg, values, u, b, rua, corr, p, flux,
                part);
  WCFD_LAUNCH(K_CONVERGENCE, s, k_conv_fold, sg.n_cases, sg, part, out);
  return check_launch();
}

bool conv_args_ok(const wcfd_mesh* m, const real* values, const real* u, const real* b,
                  const real* rua, const real* corr, const real* flux, const void* workspace,
                  const double* out) {
  if (!mesh_ok(m) || !workspace || !out) return false;
  if (values && (!u || !b || !m->slot_nbr)) return false;
  if (flux && !m->slot_side) return false;
  if (rua && !corr) return false;
  return true;
}

}  // namespace
}  // namespace wcfd

extern "C" {

size_t wcfd_convergence_workspace_bytes(void) {
  return sizeof(double) * (size_t)kConvN * (size_t)kConvMaxBlocks;
}

int wcfd_convergence_report(const wcfd_mesh* m, const real* values, const real* u,
                            const real* b, const real* rua, const real* corr, const real* p,
                            const real* flux, void* workspace, double* out, void* stream) {
  if (!conv_args_ok(m, values, u, b, rua, corr, flux, workspace, out)) return WCFD_EINVAL;
  const Seg sg = single_seg(m->n_cells, kBlock, kConvMaxBlocks);
  return launch_conv(*m, sg, values, u, b, rua, corr, p, flux, (double*)workspace, out,
                     (hipStream_t)stream);
}

size_t wcfd_batch_convergence_workspace_bytes(int32_t n_blocks) {
  return sizeof(double) * (size_t)kConvN * (size_t)(n_blocks > 0 ? n_blocks : 1);
}

int wcfd_batch_convergence_report(const wcfd_mesh* m, const wcfd_batch* bt, const real* values,
                                  const real* u, const real* b, const real* rua,
                                  const real* corr, const real* p, const real* flux,
                                  void* workspace, double* out, void* stream) {
  if (!conv_args_ok(m, values, u, b, rua, corr, flux, workspace, out)) return WCFD_EINVAL;
  if (!batch_ok(m, bt)) return WCFD_EINVAL;
  const Seg sg = batch_seg(*bt, m->n_cells);
  return launch_conv(*m, sg, values, u, b, rua, corr, p, flux, (double*)workspace, out,
                     (hipStream_t)stream);
}

}  // extern "C"

// ============================================ report without an allocation
// wcfd_batch_convergence_report_noalloc keeps its partials in this array: one
// copy per GPU, from the first call on that GPU until the process ends.  Host
// code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory); it passes the device address that
// hipGetSymbolAddress returns for the current GPU.  k_conv_cells writes every
// partial that k_conv_fold then reads, so nothing needs clearing between calls.
static __device__ __align__(256) double scratch[wcfd::kConvN * WCFD_NOALLOC_MAX_BLOCKS];

namespace wcfd {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per GPU.
int scratch_address(double** out) {
  int dev = -1;
  if (const int e = check_hip(hipGetDevice(&dev))) return e;
  if (dev < 0 || dev >= kMaxDevices) return WCFD_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (const int e = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return e;
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return WCFD_OK;
}

}  // namespace
}  // namespace wcfd

extern "C" {

int wcfd_batch_convergence_report_noalloc(const wcfd_mesh* m, const wcfd_batch* bt,
                                          const real* values, const real* u, const real* b,
                                          const real* rua, const real* corr, const real* p,
                                          const real* flux, double* out, void* stream) {
  // the workspace check of conv_args_ok is satisfied by a placeholder: the
  // partials go to the static array
  if (!conv_args_ok(m, values, u, b, rua, corr, flux, out, out)) return WCFD_EINVAL;
  if (!batch_ok(m, bt)) return WCFD_EINVAL;
  if (bt->n_blocks > WCFD_NOALLOC_MAX_BLOCKS) return WCFD_ESCRATCH;
  double* part = nullptr;
  if (const int e = scratch_address(&part)) return e;
  const Seg sg = batch_seg(*bt, m->n_cells);
  return launch_conv(*m, sg, values, u, b, rua, corr, p, flux, part, out, (hipStream_t)stream);
}

}  // extern "C"
