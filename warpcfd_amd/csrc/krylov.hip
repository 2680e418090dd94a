// Krylov solvers on CSR matrices (gfx950): Jacobi-preconditioned CG and
// right-preconditioned BiCGSTAB, the two solvers the reference takes from
// warp.optim.linear (equation.py:195-200; incompressible.py:32).
//
// Vectors are float32; every dot product and norm accumulates in float64 and
// is folded deterministically (fixed block partition, last-block fold), so a
// solve gives bitwise the same result every run.  Each iteration is a short
// chain of fused kernels:
//   CG:        [p = z + beta p (on the fly) ; q = A p ; p.q]  [x, r update ; r.z, r.r]
//   BiCGSTAB:  [p = r + beta(p - omega v) ; v = A M p ; r0.v]
//              [s = r - alpha v ; t = A M s ; t.s, t.t, s.s]
//              [x += alpha M p + omega M s ; r = s - omega t ; r0.r, r.r]
// The scalars live in device memory; the host reads ||r|| only every
// check_every iterations (the same check interval as the reference's solver).
#include "fv_kernels.h"

#include <math.h>

namespace wcfd {

struct KState {
  double alpha, beta, omega, rho, rz, rr, ss, thr2;
  double atol, tol, bb, pad;
  unsigned ticket[8];
};

__device__ inline double safe_div(double a, double b) { return b != 0.0 ? a / b : 0.0; }

__global__ void k_inv_diag(int n, const float* __restrict__ values, const int32_t* __restrict__ dpos,
                           float* __restrict__ inv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = values[dpos[i]];
  inv[i] = d != 0.0f ? 1.0f / d : 1.0f;
}

template <int NV>
__global__ void k_spmv(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                       const float* __restrict__ val, const float* __restrict__ x,
                       float* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.f;
  for (int j = rp[i]; j < rp[i + 1]; ++j) {
    const float a = val[j];
    const int c = col[j];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += a * x[(size_t)q * n + c];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) y[(size_t)q * n + i] = acc[q];
}

__global__ void k_dot(int n, const float* __restrict__ a, const float* __restrict__ b,
                      double* partials, unsigned* ticket, double* out) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc[0] += (double)a[i] * (double)b[i];
  double tot[1];
  if (grid_sum<1>(acc, partials, ticket, tot) && threadIdx.x == 0) out[0] = tot[0];
}

// ------------------------------------------------------------------ CG
__global__ void k_cg_init(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const float* __restrict__ val, const float* __restrict__ b,
                          const float* __restrict__ x, float* __restrict__ r,
                          const float* __restrict__ inv, double* partials, KState* st) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float ax = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j) ax += val[j] * x[col[j]];
    const float ri = b[i] - ax;
    r[i] = ri;
    const float zi = ri * inv[i];
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
    acc[2] += (double)b[i] * (double)b[i];
  }
  double tot[3];
  if (grid_sum<3>(acc, partials, &st->ticket[0], tot) && threadIdx.x == 0) {
    st->rz = tot[0];
    st->rr = tot[1];
    st->bb = tot[2];
    const double thr = fmax(st->tol * sqrt(tot[2]), st->atol);
    st->thr2 = thr * thr;
    st->beta = 0.0;
    st->alpha = 0.0;
  }
}

__global__ void k_cg_1(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                       const float* __restrict__ val, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p_in,
                       float* __restrict__ p_out, float* __restrict__ q, double* partials,
                       KState* st) {
  const double beta_d = st->beta;
  const float beta = (float)beta_d;
  const bool first = beta_d == 0.0;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float qi = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      const int c = col[j];
      const float z = r[c] * inv[c];
      const float pc = first ? z : z + beta * p_in[c];
      qi += val[j] * pc;
    }
    const float zi = r[i] * inv[i];
    const float pi = first ? zi : zi + beta * p_in[i];
    p_out[i] = pi;
    q[i] = qi;
    acc[0] += (double)pi * (double)qi;
  }
  double tot[1];
  if (grid_sum<1>(acc, partials, &st->ticket[1], tot) && threadIdx.x == 0)
    st->alpha = safe_div(st->rz, tot[0]);
}

__global__ void k_cg_2(int n, float* __restrict__ x, float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p,
                       const float* __restrict__ q, double* partials, KState* st) {
  const float alpha = (float)st->alpha;
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    const float zi = ri * inv[i];
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
  }
  double tot[2];
  if (grid_sum<2>(acc, partials, &st->ticket[2], tot) && threadIdx.x == 0) {
    st->beta = safe_div(tot[0], st->rz);
    st->rz = tot[0];
    st->rr = tot[1];
  }
}

// ------------------------------------------------------------- BiCGSTAB
template <int NV>
__global__ void k_bi_init(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const float* __restrict__ val, const float* __restrict__ b,
                          const float* __restrict__ x, float* __restrict__ r,
                          float* __restrict__ r0, double* partials, KState* st) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float ax[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) ax[q] = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      const float a = val[j];
      const int c = col[j];
#pragma unroll
      for (int q = 0; q < NV; ++q) ax[q] += a * x[(size_t)q * n + c];
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float bi = b[k];
      const float ri = bi - ax[q];
      r[k] = ri;
      r0[k] = ri;
      acc[0] += (double)ri * (double)ri;
      acc[1] += (double)bi * (double)bi;
    }
  }
  double tot[2];
  if (grid_sum<2>(acc, partials, &st->ticket[0], tot) && threadIdx.x == 0) {
    st->rho = tot[0];
    st->rr = tot[0];
    st->bb = tot[1];
    const double thr = fmax(st->tol * sqrt(tot[1]), st->atol);
    st->thr2 = thr * thr;
    st->beta = 0.0;
    st->omega = 0.0;
    st->alpha = 0.0;
  }
}

template <int NV>
__global__ void k_bi_1(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                       const float* __restrict__ val, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p_in,
                       const float* __restrict__ v_in, float* __restrict__ p_out,
                       float* __restrict__ v_out, const float* __restrict__ r0,
                       double* partials, KState* st) {
  const bool first = st->beta == 0.0;
  const float beta = (float)st->beta;
  const float omega = (float)st->omega;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      const float a = val[j];
      const int c = col[j];
      const float ic = inv[c];
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + c;
        const float pc = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
        v[q] += a * (pc * ic);
      }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float pi = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
      p_out[k] = pi;
      v_out[k] = v[q];
      acc[0] += (double)r0[k] * (double)v[q];
    }
  }
  double tot[1];
  if (grid_sum<1>(acc, partials, &st->ticket[1], tot) && threadIdx.x == 0)
    st->alpha = safe_div(st->rho, tot[0]);
}

template <int NV>
__global__ void k_bi_2(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                       const float* __restrict__ val, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ v,
                       float* __restrict__ s, float* __restrict__ t, double* partials,
                       KState* st) {
  const float alpha = (float)st->alpha;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float tv[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) tv[q] = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      const float a = val[j];
      const int c = col[j];
      const float ic = inv[c];
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + c;
        const float sc = r[k] - alpha * v[k];
        tv[q] += a * (sc * ic);
      }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float si = r[k] - alpha * v[k];
      s[k] = si;
      t[k] = tv[q];
      acc[0] += (double)tv[q] * (double)si;
      acc[1] += (double)tv[q] * (double)tv[q];
      acc[2] += (double)si * (double)si;
    }
  }
  double tot[3];
  if (grid_sum<3>(acc, partials, &st->ticket[2], tot) && threadIdx.x == 0) {
    st->ss = tot[2];
    st->omega = tot[2] > st->thr2 ? safe_div(tot[0], tot[1]) : 0.0;
  }
}

template <int NV>
__global__ void k_bi_3(int n, float* __restrict__ x, float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p,
                       const float* __restrict__ s, const float* __restrict__ t,
                       const float* __restrict__ r0, double* partials, KState* st) {
  const float alpha = (float)st->alpha;
  const float omega = (float)st->omega;
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float ii = inv[i];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float y = p[k] * ii;
      const float z = s[k] * ii;
      x[k] = x[k] + (alpha * y + omega * z);
      const float rk = s[k] - omega * t[k];
      r[k] = rk;
      acc[0] += (double)r0[k] * (double)rk;
      acc[1] += (double)rk * (double)rk;
    }
  }
  double tot[2];
  if (grid_sum<2>(acc, partials, &st->ticket[3], tot) && threadIdx.x == 0) {
    st->beta = safe_div(tot[0], st->rho) * safe_div(st->alpha, st->omega);
    st->rho = tot[0];
    st->rr = tot[1];
  }
}

// ------------------------------------------------------------- host side
struct KWork {
  float *r, *r0, *p0, *p1, *v0, *v1, *s, *t, *inv;
  double* partials;
  KState* st;
};

static KWork carve_krylov(void* ws, int32_t n, int nvec) {
  Carver c(ws);
  KWork w;
  const size_t nv = (size_t)n * nvec;
  w.r = c.take<float>(nv);
  w.r0 = c.take<float>(nv);
  w.p0 = c.take<float>(nv);
  w.p1 = c.take<float>(nv);
  w.v0 = c.take<float>(nv);
  w.v1 = c.take<float>(nv);
  w.s = c.take<float>(nv);
  w.t = c.take<float>(nv);
  w.inv = c.take<float>(n);
  w.partials = c.take<double>(4 * kMaxReduceBlocks);
  w.st = c.take<KState>(1);
  return w;
}

size_t krylov_workspace_bytes(int32_t n, int nvec) {
  Carver c(nullptr);
  const size_t nv = (size_t)n * nvec;
  for (int i = 0; i < 8; ++i) c.take<float>(nv);
  c.take<float>(n);
  c.take<double>(4 * kMaxReduceBlocks);
  c.take<KState>(1);
  return c.off;
}

static int read_state(KState* dev, KState* host, hipStream_t s) {
  int e = check_hip(hipMemcpyAsync(host, dev, sizeof(KState), hipMemcpyDeviceToHost, s));
  if (e) return e;
  return check_hip(hipStreamSynchronize(s));
}

static int set_tolerances(KState* dev, double atol, double tol, hipStream_t s) {
  double t[2] = {atol, tol};
  int e = check_hip(hipMemcpyAsync(&dev->atol, t, sizeof(t), hipMemcpyHostToDevice, s));
  if (e) return e;
  // the host array must stay alive until the copy has been consumed
  return check_hip(hipStreamSynchronize(s));
}

static void fill_info(wcfd_solve_info* info, int it, const KState& h) {
  if (!info) return;
  info->iterations = it;
  info->converged = h.rr <= h.thr2 ? 1 : 0;
  info->residual = sqrt(h.rr);
  info->threshold = sqrt(h.thr2);
}

int solve_cg(const CsrView& A, const float* b, float* x, double atol, double tol, int maxiter,
             int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s) {
  const int n = A.n;
  KWork w = carve_krylov(workspace, n, 1);
  float* q = w.s;
  int e = set_tolerances(w.st, atol, tol, s);
  if (e) return e;
  const unsigned g = grid_for(n), rg = reduce_grid(n);
  WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag, g, n, A.values, A.diag_pos, w.inv);
  WCFD_LAUNCH(K_CG_INIT, s, k_cg_init, rg, n, A.row_ptr, A.col, A.values, b, x, w.r, w.inv, w.partials, w.st);
  if ((e = check_launch())) return e;
  KState h;
  if ((e = read_state(w.st, &h, s))) return e;
  int it = 0;
  if (!(h.rr > h.thr2) || maxiter <= 0) {  // converged already (or NaN: reported by caller)
    fill_info(info, 0, h);
    return isfinite(h.rr) ? WCFD_OK : WCFD_ENAN;
  }
  float* pb[2] = {w.p0, w.p1};
  int cur = 0;
  const int ce = check_every > 0 ? check_every : 1;
  while (true) {
    for (int k = 0; k < ce; ++k) {
      WCFD_LAUNCH(K_CG_1, s, k_cg_1, rg, n, A.row_ptr, A.col, A.values, w.r, w.inv, pb[cur], pb[1 - cur],
                                   q, w.partials, w.st);
      WCFD_LAUNCH(K_CG_2, s, k_cg_2, rg, n, x, w.r, w.inv, pb[1 - cur], q, w.partials, w.st);
      cur = 1 - cur;
    }
    if ((e = check_launch())) return e;
    it += ce;
    if ((e = read_state(w.st, &h, s))) return e;
    if (!isfinite(h.rr)) {
      fill_info(info, it, h);
      return WCFD_ENAN;
    }
    if (h.rr <= h.thr2 || it >= maxiter) break;
  }
  fill_info(info, it, h);
  return WCFD_OK;
}

template <int NV>
static int bicg_impl(const CsrView& A, const float* b, float* x, double atol, double tol,
                     int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                     hipStream_t s) {
  const int n = A.n;
  KWork w = carve_krylov(workspace, n, NV);
  int e = set_tolerances(w.st, atol, tol, s);
  if (e) return e;
  const unsigned g = grid_for(n), rg = reduce_grid(n);
  WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag, g, n, A.values, A.diag_pos, w.inv);
  WCFD_LAUNCH(K_BI_INIT, s, k_bi_init<NV>, rg, n, A.row_ptr, A.col, A.values, b, x, w.r, w.r0, w.partials,
                                       w.st);
  if ((e = check_launch())) return e;
  KState h;
  if ((e = read_state(w.st, &h, s))) return e;
  int it = 0;
  if (!(h.rr > h.thr2) || maxiter <= 0) {
    fill_info(info, 0, h);
    return isfinite(h.rr) ? WCFD_OK : WCFD_ENAN;
  }
  float* pb[2] = {w.p0, w.p1};
  float* vb[2] = {w.v0, w.v1};
  int cur = 0;
  const int ce = check_every > 0 ? check_every : 1;
  while (true) {
    for (int k = 0; k < ce; ++k) {
      WCFD_LAUNCH(K_BI_1, s, k_bi_1<NV>, rg, n, A.row_ptr, A.col, A.values, w.r, w.inv, pb[cur], vb[cur],
                                        pb[1 - cur], vb[1 - cur], w.r0, w.partials, w.st);
      WCFD_LAUNCH(K_BI_2, s, k_bi_2<NV>, rg, n, A.row_ptr, A.col, A.values, w.r, w.inv, vb[1 - cur], w.s,
                                        w.t, w.partials, w.st);
      WCFD_LAUNCH(K_BI_3, s, k_bi_3<NV>, rg, n, x, w.r, w.inv, pb[1 - cur], w.s, w.t, w.r0, w.partials,
                                        w.st);
      cur = 1 - cur;
    }
    if ((e = check_launch())) return e;
    it += ce;
    if ((e = read_state(w.st, &h, s))) return e;
    if (!isfinite(h.rr)) {
      fill_info(info, it, h);
      return WCFD_ENAN;
    }
    if (h.rr <= h.thr2 || it >= maxiter) break;
  }
  fill_info(info, it, h);
  return WCFD_OK;
}

int solve_bicgstab(const CsrView& A, int nvec, const float* b, float* x, double atol, double tol,
                   int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                   hipStream_t s) {
  if (nvec == 1) return bicg_impl<1>(A, b, x, atol, tol, maxiter, check_every, workspace, info, s);
  if (nvec == 3) return bicg_impl<3>(A, b, x, atol, tol, maxiter, check_every, workspace, info, s);
  return WCFD_EINVAL;
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

int wcfd_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const float* values, int nvec,
              const float* x, float* y, void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !x || !y) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (nvec == 1)
    WCFD_LAUNCH(K_SPMV, s, k_spmv<1>, grid_for(n), n, row_ptr, col, values, x, y);
  else if (nvec == 3)
    WCFD_LAUNCH(K_SPMV, s, k_spmv<3>, grid_for(n), n, row_ptr, col, values, x, y);
  else
    return WCFD_EINVAL;
  return check_launch();
}

int wcfd_dot_f64(int32_t n, const float* a, const float* b, double* out, void* workspace,
                 void* stream) {
  if (n <= 0 || !a || !b || !out || !workspace) return WCFD_EINVAL;
  Carver c(workspace);
  double* partials = c.take<double>(2 * kMaxReduceBlocks);
  unsigned* ticket = c.take<unsigned>(64);
  WCFD_LAUNCH(K_DOT, (hipStream_t)stream, k_dot, reduce_grid(n), n, a, b, partials, ticket, out);
  return check_launch();
}

size_t wcfd_krylov_workspace_bytes(int32_t n, int nvec) {
  if (n <= 0 || (nvec != 1 && nvec != 3)) return 0;
  return krylov_workspace_bytes(n, nvec);
}

int wcfd_cg(int32_t n, const int32_t* row_ptr, const int32_t* col, const float* values,
            const int32_t* diag_pos, const float* b, float* x, double atol, double tol, int maxiter,
            int check_every, void* workspace, wcfd_solve_info* info, void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !diag_pos || !b || !x || !workspace)
    return WCFD_EINVAL;
  CsrView A{n, row_ptr, col, values, diag_pos};
  return solve_cg(A, b, x, atol, tol, maxiter, check_every, workspace, info, (hipStream_t)stream);
}

int wcfd_bicgstab(int32_t n, const int32_t* row_ptr, const int32_t* col, const float* values,
                  const int32_t* diag_pos, int nvec, const float* b, float* x, double atol,
                  double tol, int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                  void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !diag_pos || !b || !x || !workspace)
    return WCFD_EINVAL;
  CsrView A{n, row_ptr, col, values, diag_pos};
  return solve_bicgstab(A, nvec, b, x, atol, tol, maxiter, check_every, workspace, info,
                        (hipStream_t)stream);
}

}  // extern "C"
