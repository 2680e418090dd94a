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

#include <atomic>

namespace wcfd {

// Scalars of one solve.  Quantities that change every iteration sit in
// two-slot rings indexed by iteration parity: block 0 of a kernel writes the
// slot of the current iteration while the other blocks of the same kernel
// read the previous one.
// Flags the host polls (copied to pinned memory after every check).
struct KFlags {
  int32_t done;   // set once converged, non-finite or out of iterations
  int32_t iters;  // iterations run when done was set
  double rr;      // ||r||^2 at the last check
  double thr2;    // squared stopping threshold
};

struct KState {
  KFlags f;
  double rz[2], rho[2], alpha[2], omega[2];
  double atol, tol, bb, ss;
};

__device__ inline double safe_div(double a, double b) { return b != 0.0 ? a / b : 0.0; }

// Every reduction writes one float64 partial per workgroup; the NEXT kernel
// folds the partials in its prologue, every workgroup in the same fixed order,
// so all of them hold bitwise the same scalar.  No kernel waits on another
// workgroup: a cross-workgroup handshake inside a kernel costs an agent-scope
// release (an L2 write-back on every XCD) per workgroup, which measured
// 5-6x slower than this scheme on MI355X at 320 k rows.
__device__ inline double fold(const double* __restrict__ part, int G) {
  __shared__ double lds[kMaxWaves];
  __shared__ double res;
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) s += part[b];
  s = wave_sum(s);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[w];
    res = t;
  }
  __syncthreads();
  return res;
}

// Block sum of ND values; thread 0 stores them as this block's partials
// part[i * G + blockIdx.x].
template <int ND>
__device__ inline void store_partials(double (&v)[ND], double* __restrict__ part, int G) {
  __shared__ double lds[ND * kMaxWaves];
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    double s = wave_sum(v[i]);
    if (lane == 0) lds[i * kMaxWaves + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[i * kMaxWaves + w];
      part[i * G + blockIdx.x] = t;
    }
  }
}

template <class M>
__global__ void k_inv_diag(M A, int n, float* __restrict__ inv) {
  const int i = xcd_item();
  if (i >= n) return;
  const float d = A.diag(i);
  inv[i] = d != 0.0f ? 1.0f / d : 1.0f;
}

template <class M, int NV>
__global__ void k_spmv(M A, int n, const float* __restrict__ x, float* __restrict__ y) {
  const int i = xcd_item();
  if (i >= n) return;
  float acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.f;
  A.row(i, [&](float a, int c) {
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += a * x[(size_t)q * n + c];
  });
#pragma unroll
  for (int q = 0; q < NV; ++q) y[(size_t)q * n + i] = acc[q];
}

__global__ void k_dot(int n, const float* __restrict__ a, const float* __restrict__ b,
                      double* __restrict__ part) {
  double acc[1] = {0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x)
    acc[0] += (double)a[i] * (double)b[i];
  store_partials<1>(acc, part, gridDim.x);
}

__global__ void k_fold_to(const double* __restrict__ part, int G, double* out) {
  const double v = fold(part, G);
  if (threadIdx.x == 0) out[0] = v;
}

// ------------------------------------------------------------------ CG
// partial slots: PA[3][G] = {r.z, r.r, b.b} (init, k_cg_2), PB[1][G] = {p.q} (k_cg_1)
template <class M>
__global__ void k_cg_init(M A, int n, const float* __restrict__ b, const float* __restrict__ x,
                          float* __restrict__ r, const float* __restrict__ inv,
                          double* __restrict__ PA) {
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float ax = 0.f;
    A.row(i, [&](float a, int c) { ax += a * x[c]; });
    const float ri = b[i] - ax;
    r[i] = ri;
    const float zi = ri * inv[i];
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
    acc[2] += (double)b[i] * (double)b[i];
  }
  store_partials<3>(acc, PA, gridDim.x);
}

// one workgroup: fold the init partials, set the threshold and rz[0] / rho[0]
__global__ void k_init_fold(const double* __restrict__ PA, int G, KState* st, int is_cg,
                            int maxiter) {
  const double a = fold(PA, G), rr = fold(PA + G, G), bb = fold(PA + 2 * G, G);
  if (threadIdx.x == 0) {
    if (is_cg) st->rz[0] = a; else st->rho[0] = a;
    st->bb = bb;
    const double thr = fmax(st->tol * sqrt(bb), st->atol);
    st->f.rr = rr;
    st->f.thr2 = thr * thr;
    st->f.iters = 0;
    // converged already, non-finite, or no iterations allowed
    st->f.done = (!(rr > thr * thr) || maxiter <= 0) ? 1 : 0;
    st->alpha[0] = st->alpha[1] = 0.0;
    st->omega[0] = st->omega[1] = 0.0;
  }
}

// one workgroup, after every check_every iterations: fold ||r||^2 and set
// the done flag the iteration kernels test at entry
__global__ void k_check(const double* __restrict__ part, int G, KState* st, int it_end,
                        int maxiter) {
  if (st->f.done) return;
  const double rr = fold(part, G);
  if (threadIdx.x == 0) {
    st->f.rr = rr;
    if (!(rr > st->f.thr2) || it_end >= maxiter) {
      st->f.done = 1;
      st->f.iters = it_end;
    }
  }
}

template <class M>
__global__ void k_cg_1(M A, int n, int it, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p_in,
                       float* __restrict__ p_out, float* __restrict__ q,
                       const double* __restrict__ PA, double* __restrict__ PB, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  bool first = (it == 0);
  double beta_d = 0.0;
  if (!first) {
    const double rz = fold(PA, G);                 // r.z of the previous iteration
    beta_d = safe_div(rz, st->rz[(it - 1) & 1]);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rz[it & 1] = rz;
    first = (beta_d == 0.0);
  }
  const float beta = (float)beta_d;
  double acc[1] = {0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float qi = 0.f;
    A.row(i, [&](float a, int c) {
      const float z = r[c] * inv[c];
      const float pc = first ? z : z + beta * p_in[c];
      qi += a * pc;
    });
    const float zi = r[i] * inv[i];
    const float pi = first ? zi : zi + beta * p_in[i];
    p_out[i] = pi;
    q[i] = qi;
    acc[0] += (double)pi * (double)qi;
  }
  store_partials<1>(acc, PB, G);
}

__global__ void k_cg_2(int n, int it, float* __restrict__ x, float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p,
                       const float* __restrict__ q, const double* __restrict__ PB,
                       double* __restrict__ PA, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  const double pq = fold(PB, G);
  const float alpha = (float)safe_div(st->rz[it & 1], pq);
  double acc[2] = {0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    x[i] = x[i] + alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    const float zi = ri * inv[i];
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
  }
  store_partials<2>(acc, PA, G);
}

// ------------------------------------------- CG with one reduction per iteration
// Chronopoulos-Gear form of the same Jacobi-preconditioned CG: the two inner
// products of an iteration (gamma = r.u, delta = w.u, u = M r, w = A u) are
// taken together, so one fused kernel per iteration updates p, s (= A p), x
// and r, forms the next w by SpMV and writes the partials of gamma, delta
// and r.r.  The p and s recurrences of iteration i are applied at the start
// of kernel i (they need beta_i, known only after the reduction of kernel
// i-1); the neighbour values of r_{i+1} = r_i - alpha_i s_i are formed on the
// fly.  Partials ping-pong between PA and PC: kernel i reads the buffer
// kernel i-1 wrote.
template <class M>
__global__ void k_cgs_init1(M A, int n, const float* __restrict__ b, const float* __restrict__ x,
                            float* __restrict__ r, double* __restrict__ PC) {
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float ax = 0.f;
    A.row(i, [&](float a, int c) { ax += a * x[c]; });
    const float ri = b[i] - ax;
    r[i] = ri;
    acc[1] += (double)ri * (double)ri;
    acc[2] += (double)b[i] * (double)b[i];
  }
  store_partials<3>(acc, PC, gridDim.x);
}

template <class M>
__global__ void k_cgs_init2(M A, int n, const float* __restrict__ r, const float* __restrict__ inv,
                            float* __restrict__ w, double* __restrict__ PA) {
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float wi = 0.f;
    A.row(i, [&](float a, int c) { wi += a * (r[c] * inv[c]); });
    w[i] = wi;
    const float ui = r[i] * inv[i];
    acc[0] += (double)r[i] * (double)ui;
    acc[1] += (double)wi * (double)ui;
    acc[2] += (double)r[i] * (double)r[i];
  }
  store_partials<3>(acc, PA, gridDim.x);
}

__global__ void k_cgs_init_fold(const double* __restrict__ PA, const double* __restrict__ PC, int G,
                                KState* st, int maxiter) {
  const double gamma = fold(PA, G), delta = fold(PA + G, G), rr = fold(PA + 2 * G, G);
  const double bb = fold(PC + 2 * G, G);
  if (threadIdx.x == 0) {
    st->bb = bb;
    const double thr = fmax(st->tol * sqrt(bb), st->atol);
    st->f.rr = rr;
    st->f.thr2 = thr * thr;
    st->f.iters = 0;
    st->f.done = (!(rr > thr * thr) || maxiter <= 0) ? 1 : 0;
    st->rz[0] = gamma;
    st->alpha[0] = safe_div(gamma, delta);
  }
}

template <class M>
__global__ void k_cgs_iter(M A, int n, int it, float* __restrict__ x, const float* __restrict__ r_in,
                           float* __restrict__ r_out, const float* __restrict__ w_in,
                           float* __restrict__ w_out, const float* __restrict__ s_in,
                           float* __restrict__ s_out, float* __restrict__ p,
                           const float* __restrict__ inv, const double* __restrict__ P_in,
                           double* __restrict__ P_out, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  const bool first = (it == 0);
  double alpha_d, beta_d = 0.0;
  if (first) {
    alpha_d = st->alpha[0];
  } else {
    const int pv = (it - 1) & 1;
    const double g = fold(P_in, G), d = fold(P_in + G, G);
    beta_d = safe_div(g, st->rz[pv]);
    alpha_d = safe_div(g, d - beta_d * safe_div(g, st->alpha[pv]));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rz[it & 1] = g;
      st->alpha[it & 1] = alpha_d;
    }
  }
  const float alpha = (float)alpha_d, beta = (float)beta_d;
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float wn = 0.f;
    A.row(i, [&](float a, int c) {
      const float sc = first ? w_in[c] : w_in[c] + beta * s_in[c];
      wn += a * (inv[c] * (r_in[c] - alpha * sc));
    });
    const float ii = inv[i];
    const float ui = r_in[i] * ii;
    const float pi = first ? ui : ui + beta * p[i];
    const float si = first ? w_in[i] : w_in[i] + beta * s_in[i];
    x[i] = x[i] + alpha * pi;
    const float rn = r_in[i] - alpha * si;
    p[i] = pi;
    s_out[i] = si;
    r_out[i] = rn;
    w_out[i] = wn;
    const float un = rn * ii;
    acc[0] += (double)rn * (double)un;
    acc[1] += (double)wn * (double)un;
    acc[2] += (double)rn * (double)rn;
  }
  store_partials<3>(acc, P_out, G);
}

// ------------------------------------------------------------- BiCGSTAB
// partial slots: PA[3][G] = {r0.r, r.r, b.b} (init, k_bi_3), PB[1][G] = {r0.v}
// (k_bi_1), PC[3][G] = {t.s, t.t, s.s} (k_bi_2)
template <class M, int NV>
__global__ void k_bi_init(M A, int n, const float* __restrict__ b, const float* __restrict__ x,
                          float* __restrict__ r, float* __restrict__ r0, double* __restrict__ PA) {
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float ax[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) ax[q] = 0.f;
    A.row(i, [&](float a, int c) {
#pragma unroll
      for (int q = 0; q < NV; ++q) ax[q] += a * x[(size_t)q * n + c];
    });
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float bi = b[k];
      const float ri = bi - ax[q];
      r[k] = ri;
      r0[k] = ri;
      acc[0] += (double)ri * (double)ri;
      acc[1] += (double)ri * (double)ri;
      acc[2] += (double)bi * (double)bi;
    }
  }
  store_partials<3>(acc, PA, gridDim.x);
}

template <class M, int NV>
__global__ void k_bi_1(M A, int n, int it, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p_in,
                       const float* __restrict__ v_in, float* __restrict__ p_out,
                       float* __restrict__ v_out, const float* __restrict__ r0,
                       const double* __restrict__ PA, double* __restrict__ PB, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  bool first = (it == 0);
  double beta_d = 0.0, omega_d = 0.0;
  if (!first) {
    const int pv = (it - 1) & 1;
    const double rho = fold(PA, G);               // r0.r after the previous iteration
    omega_d = st->omega[pv];
    beta_d = safe_div(rho, st->rho[pv]) * safe_div(st->alpha[pv], omega_d);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rho[it & 1] = rho;
    first = (beta_d == 0.0);
  }
  const float beta = (float)beta_d;
  const float omega = (float)omega_d;
  double acc[1] = {0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.f;
    A.row(i, [&](float a, int c) {
      const float ic = inv[c];
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + c;
        const float pc = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
        v[q] += a * (pc * ic);
      }
    });
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const float pi = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
      p_out[k] = pi;
      v_out[k] = v[q];
      acc[0] += (double)r0[k] * (double)v[q];
    }
  }
  store_partials<1>(acc, PB, G);
}

template <class M, int NV>
__global__ void k_bi_2(M A, int n, int it, const float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ v,
                       float* __restrict__ s, float* __restrict__ t,
                       const double* __restrict__ PB, double* __restrict__ PC, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  const double r0v = fold(PB, G);
  const double alpha_d = safe_div(st->rho[it & 1], r0v);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha[it & 1] = alpha_d;
  const float alpha = (float)alpha_d;
  double acc[3] = {0.0, 0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
    float tv[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) tv[q] = 0.f;
    A.row(i, [&](float a, int c) {
      const float ic = inv[c];
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + c;
        const float sc = r[k] - alpha * v[k];
        tv[q] += a * (sc * ic);
      }
    });
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
  store_partials<3>(acc, PC, G);
}

template <int NV>
__global__ void k_bi_3(int n, int it, float* __restrict__ x, float* __restrict__ r,
                       const float* __restrict__ inv, const float* __restrict__ p,
                       const float* __restrict__ s, const float* __restrict__ t,
                       const float* __restrict__ r0, const double* __restrict__ PC,
                       double* __restrict__ PA, KState* st) {
  if (st->f.done) return;
  const int G = gridDim.x;
  const double ts = fold(PC, G), tt = fold(PC + G, G), ss = fold(PC + 2 * G, G);
  const double omega_d = ss > st->f.thr2 ? safe_div(ts, tt) : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) { st->omega[it & 1] = omega_d; st->ss = ss; }
  const float alpha = (float)st->alpha[it & 1];
  const float omega = (float)omega_d;
  double acc[2] = {0.0, 0.0};
  const RowRange rr_ = my_rows(n);
  for (int i = rr_.begin + (int)threadIdx.x; i < rr_.end; i += blockDim.x) {
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
  store_partials<2>(acc, PA, G);
}

// ------------------------------------------------------------- host side
struct KWork {
  float *r, *r0, *p0, *p1, *v0, *v1, *s, *t, *inv;
  double *PA, *PB, *PC;
  KState* st;
};

// The scalar state and the reduction partials come first, at offsets that do
// not depend on nvec: CG (1 vector) and BiCGSTAB (3 vectors) may share one
// workspace.
static KWork carve_krylov(void* ws, int32_t n, int nvec) {
  Carver c(ws);
  KWork w;
  w.st = c.take<KState>(1);
  w.PA = c.take<double>(3 * kMaxReduceBlocks);
  w.PB = c.take<double>(kMaxReduceBlocks);
  w.PC = c.take<double>(3 * kMaxReduceBlocks);
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
  return w;
}

size_t krylov_workspace_bytes(int32_t n, int nvec) {
  Carver c(nullptr);
  c.take<KState>(1);
  c.take<double>(3 * kMaxReduceBlocks);
  c.take<double>(kMaxReduceBlocks);
  c.take<double>(3 * kMaxReduceBlocks);
  const size_t nv = (size_t)n * nvec;
  for (int i = 0; i < 8; ++i) c.take<float>(nv);
  c.take<float>(n);
  return c.off;
}

static int set_tolerances(KState* dev, double atol, double tol, hipStream_t s) {
  double t[2] = {atol, tol};
  int e = check_hip(hipMemcpyAsync(&dev->atol, t, sizeof(t), hipMemcpyHostToDevice, s));
  if (e) return e;
  // the host array must stay alive until the copy has been consumed
  return check_hip(hipStreamSynchronize(s));
}

static void fill_info(wcfd_solve_info* info, const KFlags& h) {
  if (!info) return;
  info->iterations = h.iters;
  info->converged = h.rr <= h.thr2 ? 1 : 0;
  info->residual = sqrt(h.rr);
  info->threshold = sqrt(h.thr2);
}

// Pinned host slots and events for polling the done flag without draining
// the stream (one set per host thread, created on first use).
constexpr int kSlots = 4;
struct HostSync {
  KFlags* slot = nullptr;
  hipEvent_t ev[kSlots] = {};
  hipEvent_t tail = nullptr;
  bool tail_pending = false;
};

static HostSync* host_sync() {
  static thread_local HostSync hs;
  static thread_local bool ready = false;
  if (!ready) {
    if (check_hip(hipHostMalloc((void**)&hs.slot, sizeof(KFlags) * kSlots, hipHostMallocDefault)))
      return nullptr;
    for (auto& e : hs.ev)
      if (check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming))) return nullptr;
    if (check_hip(hipEventCreateWithFlags(&hs.tail, hipEventDisableTiming))) return nullptr;
    ready = true;
  }
  return &hs;
}

// Runs the iteration groups of one solve.  Each group is `ce` iterations
// followed by k_check; its flags are copied to a pinned slot and an event is
// recorded.  The host keeps `depth` groups in flight and stops launching as
// soon as a polled group reports done; groups already queued find the flag
// set and exit at entry, so the result is the same as checking after every
// group synchronously (the reference solver's check interval).
template <class LaunchGroup, class RrPart>
static int run_groups(KState* st, RrPart&& rr_part, unsigned G, int maxiter, int ce,
                      LaunchGroup&& launch_group, wcfd_solve_info* info, hipStream_t s) {
  HostSync* hs = host_sync();
  if (!hs) return WCFD_EHIP;
  int e;
  if (hs->tail_pending) {   // copies of an earlier solve may still target the slots
    if ((e = check_hip(hipEventSynchronize(hs->tail)))) return e;
    hs->tail_pending = false;
  }
  // flags after init (synchronous: one round trip per solve)
  if ((e = check_hip(hipMemcpyAsync(&hs->slot[0], &st->f, sizeof(KFlags), hipMemcpyDeviceToHost, s))))
    return e;
  if ((e = check_hip(hipStreamSynchronize(s)))) return e;
  KFlags last = hs->slot[0];
  if (last.done) {
    fill_info(info, last);
    return isfinite(last.rr) ? WCFD_OK : WCFD_ENAN;
  }
  const int depth = 2;
  int launched = 0, polled = 0, it = 0;
  const int cap = maxiter;
  while (true) {
    if (it < cap && launched - polled < depth) {
      launch_group(it);
      it += ce;
      WCFD_LAUNCH(K_FOLD, s, k_check, 1, rr_part(it), (int)G, st, it, cap);
      const int sl = launched % kSlots;
      if ((e = check_launch())) return e;
      if ((e = check_hip(hipMemcpyAsync(&hs->slot[sl], &st->f, sizeof(KFlags),
                                        hipMemcpyDeviceToHost, s))))
        return e;
      if ((e = check_hip(hipEventRecord(hs->ev[sl], s)))) return e;
      ++launched;
      continue;
    }
    if (polled == launched) break;   // nothing left (cannot happen: the last check sets done)
    const int sl = polled % kSlots;
    if ((e = check_hip(hipEventSynchronize(hs->ev[sl])))) return e;
    last = hs->slot[sl];
    ++polled;
    if (last.done) break;
  }
  if (launched > polled) {
    if ((e = check_hip(hipEventRecord(hs->tail, s)))) return e;
    hs->tail_pending = true;
  }
  fill_info(info, last);
  return isfinite(last.rr) ? WCFD_OK : WCFD_ENAN;
}

namespace {
std::atomic<int> g_krylov_grid{kKrylovGrid};
std::atomic<int> g_krylov_block{kBlock};
}  // namespace

KrylovLaunch krylov_launch() { return {g_krylov_grid.load(), g_krylov_block.load()}; }

// workgroups of the reduction kernels: a grid-stride grid of at most the
// tuned size (every workgroup of the next kernel folds that many partials)
static unsigned krylov_grid(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  const int64_t cap = krylov_launch().grid;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g < 1 ? 1 : g);
}

template <class M>
static int cg_impl(const M& A, int n, const float* b, float* x, double atol, double tol, int maxiter,
                   int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s) {
  KWork w = carve_krylov(workspace, n, 1);
  float* q = w.s;
  int e = set_tolerances(w.st, atol, tol, s);
  if (e) return e;
  const int B = krylov_launch().block;
  const unsigned g = grid_for(n), G = krylov_grid(n, B);
  WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, g, A, n, w.inv);
  WCFD_LAUNCH_B(K_CG_INIT, s, k_cg_init<M>, G, B, A, n, b, x, w.r, w.inv, w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_init_fold, 1, w.PA, (int)G, w.st, 1, maxiter);
  if ((e = check_launch())) return e;
  float* pb[2] = {w.p0, w.p1};
  const int ce = check_every > 0 ? check_every : 1;
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) {
      const int cur = it & 1;
      WCFD_LAUNCH_B(K_CG_1, s, k_cg_1<M>, G, B, A, n, it, w.r, w.inv, pb[cur], pb[1 - cur], q,
                    w.PA, w.PB, w.st);
      WCFD_LAUNCH_B(K_CG_2, s, k_cg_2, G, B, n, it, x, w.r, w.inv, pb[1 - cur], q, w.PB, w.PA,
                    w.st);
    }
  };
  auto rr_part = [&](int) { return (const double*)(w.PA + G); };
  return run_groups(w.st, rr_part, G, maxiter, ce, group, info, s);
}

// Chronopoulos-Gear CG: one fused kernel per iteration.
template <class M>
static int cgs_impl(const M& A, int n, const float* b, float* x, double atol, double tol,
                    int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                    hipStream_t s) {
  KWork w = carve_krylov(workspace, n, 1);
  int e = set_tolerances(w.st, atol, tol, s);
  if (e) return e;
  const int B = krylov_launch().block;
  const unsigned g = grid_for(n), G = krylov_grid(n, B);
  float* rb[2] = {w.r, w.r0};
  float* wb[2] = {w.v0, w.v1};
  float* sb[2] = {w.s, w.t};
  double* pb[2] = {w.PA, w.PC};   // iteration i reads pb[i&1], writes pb[(i+1)&1]
  WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, g, A, n, w.inv);
  WCFD_LAUNCH_B(K_CG_INIT, s, k_cgs_init1<M>, G, B, A, n, b, x, rb[0], w.PC);
  WCFD_LAUNCH_B(K_CG_INIT, s, k_cgs_init2<M>, G, B, A, n, rb[0], w.inv, wb[0], w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_cgs_init_fold, 1, w.PA, w.PC, (int)G, w.st, maxiter);
  if ((e = check_launch())) return e;
  const int ce = check_every > 0 ? check_every : 1;
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) {
      const int cur = it & 1;
      WCFD_LAUNCH_B(K_CG_1, s, k_cgs_iter<M>, G, B, A, n, it, x, rb[cur], rb[1 - cur], wb[cur],
                    wb[1 - cur], sb[cur], sb[1 - cur], w.p0, w.inv, pb[cur], pb[1 - cur], w.st);
    }
  };
  // the r.r partials of the last iteration (it_end - 1) are in pb[it_end & 1], slot 2
  auto rr_part = [&](int it_end) { return (const double*)(pb[it_end & 1] + 2 * G); };
  return run_groups(w.st, rr_part, G, maxiter, ce, group, info, s);
}

template <class M, int NV>
static int bicg_impl(const M& A, int n, const float* b, float* x, double atol, double tol,
                     int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                     hipStream_t s) {
  KWork w = carve_krylov(workspace, n, NV);
  int e = set_tolerances(w.st, atol, tol, s);
  if (e) return e;
  const int B = krylov_launch().block;
  const unsigned g = grid_for(n), G = krylov_grid(n, B);
  WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, g, A, n, w.inv);
  WCFD_LAUNCH_B(K_BI_INIT, s, (k_bi_init<M, NV>), G, B, A, n, b, x, w.r, w.r0, w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_init_fold, 1, w.PA, (int)G, w.st, 0, maxiter);
  if ((e = check_launch())) return e;
  float* pb[2] = {w.p0, w.p1};
  float* vb[2] = {w.v0, w.v1};
  const int ce = check_every > 0 ? check_every : 1;
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) {
      const int cur = it & 1;
      WCFD_LAUNCH_B(K_BI_1, s, (k_bi_1<M, NV>), G, B, A, n, it, w.r, w.inv, pb[cur], vb[cur],
                    pb[1 - cur], vb[1 - cur], w.r0, w.PA, w.PB, w.st);
      WCFD_LAUNCH_B(K_BI_2, s, (k_bi_2<M, NV>), G, B, A, n, it, w.r, w.inv, vb[1 - cur], w.s, w.t,
                    w.PB, w.PC, w.st);
      WCFD_LAUNCH_B(K_BI_3, s, k_bi_3<NV>, G, B, n, it, x, w.r, w.inv, pb[1 - cur], w.s, w.t, w.r0,
                    w.PC, w.PA, w.st);
    }
  };
  auto rr_part = [&](int) { return (const double*)(w.PA + G); };
  return run_groups(w.st, rr_part, G, maxiter, ce, group, info, s);
}

template <class M>
static int bicg_dispatch(const M& A, int n, int nvec, const float* b, float* x, double atol,
                         double tol, int maxiter, int check_every, void* ws,
                         wcfd_solve_info* info, hipStream_t s) {
  if (nvec == 1) return bicg_impl<M, 1>(A, n, b, x, atol, tol, maxiter, check_every, ws, info, s);
  if (nvec == 3) return bicg_impl<M, 3>(A, n, b, x, atol, tol, maxiter, check_every, ws, info, s);
  return WCFD_EINVAL;
}

template <class M>
static int cg_dispatch(const M& A, int n, int variant, const float* b, float* x, double atol,
                       double tol, int maxiter, int check_every, void* ws, wcfd_solve_info* info,
                       hipStream_t s) {
  if (variant == WCFD_CG_SINGLE_REDUCTION)
    return cgs_impl(A, n, b, x, atol, tol, maxiter, check_every, ws, info, s);
  if (variant == WCFD_CG_STANDARD)
    return cg_impl(A, n, b, x, atol, tol, maxiter, check_every, ws, info, s);
  return WCFD_EINVAL;
}

int solve_cg(const CsrView& A, int variant, const float* b, float* x, double atol, double tol,
             int maxiter, int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s) {
  const CsrMat M{A.n, A.row_ptr, A.col, A.values, A.diag_pos};
  return cg_dispatch(M, A.n, variant, b, x, atol, tol, maxiter, check_every, workspace, info, s);
}

int solve_bicgstab(const CsrView& A, int nvec, const float* b, float* x, double atol, double tol,
                   int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                   hipStream_t s) {
  const CsrMat M{A.n, A.row_ptr, A.col, A.values, A.diag_pos};
  return bicg_dispatch(M, A.n, nvec, b, x, atol, tol, maxiter, check_every, workspace, info, s);
}

int solve_cg_ell(const wcfd_mesh& m, const float* values, int variant, const float* b, float* x,
                 double atol, double tol, int maxiter, int check_every, void* workspace,
                 wcfd_solve_info* info, hipStream_t s) {
  const int n = m.n_cells;
  if (m.slot_dnb) {
    const EllMat<true> M{n, values, m.slot_nbr, m.slot_dnb};
    return cg_dispatch(M, n, variant, b, x, atol, tol, maxiter, check_every, workspace, info, s);
  }
  const EllMat<false> M{n, values, m.slot_nbr, m.slot_dnb};
  return cg_dispatch(M, n, variant, b, x, atol, tol, maxiter, check_every, workspace, info, s);
}

int solve_bicgstab_ell(const wcfd_mesh& m, const float* values, int nvec, const float* b, float* x,
                       double atol, double tol, int maxiter, int check_every, void* workspace,
                       wcfd_solve_info* info, hipStream_t s) {
  const int n = m.n_cells;
  if (m.slot_dnb) {
    const EllMat<true> M{n, values, m.slot_nbr, m.slot_dnb};
    return bicg_dispatch(M, n, nvec, b, x, atol, tol, maxiter, check_every, workspace, info, s);
  }
  const EllMat<false> M{n, values, m.slot_nbr, m.slot_dnb};
  return bicg_dispatch(M, n, nvec, b, x, atol, tol, maxiter, check_every, workspace, info, s);
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

int wcfd_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const float* values, int nvec,
              const float* x, float* y, void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !x || !y) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const CsrMat A{n, row_ptr, col, values, nullptr};
  if (nvec == 1)
    WCFD_LAUNCH(K_SPMV, s, (k_spmv<CsrMat, 1>), grid_for(n), A, n, x, y);
  else if (nvec == 3)
    WCFD_LAUNCH(K_SPMV, s, (k_spmv<CsrMat, 3>), grid_for(n), A, n, x, y);
  else
    return WCFD_EINVAL;
  return check_launch();
}

int wcfd_spmv_ell(const wcfd_mesh* m, const float* values, int nvec, const float* x, float* y,
                  void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !values || !x || !y) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int n = m->n_cells;
  if (nvec != 1 && nvec != 3) return WCFD_EINVAL;
  if (m->slot_dnb) {
    const EllMat<true> A{n, values, m->slot_nbr, m->slot_dnb};
    if (nvec == 1) WCFD_LAUNCH(K_SPMV, s, (k_spmv<EllMat<true>, 1>), grid_for(n), A, n, x, y);
    else WCFD_LAUNCH(K_SPMV, s, (k_spmv<EllMat<true>, 3>), grid_for(n), A, n, x, y);
  } else {
    const EllMat<false> A{n, values, m->slot_nbr, m->slot_dnb};
    if (nvec == 1) WCFD_LAUNCH(K_SPMV, s, (k_spmv<EllMat<false>, 1>), grid_for(n), A, n, x, y);
    else WCFD_LAUNCH(K_SPMV, s, (k_spmv<EllMat<false>, 3>), grid_for(n), A, n, x, y);
  }
  return check_launch();
}

int wcfd_dot_f64(int32_t n, const float* a, const float* b, double* out, void* workspace,
                 void* stream) {
  if (n <= 0 || !a || !b || !out || !workspace) return WCFD_EINVAL;
  Carver c(workspace);
  double* part = c.take<double>(kMaxReduceBlocks);
  hipStream_t s = (hipStream_t)stream;
  const unsigned G = reduce_grid(n);
  WCFD_LAUNCH(K_DOT, s, k_dot, G, n, a, b, part);
  WCFD_LAUNCH(K_FOLD, s, k_fold_to, 1, part, (int)G, out);
  return check_launch();
}

int wcfd_set_krylov_launch(int grid, int block) {
  if (grid < 1 || grid > kMaxReduceBlocks || block < kWave || block > kMaxWaves * kWave ||
      block % kWave != 0)
    return WCFD_EINVAL;
  g_krylov_grid = grid;
  g_krylov_block = block;
  return WCFD_OK;
}

int wcfd_get_krylov_launch(int* grid, int* block) {
  if (!grid || !block) return WCFD_EINVAL;
  const KrylovLaunch k = krylov_launch();
  *grid = k.grid;
  *block = k.block;
  return WCFD_OK;
}

size_t wcfd_krylov_workspace_bytes(int32_t n, int nvec) {
  if (n <= 0 || (nvec != 1 && nvec != 3)) return 0;
  return krylov_workspace_bytes(n, nvec);
}

int wcfd_cg(int32_t n, const int32_t* row_ptr, const int32_t* col, const float* values,
            const int32_t* diag_pos, int variant, const float* b, float* x, double atol, double tol,
            int maxiter, int check_every, void* workspace, wcfd_solve_info* info, void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !diag_pos || !b || !x || !workspace)
    return WCFD_EINVAL;
  CsrView A{n, row_ptr, col, values, diag_pos};
  return solve_cg(A, variant, b, x, atol, tol, maxiter, check_every, workspace, info,
                  (hipStream_t)stream);
}

int wcfd_cg_ell(const wcfd_mesh* m, const float* values, int variant, const float* b, float* x,
                double atol, double tol, int maxiter, int check_every, void* workspace,
                wcfd_solve_info* info, void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !values || !b || !x || !workspace)
    return WCFD_EINVAL;
  return solve_cg_ell(*m, values, variant, b, x, atol, tol, maxiter, check_every, workspace, info,
                      (hipStream_t)stream);
}

int wcfd_bicgstab_ell(const wcfd_mesh* m, const float* values, int nvec, const float* b, float* x,
                      double atol, double tol, int maxiter, int check_every, void* workspace,
                      wcfd_solve_info* info, void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !values || !b || !x || !workspace)
    return WCFD_EINVAL;
  return solve_bicgstab_ell(*m, values, nvec, b, x, atol, tol, maxiter, check_every, workspace,
                            info, (hipStream_t)stream);
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
