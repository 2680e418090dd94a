// Krylov solvers (gfx950): Jacobi-preconditioned CG -- in the standard
// two-reduction form and in the single-reduction (Chronopoulos-Gear) form --
// and right-preconditioned BiCGSTAB, the solvers the reference takes from
// warp.optim.linear (equation.py:195-200; incompressible.py:32).
//
// One launch solves K independent systems (a batch of cases whose matrices
// sit block-diagonally in one array; K = 1 for a single case).  Every case
// owns a contiguous range of rows and of workgroups (struct Seg), so each case
// has its own dot products, scalars, stopping flag and iteration count, and
// its arithmetic does not depend on which other cases share the launch.
//
// Vectors are float32; dot products and norms accumulate in float64.  Every
// reduction writes one partial per workgroup; the NEXT kernel folds its case's
// partials in its prologue (all workgroups of the case in the same order), so
// no kernel waits on another workgroup -- a cross-workgroup handshake costs
// an agent-scope release per workgroup, measured 5-6x slower on MI355X at
// 320 k rows -- and every run is bitwise reproducible.  Folding in every
// workgroup costs G^2 reads per kernel, so above a set number of workgroups
// per case (Seg::prefold) a fold kernel with one workgroup per case first
// reduces the partials to one value, in the same order and workgroup size,
// and the readers take that value: the sums are bitwise the same either way.
// Iteration kernels:
//   CG:         [p = z + beta p ; q = A p ; p.q]  [x, r update ; r.z, r.r]
//   CG (1-red): [p, s recurrences, x, r updates, w = A M r ; r.u, w.u, r.r]
//   BiCGSTAB:   [p = r + beta(p - omega v) ; v = A M p ; r0.v]
//               [s = r - alpha v ; t = A M s ; t.s, t.t, s.s]
//               [x += alpha M p + omega M s ; r = s - omega t ; r0.r, r.r]
// A check kernel (one workgroup per case) after every `check_every`
// iterations folds ||r||^2 and sets the case's done flag; the iteration
// kernels of a done case return at entry.
#include "fv_kernels.h"

#include <math.h>

#include <atomic>
#include <mutex>
#include <vector>

namespace wcfd {

// Flags the host polls (copied to pinned memory after every check).
struct KFlags {
  int32_t done;   // converged, non-finite, out of iterations or inactive
  int32_t iters;  // iterations run when done was set
  double rr;      // ||r||^2 at the last check
  double thr2;    // squared stopping threshold
};

// Scalars of one case.  Values that change every iteration sit in two-slot
// rings indexed by iteration parity: the case's first workgroup writes the
// slot of the current iteration while the others read the previous one.
struct KState {
  KFlags f;
  double rz[2], rho[2], alpha[2], omega[2];
  double atol, tol, rtol0, bb, ss;
};

__device__ inline double safe_div(double a, double b) { return b != 0.0 ? a / b : 0.0; }

// ------------------------------------------------------------ small kernels
template <class M>
__global__ void k_inv_diag(M A, Items items, int n, real* __restrict__ inv) {
  const int i = item_index(items);
  if (i >= n) return;
  const real d = A.diag(i);
  inv[i] = d != 0.0f ? 1.0f / d : 1.0f;
}

template <class M, int NV>
__global__ void k_spmv(M A, int n, const real* __restrict__ x, real* __restrict__ y) {
  const int i = xcd_item();
  if (i >= n) return;
  real acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = 0.f;
  A.row(i, [&](real a, int c) {
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += a * x[(size_t)q * n + c];
  });
#pragma unroll
  for (int q = 0; q < NV; ++q) y[(size_t)q * n + i] = acc[q];
}

__global__ __launch_bounds__(kMaxKrylovBlock) void k_dot(Seg sg, const real* __restrict__ a, const real* __restrict__ b,
                      double* __restrict__ part) {
  const BlockRows br = block_rows(sg);
  double acc[1] = {0.0};
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x)
    acc[0] += (double)a[i] * (double)b[i];
  store_partials<1>(acc, part, sg.G, br.vb);
}

__global__ void k_fold_to(const double* __restrict__ part, int G, double* out) {
  const double v = fold(part, G);
  if (threadIdx.x == 0) out[0] = v;
}

__global__ void k_set_tolerances(KState* st, int n_cases, Tol tl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_cases) return;
  st[k].atol = tl.atol;
  st[k].tol = tl.tol;
  st[k].rtol0 = tl.rtol0;
}

// Init fold, one workgroup per case: threshold, first scalars and the done
// flag.  kind 0 (CG): P = {r.z, r.r, b.b}.  kind 1 (BiCGSTAB):
// P = {r0.r, r.r, b.b}.  kind 2 (single-reduction CG): P = {r.u, w.u, r.r}
// and Pb holds b.b in slot 2.
__global__ void k_init_fold(Seg sg, const double* __restrict__ P, const double* __restrict__ Pb,
                            KState* st, int kind, int maxiter) {
  const int k = blockIdx.x;
  int b0, nb;
  case_blocks(sg, k, b0, nb);
  const int G = sg.G;
  const double a = fold(P + b0, nb);
  const double x1 = fold(P + G + b0, nb);
  const double x2 = fold((kind == 2 ? Pb : P) + 2 * G + b0, nb);
  const double x3 = kind == 2 ? fold(P + 2 * G + b0, nb) : 0.0;
  if (threadIdx.x != 0) return;
  const double rr = kind == 2 ? x3 : x1;
  const double bb = x2;
  KState* S = st + k;
  S->bb = bb;
  const double thr = fmax(fmax(S->tol * sqrt(bb), S->rtol0 * sqrt(rr)), S->atol);
  S->f.rr = rr;
  S->f.thr2 = thr * thr;
  S->f.iters = 0;
  const bool inactive = sg.active && !sg.active[k];
  // converged already, non-finite, no iterations allowed, or not iterating
  S->f.done = (inactive || !(rr > thr * thr) || maxiter <= 0) ? 1 : 0;
  S->alpha[0] = S->alpha[1] = 0.0;
  S->omega[0] = S->omega[1] = 0.0;
  if (kind == 1) {
    S->rho[0] = a;
  } else {
    S->rz[0] = a;
    if (kind == 2) S->alpha[0] = safe_div(a, x1);
  }
}

// One workgroup per case, after every check_every iterations: fold ||r||^2,
// set the done flag, and write the case's flag record to `out` -- pinned
// host memory the host reads after an event on this stream, so no copy is
// queued (a strided device-to-host copy costs ~20 us per check on MI355X).
__global__ void k_check(Seg sg, const double* __restrict__ part, KState* st, int it_end,
                        int maxiter, KFlags* out) {
  const int k = blockIdx.x;
  KState* S = st + k;
  if (!S->f.done) {
    int b0, nb;
    case_blocks(sg, k, b0, nb);
    const double rr = fold(part + b0, sg.prefold ? 1 : nb);
    if (threadIdx.x == 0) {
      S->f.rr = rr;
      if (!(rr > S->f.thr2) || it_end >= maxiter) {
        S->f.done = 1;
        S->f.iters = it_end;
      }
    }
  }
  if (out && threadIdx.x == 0) {
    out[k] = S->f;
    __threadfence_system();   // visible to the host once the kernel has completed
  }
}

// The stop test of k_check made inside an iteration kernel (single-case
// single-reduction CG): the kernel of iteration `it` = m * check_every folds
// ||r_it||^2 with its other partials, and its case's first workgroup sets the
// flags and writes the record; the whole case stops before updating when the
// test passes.  Every workgroup of the case folds the same partials in the
// same order as k_check does, so each reaches the same decision, and the
// iterations run, the result and the record are those of a separate check.
__device__ inline void publish_flags(KFlags* out, const KFlags& f) {
  if (out) {
    *out = f;
    __threadfence_system();   // visible to the host once the kernel has completed
  }
}

// true: stop (the case's flags are set); called by every thread of the case
__device__ inline bool inline_check(const BlockRows& br, KState* S, double rr, int it,
                                    int maxiter, KFlags* out) {
  const bool stop = !(rr > S->f.thr2) || it >= maxiter;
  if (br.vb == br.b0 && threadIdx.x == 0) {
    S->f.rr = rr;
    if (stop) {
      S->f.done = 1;
      S->f.iters = it;
    }
    publish_flags(out, S->f);
  }
  return stop;
}

// Prefold, one workgroup per case (launched with the iteration kernels'
// workgroup size): the ND partial rows of a case are folded in the order of
// fold() and the sums replace the case's first partial of each row.  Every
// thread's reads finish before the first barrier inside fold_n, so thread
// 0's in-place store cannot race them.  Done cases keep their partials.
template <int ND>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_prefold(Seg sg, double* __restrict__ P,
                                                            const KState* st) {
  const int k = blockIdx.x;
  if (st[k].f.done) return;
  int b0, nb;
  case_blocks(sg, k, b0, nb);
  const int G = sg.G;
  const double* pp[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) pp[i] = P + (size_t)i * G + b0;
  double v[ND];
  fold_n<ND, WCFD_PREFOLD_UNROLL>(pp, nb, v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < ND; ++i) P[(size_t)i * G + b0] = v[i];
  }
}

// ------------------------------------------------------------------ CG
// partial slots: PA[3][G] = {r.z, r.r, b.b} (init, k_cg_2), PB[1][G] = {p.q} (k_cg_1)
// Initial residual r = b - A x.  With R64 the product and the difference are
// accumulated in float64 and rounded once (the residual of mixed-precision
// iterative refinement); otherwise in float32, as the reference's solver.
template <bool R64, class M>
__device__ inline real residual_row(const M& A, int i, const real* __restrict__ b,
                                     const real* __restrict__ x) {
  if constexpr (R64) {
    double ax = 0.0;
    A.row(i, [&](real a, int c) { ax += (double)a * (double)x[c]; });
    return (real)((double)b[i] - ax);
  } else {
    real ax = 0.f;
    A.row(i, [&](real a, int c) { ax += a * x[c]; });
    return b[i] - ax;
  }
}

template <class M, bool R64>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cg_init(M A, Seg sg, const real* __restrict__ b, const real* __restrict__ x,
                          real* __restrict__ r, const real* __restrict__ inv,
                          double* __restrict__ PA) {
  const BlockRows br = block_rows(sg);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    const real ri = residual_row<R64>(A, i, b, x);
    r[i] = ri;
    const real zi = ri * inv[i];
    acc[0] += (double)ri * (double)zi;
    acc[1] += (double)ri * (double)ri;
    acc[2] += (double)b[i] * (double)b[i];
  }
  store_partials<3>(acc, PA, sg.G, br.vb);
}

template <class M, bool PF>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cg_1(M A, Seg sg, int it, const real* __restrict__ r,
                       const real* __restrict__ inv, const real* __restrict__ p_in,
                       real* __restrict__ p_out, real* __restrict__ q,
                       const double* __restrict__ PA, double* __restrict__ PB, KState* st) {
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) return;
  // one row per thread on an ELL matrix: issue the row's loads before the fold
  constexpr bool pf = PF && M::kEll;
  const double* const fp[1] = {PA + br.b0};
  EarlyFold<1> ef;
  if constexpr (pf)
    if (it) ef.load(fp, br.nf);
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real a[M::kRow], rc[M::kRow], ic[M::kRow], pc[M::kRow];
  if constexpr (M::kEll) {
    if (has) {
      uint32_t o[M::kRow];
      A.load_off(i0, a, o);
      const uint32_t nb4 = (uint32_t)A.n * 4u;
      const FBuf R(r, nb4), I(inv, nb4), P(p_in, nb4);
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        rc[j] = R(o[j]);
        ic[j] = I(o[j]);
        pc[j] = it ? P(o[j]) : 0.f;
      }
    }
  }
  bool first = (it == 0);
  double beta_d = 0.0;
  if (!first) {
    double rz;                                   // r.z of the previous iteration
    if constexpr (pf) {
      double o1[1];
      ef.finish(fp, br.nf, o1);
      rz = o1[0];
    } else {
      rz = fold(PA + br.b0, br.nf);
    }
    beta_d = safe_div(rz, S->rz[(it - 1) & 1]);
    if (br.vb == br.b0 && threadIdx.x == 0) S->rz[it & 1] = rz;
    first = (beta_d == 0.0);
  }
  const real beta = (real)beta_d;
  double acc[1] = {0.0};
  if (pf) {
    if (has) {
      real qi = 0.f;
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        const real z = rc[j] * ic[j];
        qi += a[j] * (first ? z : z + beta * pc[j]);
      }
      const real zi = rc[0] * ic[0];
      const real pi = first ? zi : zi + beta * pc[0];
      const uint32_t o0 = (uint32_t)i0 * 4u;
      st_off(p_out, o0, pi);
      st_off(q, o0, qi);
      acc[0] += (double)pi * (double)qi;
    }
  } else {
    for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
      real qi = 0.f;
      A.row(i, [&](real a, int c) {
        const real z = r[c] * inv[c];
        const real pc = first ? z : z + beta * p_in[c];
        qi += a * pc;
      });
      const real zi = r[i] * inv[i];
      const real pi = first ? zi : zi + beta * p_in[i];
      p_out[i] = pi;
      q[i] = qi;
      acc[0] += (double)pi * (double)qi;
    }
  }
  store_partials<1>(acc, PB, sg.G, br.vb);
}

template <bool PF>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cg_2(Seg sg, int it, real* __restrict__ x, real* __restrict__ r,
                       const real* __restrict__ inv, const real* __restrict__ p,
                       const real* __restrict__ q, const double* __restrict__ PB,
                       double* __restrict__ PA, KState* st) {
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) return;
  // one row per thread: load it before the fold
  constexpr bool pf = PF;
  const double* const fp[1] = {PB + br.b0};
  EarlyFold<1> ef;
  if constexpr (pf) ef.load(fp, br.nf);
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real xv = 0.f, pv = 0.f, rv = 0.f, qv = 0.f, iv = 0.f;
  if (has) {
    xv = x[i0];
    pv = p[i0];
    rv = r[i0];
    qv = q[i0];
    iv = inv[i0];
  }
  double pq;
  if constexpr (pf) {
    double o1[1];
    ef.finish(fp, br.nf, o1);
    pq = o1[0];
  } else {
    pq = fold(PB + br.b0, br.nf);
  }
  const real alpha = (real)safe_div(S->rz[it & 1], pq);
  double acc[2] = {0.0, 0.0};
  if (pf) {
    if (has) {
      x[i0] = xv + alpha * pv;
      const real ri = rv - alpha * qv;
      r[i0] = ri;
      const real zi = ri * iv;
      acc[0] += (double)ri * (double)zi;
      acc[1] += (double)ri * (double)ri;
    }
  } else {
    for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
      x[i] = x[i] + alpha * p[i];
      const real ri = r[i] - alpha * q[i];
      r[i] = ri;
      const real zi = ri * inv[i];
      acc[0] += (double)ri * (double)zi;
      acc[1] += (double)ri * (double)ri;
    }
  }
  store_partials<2>(acc, PA, sg.G, br.vb);
}

// ------------------------------------------- CG with one reduction per iteration
// One row of the single-reduction iteration (a = the row of A; wc, sc, rc,
// ic = w, s, r, inv at the row's columns, the row itself first): the next
// w = A M r, p, s and r of the row.  M is applied to the vector, as the
// reference's Jacobi preconditioner does; a matrix scaled once by M would
// round A M once per solve, a fixed perturbation of the operator that
// breaks its zero row sums (measured: 10x further from the float64 build).
// Every multiply-add is explicit, so every kernel form rounds the same way.
template <int KR, bool FIRST>
__device__ __forceinline__ void cg_row(const real (&a)[KR], const real (&wc)[KR],
                                       const real (&sc)[KR], const real (&rc)[KR],
                                       const real (&ic)[KR], real p0, real alpha, real beta,
                                       real& wn, real& pi, real& si, real& rn) {
  wn = 0.f;
#pragma unroll
  for (int j = 0; j < KR; ++j) {
    const real scj = FIRST ? wc[j] : mad(beta, sc[j], wc[j]);
    wn = mad(a[j], ic[j] * mad(-alpha, scj, rc[j]), wn);
  }
  const real ii = ic[0];
  const real ui = rc[0] * ii;
  pi = FIRST ? ui : mad(beta, p0, ui);
  si = FIRST ? wc[0] : mad(beta, sc[0], wc[0]);
  rn = mad(-alpha, si, rc[0]);
}

// the row's shares of r.u, w.u and r.r (u = M r), accumulated in float64
__device__ __forceinline__ void cg_dots(double (&acc)[3], real rn, real wn, real ii) {
  const real un = rn * ii;
  acc[0] = __builtin_fma((double)rn, (double)un, acc[0]);
  acc[1] = __builtin_fma((double)wn, (double)un, acc[1]);
  acc[2] = __builtin_fma((double)rn, (double)rn, acc[2]);
}
// Chronopoulos-Gear form of the same Jacobi-preconditioned CG: the two inner
// products of an iteration (gamma = r.u, delta = w.u, u = M r, w = A u) are
// taken together, so one fused kernel per iteration updates p, s (= A p), x
// and r, forms the next w by SpMV and writes the partials of gamma, delta
// and r.r.  The p and s recurrences of iteration i are applied at the start
// of kernel i (they need beta_i, known only after the reduction of kernel
// i-1); the neighbour values of r_{i+1} = r_i - alpha_i s_i are formed on the
// fly.  Partials ping-pong between PA and PC: kernel i reads the buffer
// kernel i-1 wrote.
template <class M, bool R64>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cgs_init1(M A, Seg sg, const real* __restrict__ b, const real* __restrict__ x,
                            real* __restrict__ r, double* __restrict__ PC) {
  const BlockRows br = block_rows(sg);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    const real ri = residual_row<R64>(A, i, b, x);
    r[i] = ri;
    acc[1] += (double)ri * (double)ri;
    acc[2] += (double)b[i] * (double)b[i];
  }
  store_partials<3>(acc, PC, sg.G, br.vb);
}

template <class M>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cgs_init2(M A, Seg sg, const real* __restrict__ r, const real* __restrict__ inv,
                            real* __restrict__ w, double* __restrict__ PA) {
  const BlockRows br = block_rows(sg);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    real wi = 0.f;
    A.row(i, [&](real a, int c) { wi = mad(a, r[c] * inv[c], wi); });
    w[i] = wi;
    cg_dots(acc, r[i], wi, inv[i]);
  }
  store_partials<3>(acc, PA, sg.G, br.vb);
}

template <class M, bool PF, bool FIRST, bool CHECK = false>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cgs_iter(M A, Seg sg, int it, real* __restrict__ x,
                           const real* __restrict__ r_in, real* __restrict__ r_out,
                           const real* __restrict__ w_in, real* __restrict__ w_out,
                           const real* __restrict__ s_in, real* __restrict__ s_out,
                           real* __restrict__ p, const real* __restrict__ inv,
                           const double* __restrict__ P_in, double* __restrict__ P_out,
                           KState* st, int maxiter, KFlags* out) {
  static_assert(!(FIRST && CHECK), "iteration 0 is checked by the init fold");
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) {
    if (CHECK && br.vb == br.b0 && threadIdx.x == 0) publish_flags(out, S->f);
    return;
  }
  const int G = sg.G;
  // one row per thread on an ELL matrix: issue the row's loads before the
  // fold, and the fold's own loads (and the case's scalars) before those
  constexpr bool pf = PF && M::kEll;
  constexpr int ND = CHECK ? 3 : 2;   // r.u, w.u (and r.r for the stop test)
  const int pv = (it - 1) & 1;
  const double* pp[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) pp[i] = P_in + (size_t)i * G + br.b0;
  EarlyFold<ND> ef;
  double rz_prev = 0.0, alpha_prev = 0.0;
  if constexpr (!FIRST) {
    if constexpr (pf) ef.load(pp, br.nf);
    rz_prev = S->rz[pv];
    alpha_prev = S->alpha[pv];
  }
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real a[M::kRow], wc[M::kRow], sc[M::kRow], ic[M::kRow], rc[M::kRow];
  real p0 = 0.f, x0 = 0.f;
  const uint32_t o0 = (uint32_t)i0 * 4u;
  if constexpr (M::kEll) {
    if (has) {
      uint32_t o[M::kRow];
      A.load_off(i0, a, o);
      const uint32_t nb4 = (uint32_t)A.n * 4u;
      const FBuf W(w_in, nb4), Sv(s_in, nb4), I(inv, nb4), R(r_in, nb4);
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        wc[j] = W(o[j]);
        sc[j] = FIRST ? 0.f : Sv(o[j]);
        ic[j] = I(o[j]);
        rc[j] = R(o[j]);
      }
      p0 = FIRST ? 0.f : ld_off((const real*)p, o0);
      x0 = ld_off((const real*)x, o0);
    }
  }
  double alpha_d, beta_d = 0.0;
  if constexpr (FIRST) {
    alpha_d = S->alpha[0];
  } else {
    double gd[ND];
    if constexpr (pf)
      ef.finish(pp, br.nf, gd);
    else
      fold_n<ND>(pp, br.nf, gd);
    if constexpr (CHECK)
      if (inline_check(br, S, gd[2], it, maxiter, out)) return;
    const double g = gd[0], d = gd[1];
    beta_d = safe_div(g, rz_prev);
    alpha_d = safe_div(g, d - beta_d * safe_div(g, alpha_prev));
    if (br.vb == br.b0 && threadIdx.x == 0) {
      S->rz[it & 1] = g;
      S->alpha[it & 1] = alpha_d;
    }
  }
  const real alpha = (real)alpha_d, beta = (real)beta_d;
  double acc[3] = {0.0, 0.0, 0.0};
  if (pf) {
    if (has) {
      real wn, pi, si, rn;
      const real ii = ic[0];
      cg_row<M::kRow, FIRST>(a, wc, sc, rc, ic, p0, alpha, beta, wn, pi, si, rn);
      st_off(x, o0, mad(alpha, pi, x0));
      st_off(p, o0, pi);
      st_off(s_out, o0, si);
      st_off(r_out, o0, rn);
      st_off(w_out, o0, wn);
      cg_dots(acc, rn, wn, ii);
    }
  } else {
    for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
      real wn = 0.f;
      A.row(i, [&](real a, int c) {
        const real sc = FIRST ? w_in[c] : mad(beta, s_in[c], w_in[c]);
        wn = mad(a, inv[c] * mad(-alpha, sc, r_in[c]), wn);
      });
      const real ii = inv[i];
      const real ui = r_in[i] * ii;
      const real pi = FIRST ? ui : mad(beta, p[i], ui);
      const real si = FIRST ? w_in[i] : mad(beta, s_in[i], w_in[i]);
      x[i] = mad(alpha, pi, x[i]);
      const real rn = mad(-alpha, si, r_in[i]);
      p[i] = pi;
      s_out[i] = si;
      r_out[i] = rn;
      w_out[i] = wn;
      cg_dots(acc, rn, wn, ii);
    }
  }
  store_partials<3>(acc, P_out, G, br.vb);
}

// Two rows per thread (chunk <= 2 B): rows tid and tid + B of the chunk,
// the grid-stride order, with all their loads issued before the fold; the
// same sums as the grid-stride form.  Kept apart from k_cgs_iter so that
// the one-row form keeps its register count (and occupancy).  The fold's
// partials and the case's scalars are loaded before the rows (EarlyFold), so
// the fold completes while the gathers are in flight.  FIRST (iteration 0:
// no fold, s = p = 0) is a separate instance, without the branches.
template <class M, bool FIRST, bool CHECK = false>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cgs_iter2(M A, Seg sg, int it, real* __restrict__ x,
                           const real* __restrict__ r_in, real* __restrict__ r_out,
                           const real* __restrict__ w_in, real* __restrict__ w_out,
                           const real* __restrict__ s_in, real* __restrict__ s_out,
                           real* __restrict__ p, const real* __restrict__ inv,
                           const double* __restrict__ P_in, double* __restrict__ P_out,
                           KState* st, int maxiter, KFlags* out) {
  static_assert(!(FIRST && CHECK), "iteration 0 is checked by the init fold");
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) {
    if (CHECK && br.vb == br.b0 && threadIdx.x == 0) publish_flags(out, S->f);
    return;
  }
  const int G = sg.G;
  static_assert(M::kEll, "the two-row form reads ELL rows");
  constexpr int NR = 2;
  constexpr int ND = CHECK ? 3 : 2;
  const int pv = (it - 1) & 1;
  const double* pp[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i) pp[i] = P_in + (size_t)i * G + br.b0;
  EarlyFold<ND> ef;
  double rz_prev = 0.0, alpha_prev = 0.0;
  if constexpr (!FIRST) {
    ef.load(pp, br.nf);
    rz_prev = S->rz[pv];
    alpha_prev = S->alpha[pv];
  }
  real a[NR][M::kRow], wc[NR][M::kRow], sc[NR][M::kRow], ic[NR][M::kRow], rc[NR][M::kRow];
  real p0[NR], x0[NR];
  bool has[NR];
  uint32_t o0[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int i = br.begin + (int)threadIdx.x + q * (int)blockDim.x;
    has[q] = i < br.end;
    o0[q] = (uint32_t)i * 4u;
    p0[q] = 0.f;
    x0[q] = 0.f;
    if (has[q]) {
      const uint32_t nb4 = (uint32_t)A.n * 4u;
      const FBuf W(w_in, nb4), Sv(s_in, nb4), I(inv, nb4), Rv(r_in, nb4);
      uint32_t o[M::kRow];
      A.load_off(i, a[q], o);
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        wc[q][j] = W(o[j]);
        sc[q][j] = FIRST ? 0.f : Sv(o[j]);
        ic[q][j] = I(o[j]);
        rc[q][j] = Rv(o[j]);
      }
      p0[q] = FIRST ? 0.f : ld_off((const real*)p, o0[q]);
      x0[q] = ld_off((const real*)x, o0[q]);
    }
  }
  double alpha_d, beta_d = 0.0;
  if constexpr (FIRST) {
    alpha_d = S->alpha[0];
  } else {
    double gd[ND];
    ef.finish(pp, br.nf, gd);
    if constexpr (CHECK)
      if (inline_check(br, S, gd[2], it, maxiter, out)) return;
    const double g = gd[0], d = gd[1];
    beta_d = safe_div(g, rz_prev);
    alpha_d = safe_div(g, d - beta_d * safe_div(g, alpha_prev));
    if (br.vb == br.b0 && threadIdx.x == 0) {
      S->rz[it & 1] = g;
      S->alpha[it & 1] = alpha_d;
    }
  }
  const real alpha = (real)alpha_d, beta = (real)beta_d;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    if (!has[q]) continue;
    real wn, pi, si, rn;
    const real ii = ic[q][0];
    cg_row<M::kRow, FIRST>(a[q], wc[q], sc[q], rc[q], ic[q], p0[q], alpha, beta, wn, pi, si, rn);
    st_off(x, o0[q], mad(alpha, pi, x0[q]));
    st_off(p, o0[q], pi);
    st_off(s_out, o0[q], si);
    st_off(r_out, o0[q], rn);
    st_off(w_out, o0[q], wn);
    cg_dots(acc, rn, wn, ii);
  }
  store_partials<3>(acc, P_out, G, br.vb);
}

// ------------------------------------- XCD-local persistent single-reduction CG
// One cooperative launch runs the whole solve.  Every case is placed on ONE
// XCD: its workgroups all run on that XCD's CUs and share its L2, so they can
// synchronise once per iteration with a barrier in memory (about 1 us for 32
// workgroups, measured on MI355X) instead of a kernel boundary (about 4-5 us
// at these sizes).  The coherence rules this relies on, measured on MI355X
// (every combination over 5e9 reads):
//   - a plain store, followed by s_waitcnt vmcnt(0) before the barrier's
//     arrival, is seen by an agent-scope (sc1) load of another workgroup of
//     the same XCD after the barrier;
//   - plain loads can return stale data from the CU's L1 even after
//     buffer_inv sc0, and buffer_inv sc1 (the L2 as well) costs 8 us;
// so every value written inside the launch (x, p, r, s, w, the partials) is
// read back with sc1 loads, and the data written before it (the matrix and
// the Jacobi inverse) with ordinary ones.
//
// A workgroup reads its XCD from the XCC_ID register, takes a slot on that
// XCD with an atomic counter, and works out its part with the plan every
// workgroup computes alike from the solve's state: the cases still to solve,
// in order, each on the XCD with the fewest slots used so far, taking
// ceil(nb / rv) slots for its nb virtual workgroups (at most rv virtual
// workgroups each).  A virtual workgroup is the launch path's workgroup: the
// same rows, threads, per-row arithmetic, partials and fold order, so the
// solve is bitwise the one k_cgs_iter / k_cgs_iter2 make at the same launch
// shape (one row per thread).  Nothing depends on the dispatch order for its
// correctness: a slot beyond the plan, a plan that does not fit, or a barrier
// wait longer than wait_ticks sets the abort flag, every workgroup leaves at
// its next barrier, the commit kernel leaves x as it was and the host runs the
// launch path instead.
constexpr int kSyncPad = 32;     // ints between sync words (128 B)
constexpr int kSyncXcc = 0;      // [0, 8): slots taken per XCD
constexpr int kSyncAbort = 8;    // abort flag
constexpr int kSyncLbar = 9;     // [9, 17): arrivals per XCD (grid mode)
constexpr int kSyncLrel = 17;    // [17, 25): releases per XCD (grid mode)
constexpr int kSyncGbar = 25;    // arrivals of the XCDs' leaders (grid mode)
constexpr int kSyncCase = 26;    // [26, 26 + K): barrier counter of case k
constexpr unsigned kHwregXccId = 20u | (0u << 6) | ((4u - 1u) << 11);   // XCC_ID[3:0]

inline size_t xcd_sync_ints(int n_cases) { return (size_t)(kSyncCase + n_cases) * kSyncPad; }

struct XcdPlan {
  int per_xcc;           // workgroup slots per XCD in this launch
  int rv;                // most virtual workgroups one workgroup takes
  int check_every;
  long long wait_ticks;  // longest barrier wait (s_memrealtime, 100 MHz) before aborting
};

template <class T>
__device__ __forceinline__ T ld_coh(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_coh_off(const T* base, uint32_t byte_off) {
  return ld_coh(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                           scaled<T>(byte_off)));
}

template <class T>
__device__ __forceinline__ void st_coh_off(T* base, uint32_t byte_off, T v) {
  __hip_atomic_store(reinterpret_cast<T*>(reinterpret_cast<char*>(base) + scaled<T>(byte_off)), v,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one of a ping-pong pair
__device__ __forceinline__ real* wb_sel(real* a, real* b, int which) { return which ? b : a; }

// EarlyFold with agent-scope loads; the same terms in the same order, so the
// same sums
template <int ND, int E = kFoldUnroll>
struct EarlyFoldC {
  double v[E][ND];
  __device__ __forceinline__ void load(const double* const (&p)[ND], int nb) {
    const int step = blockDim.x;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int b = (int)threadIdx.x + u * step;
#pragma unroll
      for (int i = 0; i < ND; ++i) v[u][i] = b < nb ? ld_coh(p[i] + b) : 0.0;
    }
  }
  __device__ __forceinline__ void finish(const double* const (&p)[ND], int nb,
                                         double (&out)[ND]) {
    __shared__ double lds[ND * kMaxWaves];
    __shared__ double res[ND];
    const int step = blockDim.x;
    double s[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i) s[i] = 0.0;
#pragma unroll
    for (int u = 0; u < E; ++u)
      if ((int)threadIdx.x + u * step < nb) {
#pragma unroll
        for (int i = 0; i < ND; ++i) s[i] += v[u][i];
      }
    for (int b = (int)threadIdx.x + E * step; b < nb; b += step) {
#pragma unroll
      for (int i = 0; i < ND; ++i) s[i] += ld_coh(p[i] + b);
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    wave_sum_n<ND>(s);
#pragma unroll
    for (int i = 0; i < ND; ++i)
      if (lane == 0) lds[i * kMaxWaves + wid] = s[i];
    __syncthreads();
    if (threadIdx.x < ND) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[threadIdx.x * kMaxWaves + w];
      res[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ND; ++i) out[i] = res[i];
  }
};

// The part of one workgroup: case k, virtual workgroups [v_lo, v_hi), and
// the number of workgroups in the case's group.
struct XcdPart {
  int k, v_lo, v_hi, members;
};

// The plan (see above) evaluated for (xcc, slot).  false: the plan does not
// fit the launch.  part.k = -1: the slot has no work.
__device__ inline bool xcd_part(const Seg& sg, const KState* st, const XcdPlan& pl, int xcc,
                                int slot, XcdPart& part) {
  int load[kNumXcd];
#pragma unroll
  for (int j = 0; j < (int)kNumXcd; ++j) load[j] = 0;
  part.k = -1;
  for (int k = 0; k < sg.n_cases; ++k) {
    if (st[k].f.done) continue;
    int b0, nb;
    case_blocks(sg, k, b0, nb);
    if (nb <= 0) continue;
    const int m = (nb + pl.rv - 1) / pl.rv;
    int x = 0, lx = load[0];
#pragma unroll
    for (int j = 1; j < (int)kNumXcd; ++j)
      if (load[j] < lx) {
        lx = load[j];
        x = j;
      }
    if (lx + m > pl.per_xcc) return false;
    if (x == xcc && slot >= lx && slot < lx + m) {
      const int r = slot - lx;
      part.k = k;
      part.v_lo = b0 + (int)((int64_t)r * nb / m);
      part.v_hi = b0 + (int)((int64_t)(r + 1) * nb / m);
      part.members = m;
    }
#pragma unroll
    for (int j = 0; j < (int)kNumXcd; ++j)
      if (j == x) load[j] = lx + m;
  }
  return true;
}

__device__ inline void xcd_abort(int* sync) {
  __hip_atomic_store(sync + kSyncAbort * kSyncPad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Arrive at the case's barrier for the (n+1)-th time and wait for its other
// workgroups (thread 0); false on abort or when the wait passes wait_ticks.
__device__ inline bool xcd_barrier(int* cnt, int target, int* sync, long long wait_ticks) {
  __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
  for (int spins = 0; ld_coh(cnt) < target; ++spins) {
    if (ld_coh(sync + kSyncAbort * kSyncPad)) return false;
    if ((long long)__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks || spins > (1 << 20)) {
      xcd_abort(sync);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// The barrier of a case spread over every XCD (thread 0): each workgroup
// arrives at its XCD's counter; the XCD's slot-0 workgroup waits for its
// XCD, arrives at the grid counter, waits for the other XCDs' leaders and
// releases its XCD; the others wait for that release.  round = n + 1 for
// the n-th barrier.  false on abort or a wait past wait_ticks.
__device__ inline bool xcd_wait_ge(const int* p, int target, int* sync, long long wait_ticks) {
  const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
  for (int spins = 0; ld_coh(p) < target; ++spins) {
    if (ld_coh(sync + kSyncAbort * kSyncPad)) return false;
    if ((long long)__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks || spins > (1 << 20)) {
      xcd_abort(sync);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__device__ inline bool grid_barrier(int* sync, int xcc, int slot, int per_xcc, int round,
                                    long long wait_ticks) {
  int* lbar = sync + (kSyncLbar + xcc) * kSyncPad;
  int* lrel = sync + (kSyncLrel + xcc) * kSyncPad;
  __hip_atomic_fetch_add(lbar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (slot != 0) return xcd_wait_ge(lrel, round, sync, wait_ticks);
  if (!xcd_wait_ge(lbar, per_xcc * round, sync, wait_ticks)) return false;
  int* gbar = sync + kSyncGbar * kSyncPad;
  __hip_atomic_fetch_add(gbar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!xcd_wait_ge(gbar, (int)kNumXcd * round, sync, wait_ticks)) return false;
  __hip_atomic_fetch_add(lrel, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// The persistent kernel: every workgroup takes RV (1 or 2) of its case's
// virtual workgroups, one row per thread each, and keeps the rows'
// coefficients, column offsets, Jacobi inverses at the columns, x and p in
// registers for the whole solve; an iteration reads only the gathers of w,
// s and r (sc1) and the case's partials, and writes w, s, r and the
// partials.  The fold's loads are issued before the gathers (as EarlyFold
// does in k_cgs_iter), so both are in flight together.
// GRID: one case (the single-case solve) spread over every XCD: XCD x takes
// the x-th eighth of the case's virtual workgroups as a contiguous band and
// its slots split the band; the case's workgroups meet at grid_barrier, and
// every value another workgroup reads is stored at agent scope (sc1, which
// other XCDs' sc1 loads see; measured on MI355X).  Otherwise each case sits
// on one XCD, as planned by xcd_part, and meets at its own counter.
// RR: rows per thread of a virtual workgroup (its chunk <= RR x block: rows
// tid, tid + block, ... as the per-iteration kernels stride), RV / RR virtual
// workgroups per workgroup; RV rows per thread in all, cached in registers.
template <class M, int RV, bool GRID, int RR = 1>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_cgs_xcd(
    M A, Seg sg, XcdPlan pl, const real* __restrict__ x_in, real* x_out, real* r0, real* r1,
    real* w0, real* w1, real* s0, real* s1, const real* __restrict__ inv, double* P0,
    double* P1, KState* st, int maxiter, int* sync) {
  constexpr int KR = M::kRow;
  __shared__ int sh[8];
  __shared__ int s_go;
  if (threadIdx.x == 0) {
    const int xcc = (int)(__builtin_amdgcn_s_getreg(kHwregXccId) & (kNumXcd - 1));
    const int slot = __hip_atomic_fetch_add(sync + (kSyncXcc + xcc) * kSyncPad, 1,
                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    XcdPart part;
    bool ok = slot < pl.per_xcc;
    if (GRID) {
      // one case: its band on this XCD, split over the XCD's slots
      part.k = st[0].f.done ? -1 : 0;
      int b0g, nbg;
      case_blocks(sg, 0, b0g, nbg);
      const int lo = b0g + (int)((int64_t)xcc * nbg / kNumXcd);
      const int nbx = b0g + (int)((int64_t)(xcc + 1) * nbg / kNumXcd) - lo;
      part.v_lo = lo + (int)((int64_t)slot * nbx / pl.per_xcc);
      part.v_hi = lo + (int)((int64_t)(slot + 1) * nbx / pl.per_xcc);
      part.members = (int)kNumXcd * pl.per_xcc;
      ok = ok && sg.n_cases == 1;
      sh[6] = lo;          // the XCD's band of virtual workgroups
      sh[7] = lo + nbx;
    } else {
      ok = ok && xcd_part(sg, st, pl, xcc, slot, part);
    }
    ok = ok && (part.k < 0 || part.v_hi - part.v_lo <= RV / RR);
    if (!ok) {
      xcd_abort(sync);
      part.k = -1;
    }
    sh[0] = part.k;
    sh[1] = part.v_lo;
    sh[2] = part.v_hi;
    sh[3] = part.members;
    sh[4] = xcc;
    sh[5] = slot;
  }
  __syncthreads();
  // Values broadcast through LDS are the same in every lane; readfirstlane
  // says so to the compiler, so that the exits below are scalar branches.
  const int k = __builtin_amdgcn_readfirstlane(sh[0]);
  if (k < 0) return;
  const int v_lo = __builtin_amdgcn_readfirstlane(sh[1]);
  const int nv = __builtin_amdgcn_readfirstlane(sh[2]) - v_lo;
  const int members = __builtin_amdgcn_readfirstlane(sh[3]);
  const int my_xcc = __builtin_amdgcn_readfirstlane(sh[4]);
  const int my_slot = __builtin_amdgcn_readfirstlane(sh[5]);
  KState* S = st + k;
  int b0, nb;
  case_blocks(sg, k, b0, nb);
  const int G = sg.G;
  const int ce = pl.check_every > 0 ? pl.check_every : 1;
  const double thr2 = S->f.thr2;
  double rz_prev = S->rz[0], alpha_prev = S->alpha[0];
  // records the case's flags: the workgroup holding the case's first virtual
  // workgroup (in grid mode a workgroup may hold none: XCD 0, slot 0)
  const bool leader = threadIdx.x == 0 && (GRID ? my_xcc == 0 && my_slot == 0 : v_lo == b0);
  int* cnt = sync + (kSyncCase + k) * kSyncPad;
  // the rows, cached
  const int e = sg.cell_end ? sg.cell_end[k] : (sg.cell_off ? sg.cell_off[k + 1] : sg.n);
  const int c0 = sg.cell_off ? sg.cell_off[k] : 0;
  const uint32_t nb4 = (uint32_t)A.n * 4u;
  real a[RV][KR], ic[RV][KR], xr[RV], pr[RV];
  uint32_t o[RV][KR];
  bool has[RV];
  // GRID: rows of the XCD's band, [band_lo, band_hi) (its last one may end
  // past the case: no column reaches there)
  const uint32_t band_lo = GRID ? (uint32_t)(c0 + (sh[6] - b0) * sg.chunk) : 0u;
  const uint32_t band_hi = GRID ? (uint32_t)(c0 + (sh[7] - b0) * sg.chunk) : 0u;
  // GRID: the row is read by another XCD.  The pattern is symmetric (mesh
  // adjacency), so that is when one of the row's own columns lies outside
  // the band.  Those rows are stored at agent scope (other XCDs' agent-scope
  // loads see them); the rest are read only on this XCD and stored plainly,
  // as the one-XCD form stores every row.
  bool remote[RV];
#pragma unroll
  for (int q = 0; q < RV; ++q) {
    const int vq = q / RR, rq = q % RR;   // virtual workgroup, row of the thread in it
    const int begin = c0 + (v_lo + vq - b0) * sg.chunk;
    const int end = begin + sg.chunk < e ? begin + sg.chunk : e;
    const int i = begin + (int)threadIdx.x + rq * (int)blockDim.x;
    has[q] = vq < nv && i < end;
    xr[q] = pr[q] = 0.f;
    remote[q] = false;
    if (has[q]) {
      A.load_off(i, a[q], o[q]);
      if constexpr (GRID) {
#pragma unroll
        for (int j = 1; j < KR; ++j) {
          const uint32_t cj = o[q][j] >> 2;
          remote[q] = remote[q] || cj < band_lo || cj >= band_hi;
        }
      }
      const FBuf I(inv, nb4);
#pragma unroll
      for (int j = 0; j < KR; ++j) ic[q][j] = I(o[q][j]);
      xr[q] = ld_off(x_in, o[q][0]);
    }
  }
  for (int it = 0;; ++it) {
    const int cur = it & 1;
    const real* W = wb_sel(w0, w1, cur);
    const real* Sv = wb_sel(s0, s1, cur);
    const real* R = wb_sel(r0, r1, cur);
    real* w_out = wb_sel(w0, w1, 1 - cur);
    real* s_out = wb_sel(s0, s1, 1 - cur);
    real* r_out = wb_sel(r0, r1, 1 - cur);
    double* P_in = cur ? P1 : P0;
    double* P_out = cur ? P0 : P1;
    const double* pp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) pp[i] = P_in + (size_t)i * G + b0;
    EarlyFoldC<3> ef;
    if (it > 0) ef.load(pp, nb);   // the prefold kernel does not run here: fold all of them
    real wc[RV][KR], sc[RV][KR], rc[RV][KR];
#pragma unroll
    for (int q = 0; q < RV; ++q)
      if (has[q]) {
#pragma unroll
        for (int j = 0; j < KR; ++j) {
          wc[q][j] = ld_coh_off(W, o[q][j]);
          sc[q][j] = it > 0 ? ld_coh_off(Sv, o[q][j]) : (real)0;
          rc[q][j] = ld_coh_off(R, o[q][j]);
        }
      }
    // The stop test is a flag, and the loop's only exits are at the top
    // level of its body: an exit nested in the check branch was compiled so
    // that a check iteration that went on kept the previous rz and alpha.
    real alpha = (real)alpha_prev, beta = 0;
    int stop = 0;
    if (it > 0) {
      double gd[3];
      ef.finish(pp, nb, gd);
      const bool check = it % ce == 0;
      // the stop test of k_check / inline_check, on every workgroup alike
      stop = check && (!(gd[2] > thr2) || it >= maxiter) ? 1 : 0;
      if (leader && check) {
        S->f.rr = gd[2];
        if (stop) {
          S->f.done = 1;
          S->f.iters = it;
        }
      }
      const double g = gd[0], d = gd[1];
      const double beta_d = safe_div(g, rz_prev);
      const double alpha_d = safe_div(g, d - beta_d * safe_div(g, alpha_prev));
      rz_prev = g;
      alpha_prev = alpha_d;
      alpha = (real)alpha_d;
      beta = (real)beta_d;
    }
    if (__builtin_amdgcn_readfirstlane(stop)) break;
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < RV; ++q) {
      if (q / RR >= nv) continue;   // uniform across the workgroup
      if (has[q]) {
        real wn, pi, si, rn;
        if (it == 0)
          cg_row<KR, true>(a[q], wc[q], sc[q], rc[q], ic[q], pr[q], alpha, beta, wn, pi, si, rn);
        else
          cg_row<KR, false>(a[q], wc[q], sc[q], rc[q], ic[q], pr[q], alpha, beta, wn, pi, si, rn);
        xr[q] = mad(alpha, pi, xr[q]);
        pr[q] = pi;
        if (GRID && remote[q]) {
          st_coh_off(s_out, o[q][0], si);
          st_coh_off(r_out, o[q][0], rn);
          st_coh_off(w_out, o[q][0], wn);
        } else {
          st_plain(s_out, o[q][0], si);
          st_plain(r_out, o[q][0], rn);
          st_plain(w_out, o[q][0], wn);
        }
        cg_dots(acc, rn, wn, ic[q][0]);
      }
      if (q % RR == RR - 1) {   // the virtual workgroup's last row of the thread
        store_partials<3, GRID>(acc, P_out, G, v_lo + q / RR);
        // thread 0 has read the wave sums from LDS before the next
        // store_partials overwrites them
        __syncthreads();
        acc[0] = acc[1] = acc[2] = 0.0;
      }
    }
    // every store of this workgroup has reached the L2 before it arrives
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0)
      s_go = GRID ? grid_barrier(sync, my_xcc, my_slot, pl.per_xcc, it + 1, pl.wait_ticks)
                  : xcd_barrier(cnt, members * (it + 1), sync, pl.wait_ticks);
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_go)) return;
  }
#pragma unroll
  for (int q = 0; q < RV; ++q)
    if (has[q]) st_plain(x_out, o[q][0], xr[q]);
}

// Zero the sync words of a launch (one workgroup).
__global__ void k_xcd_reset(int* sync, int n_words) {
  for (int i = threadIdx.x; i < n_words; i += blockDim.x) sync[i * kSyncPad] = 0;
}

// After k_cgs_xcd, over the launch path's virtual workgroups: unless the
// launch aborted, copy the iterate of every case that ran (f.iters > 0: a
// case stops only at a check, iteration check_every or later) into x.
__global__ __launch_bounds__(kMaxKrylovBlock) void k_xcd_commit(Seg sg, const KState* st,
                                                                 const int* sync, const real* xs,
                                                                 real* x) {
  if (sync[kSyncAbort * kSyncPad]) return;
  const BlockRows br = block_rows(sg);
  if (st[br.k].f.iters <= 0) return;
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) x[i] = xs[i];
}

// The flag records of every case to `out` (pinned host memory), and the
// abort flag as record n_cases (its `done`).
__global__ void k_xcd_publish(const KState* st, int n_cases, const int* sync, KFlags* out) {
  for (int k = threadIdx.x; k <= n_cases; k += blockDim.x) {
    if (k < n_cases) {
      out[k] = st[k].f;
    } else {
      KFlags f{};
      f.done = sync[kSyncAbort * kSyncPad];
      out[k] = f;
    }
  }
  __threadfence_system();
}

// ------------------------------------------------------------- BiCGSTAB
// partial slots: PA[3][G] = {r0.r, r.r, b.b} (init, k_bi_3), PB[1][G] = {r0.v}
// (k_bi_1), PC[3][G] = {t.s, t.t, s.s} (k_bi_2)
template <class M, int NV>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_bi_init(M A, Seg sg, const real* __restrict__ b, const real* __restrict__ x,
                          real* __restrict__ r, real* __restrict__ r0, double* __restrict__ PA) {
  const BlockRows br = block_rows(sg);
  const int n = sg.n;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    real ax[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) ax[q] = 0.f;
    A.row(i, [&](real a, int c) {
#pragma unroll
      for (int q = 0; q < NV; ++q) ax[q] += a * x[(size_t)q * n + c];
    });
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const real bi = b[k];
      const real ri = bi - ax[q];
      r[k] = ri;
      r0[k] = ri;
      acc[0] += (double)ri * (double)ri;
      acc[1] += (double)ri * (double)ri;
      acc[2] += (double)bi * (double)bi;
    }
  }
  store_partials<3>(acc, PA, sg.G, br.vb);
}

template <class M, int NV, bool PF>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_bi_1(M A, Seg sg, int it, const real* __restrict__ r,
                       const real* __restrict__ inv, const real* __restrict__ p_in,
                       const real* __restrict__ v_in, real* __restrict__ p_out,
                       real* __restrict__ v_out, const real* __restrict__ r0,
                       const double* __restrict__ PA, double* __restrict__ PB, KState* st) {
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) return;
  const int n = sg.n;
  // one row per thread on an ELL matrix: issue the row's loads before the fold
  constexpr bool pf = PF && M::kEll;
  const double* const fp[1] = {PA + br.b0};
  EarlyFold<1> ef;
  if constexpr (pf)
    if (it) ef.load(fp, br.nf);
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real a[M::kRow], ic[M::kRow], rr[NV][M::kRow], pp[NV][M::kRow], vv[NV][M::kRow], r0i[NV];
  if constexpr (M::kEll) {
    if (has) {
      uint32_t o[M::kRow];
      A.load_off(i0, a, o);
      const uint32_t nb4 = (uint32_t)n * 4u * NV;
      const FBuf I(inv, (uint32_t)n * 4u), R(r, nb4), P(p_in, nb4), V(v_in, nb4);
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        ic[j] = I(o[j]);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const uint32_t oq = o[j] + (uint32_t)q * (uint32_t)n * 4u;
          rr[q][j] = R(oq);
          pp[q][j] = it ? P(oq) : 0.f;
          vv[q][j] = it ? V(oq) : 0.f;
        }
      }
#pragma unroll
      for (int q = 0; q < NV; ++q) r0i[q] = ld_off(r0 + (size_t)q * n, o[0]);
    }
  }
  bool first = (it == 0);
  double beta_d = 0.0, omega_d = 0.0;
  if (!first) {
    const int pv = (it - 1) & 1;
    double rho;                                   // r0.r after the previous iteration
    if constexpr (pf) {
      double o1[1];
      ef.finish(fp, br.nf, o1);
      rho = o1[0];
    } else {
      rho = fold(PA + br.b0, br.nf);
    }
    omega_d = S->omega[pv];
    beta_d = safe_div(rho, S->rho[pv]) * safe_div(S->alpha[pv], omega_d);
    if (br.vb == br.b0 && threadIdx.x == 0) S->rho[it & 1] = rho;
    first = (beta_d == 0.0);
  }
  const real beta = (real)beta_d;
  const real omega = (real)omega_d;
  double acc[1] = {0.0};
  if (pf) {
    if (has) {
      real v[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] = 0.f;
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const real pc = first ? rr[q][j] : rr[q][j] + beta * (pp[q][j] - omega * vv[q][j]);
          v[q] += a[j] * (pc * ic[j]);
        }
      }
      const uint32_t o0 = (uint32_t)i0 * 4u;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t qn = (size_t)q * n;
        const real pi = first ? rr[q][0] : rr[q][0] + beta * (pp[q][0] - omega * vv[q][0]);
        st_off(p_out + qn, o0, pi);
        st_off(v_out + qn, o0, v[q]);
        acc[0] += (double)r0i[q] * (double)v[q];
      }
    }
  } else {
    for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
      real v[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] = 0.f;
      A.row(i, [&](real a, int c) {
        const real ic = inv[c];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const size_t k = (size_t)q * n + c;
          const real pc = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
          v[q] += a * (pc * ic);
        }
      });
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + i;
        const real pi = first ? r[k] : r[k] + beta * (p_in[k] - omega * v_in[k]);
        p_out[k] = pi;
        v_out[k] = v[q];
        acc[0] += (double)r0[k] * (double)v[q];
      }
    }
  }
  store_partials<1>(acc, PB, sg.G, br.vb);
}

template <class M, int NV, bool PF>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_bi_2(M A, Seg sg, int it, const real* __restrict__ r,
                       const real* __restrict__ inv, const real* __restrict__ v,
                       real* __restrict__ s, real* __restrict__ t,
                       const double* __restrict__ PB, double* __restrict__ PC, KState* st) {
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) return;
  const int n = sg.n;
  // one row per thread on an ELL matrix: issue the row's loads before the fold
  constexpr bool pf = PF && M::kEll;
  const double* const fp[1] = {PB + br.b0};
  EarlyFold<1> ef;
  if constexpr (pf) ef.load(fp, br.nf);
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real a[M::kRow], ic[M::kRow], rr[NV][M::kRow], vv[NV][M::kRow];
  if constexpr (M::kEll) {
    if (has) {
      uint32_t o[M::kRow];
      A.load_off(i0, a, o);
      const uint32_t nb4 = (uint32_t)n * 4u * NV;
      const FBuf I(inv, (uint32_t)n * 4u), R(r, nb4), V(v, nb4);
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
        ic[j] = I(o[j]);
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const uint32_t oq = o[j] + (uint32_t)q * (uint32_t)n * 4u;
          rr[q][j] = R(oq);
          vv[q][j] = V(oq);
        }
      }
    }
  }
  double r0v;
  if constexpr (pf) {
    double o1[1];
    ef.finish(fp, br.nf, o1);
    r0v = o1[0];
  } else {
    r0v = fold(PB + br.b0, br.nf);
  }
  const double alpha_d = safe_div(S->rho[it & 1], r0v);
  if (br.vb == br.b0 && threadIdx.x == 0) S->alpha[it & 1] = alpha_d;
  const real alpha = (real)alpha_d;
  double acc[3] = {0.0, 0.0, 0.0};
  if (pf) {
    if (has) {
      real tv[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) tv[q] = 0.f;
#pragma unroll
      for (int j = 0; j < M::kRow; ++j) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const real sc = rr[q][j] - alpha * vv[q][j];
          tv[q] += a[j] * (sc * ic[j]);
        }
      }
      const uint32_t o0 = (uint32_t)i0 * 4u;
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t qn = (size_t)q * n;
        const real si = rr[q][0] - alpha * vv[q][0];
        st_off(s + qn, o0, si);
        st_off(t + qn, o0, tv[q]);
        acc[0] += (double)tv[q] * (double)si;
        acc[1] += (double)tv[q] * (double)tv[q];
        acc[2] += (double)si * (double)si;
      }
    }
  } else {
    for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
      real tv[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) tv[q] = 0.f;
      A.row(i, [&](real a, int c) {
        const real ic = inv[c];
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const size_t k = (size_t)q * n + c;
          const real sc = r[k] - alpha * v[k];
          tv[q] += a * (sc * ic);
        }
      });
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + i;
        const real si = r[k] - alpha * v[k];
        s[k] = si;
        t[k] = tv[q];
        acc[0] += (double)tv[q] * (double)si;
        acc[1] += (double)tv[q] * (double)tv[q];
        acc[2] += (double)si * (double)si;
      }
    }
  }
  store_partials<3>(acc, PC, sg.G, br.vb);
}

template <int NV, bool PF>
__global__ __launch_bounds__(kMaxKrylovBlock) void k_bi_3(Seg sg, int it, real* __restrict__ x, real* __restrict__ r,
                       const real* __restrict__ inv, const real* __restrict__ p,
                       const real* __restrict__ s, const real* __restrict__ t,
                       const real* __restrict__ r0, const double* __restrict__ PC,
                       double* __restrict__ PA, KState* st) {
  const BlockRows br = block_rows(sg);
  KState* S = st + br.k;
  if (S->f.done) return;
  const int n = sg.n, G = sg.G;
  // one row per thread: load it before the fold
  constexpr bool pf = PF;
  const double* pp[3] = {PC + br.b0, PC + G + br.b0, PC + 2 * G + br.b0};
  EarlyFold<3> ef;
  if constexpr (pf) ef.load(pp, br.nf);
  const int i0 = br.begin + (int)threadIdx.x;
  const bool has = pf && i0 < br.end;
  real iv = 0.f, pv[NV], sv[NV], tv[NV], xv[NV], r0v[NV];
  if (has) {
    iv = inv[i0];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i0;
      pv[q] = p[k];
      sv[q] = s[k];
      tv[q] = t[k];
      xv[q] = x[k];
      r0v[q] = r0[k];
    }
  }
  double tv3[3];
  if constexpr (pf)
    ef.finish(pp, br.nf, tv3);
  else
    fold_n<3>(pp, br.nf, tv3);
  const double ts = tv3[0], tt = tv3[1], ss = tv3[2];
  const double omega_d = ss > S->f.thr2 ? safe_div(ts, tt) : 0.0;
  if (br.vb == br.b0 && threadIdx.x == 0) {
    S->omega[it & 1] = omega_d;
    S->ss = ss;
  }
  const real alpha = (real)S->alpha[it & 1];
  const real omega = (real)omega_d;
  double acc[2] = {0.0, 0.0};
  if (pf) {
    if (has) {
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const size_t k = (size_t)q * n + i0;
        const real y = pv[q] * iv;
        const real z = sv[q] * iv;
        x[k] = xv[q] + (alpha * y + omega * z);
        const real rk = sv[q] - omega * tv[q];
        r[k] = rk;
        acc[0] += (double)r0v[q] * (double)rk;
        acc[1] += (double)rk * (double)rk;
      }
    }
    store_partials<2>(acc, PA, G, br.vb);
    return;
  }
  for (int i = br.begin + (int)threadIdx.x; i < br.end; i += blockDim.x) {
    const real ii = inv[i];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const size_t k = (size_t)q * n + i;
      const real y = p[k] * ii;
      const real z = s[k] * ii;
      x[k] = x[k] + (alpha * y + omega * z);
      const real rk = s[k] - omega * t[k];
      r[k] = rk;
      acc[0] += (double)r0[k] * (double)rk;
      acc[1] += (double)rk * (double)rk;
    }
  }
  store_partials<2>(acc, PA, G, br.vb);
}

// ------------------------------------------------------------- host side
struct KWork {
  real *r, *r0, *p0, *p1, *v0, *v1, *s, *t, *inv;
  double *PA, *PB, *PC;
  KState* st;
  int* sync;   // k_cgs_xcd's slot, abort and barrier words
};

// State and partials first, at offsets that depend only on (n_cases,
// max_blocks): CG (1 vector) and BiCGSTAB (3 vectors) may share a workspace.
static KWork carve_krylov(void* ws, int32_t n, int nvec, int n_cases, int max_blocks) {
  Carver c(ws);
  KWork w;
  w.st = c.take<KState>(n_cases);
  w.PA = c.take<double>(3 * (size_t)max_blocks);
  w.PB = c.take<double>((size_t)max_blocks);
  w.PC = c.take<double>(3 * (size_t)max_blocks);
  w.sync = c.take<int>(xcd_sync_ints(n_cases));
  const size_t nv = (size_t)n * nvec;
  w.r = c.take<real>(nv);
  w.r0 = c.take<real>(nv);
  w.p0 = c.take<real>(nv);
  w.p1 = c.take<real>(nv);
  w.v0 = c.take<real>(nv);
  w.v1 = c.take<real>(nv);
  w.s = c.take<real>(nv);
  w.t = c.take<real>(nv);
  w.inv = c.take<real>(n);
  return w;
}

size_t krylov_workspace_bytes_batch(int32_t n, int nvec, int n_cases, int max_blocks) {
  Carver c(nullptr);
  c.take<KState>(n_cases);
  c.take<double>(3 * (size_t)max_blocks);
  c.take<double>((size_t)max_blocks);
  c.take<double>(3 * (size_t)max_blocks);
  c.take<int>(xcd_sync_ints(n_cases));
  const size_t nv = (size_t)n * nvec;
  for (int i = 0; i < 8; ++i) c.take<real>(nv);
  c.take<real>(n);
  return c.off;
}

// Most workgroups a single-case launch over n rows can have (one row per
// thread at the smallest workgroup): the partial rows of its workspace.
static int single_max_blocks(int32_t n) {
  const int64_t g = ((int64_t)n + kWave - 1) / kWave;
  return (int)(g < kMaxKrylovBlocks ? (g < 1 ? 1 : g) : kMaxKrylovBlocks);
}

size_t krylov_workspace_bytes(int32_t n, int nvec) {
  return krylov_workspace_bytes_batch(n, nvec, 1, single_max_blocks(n));
}

static void fill_info(wcfd_solve_info* info, const KFlags& h) {
  if (!info) return;
  info->iterations = h.iters;
  info->converged = h.rr <= h.thr2 ? 1 : 0;
  info->residual = sqrt(h.rr);
  info->threshold = sqrt(h.thr2);
}

// Pinned host slots and events for polling the done flags without draining
// the stream (one set per host thread, grown to the largest batch seen).
// k_check writes the slots directly through their device address.
constexpr int kSlots = 4;
struct HostSync {
  KFlags* slot = nullptr;   // [kSlots][cap], host address
  KFlags* dslot = nullptr;  // the same memory, device address
  int cap = 0;
  hipEvent_t ev[kSlots] = {};
  hipEvent_t tail = nullptr;
  bool tail_pending = false;
  bool ready = false;
};

static HostSync* host_sync(int n_cases) {
  static thread_local HostSync hs;
  if (!hs.ready) {
    for (auto& e : hs.ev)
      if (check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming))) return nullptr;
    if (check_hip(hipEventCreateWithFlags(&hs.tail, hipEventDisableTiming))) return nullptr;
    hs.ready = true;
  }
  if (hs.tail_pending) {   // check kernels of an earlier solve may still write the slots
    if (check_hip(hipEventSynchronize(hs.tail))) return nullptr;
    hs.tail_pending = false;
  }
  if (hs.cap < n_cases) {
    if (hs.slot && check_hip(hipHostFree(hs.slot))) return nullptr;
    hs.slot = nullptr;
    hs.cap = 0;
    const int cap = n_cases > 64 ? n_cases : 64;
    if (check_hip(hipHostMalloc((void**)&hs.slot, sizeof(KFlags) * kSlots * (size_t)cap,
                                hipHostMallocMapped)))
      return nullptr;
    if (check_hip(hipHostGetDevicePointer((void**)&hs.dslot, hs.slot, 0))) return nullptr;
    hs.cap = cap;
  }
  return &hs;
}

static bool all_done(const KFlags* f, int n) {
  for (int k = 0; k < n; ++k)
    if (!f[k].done) return false;
  return true;
}

static int finish(const KFlags* f, int n, wcfd_solve_info* info) {
  bool finite = true;
  for (int k = 0; k < n; ++k) {
    if (info) fill_info(info + k, f[k]);
    if (!isfinite(f[k].rr)) finite = false;
  }
  return finite ? WCFD_OK : WCFD_ENAN;
}

// Runs the iteration groups of one (batched) solve.  Each group is `ce`
// iterations followed by the check kernel, which writes the K flag records
// to a pinned slot; an event is recorded after it.  The host keeps `depth` groups in
// flight and stops launching once a polled group reports every case done;
// groups already queued find the flags set and exit at entry, so the result
// is the same as checking after every group synchronously (the reference
// solver's check interval).
// The check kernels run with the iteration kernels' workgroup size B, so
// that their fold of ||r||^2 has the order of a prefold.
template <class LaunchGroup, class RrPart>
static int run_groups(const Seg& sg, int B, KState* st, RrPart&& rr_part, int maxiter, int ce,
                      int fixed_groups, LaunchGroup&& launch_group, wcfd_solve_info* info,
                      hipStream_t s) {
  const int K = sg.n_cases;
  if (fixed_groups > 0) {
    // Fixed launch budget (stream capture): exactly fixed_groups groups, no
    // event, no host read; the check kernels still stop every case at the
    // same iteration as the polled loop, and maxiter was clipped to the
    // budget by the caller.  info is left untouched.
    for (int g = 0, it = 0; g < fixed_groups; ++g) {
      launch_group(it);
      it += ce;
      WCFD_LAUNCH_B(K_FOLD, s, k_check, K, B, sg, rr_part(it), st, it, maxiter,
                    (KFlags*)nullptr);
    }
    return check_launch();
  }
  HostSync* hs = host_sync(K);
  if (!hs) return WCFD_EHIP;
  const size_t prof_from = prof_mark();
  int e;
  // No round trip after init: the first group is launched at once (its
  // kernels exit at entry for cases the init fold already marked done) and
  // the first poll reports those cases like any other.
  const int depth = 2;
  int launched = 0, polled = 0, it = 0;
  const KFlags* last = nullptr;
  while (true) {
    if ((it < maxiter || launched == 0) && launched - polled < depth) {
      launch_group(it);
      it += ce;
      const int sl = launched % kSlots;
      WCFD_LAUNCH_B(K_FOLD, s, k_check, K, B, sg, rr_part(it), st, it, maxiter,
                    hs->dslot + (size_t)sl * hs->cap);
      if ((e = check_launch())) return e;
      if ((e = check_hip(hipEventRecord(hs->ev[sl], s)))) return e;
      ++launched;
      continue;
    }
    if (polled == launched) break;   // cannot happen: the last check sets every flag
    // (maxiter <= 0: the init fold set every flag and the one group is empty)
    const int sl = polled % kSlots;
    if ((e = check_hip(hipEventSynchronize(hs->ev[sl])))) return e;
    last = hs->slot + (size_t)sl * hs->cap;
    ++polled;
    if (all_done(last, K)) break;
  }
  if (launched > polled) {
    if ((e = check_hip(hipEventRecord(hs->tail, s)))) return e;
    hs->tail_pending = true;
  }
  if (last) {
    int ran = 0;
    for (int k = 0; k < K; ++k) ran = last[k].iters > ran ? last[k].iters : ran;
    prof_retire(prof_from, ran);
  }
  return finish(last, K, info);
}

// run_groups for one case whose iteration kernels make the stop test
// themselves: iterations 0 .. ce-1, then units of [iteration it with the
// test of r_it (its record goes to a pinned slot, an event follows),
// iterations it+1 .. it+ce-1], it = ce, 2 ce, ...  The unit with it >=
// maxiter holds the test alone.  The host keeps `depth` units in flight and
// stops launching once a polled record reports the case done, as in
// run_groups; the iterations run and the result are the same, and no check
// kernel is launched.
template <class LaunchIter>
static int run_groups_inline(int maxiter, int ce, LaunchIter&& iteration,
                             wcfd_solve_info* info, hipStream_t s) {
  HostSync* hs = host_sync(1);
  if (!hs) return WCFD_EHIP;
  const size_t prof_from = prof_mark();
  int e;
  for (int it = 0; it < ce; ++it) iteration(it, false, nullptr);
  if ((e = check_launch())) return e;
  const int depth = 2;
  int launched = 0, polled = 0, it = ce;
  bool more = true;
  const KFlags* last = nullptr;
  while (true) {
    if (more && launched - polled < depth) {
      const int sl = launched % kSlots;
      iteration(it, true, hs->dslot + (size_t)sl * hs->cap);
      if ((e = check_launch())) return e;
      if ((e = check_hip(hipEventRecord(hs->ev[sl], s)))) return e;
      ++launched;
      if (it >= maxiter) {
        more = false;   // this test stops the case
      } else {
        for (int j = it + 1; j < it + ce; ++j) iteration(j, false, nullptr);
        if ((e = check_launch())) return e;
      }
      it += ce;
      continue;
    }
    if (polled == launched) break;   // cannot happen: the last test sets the flag
    const int sl = polled % kSlots;
    if ((e = check_hip(hipEventSynchronize(hs->ev[sl])))) return e;
    last = hs->slot + (size_t)sl * hs->cap;
    ++polled;
    if (all_done(last, 1)) break;
  }
  if (launched > polled) {   // queued tests still write their slots
    if ((e = check_hip(hipEventRecord(hs->tail, s)))) return e;
    hs->tail_pending = true;
  }
  if (last) prof_retire(prof_from, last[0].iters);
  return finish(last, 1, info);
}

namespace {
std::atomic<int> g_krylov_grid{kKrylovGrid};
std::atomic<int> g_inline_check{1};
std::atomic<int> g_krylov_block{kBlock};
std::atomic<int> g_krylov_fold{kMaxInlineFold};
std::atomic<int> g_persistent{0};   // most rows per case of a persistent solve (0: off)
std::atomic<long long> g_xcd_solves{0}, g_xcd_fallbacks{0};
std::atomic<int> g_xcd_shrink{0};   // test hook: slots per XCD taken from the plan's launch
}  // namespace

KrylovLaunch krylov_launch() { return {g_krylov_grid.load(), g_krylov_block.load()}; }

// Launch of one solve: the segmentation, the workgroup size and the
// compacted row ranges of the batch's cell launches (for k_inv_diag).
struct KLaunch {
  Seg sg;
  int B;
  Items rows;
  int max_case_blocks;   // most virtual workgroups of one case
};

// One row per thread (every workgroup's chunk fits its threads): the
// iteration kernels' prefetching form, compiled without the grid-stride
// loop, whose 64-bit gather addresses would otherwise set the register
// count (and so the occupancy) of the whole kernel.  Both forms evaluate a
// row with the same expression, so the choice does not change results.
template <class M>
static bool row_per_thread(const KLaunch& L) {
  return M::kEll && L.sg.chunk <= L.B;
}

// rows per thread of the prefetching single-reduction CG kernel: 1 or 2
// when every workgroup's chunk fits, else 0 (grid-stride)
template <class M>
static int rows_per_thread(const KLaunch& L) {
  if (!M::kEll) return 0;
  if (L.sg.chunk <= L.B) return 1;
  if (L.sg.chunk <= 2 * L.B) return 2;
  return 0;
}

// Reduces the ND partial rows at P to one value per case (Seg::prefold).
template <int ND>
static void prefold(const KLaunch& L, double* P, KState* st, hipStream_t s) {
  if (L.sg.prefold)
    WCFD_LAUNCH_B(K_FOLD, s, k_prefold<ND>, L.sg.n_cases, L.B, L.sg, P, st);
}

static int prepare(KState* st, const Seg& sg, const Tol& tl, hipStream_t s) {
  hipLaunchKernelGGL(k_set_tolerances, dim3(grid_for(sg.n_cases)), dim3(kBlock), 0, s, st,
                     sg.n_cases, tl);
  return check_launch();
}

template <class M>
static int cg_impl(const M& A, const KLaunch& L, const real* b, real* x, const Tol& tl,
                   bool r64, int maxiter, int check_every, KWork& w, wcfd_solve_info* info,
                   hipStream_t s) {
  const Seg& sg = L.sg;
  const int n = sg.n, B = L.B, G = sg.G, K = sg.n_cases;
  const bool pf = row_per_thread<M>(L);
  real* q = w.s;
  int e = prepare(w.st, sg, tl, s);
  if (e) return e;
  if (const unsigned gi = items_grid(L.rows, n))
    WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, gi, A, L.rows, n, w.inv);
  if (r64)
    WCFD_LAUNCH_B(K_CG_INIT, s, (k_cg_init<M, true>), G, B, A, sg, b, x, w.r, w.inv, w.PA);
  else
    WCFD_LAUNCH_B(K_CG_INIT, s, (k_cg_init<M, false>), G, B, A, sg, b, x, w.r, w.inv, w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_init_fold, K, sg, w.PA, w.PA, w.st, 0, maxiter);
  if ((e = check_launch())) return e;
  real* pb[2] = {w.p0, w.p1};
  const int ce = check_every > 0 ? check_every : 1;
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) {
      prof_set_iter(it);
      const int cur = it & 1;
      if (pf)
        WCFD_LAUNCH_B(K_CG_1, s, (k_cg_1<M, true>), G, B, A, sg, it, w.r, w.inv, pb[cur],
                      pb[1 - cur], q, w.PA, w.PB, w.st);
      else
        WCFD_LAUNCH_B(K_CG_1, s, (k_cg_1<M, false>), G, B, A, sg, it, w.r, w.inv, pb[cur],
                      pb[1 - cur], q, w.PA, w.PB, w.st);
      prefold<1>(L, w.PB, w.st, s);
      if (pf)
        WCFD_LAUNCH_B(K_CG_2, s, k_cg_2<true>, G, B, sg, it, x, w.r, w.inv, pb[1 - cur], q, w.PB,
                      w.PA, w.st);
      else
        WCFD_LAUNCH_B(K_CG_2, s, k_cg_2<false>, G, B, sg, it, x, w.r, w.inv, pb[1 - cur], q, w.PB,
                      w.PA, w.st);
      prefold<2>(L, w.PA, w.st, s);
    }
    prof_set_iter(-1);
  };
  auto rr_part = [&](int) { return (const double*)(w.PA + G); };
  return run_groups(sg, B, w.st, rr_part, maxiter, ce, tl.fixed_groups, group, info, s);
}

template <class M>
static int cgs_impl(const M& A, const KLaunch& L, const real* b, real* x, const Tol& tl,
                    bool r64, int maxiter, int check_every, KWork& w, wcfd_solve_info* info,
                    hipStream_t s) {
  const Seg& sg = L.sg;
  const int n = sg.n, B = L.B, G = sg.G, K = sg.n_cases;
  const int rows = rows_per_thread<M>(L);
  int e = prepare(w.st, sg, tl, s);
  if (e) return e;
  real* rb[2] = {w.r, w.r0};
  real* wb[2] = {w.v0, w.v1};
  real* sb[2] = {w.s, w.t};
  double* pb[2] = {w.PA, w.PC};   // iteration i reads pb[i&1], writes pb[(i+1)&1]
  if (const unsigned gi = items_grid(L.rows, n))
    WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, gi, A, L.rows, n, w.inv);
  if (r64)
    WCFD_LAUNCH_B(K_CG_INIT, s, (k_cgs_init1<M, true>), G, B, A, sg, b, x, rb[0], w.PC);
  else
    WCFD_LAUNCH_B(K_CG_INIT, s, (k_cgs_init1<M, false>), G, B, A, sg, b, x, rb[0], w.PC);
  WCFD_LAUNCH_B(K_CG_INIT, s, k_cgs_init2<M>, G, B, A, sg, rb[0], w.inv, wb[0], w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_init_fold, K, sg, w.PA, w.PC, w.st, 2, maxiter);
  if ((e = check_launch())) return e;
  const int ce = check_every > 0 ? check_every : 1;
  // one iteration; `check`: the kernel first makes the stop test of the
  // iterations before it (writing the flag record to `out`)
  auto iteration = [&](int it, bool check, KFlags* out) {
    prof_set_iter(it);
    const int cur = it & 1;
#define WCFD_CGS_ITER(KERNEL)                                                             \
  WCFD_LAUNCH_B(K_CG_1, s, KERNEL, G, B, A, sg, it, x, rb[cur], rb[1 - cur], wb[cur],       \
                wb[1 - cur], sb[cur], sb[1 - cur], w.p0, w.inv, pb[cur], pb[1 - cur], w.st,  \
                maxiter, out)
    const bool first = it == 0;
    if (rows == 1) {
      if (first)
        WCFD_CGS_ITER((k_cgs_iter<M, true, true>));
      else if (check)
        WCFD_CGS_ITER((k_cgs_iter<M, true, false, true>));
      else
        WCFD_CGS_ITER((k_cgs_iter<M, true, false>));
    } else if constexpr (M::kEll) {
      if (rows == 2 && first)
        WCFD_CGS_ITER((k_cgs_iter2<M, true>));
      else if (rows == 2 && check)
        WCFD_CGS_ITER((k_cgs_iter2<M, false, true>));
      else if (rows == 2)
        WCFD_CGS_ITER((k_cgs_iter2<M, false>));
      else if (first)
        WCFD_CGS_ITER((k_cgs_iter<M, false, true>));
      else if (check)
        WCFD_CGS_ITER((k_cgs_iter<M, false, false, true>));
      else
        WCFD_CGS_ITER((k_cgs_iter<M, false, false>));
    } else if (first) {
      WCFD_CGS_ITER((k_cgs_iter<M, false, true>));
    } else if (check) {
      WCFD_CGS_ITER((k_cgs_iter<M, false, false, true>));
    } else {
      WCFD_CGS_ITER((k_cgs_iter<M, false, false>));
    }
#undef WCFD_CGS_ITER
    prefold<3>(L, pb[1 - cur], w.st, s);
    prof_set_iter(-1);
  };
  if (K == 1 && tl.fixed_groups <= 0 && g_inline_check.load())
    return run_groups_inline(maxiter, ce, iteration, info, s);
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) iteration(it, false, nullptr);
  };
  // the r.r partials of the last iteration (it_end - 1) are in pb[it_end & 1], slot 2
  auto rr_part = [&](int it_end) { return (const double*)(pb[it_end & 1] + 2 * G); };
  return run_groups(sg, B, w.st, rr_part, maxiter, ce, tl.fixed_groups, group, info, s);
}

// Workgroup slots per XCD for k_cgs_xcd<M, RV> at B threads per workgroup on
// the current device: co-resident workgroups per CU times CUs per XCD (0: no
// cooperative launch).
template <class M, int RV, bool GRID, int RR = 1>
static int xcd_slots(int B) {
  struct Entry {
    int dev, B, slots;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lk(mu);
  for (const Entry& e : cache)
    if (e.dev == dev && e.B == B) return e.slots;
  int coop = 0, cus = 0, occ = 0, slots = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess &&
      coop &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cgs_xcd<M, RV, GRID, RR>, B, 0) == hipSuccess)
    slots = occ * (cus / (int)kNumXcd);
  (void)hipGetLastError();
  cache.push_back({dev, B, slots});
  return slots;
}

// Whether virtual workgroups taken rv at a time fit `slots` workgroups per
// XCD, and the slots per XCD the launch then needs: one case takes
// ceil(nb / rv) slots of one XCD; several are placed greedily on the
// least-used XCD, which bounds every XCD's use by (sum of the cases' slots)
// / 8 + the largest case's slots.
constexpr int64_t kXcdLocalRows = 8192;   // the measured crossover (design notes)

static bool xcd_plan_fits(const KLaunch& L, int rv, int slots, XcdPlan& pl) {
  const int G = L.sg.G, K = L.sg.n_cases;
  const int widest = L.max_case_blocks > 0 ? L.max_case_blocks : G;
  const int64_t m_case = (widest + rv - 1) / rv;
  const int64_t need =
      K == 1 ? m_case : ((int64_t)(G + rv - 1) / rv + K + kNumXcd - 1) / kNumXcd + m_case;
  if (slots <= 0 || need > slots) return false;
  pl.per_xcc = (int)need;
  pl.rv = rv;
  return true;
}

// The single-reduction CG solve as one k_cgs_xcd launch, with the init
// kernels of cgs_impl before it.  ran = false: the solve is not eligible,
// or the launch aborted with x untouched; the caller runs cgs_impl.
template <class M>
static int cgs_xcd_impl(const M& A, const KLaunch& L, const real* b, real* x, const Tol& tl,
                        bool r64, int maxiter, int check_every, KWork& w,
                        wcfd_solve_info* info, hipStream_t s, bool& ran) {
  ran = false;
  if constexpr (!M::kEll) {
    return WCFD_OK;
  } else {
    const Seg& sg = L.sg;
    const int n = sg.n, B = L.B, G = sg.G, K = sg.n_cases;
    const int rr = (sg.chunk + B - 1) / B;   // rows per thread of a virtual workgroup
    if (tl.fixed_groups > 0 || rr > 4 || (rr > 1 && K != 1)) return WCFD_OK;
    const int widest = L.max_case_blocks > 0 ? L.max_case_blocks : G;
    if ((int64_t)widest * sg.chunk > (int64_t)g_persistent.load()) return WCFD_OK;
    XcdPlan pl{};
    // A case of more than kXcdLocalRows rows alone in the solve spreads over
    // every XCD (one XCD gathers at an eighth of the chip's rate); smaller
    // cases, and every case of a batch, sit on one XCD each.
    bool grid = K == 1 && ((int64_t)widest * sg.chunk > kXcdLocalRows || rr > 1);
    if (!grid && !xcd_plan_fits(L, 1, xcd_slots<M, 1, false>(B), pl) &&
        !xcd_plan_fits(L, 2, xcd_slots<M, 2, false>(B), pl))
      grid = K == 1;   // one XCD's registers do not hold it
    if (grid) {
      const int per_band = (G + (int)kNumXcd - 1) / (int)kNumXcd;
      pl.rv = 0;
      if (rr == 1) {
        // one row per thread: 1-4 virtual workgroups per workgroup
        const int slots[5] = {0, xcd_slots<M, 1, true>(B), xcd_slots<M, 2, true>(B),
                              xcd_slots<M, 3, true>(B), xcd_slots<M, 4, true>(B)};
        for (int rv = 1; rv <= 4 && !pl.rv; ++rv) {
          const int need = (per_band + rv - 1) / rv;
          if (need <= slots[rv]) {
            pl.rv = rv;
            pl.per_xcc = need;
          }
        }
      } else {
        // rr rows per thread: one virtual workgroup per workgroup
        const int slots = rr == 2 ? xcd_slots<M, 2, true, 2>(B)
                          : rr == 3 ? xcd_slots<M, 3, true, 3>(B)
                                    : xcd_slots<M, 4, true, 4>(B);
        if (per_band <= slots) {
          pl.rv = rr;
          pl.per_xcc = per_band;
        }
      }
      if (!pl.rv) return WCFD_OK;   // more rows than the chip's registers hold
    } else if (!pl.rv) {
      return WCFD_OK;
    }
    pl.check_every = check_every > 0 ? check_every : 1;
    pl.wait_ticks = 10000000;   // 100 ms at 100 MHz
    if (const int cut = g_xcd_shrink.load()) {
      // the plan's own bound is unchanged; the launch has fewer slots, so the
      // plan does not fit and every workgroup aborts at entry
      pl.per_xcc = pl.per_xcc > cut ? pl.per_xcc - cut : 0;
      if (pl.per_xcc == 0) pl.per_xcc = 1;
    }
    HostSync* hs = host_sync(K + 1);
    if (!hs) return WCFD_EHIP;
    int e = prepare(w.st, sg, tl, s);
    if (e) return e;
    real* rb[2] = {w.r, w.r0};
    real* wb[2] = {w.v0, w.v1};
    real* sb[2] = {w.s, w.t};
    double* pb[2] = {w.PA, w.PC};
    if (const unsigned gi = items_grid(L.rows, n))
      WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, gi, A, L.rows, n, w.inv);
    if (r64)
      WCFD_LAUNCH_B(K_CG_INIT, s, (k_cgs_init1<M, true>), G, B, A, sg, b, x, rb[0], w.PC);
    else
      WCFD_LAUNCH_B(K_CG_INIT, s, (k_cgs_init1<M, false>), G, B, A, sg, b, x, rb[0], w.PC);
    WCFD_LAUNCH_B(K_CG_INIT, s, k_cgs_init2<M>, G, B, A, sg, rb[0], w.inv, wb[0], w.PA);
    WCFD_LAUNCH(K_FOLD, s, k_init_fold, K, sg, w.PA, w.PC, w.st, 2, maxiter);
    hipLaunchKernelGGL(k_xcd_reset, dim3(1), dim3(kWave), 0, s, w.sync, kSyncCase + K);
    if ((e = check_launch())) return e;
    M a_ = A;
    Seg sg_ = sg;
    const real* x_in = x;
    real* x_out = w.p1;
    int maxiter_ = maxiter;
    void* args[] = {&a_,    &sg_,   &pl,    &x_in,  &x_out, &rb[0],    &rb[1], &wb[0],
                    &wb[1], &sb[0], &sb[1], &w.inv, &pb[0], &pb[1], &w.st, &maxiter_,
                    &w.sync};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = prof_active() && prof_events(K_CG_XCD, &e0, &e1);
    if (timed && (e = check_hip(hipEventRecord(e0, s)))) return e;
    const void* kern =
        !grid ? (pl.rv == 1 ? (const void*)k_cgs_xcd<M, 1, false> : (const void*)k_cgs_xcd<M, 2, false>)
        : rr == 2 ? (const void*)k_cgs_xcd<M, 2, true, 2>
        : rr == 3 ? (const void*)k_cgs_xcd<M, 3, true, 3>
        : rr == 4 ? (const void*)k_cgs_xcd<M, 4, true, 4>
        : pl.rv == 1 ? (const void*)k_cgs_xcd<M, 1, true>
        : pl.rv == 2 ? (const void*)k_cgs_xcd<M, 2, true>
        : pl.rv == 3 ? (const void*)k_cgs_xcd<M, 3, true>
                     : (const void*)k_cgs_xcd<M, 4, true>;
    if ((e = check_hip(hipLaunchCooperativeKernel(kern, dim3(kNumXcd * pl.per_xcc), dim3(B),
                                                  args, 0, s))))
      return e;
    if (timed && (e = check_hip(hipEventRecord(e1, s)))) return e;
    WCFD_LAUNCH_B(K_FOLD, s, k_xcd_commit, G, B, sg, w.st, w.sync, w.p1, x);
    hipLaunchKernelGGL(k_xcd_publish, dim3(1), dim3(kWave), 0, s, w.st, K, w.sync, hs->dslot);
    if ((e = check_launch())) return e;
    if ((e = check_hip(hipEventRecord(hs->ev[0], s)))) return e;
    if ((e = check_hip(hipEventSynchronize(hs->ev[0])))) return e;
    const KFlags* f = hs->slot;
    if (f[K].done) {   // aborted: x is as it was; the setting turns itself off
      g_persistent = 0;
      ++g_xcd_fallbacks;
      return WCFD_OK;
    }
    ++g_xcd_solves;
    ran = true;
    return finish(f, K, info);
  }
}

template <class M, int NV>
static int bicg_impl(const M& A, const KLaunch& L, const real* b, real* x, const Tol& tl,
                     int maxiter, int check_every, KWork& w, wcfd_solve_info* info,
                     hipStream_t s) {
  const Seg& sg = L.sg;
  const int n = sg.n, B = L.B, G = sg.G, K = sg.n_cases;
  const bool pf = row_per_thread<M>(L);
  int e = prepare(w.st, sg, tl, s);
  if (e) return e;
  if (const unsigned gi = items_grid(L.rows, n))
    WCFD_LAUNCH(K_INV_DIAG, s, k_inv_diag<M>, gi, A, L.rows, n, w.inv);
  WCFD_LAUNCH_B(K_BI_INIT, s, (k_bi_init<M, NV>), G, B, A, sg, b, x, w.r, w.r0, w.PA);
  WCFD_LAUNCH(K_FOLD, s, k_init_fold, K, sg, w.PA, w.PA, w.st, 1, maxiter);
  if ((e = check_launch())) return e;
  real* pb[2] = {w.p0, w.p1};
  real* vb[2] = {w.v0, w.v1};
  const int ce = check_every > 0 ? check_every : 1;
  auto group = [&](int it0) {
    for (int it = it0; it < it0 + ce; ++it) {
      prof_set_iter(it);
      const int cur = it & 1;
      if (pf)
        WCFD_LAUNCH_B(K_BI_1, s, (k_bi_1<M, NV, true>), G, B, A, sg, it, w.r, w.inv, pb[cur],
                      vb[cur], pb[1 - cur], vb[1 - cur], w.r0, w.PA, w.PB, w.st);
      else
        WCFD_LAUNCH_B(K_BI_1, s, (k_bi_1<M, NV, false>), G, B, A, sg, it, w.r, w.inv, pb[cur],
                      vb[cur], pb[1 - cur], vb[1 - cur], w.r0, w.PA, w.PB, w.st);
      prefold<1>(L, w.PB, w.st, s);
      if (pf)
        WCFD_LAUNCH_B(K_BI_2, s, (k_bi_2<M, NV, true>), G, B, A, sg, it, w.r, w.inv, vb[1 - cur],
                      w.s, w.t, w.PB, w.PC, w.st);
      else
        WCFD_LAUNCH_B(K_BI_2, s, (k_bi_2<M, NV, false>), G, B, A, sg, it, w.r, w.inv, vb[1 - cur],
                      w.s, w.t, w.PB, w.PC, w.st);
      prefold<3>(L, w.PC, w.st, s);
      if (pf)
        WCFD_LAUNCH_B(K_BI_3, s, (k_bi_3<NV, true>), G, B, sg, it, x, w.r, w.inv, pb[1 - cur], w.s,
                      w.t, w.r0, w.PC, w.PA, w.st);
      else
        WCFD_LAUNCH_B(K_BI_3, s, (k_bi_3<NV, false>), G, B, sg, it, x, w.r, w.inv, pb[1 - cur],
                      w.s, w.t, w.r0, w.PC, w.PA, w.st);
      prefold<2>(L, w.PA, w.st, s);
    }
    prof_set_iter(-1);
  };
  auto rr_part = [&](int) { return (const double*)(w.PA + G); };
  return run_groups(sg, B, w.st, rr_part, maxiter, ce, tl.fixed_groups, group, info, s);
}

// maxiter under a fixed launch budget: the budget's iterations at most
static int budget_maxiter(const Tol& tl, int maxiter, int check_every) {
  if (tl.fixed_groups <= 0) return maxiter;
  const long cap = (long)tl.fixed_groups * (check_every > 0 ? check_every : 1);
  return maxiter < cap ? maxiter : (int)cap;
}

template <class M>
static int cg_dispatch(const M& A, const KLaunch& L, int variant, const real* b, real* x,
                       const Tol& tl, int maxiter, int check_every, KWork& w,
                       wcfd_solve_info* info, hipStream_t s) {
  if (variant != WCFD_CG_SINGLE_REDUCTION && variant != WCFD_CG_STANDARD) return WCFD_EINVAL;
  maxiter = budget_maxiter(tl, maxiter, check_every);
  // Mixed-precision iterative refinement: with tl.refine = n > 0 the solve
  // runs 1 + n warm-started sweeps, each from a residual b - A x formed in
  // float64; iteration counts add up over the sweeps.
  const int sweeps = 1 + (tl.refine > 0 ? tl.refine : 0);
  const bool r64 = tl.refine > 0;
  const int K = L.sg.n_cases;
  std::vector<wcfd_solve_info> tmp((info && sweeps > 1) ? K : 0);
  for (int sw = 0; sw < sweeps; ++sw) {
    wcfd_solve_info* out = (info && sweeps > 1) ? tmp.data() : info;
    bool ran = false;
    int e = WCFD_OK;
    if (variant == WCFD_CG_SINGLE_REDUCTION && g_persistent.load() > 0)
      e = cgs_xcd_impl(A, L, b, x, tl, r64, maxiter, check_every, w, out, s, ran);
    if (e) return e;
    if (!ran)
      e = variant == WCFD_CG_SINGLE_REDUCTION
              ? cgs_impl(A, L, b, x, tl, r64, maxiter, check_every, w, out, s)
              : cg_impl(A, L, b, x, tl, r64, maxiter, check_every, w, out, s);
    if (e) return e;
    if (info && sweeps > 1) {
      for (int k = 0; k < K; ++k) {
        const int prev = sw ? info[k].iterations : 0;
        info[k] = tmp[k];
        info[k].iterations += prev;
      }
    }
  }
  return WCFD_OK;
}

template <class M>
static int bicg_dispatch(const M& A, const KLaunch& L, int nvec, const real* b, real* x,
                         const Tol& tl, int maxiter, int check_every, KWork& w,
                         wcfd_solve_info* info, hipStream_t s) {
  maxiter = budget_maxiter(tl, maxiter, check_every);
  if (nvec == 1) return bicg_impl<M, 1>(A, L, b, x, tl, maxiter, check_every, w, info, s);
  if (nvec == 3) return bicg_impl<M, 3>(A, L, b, x, tl, maxiter, check_every, w, info, s);
  return WCFD_EINVAL;
}

// ELL matrix of a mesh, with or without int16 neighbour offsets
template <class F>
static int with_ell(const wcfd_mesh& m, const real* values, F&& f) {
  // tetrahedra and wedges read 4 or 5 slots (int16 offsets only; the int32
  // form reads all six, the unused ones zero)
  if (m.slot_dnb) {
    if (m.n_slots == 4) return f(EllMat<true, 4>{m.n_cells, values, m.slot_nbr, m.slot_dnb});
    if (m.n_slots == 5) return f(EllMat<true, 5>{m.n_cells, values, m.slot_nbr, m.slot_dnb});
    return f(EllMat<true>{m.n_cells, values, m.slot_nbr, m.slot_dnb});
  }
  return f(EllMat<false>{m.n_cells, values, m.slot_nbr, m.slot_dnb});
}

// grid 0: one row per thread (single_seg caps the grid at ceil(n / block))
static KLaunch single_launch(int n) {
  const KrylovLaunch kl = krylov_launch();
  Seg sg = single_seg(n, kl.block, kl.grid > 0 ? kl.grid : kMaxKrylovBlocks);
  sg.prefold = sg.G > g_krylov_fold.load() ? 1 : 0;
  return {sg, kl.block, Items{}, sg.G};
}

static KLaunch launch_of(const wcfd_batch* bt, int n) {
  if (!bt) return single_launch(n);
  Seg sg = batch_seg(*bt, n);
  const int widest = bt->max_case_blocks > 0 ? bt->max_case_blocks : bt->n_blocks;
  sg.prefold = widest > g_krylov_fold.load() ? 1 : 0;
  return {sg, bt->block, Items{bt->cell_blk, bt->n_cell_blk}, widest};
}

static KWork work_of(void* ws, const wcfd_batch* bt, int n, int nvec) {
  return bt ? carve_krylov(ws, n, nvec, bt->n_cases, bt->n_blocks)
            : carve_krylov(ws, n, nvec, 1, single_max_blocks(n));
}

int solve_cg(const CsrView& A, int variant, const real* b, real* x, const Tol& tl,
             int maxiter, int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s) {
  const CsrMat M{A.n, A.row_ptr, A.col, A.values, A.diag_pos};
  KWork w = work_of(workspace, nullptr, A.n, 1);
  return cg_dispatch(M, single_launch(A.n), variant, b, x, tl, maxiter, check_every, w,
                     info, s);
}

int solve_bicgstab(const CsrView& A, int nvec, const real* b, real* x, const Tol& tl,
                   int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                   hipStream_t s) {
  const CsrMat M{A.n, A.row_ptr, A.col, A.values, A.diag_pos};
  KWork w = work_of(workspace, nullptr, A.n, nvec);
  return bicg_dispatch(M, single_launch(A.n), nvec, b, x, tl, maxiter, check_every, w,
                       info, s);
}

int solve_cg_ell(const wcfd_mesh& m, const wcfd_batch* bt, const real* values, int variant,
                 const real* b, real* x, const Tol& tl, int maxiter, int check_every,
                 void* workspace, wcfd_solve_info* info, hipStream_t s) {
  const int n = m.n_cells;
  const KLaunch L = launch_of(bt, n);
  KWork w = work_of(workspace, bt, n, 1);
  return with_ell(m, values, [&](auto A) {
    return cg_dispatch(A, L, variant, b, x, tl, maxiter, check_every, w, info, s);
  });
}

int solve_bicgstab_ell(const wcfd_mesh& m, const wcfd_batch* bt, const real* values, int nvec,
                       const real* b, real* x, const Tol& tl, int maxiter,
                       int check_every, void* workspace, wcfd_solve_info* info, hipStream_t s) {
  const int n = m.n_cells;
  const KLaunch L = launch_of(bt, n);
  KWork w = work_of(workspace, bt, n, nvec);
  return with_ell(m, values, [&](auto A) {
    return bicg_dispatch(A, L, nvec, b, x, tl, maxiter, check_every, w, info, s);
  });
}

bool batch_ok(const wcfd_mesh* m, const wcfd_batch* bt) {
  return m && bt && m->n_cells > 0 && bt->n_cases > 0 && bt->n_blocks > 0 && bt->chunk > 0 &&
         bt->block >= kWave && bt->block <= kMaxKrylovBlock && bt->block % kWave == 0 &&
         bt->cell_off && bt->blk_off && bt->blk_case && bt->active;
}

}  // namespace wcfd

using namespace wcfd;

extern "C" {

int wcfd_spmv(int32_t n, const int32_t* row_ptr, const int32_t* col, const real* values, int nvec,
              const real* x, real* y, void* stream) {
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

int wcfd_spmv_ell(const wcfd_mesh* m, const real* values, int nvec, const real* x, real* y,
                  void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !slots_ok(m) || !values || !x || !y)
    return WCFD_EINVAL;
  if (nvec != 1 && nvec != 3) return WCFD_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const int n = m->n_cells;
  return with_ell(*m, values, [&](auto A) {
    using M = decltype(A);
    if (nvec == 1) WCFD_LAUNCH(K_SPMV, s, (k_spmv<M, 1>), grid_for(n), A, n, x, y);
    else WCFD_LAUNCH(K_SPMV, s, (k_spmv<M, 3>), grid_for(n), A, n, x, y);
    return check_launch();
  });
}

int wcfd_dot_f64(int32_t n, const real* a, const real* b, double* out, void* workspace,
                 void* stream) {
  if (n <= 0 || !a || !b || !out || !workspace) return WCFD_EINVAL;
  Carver c(workspace);
  double* part = c.take<double>(kMaxReduceBlocks);
  hipStream_t s = (hipStream_t)stream;
  const Seg sg = single_seg(n, kBlock, kMaxReduceBlocks);
  WCFD_LAUNCH(K_DOT, s, k_dot, sg.G, sg, a, b, part);
  WCFD_LAUNCH(K_FOLD, s, k_fold_to, 1, part, sg.G, out);
  return check_launch();
}

int wcfd_set_krylov_launch(int grid, int block) {
  if (grid < 0 || grid > kMaxKrylovBlocks || block < kWave || block > kMaxKrylovBlock ||
      block % kWave != 0)
    return WCFD_EINVAL;
  g_krylov_grid = grid;
  g_krylov_block = block;
  return WCFD_OK;
}

int wcfd_set_krylov_fold(int max_inline) {
  if (max_inline < 0) return WCFD_EINVAL;
  g_krylov_fold = max_inline;
  return WCFD_OK;
}

int wcfd_get_krylov_fold(void) { return g_krylov_fold.load(); }

int wcfd_set_krylov_inline_check(int on) {
  if (on != 0 && on != 1) return WCFD_EINVAL;
  g_inline_check = on;
  return WCFD_OK;
}

int wcfd_get_krylov_inline_check(void) { return g_inline_check.load(); }

int wcfd_set_krylov_persistent(int max_rows) {
  if (max_rows < 0) return WCFD_EINVAL;
  g_persistent = max_rows;
  return WCFD_OK;
}

int wcfd_get_krylov_persistent(void) { return g_persistent.load(); }

int wcfd_debug_krylov_persistent_shrink(int slots) {
  if (slots < 0) return WCFD_EINVAL;
  g_xcd_shrink = slots;
  return WCFD_OK;
}

int wcfd_krylov_persistent_stats(int64_t* solves, int64_t* fallbacks) {
  if (!solves || !fallbacks) return WCFD_EINVAL;
  *solves = g_xcd_solves.load();
  *fallbacks = g_xcd_fallbacks.load();
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

size_t wcfd_batch_krylov_workspace_bytes(int32_t n, int32_t n_cases, int32_t n_blocks) {
  if (n <= 0 || n_cases <= 0 || n_blocks <= 0) return 0;
  return krylov_workspace_bytes_batch(n, 3, n_cases, n_blocks);
}

int wcfd_cg(int32_t n, const int32_t* row_ptr, const int32_t* col, const real* values,
            const int32_t* diag_pos, int variant, const real* b, real* x, double atol, double tol,
            int maxiter, int check_every, void* workspace, wcfd_solve_info* info, void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !diag_pos || !b || !x || !workspace)
    return WCFD_EINVAL;
  CsrView A{n, row_ptr, col, values, diag_pos};
  return solve_cg(A, variant, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every, workspace, info,
                  (hipStream_t)stream);
}

int wcfd_bicgstab(int32_t n, const int32_t* row_ptr, const int32_t* col, const real* values,
                  const int32_t* diag_pos, int nvec, const real* b, real* x, double atol,
                  double tol, int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                  void* stream) {
  if (n <= 0 || !row_ptr || !col || !values || !diag_pos || !b || !x || !workspace)
    return WCFD_EINVAL;
  CsrView A{n, row_ptr, col, values, diag_pos};
  return solve_bicgstab(A, nvec, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every, workspace, info,
                        (hipStream_t)stream);
}

int wcfd_cg_ell(const wcfd_mesh* m, const real* values, int variant, const real* b, real* x,
                double atol, double tol, int maxiter, int check_every, void* workspace,
                wcfd_solve_info* info, void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !slots_ok(m) || !values || !b || !x ||
      !workspace)
    return WCFD_EINVAL;
  return solve_cg_ell(*m, nullptr, values, variant, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every,
                      workspace, info, (hipStream_t)stream);
}

int wcfd_bicgstab_ell(const wcfd_mesh* m, const real* values, int nvec, const real* b, real* x,
                      double atol, double tol, int maxiter, int check_every, void* workspace,
                      wcfd_solve_info* info, void* stream) {
  if (!m || m->n_cells <= 0 || !m->slot_nbr || !slots_ok(m) || !values || !b || !x ||
      !workspace)
    return WCFD_EINVAL;
  return solve_bicgstab_ell(*m, nullptr, values, nvec, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every,
                            workspace, info, (hipStream_t)stream);
}

int wcfd_batch_cg_ell(const wcfd_mesh* m, const wcfd_batch* bt, const real* values, int variant,
                      const real* b, real* x, double atol, double tol, int maxiter,
                      int check_every, void* workspace, wcfd_solve_info* info, void* stream) {
  if (!batch_ok(m, bt) || !m->slot_nbr || !slots_ok(m) || !values || !b || !x || !workspace)
    return WCFD_EINVAL;
  return solve_cg_ell(*m, bt, values, variant, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every, workspace,
                      info, (hipStream_t)stream);
}

int wcfd_batch_bicgstab_ell(const wcfd_mesh* m, const wcfd_batch* bt, const real* values,
                            int nvec, const real* b, real* x, double atol, double tol,
                            int maxiter, int check_every, void* workspace, wcfd_solve_info* info,
                            void* stream) {
  if (!batch_ok(m, bt) || !m->slot_nbr || !slots_ok(m) || !values || !b || !x || !workspace)
    return WCFD_EINVAL;
  return solve_bicgstab_ell(*m, bt, values, nvec, b, x, Tol{atol, tol, 0.0, 0, 0}, maxiter, check_every,
                            workspace, info, (hipStream_t)stream);
}

}  // extern "C"
