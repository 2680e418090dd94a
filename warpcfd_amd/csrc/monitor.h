// Per-case convergence monitor of a steady SIMPLE run.  No counterpart in
// the reference, whose check_convergence (model.py:278-315,
// Convergence.py:7-106) prints norms on the host and never stops a run.
//
// Inputs per outer iteration: the continuity residual c (sum_c |sum_f F_c| /
// sum_c sum_f |F_c| of the corrected flux) and the relative velocity change
// d = ||u^n - u^(n-1)||_2 / ||u^n||_2, both accumulated in float64.
//
// SIMPLE converges linearly, e_n ~ rho^n, so the distance to the fixed point
// is estimated a posteriori as  E = dbar rho / (1 - rho),  where dbar is the
// mean of d over the last W iterations and rho = (D_now / D_prev)^(1/W) from
// the sums of d over the last two windows of W iterations (E = inf while
// rho >= 1 or fewer than 2W iterations are known).  A case is converged when
//      c <= stop_c   and   (E <= stop_u   or   dbar <= floor_u).
// floor_u covers runs that sit at the float32 rounding floor of d, where
// rho is no longer measurable.
#pragma once

#include <math.h>

namespace wcfd {

constexpr int kMonWindow = 20;
// monitor record of one case (doubles), zeroed when the case starts:
// [0] iterations seen  [1] c  [2] d  [3] dbar  [4] rho  [5] E
// [8, 8 + 2W) ring of the last 2W values of d
constexpr int kMonHead = 8;

__host__ __device__ inline bool monitor_update(double* mon, double c, double d, double stop_c,
                                               double stop_u, double floor_u) {
  const int W = kMonWindow;
  const long n = (long)mon[0];
  double* ring = mon + kMonHead;
  ring[n % (2 * W)] = d;
  const long m = n + 1;
  mon[0] = (double)m;
  mon[1] = c;
  mon[2] = d;
  double E = INFINITY, rho = INFINITY, dbar = INFINITY;
  if (m >= 2 * W) {
    double now = 0.0, prev = 0.0;
    for (int j = 0; j < W; ++j) {
      now += ring[(m - 1 - j) % (2 * W)];
      prev += ring[(m - 1 - W - j) % (2 * W)];
    }
    dbar = now / W;
    if (prev > 0.0) {
      rho = pow(now / prev, 1.0 / W);
      if (rho < 1.0) E = dbar * rho / (1.0 - rho);
    } else {
      rho = 0.0;
      E = 0.0;
    }
  }
  mon[3] = dbar;
  mon[4] = rho;
  mon[5] = E;
  return c <= stop_c && (E <= stop_u || dbar <= floor_u);
}

}  // namespace wcfd
