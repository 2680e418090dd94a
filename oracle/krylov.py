"""Oracle restatement of the Krylov solvers the reference calls (test infrastructure).

The reference calls warp.optim.linear.cg / bicgstab with a Jacobi
preconditioner, atol = 1e-5 (float32) or 1e-6 (float64), maxiter = 500
(equation.py:35-39, 195-200; incompressible.py:32).  Warp is third party and
absent here, so this is our statement of those published algorithms:

* stopping threshold  thr = max(tol * ||b||_2, atol) with tol = atol (the
  reference passes atol only).  ``rtol0`` > 0 adds rtol0 * ||r_0||_2 (the
  warm-start residual) to the max -- not in the reference; the HIP path
  offers it for runs to a continuity target;
* the residual checked is the recurrence residual r;
* the check happens before the first iteration and after every
  ``check_every`` = 10 iterations; at most ``maxiter`` iterations;
* x is updated in place (warm start from the caller's buffer);
* BiCGSTAB is right-preconditioned (Warp's default);
* dot products and norms accumulate in float64 (the HIP path does the same);
  vector updates stay in the vector dtype with the scalar rounded to it.

Iterate-level behaviour is *parity unpinned* (no reference vector or test
fixes it).  The solutions are pinned by ||b - A x|| and scipy spsolve.
"""
import numpy as np


def _dot(a, b):
    return float(np.dot(a.astype(np.float64), b.astype(np.float64)))


def _safe_div(a, b):
    return a / b if b != 0.0 else 0.0


def threshold(b, atol, tol=None, r0=None, rtol0=0.0):
    tol = atol if tol is None else tol
    t = max(tol * np.sqrt(_dot(b, b)), atol)
    if rtol0 and r0 is not None:
        t = max(t, rtol0 * np.sqrt(_dot(r0, r0)))
    return t


def _residual(A, b, x, r64):
    """b - A x in the vector dtype, or formed in float64 and rounded once."""
    if r64:
        return (b.astype(np.float64) - A.astype(np.float64) @ x.astype(np.float64)).astype(x.dtype)
    return (b - A @ x).astype(x.dtype)


def refined(solver, A, b, x, inv_diag, refine=0, **kw):
    """Mixed-precision iterative refinement (the HIP path's ``p_refine``; not
    in the reference): 1 + refine warm-started sweeps of ``solver``, each
    from a float64 residual.  Returns (total iterations, last ||r||)."""
    if refine <= 0:
        return solver(A, b, x, inv_diag, **kw)
    total = 0
    res = None
    for _ in range(1 + refine):
        it, res = solver(A, b, x, inv_diag, r64=True, **kw)
        total += it
    return total, res


def cg(A, b, x, inv_diag, atol=1e-5, tol=None, maxiter=500, check_every=10, rtol0=0.0,
       r64=False):
    """Jacobi-preconditioned conjugate gradients on CSR ``A``; returns
    (iterations, ||r||)."""
    dt = x.dtype.type
    r = _residual(A, b, x, r64)
    thr2 = threshold(b, atol, tol, r, rtol0) ** 2
    z = (r * inv_diag).astype(x.dtype)
    p = z.copy()
    rz = _dot(r, z)
    rr = _dot(r, r)
    it = 0
    if rr <= thr2 or maxiter == 0:
        return it, np.sqrt(rr)
    while True:
        for _ in range(check_every):
            q = (A @ p).astype(x.dtype)
            alpha = _safe_div(rz, _dot(p, q))
            x += dt(alpha) * p
            r -= dt(alpha) * q
            z = (r * inv_diag).astype(x.dtype)
            rz_new = _dot(r, z)
            rr = _dot(r, r)
            beta = _safe_div(rz_new, rz)
            rz = rz_new
            p = (z + dt(beta) * p).astype(x.dtype)
        it += check_every
        if rr <= thr2 or it >= maxiter:
            return it, np.sqrt(rr)


def bicgstab(A, b, x, inv_diag, atol=1e-5, tol=None, maxiter=500, check_every=10, rtol0=0.0):
    """Right-Jacobi-preconditioned BiCGSTAB on CSR ``A`` (all components of
    the interleaved momentum vector share one set of scalars, as in the
    reference's single 3C solve); returns (iterations, ||r||)."""
    dt = x.dtype.type
    r = (b - A @ x).astype(x.dtype)
    thr2 = threshold(b, atol, tol, r, rtol0) ** 2
    r0 = r.copy()
    p = r.copy()
    v = np.zeros_like(r)
    rho = _dot(r0, r)
    rr = _dot(r, r)
    it = 0
    if rr <= thr2 or maxiter == 0:
        return it, np.sqrt(rr)
    while True:
        for _ in range(check_every):
            y = (p * inv_diag).astype(x.dtype)
            v = (A @ y).astype(x.dtype)
            alpha = _safe_div(rho, _dot(r0, v))
            s = (r - dt(alpha) * v).astype(x.dtype)
            z = (s * inv_diag).astype(x.dtype)
            t = (A @ z).astype(x.dtype)
            ss = _dot(s, s)
            omega = _safe_div(_dot(t, s), _dot(t, t)) if ss > thr2 else 0.0
            x += dt(alpha) * y + dt(omega) * z
            r = (s - dt(omega) * t).astype(x.dtype)
            rho_new = _dot(r0, r)
            rr = _dot(r, r)
            beta = _safe_div(rho_new, rho) * _safe_div(alpha, omega)
            rho = rho_new
            p = (r + dt(beta) * (p - dt(omega) * v)).astype(x.dtype)
        it += check_every
        if rr <= thr2 or it >= maxiter:
            return it, np.sqrt(rr)


def cg_single_reduction(A, b, x, inv_diag, atol=1e-5, tol=None, maxiter=500, check_every=10,
                        rtol0=0.0, r64=False):
    """Chronopoulos-Gear PCG: the same iterates as ``cg`` in exact arithmetic,
    with both inner products of an iteration taken together (one global
    reduction per iteration).  Restates the variant the HIP path can run as
    one fused kernel per iteration; same stopping rule as ``cg``."""
    dt = x.dtype.type
    r = _residual(A, b, x, r64)
    thr2 = threshold(b, atol, tol, r, rtol0) ** 2
    u = (r * inv_diag).astype(x.dtype)
    w = (A @ u).astype(x.dtype)
    gamma = _dot(r, u)
    delta = _dot(w, u)
    rr = _dot(r, r)
    it = 0
    if rr <= thr2 or maxiter == 0:
        return it, np.sqrt(rr)
    alpha = _safe_div(gamma, delta)
    beta = 0.0
    p = np.zeros_like(r)
    s = np.zeros_like(r)
    first = True
    while True:
        for _ in range(check_every):
            if first:
                p = u.copy()
                s = w.copy()
                first = False
            else:
                p = (u + dt(beta) * p).astype(x.dtype)
                s = (w + dt(beta) * s).astype(x.dtype)
            x += dt(alpha) * p
            r = (r - dt(alpha) * s).astype(x.dtype)
            u = (r * inv_diag).astype(x.dtype)
            w = (A @ u).astype(x.dtype)
            g_new = _dot(r, u)
            d_new = _dot(w, u)
            rr = _dot(r, r)
            beta = _safe_div(g_new, gamma)
            alpha = _safe_div(g_new, d_new - beta * _safe_div(g_new, alpha))
            gamma = g_new
        it += check_every
        if rr <= thr2 or it >= maxiter:
            return it, np.sqrt(rr)
