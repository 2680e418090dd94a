"""Pins oracle.krylov (our statement of warp.optim.linear.cg / bicgstab) by
the true residual and scipy's direct solver, and fixes the stopping rule the
HIP solvers follow: check before the first iteration and every check_every
iterations, threshold max(tol ||b||, atol)."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import krylov


def _laplacian(n, seed=0, nonsym=0.0):
    rng = np.random.default_rng(seed)
    g = sp.random(n, n, density=4.0 / n, random_state=seed).tocsr()
    g.data[:] = 0.1 + rng.random(g.nnz)
    g = (g + g.T).tocsr()
    d = np.asarray(g.sum(axis=1)).ravel() + 0.5
    A = (sp.diags(d) - g).tocsr()
    if nonsym:
        Nn = sp.random(n, n, density=2.0 / n, random_state=seed + 1) * nonsym
        A = (A - Nn + sp.diags(np.asarray(Nn.sum(axis=1)).ravel())).tocsr()
    return A


def test_cg_matches_direct_solve_and_negative_definite():
    A = _laplacian(300)
    b = np.random.default_rng(1).standard_normal(300)
    for sign in (1.0, -1.0):           # the reference's pressure matrix is negative definite
        As = (sign * A).astype(np.float64)
        x = np.zeros(300)
        it, res = krylov.cg(As, sign * b, x, 1.0 / As.diagonal(), atol=1e-12, maxiter=2000)
        xs = spla.spsolve(As.tocsc(), sign * b)
        assert it % 10 == 0 and it > 0
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
        assert np.linalg.norm(sign * b - As @ x) <= 1e-12 * 1.0001 * max(1.0, np.linalg.norm(b))


def test_bicgstab_matches_direct_solve():
    A = _laplacian(300, seed=3, nonsym=0.5)
    b = np.random.default_rng(2).standard_normal(900)
    K = sp.kron(A, sp.eye(3)).tocsr()        # interleaved 3-component system
    x = np.zeros(900)
    it, res = krylov.bicgstab(K, b, x, 1.0 / K.diagonal(), atol=1e-11, maxiter=2000)
    xs = spla.spsolve(K.tocsc(), b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    assert it % 10 == 0


def test_stopping_rule():
    A = _laplacian(200, seed=5)
    b = np.random.default_rng(3).standard_normal(200)
    x = spla.spsolve(A.tocsc(), b)
    # converged start: zero iterations, x untouched
    x0 = x.copy()
    it, res = krylov.cg(A, b, x0, 1.0 / A.diagonal(), atol=1e-6)
    assert it == 0 and np.array_equal(x0, x)
    # threshold is max(tol ||b||, atol) with tol = atol by default
    assert krylov.threshold(b, 1e-5) == max(1e-5 * np.linalg.norm(b), 1e-5)
    assert krylov.threshold(b * 1e-9, 1e-5) == 1e-5
    # maxiter caps the count (multiple of check_every)
    x1 = np.zeros(200)
    it, res = krylov.cg(A, b, x1, 1.0 / A.diagonal(), atol=1e-30, maxiter=30)
    assert it == 30


def test_single_reduction_cg_matches_cg():
    """Chronopoulos-Gear CG: same solution and (in this well-conditioned
    case) the same number of check intervals as standard CG."""
    A = _laplacian(400, seed=7)
    b = np.random.default_rng(5).standard_normal(400)
    inv = 1.0 / A.diagonal()
    x1 = np.zeros(400); x2 = np.zeros(400)
    it1, _ = krylov.cg(A, b, x1, inv, atol=1e-10, maxiter=2000)
    it2, _ = krylov.cg_single_reduction(A, b, x2, inv, atol=1e-10, maxiter=2000)
    xs = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x2 - xs) / np.linalg.norm(xs) < 1e-8
    assert abs(it1 - it2) <= 10
    # float32 vectors with float64 dots: still converges to the float32 floor
    A32 = A.astype(np.float32)
    x3 = np.zeros(400, np.float32)
    it3, res3 = krylov.cg_single_reduction(A32, b.astype(np.float32), x3, inv.astype(np.float32),
                                           atol=1e-4, maxiter=2000)
    assert res3 <= 1e-4 * max(1.0, np.linalg.norm(b))


def test_float64_refinement_lowers_the_true_residual():
    """Mixed-precision iterative refinement (oracle.krylov.refined): the true
    residual of the float32 solution, measured in float64, is lower after
    extra sweeps from a float64 residual than with the float32 recurrence
    alone at the same (unreachable) tolerance."""
    import scipy.sparse as sp
    from oracle import krylov
    n = 400
    rng = np.random.default_rng(5)
    main = np.full(n, 2.0) + rng.uniform(0, 1e-3, n)
    A64 = sp.diags([-np.ones(n - 1), main, -np.ones(n - 1)], [-1, 0, 1], format='csr')
    A = A64.astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    inv = (1.0 / A.diagonal()).astype(np.float32)

    def true_res(x):
        return np.linalg.norm(b.astype(np.float64) - A.astype(np.float64) @ x.astype(np.float64))

    x0 = np.zeros(n, np.float32)
    it0, _ = krylov.cg(A, b, x0, inv, atol=1e-12, tol=0.0, maxiter=2000, check_every=10)
    x2 = np.zeros(n, np.float32)
    it2, _ = krylov.refined(krylov.cg, A, b, x2, inv, refine=2, atol=1e-12, tol=0.0,
                            maxiter=2000, check_every=10)
    assert true_res(x2) < true_res(x0)
    assert it2 > it0
