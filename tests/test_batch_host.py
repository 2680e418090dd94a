"""Host logic of the batched path (no GPU): the arena's range allocator,
the batch workload draws, the oracle's warm-start-relative Krylov stop, the
split of a batch over arenas and the worker threads' failure path."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import krylov
from warpcfd_amd.batch import _Ranges, item_blocks
from warpcfd_amd.workloads import batch64_draws


def test_ranges_first_fit_and_coalesce():
    r = _Ranges(100)
    a = r.alloc(30)
    b = r.alloc(30)
    c = r.alloc(30)
    assert (a, b, c) == (0, 30, 60)
    assert r.alloc(20) is None            # only 10 left
    r.release(b, 30)
    assert r.alloc(25) == 30              # first fit into the hole
    r.release(a, 30)
    r.release(30, 25)
    assert r.free == [(0, 60), (90, 10)]  # coalesced
    assert r.largest() == 60
    r.release(c, 30)
    assert r.free == [(0, 100)]
    assert r.alloc(0) == 0


def test_item_blocks_cover_each_range_exactly_once():
    """Compacted launch ranges: every item of every range in exactly one
    workgroup, none outside, at most `per` per workgroup, and ranges never
    share a workgroup."""
    assert item_blocks([]).shape == (2, 0)
    assert item_blocks([(5, 5)]).shape == (2, 0)
    ranges = [(0, 256), (300, 301), (1000, 1600), (2000, 2513)]
    blk = item_blocks(ranges, per=256)
    st, en = blk
    assert blk.dtype == np.int32 and np.all(en > st) and np.all(en - st <= 256)
    covered = np.concatenate([np.arange(a, b) for a, b in zip(st, en)])
    want = np.concatenate([np.arange(a, b) for a, b in ranges])
    assert np.array_equal(np.sort(covered), want)
    assert blk.shape[1] == 1 + 1 + 3 + 3
    for a, b in zip(st, en):
        assert any(lo <= a and b <= hi for lo, hi in ranges)


def test_batch64_draws_match_the_stated_distribution():
    d = batch64_draws(0)
    assert len(d) == 64
    N = np.array([n for n, _ in d])
    Re = np.array([r for _, r in d])
    assert N.min() >= 32 and N.max() <= 256 and Re.min() >= 100 and Re.max() <= 1000
    # the seed-0 batch: sum of N^2 as documented for the workload
    assert int((N * N).sum()) == 976073
    assert d == batch64_draws(0)


def _spd(n=120, seed=0):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.05, random_state=seed, format='csr')
    A = (A + A.T) * 0.5
    A = A + sp.eye(n) * (np.abs(A).sum(axis=1).max() + 0.5)
    return A.tocsr().astype(np.float64), rng.standard_normal(n)


def test_rtol0_stops_relative_to_the_warm_start_residual():
    A, b = _spd()
    inv = 1.0 / A.diagonal()
    x0 = np.zeros_like(b)
    r0 = np.linalg.norm(b - A @ x0)
    for fn in (krylov.cg, krylov.cg_single_reduction, krylov.bicgstab):
        x = x0.copy()
        it, res = fn(A, b, x, inv, atol=0.0, tol=0.0, rtol0=1e-3, maxiter=500, check_every=1)
        assert res <= 1e-3 * r0 * (1 + 1e-12)
        assert np.linalg.norm(b - A @ x) <= 2e-3 * r0
        # a stricter relative stop needs more iterations
        x2 = x0.copy()
        it2, _ = fn(A, b, x2, inv, atol=0.0, tol=0.0, rtol0=1e-8, maxiter=500, check_every=1)
        assert it2 > it
    # warm start near the solution: the stop is relative to the warm-start
    # residual, not to ||b|| (a tol on ||b|| of the same size stops at once)
    xs = sp.linalg.spsolve(A.tocsc(), b)
    xw = xs + 1e-4 * np.random.default_rng(1).standard_normal(b.size)
    rw = np.linalg.norm(b - A @ xw)
    x = xw.copy()
    it, res = krylov.cg(A, b, x, inv, atol=0.0, tol=0.0, rtol0=1e-3, maxiter=500, check_every=1)
    assert it > 0 and res <= 1e-3 * rw
    x = xw.copy()
    it_b, _ = krylov.cg(A, b, x, inv, atol=0.0, tol=1e-3, maxiter=500, check_every=1)
    assert it_b == 0


def _geometric_run(mon, rho, d0, n, c=1e-7):
    out = None
    for i in range(n):
        out = mon.update(c, d0 * rho ** i)
    return out


def test_monitor_error_estimate_on_a_geometric_sequence():
    """For d_n = d0 rho^n the exact distance to the limit of the iterates is
    d_n rho / (1 - rho); the monitor's estimate must reproduce it."""
    from warpcfd_amd.FV.implicit_Solvers.incompressible import Monitor
    for rho in (0.9, 0.99, 0.998):
        mon = Monitor(1e-6, 1e-300, 0.0)
        _geometric_run(mon, rho, 1e-3, 200)
        c, d, dbar, r, E = mon.estimate
        assert abs(r - rho) < 1e-9
        W = 20
        dbar_exact = 1e-3 * rho ** 199 * sum(rho ** -j for j in range(W)) / W
        assert abs(dbar - dbar_exact) <= 1e-12 * dbar_exact + 1e-30
        assert abs(E - dbar_exact * rho / (1 - rho)) <= 1e-9 * E


def test_monitor_rule():
    from warpcfd_amd.FV.implicit_Solvers.incompressible import Monitor
    # continuity above the stop: never converged, however small the change
    mon = Monitor(1e-6, 5e-5, 1e-7)
    assert not _geometric_run(mon, 0.5, 1e-3, 100, c=2e-6)
    # fewer than 2 windows: no estimate
    mon = Monitor(1e-6, 5e-5, 0.0)
    assert not _geometric_run(mon, 0.5, 1e-3, 39)
    assert mon.estimate[4] == float('inf')
    assert mon.update(1e-7, 1e-3 * 0.5 ** 39)      # 40th value: E tiny
    # non-decaying change above the floor: not converged; at the floor: converged
    mon = Monitor(1e-6, 5e-5, 1e-7)
    assert not _geometric_run(mon, 1.0, 1e-6, 100)
    mon = Monitor(1e-6, 5e-5, 1e-7)
    assert _geometric_run(mon, 1.0, 5e-8, 100)
    # slow decay: E = dbar rho/(1-rho) above the target until d is small enough
    mon = Monitor(1e-6, 5e-5, 0.0)
    assert not _geometric_run(mon, 0.999, 1e-6, 100)   # E ~ 1e-3


def test_arena_tuner_ties_take_the_first_candidate():
    from warpcfd_amd.tune import pick_with_ties
    r = [{'n': 0, 't': 6.364}, {'n': 1, 't': 6.618}, {'n': 2, 't': 6.353}, {'n': 3, 't': 7.2}]
    assert pick_with_ties(r, 't', 0.02)['n'] == 0        # 0.2 % slower than the fastest: a tie
    assert pick_with_ties(r, 't', 0.0)['n'] == 2         # no tolerance: the fastest
    r[0]['t'] = 6.6
    assert pick_with_ties(r, 't', 0.02)['n'] == 2        # 3.9 % slower: not a tie


def test_krylov_tuner_candidates():
    """Every fixed grid and the 2 / 1.5 / 1.25 rows-per-thread grids per block
    are tried; two-row grids fit 2 block rows per workgroup; each shape is
    listed once; grids stay within the library's limit."""
    from warpcfd_amd.tune import MAX_GRID, krylov_candidates
    C = 320000
    cand = krylov_candidates(C)
    assert len(cand) == len(set(cand))
    shapes = {(g, b) for g, b, _ in cand}
    for b in (128, 256, 512):
        assert (0, b) in shapes
        for fill in (2.0, 1.5, 1.25):
            g = -(-C // int(fill * b))
            assert (g, b) in shapes
            chunk = -(-C // g)
            assert b < chunk <= 2 * b          # the two-rows-per-thread kernel
    assert all(g <= MAX_GRID for g, _, _ in cand)
    assert (500, 512, 4096) in cand and (500, 512, 0) in cand
    # a case smaller than a workgroup: no two-row grids, no grids far beyond its rows
    assert {g for g, _, _ in krylov_candidates(100)} <= {0, 128}


def test_split_groups_deal_back_and_forth():
    """Cases in decreasing size dealt 0, 1, ..., n-1, n-1, ..., 0, 0, ...:
    here every group's total is 12; empty groups are dropped; ties keep the
    index order."""
    from warpcfd_amd.batch import split_groups
    sizes = [5, 9, 1, 9, 3, 7, 2]
    g = split_groups(sizes, 3)
    assert g == [[1, 6, 2], [3, 4], [5, 0]]
    assert [sum(sizes[k] for k in x) for x in g] == [12, 12, 12]
    assert split_groups([1, 2], 4) == [[1], [0]]
    assert split_groups([4, 4, 4, 4], 2) == [[0, 3], [1, 2]]
    assert split_groups(sizes, 1) == [[1, 3, 5, 0, 4, 6, 2]]
    assert split_groups([], 3) == []


def test_run_workers_stops_the_others_and_raises():
    """One worker fails: the abort event stops the others at their next
    check, and the failure is raised on the calling thread."""
    import time
    from warpcfd_amd.batch import run_workers
    seen = {}

    def work(j, abort):
        if j == 1:
            time.sleep(0.05)
            raise RuntimeError('worker 1 failed')
        n = 0
        while not abort.is_set() and n < 10_000:
            time.sleep(0.001)
            n += 1
        seen[j] = n

    t = time.perf_counter()
    with pytest.raises(RuntimeError, match='worker 1 failed'):
        run_workers(work, 3)
    assert time.perf_counter() - t < 5.0
    assert sorted(seen) == [0, 2] and all(n < 10_000 for n in seen.values())


def test_run_workers_all_succeed():
    from warpcfd_amd.batch import run_workers
    out = []
    run_workers(lambda j, abort: out.append(j), 4)
    assert sorted(out) == [0, 1, 2, 3]
