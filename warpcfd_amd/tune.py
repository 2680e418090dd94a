"""Autotuner for the Krylov launch shape (workgroups x threads, fold mode).

The CG and BiCGSTAB kernels are grid-stride loops whose workgroups each write
one float64 partial per reduction; the next kernel folds all of them in its
prologue, or, above ``wcfd_set_krylov_fold``'s limit, a fold kernel reduces
them to one value first.  More workgroups hide memory latency better (one row
per thread lets a row's loads issue before the fold) but make inline folds
longer, so the best shape depends on the mesh size.  ``autotune_krylov``
times fixed-iteration solves on a pressure-like matrix built on the mesh's
own pattern for every candidate, installs the fastest and caches it by cell
count.  Grid 0 means one row per thread; for each block size the grids
that give 2, 1.5 and 1.25 rows per thread (chunk <= 2 block) are tried
too, where the single-reduction CG kernel takes two rows per thread (rows t
and t + block of the workgroup's chunk) and folds fewer partials.
"""
import ctypes

import torch

from . import _lib
from .FV.device import ptr, stream_ptr

GRIDS = (128, 256, 512, 1024, 2048, 0)
BLOCKS = (128, 256, 512)
# the library's default inline-fold limit, "always prefold", and higher
# limits that keep the fold inline for one row per thread on mid-size meshes
FOLDS = (2048, 0, 4096, 8192)
PREFOLD_GRIDS = (1024, 2048, 0)    # grids on which "always prefold" is tried
WIDE_FOLD_GRIDS = (0,)             # grids on which the higher limits are tried
MAX_GRID = 65536                   # the library's workgroup limit for a Krylov launch
TWO_ROW_FILL = (2.0, 1.5, 1.25)    # rows per thread of the extra two-row grids

_cache = {}


def _laplacian_ell(dm):
    """Negative-definite ELL matrix with the geometric Laplacian weights
    A_f / |d| on the mesh pattern (the pressure matrix with rUA = 1)."""
    C = dm.n_cells
    sf = dm.slot_face.long()
    w = dm.area[sf] / dm.dmag[sf]
    w = torch.where(dm.slot_pos >= 0, w, torch.zeros_like(w))
    ell = torch.empty(7, C, dtype=dm.dtype, device=w.device)
    ell[1:] = w
    ell[0] = -(w.sum(dim=0)) * 1.001
    return ell.contiguous()


def install(grid, block, fold=FOLDS[0], lib=None, persistent=None):
    """Launch shape, inline-fold limit and, unless None, the persistent-CG
    row limit (0: off)."""
    lib = lib if lib is not None else _lib.load()
    _lib.check(lib.wcfd_set_krylov_launch(grid, block), 'set_krylov_launch')
    _lib.check(lib.wcfd_set_krylov_fold(fold), 'set_krylov_fold')
    if persistent is not None:
        _lib.check(lib.wcfd_set_krylov_persistent(persistent), 'set_krylov_persistent')


def time_solvers(dm, grid, block, iters=60, reps=2, cg_variant=0, fold=FOLDS[0], persistent=0):
    """Milliseconds of `iters` CG iterations and `iters` 3-vector BiCGSTAB
    iterations at launch shape (grid, block), inline-fold limit ``fold`` and
    persistent-CG row limit ``persistent``."""
    lib = dm.lib
    install(grid, block, fold, lib, persistent)
    C = dm.n_cells
    dev = dm.device
    A = _laplacian_ell(dm)
    g = torch.Generator(device='cpu').manual_seed(0)
    b = torch.randn(C, generator=g).to(dev, dm.dtype)
    b3 = torch.randn(3, C, generator=g).to(dev, dm.dtype)
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(C, 3), dtype=torch.uint8, device=dev)
    info = _lib.SolveInfo()
    best = [float('inf'), float('inf')]
    for _ in range(reps):
        for j, nv in enumerate((1, 3)):
            x = torch.zeros(nv, C, dtype=dm.dtype, device=dev)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            if nv == 1:
                code = lib.wcfd_cg_ell(dm.desc_ref, ptr(A), cg_variant, ptr(b), ptr(x), 0.0, 0.0, iters,
                                       iters, ptr(ws), ctypes.byref(info), stream_ptr())
            else:
                code = lib.wcfd_bicgstab_ell(dm.desc_ref, ptr(-A), 3, ptr(b3), ptr(x), 0.0, 0.0, iters,
                                             iters, ptr(ws), ctypes.byref(info), stream_ptr())
            e1.record()
            torch.cuda.synchronize()
            _lib.check(code, 'tuning solve')
            best[j] = min(best[j], e0.elapsed_time(e1))
    return best


def krylov_candidates(C, grids=GRIDS, blocks=BLOCKS):
    """(grid, block, inline-fold limit) shapes tried for a case of C rows:
    the fixed grids, and per block the grids that give TWO_ROW_FILL rows per
    thread; prefolding on PREFOLD_GRIDS and the two-row grids, the higher
    inline limits on WIDE_FOLD_GRIDS and the two-row grids."""
    out = []
    for blk in blocks:
        two_rows = []
        for fill in TWO_ROW_FILL:
            g = -(-C // int(fill * blk))
            if C > blk and g <= MAX_GRID and g not in grids and g not in two_rows:
                two_rows.append(g)
        for grid in list(grids) + two_rows:
            if grid * blk > 4 * C and grid > grids[0]:
                continue            # more threads than rows: no gain
            wide = grid in WIDE_FOLD_GRIDS or grid in two_rows
            for fold in FOLDS:
                if fold == 0 and grid not in PREFOLD_GRIDS and grid not in two_rows:
                    continue
                if fold > FOLDS[0] and not wide:
                    continue
                out.append((grid, blk, fold))
    return out


# Largest case for which the persistent CG at one row per thread is among
# the screened candidates (on one XCD up to 8 192 rows, spread over every XCD
# above); larger cases meet it in the refinement stage, at the launch shapes
# that screened fastest.
PERSISTENT_MAX_ROWS = 16384


def persistent_candidates(C, blocks=BLOCKS):
    """(grid, block, fold, persistent rows) shapes that run the persistent
    single-reduction CG for a case of C rows: one row per thread, every
    block size, when the case is small enough to be worth trying."""
    if C > PERSISTENT_MAX_ROWS:
        return []
    return [(0, blk, FOLDS[0], C) for blk in blocks]


def autotune_krylov(dm, grids=GRIDS, blocks=BLOCKS, iters=60, cg_weight=1.0, bicg_weight=0.1,
                    cg_variant=0, refine=6, refine_iters=300, refine_reps=3,
                    try_persistent=True):
    """Pick and install the fastest (grid, block, fold, persistent) for mesh
    ``dm`` (a DeviceMesh).  The weights reflect a SIMPLE iteration, where the
    pressure CG runs far more iterations than the momentum BiCGSTAB.  Every
    candidate is timed on ``iters`` iterations; the ``refine`` fastest are
    then timed again on ``refine_iters`` iterations, ``refine_reps`` rounds
    in turn (each also with the persistent CG when ``try_persistent``), and
    the smallest median score wins, so that the pick does not follow the
    timing noise of one short solve.  The persistent CG is a candidate only
    with the single-reduction form."""
    key = (dm.n_cells, cg_variant, dm.lib.wcfd_real_bytes(), bool(try_persistent))
    if key in _cache:
        c = _cache[key]
        install(c['grid'], c['block'], c['fold'], dm.lib, c['persistent'])
        return c
    cands = [c + (0,) for c in krylov_candidates(dm.n_cells, grids, blocks)]
    try_persistent = try_persistent and cg_variant == _lib.CG_SINGLE_REDUCTION
    if try_persistent:
        cands += persistent_candidates(dm.n_cells, blocks)

    def score(grid, blk, fold, pers, n_it, reps):
        cg_ms, bi_ms = time_solvers(dm, grid, blk, n_it, reps=reps, cg_variant=cg_variant,
                                    fold=fold, persistent=pers)
        return {'grid': grid, 'block': blk, 'fold': fold, 'persistent': pers, 'cg_ms': cg_ms,
                'bicg_ms': bi_ms, 'score': cg_weight * cg_ms + bicg_weight * bi_ms}

    results = [score(*c, iters, 2) for c in cands]
    top = sorted(results, key=lambda r: r['score'])[:max(1, refine)]
    if try_persistent:
        # each refined launch shape also with the persistent CG (a single case
        # larger than one XCD serves spreads over every XCD; the library takes
        # the per-iteration kernels where the shape is not eligible)
        seen = {(r['grid'], r['block'], r['fold'], r['persistent']) for r in top}
        for r in list(top):
            key2 = (r['grid'], r['block'], r['fold'], dm.n_cells)
            if r['persistent'] == 0 and key2 not in seen:
                seen.add(key2)
                top.append(dict(r, persistent=dm.n_cells))
    rounds = {i: [] for i in range(len(top))}
    for _ in range(refine_reps):
        for i, r in enumerate(top):
            rounds[i].append(score(r['grid'], r['block'], r['fold'], r['persistent'],
                                   refine_iters, 1))
    refined = []
    for i, r in enumerate(top):
        med = sorted(rounds[i], key=lambda x: x['score'])[len(rounds[i]) // 2]
        refined.append(dict(med, cg_ms=med['cg_ms'] * iters / refine_iters,
                            bicg_ms=med['bicg_ms'] * iters / refine_iters,
                            score=med['score'] * iters / refine_iters))
    best = min(refined, key=lambda r: r['score'])
    install(best['grid'], best['block'], best['fold'], dm.lib, best['persistent'])
    out = dict(best, iters=iters, candidates=results, refined=refined)
    _cache[key] = out
    return out


ARENA_CANDIDATES = (
    # (pressure CG form, rows per workgroup, threads per workgroup)
    (1, 256, 256), (1, 512, 512), (1, 128, 128), (1, 512, 256),
    (0, 256, 256), (0, 512, 512),
)


def autotune_arena(specs, arena_kw, candidates=ARENA_CANDIDATES, warm=2, steps=4, tie=0.02):
    """Pick the launch shape of a batch arena: for every candidate
    (pressure CG variant, chunk, block) build the arena for ``specs``, run
    ``warm`` untimed and ``steps`` timed SIMPLE iterations (the wide phase,
    every case active) and return the fastest as a dict of arena keywords
    plus the timings.  Candidates within ``tie`` (relative) of the fastest
    count as equal and the first of them in ``candidates`` order is taken,
    so timing noise does not switch shapes between runs (a shape fixes the
    reduction order, hence the last bits of the results).  The shape is
    fixed for the whole run afterwards, so results stay independent of the
    mix."""
    import time
    import torch
    from .batch import CaseArena
    results = []
    for variant, chunk, block in candidates:
        kw = dict(arena_kw, cg_variant=variant, chunk=chunk, block=block)
        a = CaseArena.for_cases(specs, **kw)
        for _ in range(warm):
            a.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            a.step()
        torch.cuda.synchronize()
        results.append({'cg_variant': variant, 'chunk': chunk, 'block': block,
                        'ms_per_iteration': 1e3 * (time.perf_counter() - t0) / steps})
        del a
        torch.cuda.empty_cache()
    best = pick_with_ties(results, 'ms_per_iteration', tie)
    return {'cg_variant': best['cg_variant'], 'chunk': best['chunk'], 'block': best['block'],
            'candidates': results}


def pick_with_ties(results, key, tie):
    """The first entry of ``results`` whose ``key`` is within ``tie``
    (relative) of the smallest."""
    fastest = min(r[key] for r in results)
    return next(r for r in results if r[key] <= (1.0 + tie) * fastest)
