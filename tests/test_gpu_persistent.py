"""The persistent single-reduction CG (wcfd_set_krylov_persistent) against
the per-iteration kernels.

With the setting on, a CG solve with one row per thread runs as one
cooperative launch, every case on the CUs of one XCD, its workgroups meeting
at a barrier in memory once per iteration.  Its workgroups take the launch
path's virtual workgroups (the same rows, threads, partials and fold order),
so the iterations run, the solution and the reported residual must be
bitwise those of the per-iteration kernels at the same launch shape.
Covered: single cases on one XCD and spread over every XCD (grid mode, a
barrier across XCDs), workgroups of 64 to 512 threads, one to three virtual
workgroups per workgroup, the full 320k-cell channel, prefolded partials (which the persistent kernel folds itself),
convergence, running out of iterations with and without a multiple of
check_every, maxiter 0, a zero right side, the CSR entry point (not eligible:
the per-iteration kernels), batched solves in an arena of several cases, and
whole SIMPLE runs in float32 and float64.  Every run also checks that the
persistent kernel did run and that no launch aborted.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cases import device_case, jittered_box, rel_l2
from test_gpu_inline_check import _system

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore(gpu):
    from warpcfd_amd import _lib
    libs = [_lib.load(), _lib.load64()]
    saved = []
    for lib in libs:
        g, b = ctypes.c_int(), ctypes.c_int()
        assert lib.wcfd_get_krylov_launch(ctypes.byref(g), ctypes.byref(b)) == 0
        saved.append((lib, g.value, b.value, lib.wcfd_get_krylov_fold(),
                      lib.wcfd_get_krylov_persistent()))
    yield libs[0]
    for lib, g, b, f, pe in saved:
        assert lib.wcfd_set_krylov_launch(g, b) == 0
        assert lib.wcfd_set_krylov_fold(f) == 0
        assert lib.wcfd_set_krylov_persistent(pe) == 0


def _stats(lib):
    s, f = ctypes.c_int64(), ctypes.c_int64()
    assert lib.wcfd_krylov_persistent_stats(ctypes.byref(s), ctypes.byref(f)) == 0
    return s.value, f.value


ON = 1 << 30   # wcfd_set_krylov_persistent: every eligible solve


def _cg(lib, model, A, ell, b, persistent, atol=1e-7, tol=1e-7, maxiter=2000, check_every=10,
        csr=False):
    import torch
    from warpcfd_amd._lib import SolveInfo, CG_SINGLE_REDUCTION
    from warpcfd_amd.FV.device import ptr, stream_ptr
    assert lib.wcfd_set_krylov_persistent(ON if persistent else 0) == 0
    L = model.layout
    n = L.num_cells
    dev = ell.device
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 1), dtype=torch.uint8, device=dev)
    x = torch.zeros(n, dtype=torch.float32, device=dev)
    info = SolveInfo()
    if csr:
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        rp, col, val, dp = t(A.indptr.astype(np.int32)), t(A.indices.astype(np.int32)), \
            t(A.data.astype(np.float32)), t(L.diag_pos.astype(np.int32))
        rc = lib.wcfd_cg(n, ptr(rp), ptr(col), ptr(val), ptr(dp), CG_SINGLE_REDUCTION, ptr(b),
                         ptr(x), atol, tol, maxiter, check_every, ptr(ws), ctypes.byref(info),
                         stream_ptr())
    else:
        rc = lib.wcfd_cg_ell(model.dm.desc_ref, ptr(ell), CG_SINGLE_REDUCTION, ptr(b), ptr(x),
                             atol, tol, maxiter, check_every, ptr(ws), ctypes.byref(info),
                             stream_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    return x.cpu().numpy(), (info.iterations, info.converged, info.residual, info.threshold)


def _same(lib, model, A, ell, b, ran=True, **kw):
    """ran: whether the persistent kernel must run (True), must not (False:
    the solve is not eligible), or may (None)."""
    s0, f0 = _stats(lib)
    xa, ia = _cg(lib, model, A, ell, b, 0, **kw)
    xb, ib = _cg(lib, model, A, ell, b, 1, **kw)
    s1, f1 = _stats(lib)
    assert f1 == f0, 'a persistent launch aborted'
    print(f'persistent kernel ran: {s1 > s0}', flush=True)
    if ran is not None:
        assert s1 == s0 + (1 if ran else 0)
    assert ia == ib, (kw, ia, ib)
    assert np.array_equal(xa, xb), kw
    return xb, ib


# Single cases above 8 192 rows spread over every XCD (grid mode): 9 600,
# 46 080 and 81 920 rows at 64 to 512 threads, one or more virtual
# workgroups per workgroup; 1 920 rows (the limit tests below) sit on one XCD.
@pytest.mark.parametrize('shape,grid,block,ran', [
    ((30, 20, 16), 0, 64, True), ((30, 20, 16), 0, 256, True), ((30, 20, 16), 0, 512, True),
    ((30, 20, 16), 40, 512, True), ((40, 36, 32), 0, 256, True),
    ((64, 40, 32), 0, 256, True), ((64, 40, 32), 0, 512, True),
    # grid-stride shapes (chunk > block): 2, 3 and 4 rows per thread
    ((30, 20, 16), 38, 128, True), ((64, 40, 32), 64, 512, True), ((30, 20, 16), 24, 128, True)])
@pytest.mark.parametrize('fold', [2048, 0])
def test_persistent_bitwise_equals_launch_path(restore, shape, grid, block, ran, fold):
    lib = restore
    model, A, ell, b = _system(shape)
    assert lib.wcfd_set_krylov_launch(grid, block) == 0
    assert lib.wcfd_set_krylov_fold(fold) == 0
    print(f'shape {shape} grid {grid} block {block} fold {fold}', flush=True)
    x, (it, conv, res, thr) = _same(lib, model, A, ell, b, ran=ran)
    assert conv and it % 10 == 0 and res <= thr
    xs = spla.spsolve(A.astype(np.float64).tocsc(), b.cpu().numpy().astype(np.float64))
    assert rel_l2(x, xs) <= 1e-5


@pytest.mark.parametrize('maxiter,check_every', [(40, 10), (25, 10), (0, 10), (7, 1), (13, 4)])
def test_persistent_iteration_limits(restore, maxiter, check_every):
    """Solves stopped by maxiter (tolerances out of reach): the same
    iterations, iterate and unconverged report either way."""
    lib = restore
    model, A, ell, b = _system((16, 12, 10), 43)
    assert lib.wcfd_set_krylov_launch(0, 128) == 0
    print(f'maxiter {maxiter} check_every {check_every}', flush=True)
    x, (it, conv, res, thr) = _same(lib, model, A, ell, b, atol=1e-30, tol=1e-30,
                                    maxiter=maxiter, check_every=check_every)
    assert not conv
    ce = check_every
    assert it == (0 if maxiter <= 0 else -(-maxiter // ce) * ce)


def test_persistent_zero_rhs_and_csr(restore):
    """b = 0 stops at the init fold (no case to place: the launch does
    nothing); the CSR entry point is not eligible and takes the
    per-iteration kernels."""
    import torch
    lib = restore
    model, A, ell, b = _system((16, 12, 10), 44)
    assert lib.wcfd_set_krylov_launch(0, 256) == 0
    x, (it, conv, _, _) = _same(lib, model, A, ell, torch.zeros_like(b))
    assert it == 0 and conv and not x.any()
    _same(lib, model, A, ell, b, ran=False, csr=True)


def _cavity(n, re):
    from warpcfd_amd.workloads import cavity_model
    return cavity_model(n, re)


@pytest.mark.parametrize('chunk,block', [(128, 128), (256, 256), (64, 256)])
def test_persistent_batch_bitwise(restore, chunk, block):
    """Eight batched SIMPLE iterations of an arena of four cavities and a
    channel (cases placed on different XCDs, several sharing one): bitwise
    the same fields and Krylov counts as the per-iteration kernels."""
    from warpcfd_amd import _lib
    from warpcfd_amd.batch import CaseArena, CaseSpec
    from test_gpu_batch import channel_spec
    lib = restore
    out = []
    for persistent in (0, 1):
        assert lib.wcfd_set_krylov_persistent(ON if persistent else 0) == 0
        specs = [CaseSpec(_cavity(n, re), stop=0.0, max_outer=100)
                 for n, re in ((24, 100.0), (40, 400.0), (64, 1000.0), (33, 250.0))]
        specs.append(channel_spec(stop=0.0, max_outer=100))
        a = CaseArena.for_cases(specs, chunk=chunk, block=block, atol=1e-9, tol=1e-9,
                                maxiter=2000, cg_variant=_lib.CG_SINGLE_REDUCTION)
        s0, f0 = _stats(lib)
        its = [a.step(info=True) for _ in range(8)]
        s1, f1 = _stats(lib)
        assert f1 == f0
        assert (s1 > s0) == bool(persistent)
        out.append(([a.cell_values(k) for k in range(len(specs))], its))
        del a
    for k in range(len(out[0][0])):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize('float_dtype', [np.float32, np.float64])
def test_persistent_simple_run_bitwise(restore, float_dtype):
    """Twenty SIMPLE iterations of a jittered channel with the reference
    Krylov settings: bitwise the same fields and Krylov counts."""
    from warpcfd_amd import _lib
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = jittered_box((40, 12, 12), 0.1, 5, lengths=(4.0, 1.0, 1.0))
    out = []
    for persistent in (0, 1):
        model = device_case(nodes, cells, 'channel', float_dtype=float_dtype)
        lib = model.lib
        assert lib.wcfd_set_krylov_launch(0, 256) == 0
        assert lib.wcfd_set_krylov_persistent(ON if persistent else 0) == 0
        s0, f0 = _stats(lib)
        sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
        sv.cg_variant = _lib.CG_SINGLE_REDUCTION
        hist = [sv.outer_iteration() for _ in range(20)]
        s1, f1 = _stats(lib)
        assert f1 == f0 and (s1 > s0) == bool(persistent)
        out.append((model.cell_values.cpu().numpy(),
                    [(h['p_iters'], h['u_iters']) for h in hist]))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_persistent_row_threshold(restore):
    """wcfd_set_krylov_persistent(max_rows) takes the persistent kernel for
    solves whose widest case has at most max_rows rows (virtual workgroups
    times chunk: 1 920 rows over 8 workgroups of 256 threads take chunks of
    240, so 8 x 240 = 1 920 here) and the per-iteration kernels above."""
    lib = restore
    model, A, ell, b = _system((16, 12, 10), 45)   # 1 920 rows
    assert lib.wcfd_set_krylov_launch(0, 256) == 0
    ref = _cg(lib, model, A, ell, b, 0)
    for limit, ran in ((1920, True), (1919, False), (0, False)):
        s0, f0 = _stats(lib)
        x, info = _cg(lib, model, A, ell, b, 0)
        assert lib.wcfd_set_krylov_persistent(limit) == 0
        import torch
        from warpcfd_amd._lib import SolveInfo, CG_SINGLE_REDUCTION
        from warpcfd_amd.FV.device import ptr, stream_ptr
        n = model.layout.num_cells
        ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 1), dtype=torch.uint8, device=ell.device)
        xt = torch.zeros(n, dtype=torch.float32, device=ell.device)
        si = SolveInfo()
        assert lib.wcfd_cg_ell(model.dm.desc_ref, ptr(ell), CG_SINGLE_REDUCTION, ptr(b), ptr(xt),
                               1e-7, 1e-7, 2000, 10, ptr(ws), ctypes.byref(si), stream_ptr()) == 0
        torch.cuda.synchronize()
        s1, f1 = _stats(lib)
        assert f1 == f0 and (s1 > s0) == ran, (limit, ran)
        assert np.array_equal(xt.cpu().numpy(), ref[0]) and si.iterations == ref[1][0]


def test_persistent_abort_falls_back(restore):
    """The abort path, forced from the host: the persistent launch gets fewer
    slots per XCD than its plan needs, so every workgroup aborts at entry
    (no barrier is waited on).  x is left as it was, the solve reruns on the
    per-iteration kernels -- bitwise their result -- the fallback counter
    rises and the setting turns itself off."""
    lib = restore
    model, A, ell, b = _system((16, 12, 10), 46)
    assert lib.wcfd_set_krylov_launch(0, 64) == 0
    ref = _cg(lib, model, A, ell, b, 0)
    s0, f0 = _stats(lib)
    assert lib.wcfd_debug_krylov_persistent_shrink(1 << 20) == 0
    try:
        got = _cg(lib, model, A, ell, b, 1)
    finally:
        assert lib.wcfd_debug_krylov_persistent_shrink(0) == 0
    s1, f1 = _stats(lib)
    assert (s1, f1) == (s0, f0 + 1)
    assert lib.wcfd_get_krylov_persistent() == 0
    assert got[1] == ref[1] and np.array_equal(got[0], ref[0])
    # the next solve with the setting on again runs the persistent kernel
    again = _cg(lib, model, A, ell, b, 1)
    assert _stats(lib)[0] == s1 + 1 and np.array_equal(again[0], ref[0])


@pytest.mark.parametrize('grid,block', [(0, 512), (0, 256), (256, 512)])
def test_persistent_channel_320k_bitwise(restore, grid, block):
    """Three SIMPLE iterations of the full 320k-cell channel (reference Krylov
    settings), the pressure CG in grid mode (0 x 512: three virtual
    workgroups of one row per thread per workgroup; 256 x 512, the tuner's
    shape: one virtual workgroup of three rows per thread): bitwise the same
    fields and Krylov counts as the per-iteration kernels."""
    from warpcfd_amd import _lib
    from warpcfd_amd.workloads import channel_model
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    lib = restore
    out = []
    for persistent in (0, 1):
        assert lib.wcfd_set_krylov_launch(grid, block) == 0
        assert lib.wcfd_set_krylov_persistent(ON if persistent else 0) == 0
        model = channel_model()
        model.finalize()
        sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
        sv.cg_variant = _lib.CG_SINGLE_REDUCTION
        s0, f0 = _stats(lib)
        hist = [sv.outer_iteration() for _ in range(3)]
        s1, f1 = _stats(lib)
        assert f1 == f0 and (s1 > s0) == bool(persistent)
        out.append((model.cell_values.cpu().numpy(), [(h['p_iters'], h['u_iters']) for h in hist]))
        del sv, model
    print(f'{grid} x {block}: Krylov counts {out[1][1]}')
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
