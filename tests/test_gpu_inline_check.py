"""The stop test made inside the first CG kernel of each check group
(wcfd_set_krylov_inline_check) against the separate check kernel.

A single-case single-reduction CG solve either launches a check kernel
after every `check_every` iterations or lets the next iteration kernel fold
||r||^2 with its own partials and stop the case there.  Both fold the same
partials in the same order, so the iterations run, the solution and the
reported residual must be bitwise the same.  Covered: one and two rows per
thread, the grid-stride form and the CSR matrix; inline and prefolded
partials; convergence, running out of iterations (maxiter a multiple of
check_every or not), maxiter 0, a zero right side and check_every 1; a whole
SIMPLE run of the channel in float32 and in float64.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cases import device_case, jittered_box, rel_l2
from test_gpu_parity import _csr_to_ell, _spd_system

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
                      lib.wcfd_get_krylov_inline_check()))
    yield libs[0]
    for lib, g, b, f, ic in saved:
        assert lib.wcfd_set_krylov_launch(g, b) == 0
        assert lib.wcfd_set_krylov_fold(f) == 0
        assert lib.wcfd_set_krylov_inline_check(ic) == 0


def _system(shape=(30, 20, 16), seed=41):
    import torch
    nodes, cells = jittered_box(shape, 0.1, seed)
    model = device_case(nodes, cells, 'channel')
    L = model.layout
    A = _spd_system(model, 9)
    dev = model._cv.device
    ell = torch.from_numpy(_csr_to_ell(L, A.data)).to(dev)
    rng = np.random.default_rng(10)
    b = torch.from_numpy(rng.standard_normal(L.num_cells).astype(np.float32)).to(dev)
    return model, A, ell, b


def _cg(lib, model, A, ell, b, inline, atol=1e-7, tol=1e-7, maxiter=2000, check_every=10,
        csr=False):
    import torch
    from warpcfd_amd._lib import SolveInfo, CG_SINGLE_REDUCTION
    from warpcfd_amd.FV.device import ptr, stream_ptr
    assert lib.wcfd_set_krylov_inline_check(inline) == 0
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


def _same(lib, model, A, ell, b, **kw):
    xa, ia = _cg(lib, model, A, ell, b, 0, **kw)
    xb, ib = _cg(lib, model, A, ell, b, 1, **kw)
    assert ia == ib, (kw, ia, ib)
    assert np.array_equal(xa, xb), kw
    return xb, ib


# (0, b): one row per thread; (38, 128): two rows per thread; (24, 128): grid-stride
@pytest.mark.parametrize('grid,block', [(0, 64), (0, 256), (38, 128), (24, 128)])
@pytest.mark.parametrize('fold', [2048, 0])
def test_inline_check_bitwise_equals_check_kernel(restore, grid, block, fold):
    lib = restore
    model, A, ell, b = _system()
    assert lib.wcfd_set_krylov_launch(grid, block) == 0
    assert lib.wcfd_set_krylov_fold(fold) == 0
    print(f'grid {grid} block {block} fold {fold}', flush=True)
    x, (it, conv, res, thr) = _same(lib, model, A, ell, b)
    assert conv and it % 10 == 0 and res <= thr
    xs = spla.spsolve(A.astype(np.float64).tocsc(), b.cpu().numpy().astype(np.float64))
    assert rel_l2(x, xs) <= 1e-5


@pytest.mark.parametrize('maxiter,check_every', [(40, 10), (25, 10), (0, 10), (7, 1), (13, 4)])
def test_inline_check_iteration_limits(restore, maxiter, check_every):
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


def test_inline_check_zero_rhs_and_csr(restore):
    """b = 0 stops at the init fold; the CSR entry point takes the same path."""
    import torch
    lib = restore
    model, A, ell, b = _system((16, 12, 10), 44)
    assert lib.wcfd_set_krylov_launch(0, 256) == 0
    x, (it, conv, _, _) = _same(lib, model, A, ell, torch.zeros_like(b))
    assert it == 0 and conv and not x.any()
    _same(lib, model, A, ell, b, csr=True)


@pytest.mark.parametrize('float_dtype', [np.float32, np.float64])
def test_inline_check_simple_run_bitwise(restore, float_dtype):
    """Twenty SIMPLE iterations of a jittered channel with the reference
    Krylov settings: bitwise the same fields and Krylov counts."""
    from warpcfd_amd import _lib
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = jittered_box((40, 12, 12), 0.1, 5, lengths=(4.0, 1.0, 1.0))
    out = []
    for inline in (0, 1):
        model = device_case(nodes, cells, 'channel', float_dtype=float_dtype)
        lib = model.lib
        assert lib.wcfd_set_krylov_launch(0, 256) == 0
        assert lib.wcfd_set_krylov_inline_check(inline) == 0
        sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
        sv.cg_variant = _lib.CG_SINGLE_REDUCTION
        hist = [sv.outer_iteration() for _ in range(20)]
        out.append((model.cell_values.cpu().numpy(),
                    [(h['p_iters'], h['u_iters']) for h in hist]))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
