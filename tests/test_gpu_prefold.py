"""Prefolded Krylov partials (wcfd_set_krylov_fold) on the GPU.

Above the inline-fold limit a fold kernel reduces each case's workgroup
partials to one value before the next kernel reads them.  It sums in the
order and workgroup size of the inline fold, so a solve gives bitwise the
same iterate and iteration count either way.  Checked for both CG forms and
for the 3-vector BiCGSTAB:
* at one row per thread (grid 0), where rows are prefetched before the fold;
* on a short grid that strides over rows;
* for a batch arena running SIMPLE iterations.
The prefolded solves also match scipy's direct solve.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cases import device_case, jittered_box, rel_l2
from test_gpu_parity import _csr_to_ell, _spd_system

pytestmark = pytest.mark.gpu

INLINE = 2048     # the library's default limit
ALWAYS = 0        # prefold at any grid


@pytest.fixture
def restore_launch(gpu):
    from warpcfd_amd import _lib
    lib = _lib.load()
    g, b = ctypes.c_int(), ctypes.c_int()
    assert lib.wcfd_get_krylov_launch(ctypes.byref(g), ctypes.byref(b)) == 0
    f = lib.wcfd_get_krylov_fold()
    yield lib
    assert lib.wcfd_set_krylov_launch(g.value, b.value) == 0
    assert lib.wcfd_set_krylov_fold(f) == 0


def _solve(lib, model, ell, negell, b1, b3, grid, block, fold):
    import torch
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    assert lib.wcfd_set_krylov_launch(grid, block) == 0
    assert lib.wcfd_set_krylov_fold(fold) == 0
    n = model.layout.num_cells
    dev = ell.device
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 3), dtype=torch.uint8, device=dev)
    out = {}
    for variant in (0, 1):
        x = torch.zeros(n, dtype=torch.float32, device=dev)
        info = SolveInfo()
        assert lib.wcfd_cg_ell(model.dm.desc_ref, ptr(ell), variant, ptr(b1), ptr(x), 1e-7, 1e-7,
                               2000, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
        out[f'cg{variant}'] = (x.cpu().numpy(), info.iterations, info.converged)
    x3 = torch.zeros(3, n, dtype=torch.float32, device=dev)
    info = SolveInfo()
    assert lib.wcfd_bicgstab_ell(model.dm.desc_ref, ptr(negell), 3, ptr(b3), ptr(x3), 1e-7, 1e-7,
                                 2000, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
    out['bicg'] = (x3.cpu().numpy(), info.iterations, info.converged)
    return out


# (0, b): one row per thread; (38, 128): chunks of 253 rows, two per thread;
# (24, 128): the grid-stride form
@pytest.mark.parametrize('grid,block', [(0, 64), (0, 256), (38, 128), (24, 128)])
def test_prefold_bitwise_equals_inline(restore_launch, grid, block):
    import torch
    lib = restore_launch
    nodes, cells = jittered_box((30, 20, 16), 0.1, 41)
    model = device_case(nodes, cells, 'channel')
    L = model.layout
    A = _spd_system(model, 9)
    n = L.num_cells
    dev = model._cv.device
    ell = torch.from_numpy(_csr_to_ell(L, A.data)).to(dev)
    negell = -ell
    rng = np.random.default_rng(10)
    b1 = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).to(dev)
    b3 = torch.from_numpy(rng.standard_normal((3, n)).astype(np.float32)).to(dev)
    a = _solve(lib, model, ell, negell, b1, b3, grid, block, INLINE)
    p = _solve(lib, model, ell, negell, b1, b3, grid, block, ALWAYS)
    for key in a:
        xa, ia, ca = a[key]
        xp, ip, cp = p[key]
        assert ca and cp, key
        assert ia == ip, (key, ia, ip)
        assert np.array_equal(xa, xp), key
    Ad = A.astype(np.float64).tocsc()
    xs = spla.spsolve(Ad, b1.cpu().numpy().astype(np.float64))
    for v in ('cg0', 'cg1'):
        assert rel_l2(p[v][0], xs) <= 1e-5, v
    b3h = b3.cpu().numpy().astype(np.float64)
    xs3 = np.stack([spla.spsolve(-Ad, b3h[q]) for q in range(3)])
    assert rel_l2(p['bicg'][0], xs3) <= 1e-5


def test_prefold_batch_simple_bitwise(restore_launch):
    """Twelve SIMPLE iterations of a two-case arena, partials folded inline
    and prefolded: bitwise equal fields."""
    from test_gpu_batch import cavity_spec, channel_spec
    from warpcfd_amd.batch import CaseArena
    lib = restore_launch
    fields = []
    for fold in (INLINE, ALWAYS):
        assert lib.wcfd_set_krylov_fold(fold) == 0
        a = CaseArena.for_cases([cavity_spec(20, 400.0), channel_spec()], chunk=128, block=128,
                                atol=1e-9, tol=1e-9, maxiter=2000)
        for _ in range(12):
            a.step()
        fields.append([a.cell_values(k) for k in range(2)])
    for k in range(2):
        assert np.array_equal(fields[0][k], fields[1][k]), k
