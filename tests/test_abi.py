"""The C-ABI library loads, exports every function include/wcfd.h declares,
and the ctypes structs match the C layouts (no GPU call is made)."""
import ctypes
import os
import shutil
import subprocess

import pytest

from warpcfd_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing
    assert set(_lib._SIGS) == set(syms)
    assert lib.wcfd_abi_version() == 3
    assert lib.wcfd_error_string(_lib.WCFD_EINVAL) == b'invalid argument'
    assert lib.wcfd_kernel_count() > 10
    assert lib.wcfd_krylov_workspace_bytes(1000, 3) > 8 * 3 * 1000 * 4
    assert lib.wcfd_krylov_workspace_bytes(1000, 2) == 0
    assert lib.wcfd_simple_workspace_bytes(1000, 3000) > lib.wcfd_krylov_workspace_bytes(1000, 3)
    assert lib.wcfd_real_bytes() == 4


def test_double_library_exports_the_same_symbols():
    """libwcfd64.so: the same entry points with double fields; its
    workspaces hold the same vectors at twice the bytes."""
    lib, lib64 = _lib.load(), _lib.load64()
    assert lib64 is not lib
    missing = [s for s in _lib.header_symbols() if not hasattr(lib64, s)]
    assert not missing
    assert lib64.wcfd_real_bytes() == 8
    assert lib64.wcfd_abi_version() == lib.wcfd_abi_version()
    n = 100000
    assert lib64.wcfd_krylov_workspace_bytes(n, 3) > 1.9 * lib.wcfd_krylov_workspace_bytes(n, 3)
    assert _lib.for_dtype('float64') is lib64 and _lib.for_dtype('float32') is lib


def test_invalid_arguments_are_rejected_without_gpu_work():
    lib = _lib.load()
    assert lib.wcfd_spmv(0, None, None, None, 1, None, None, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_cg(10, None, None, None, None, 0, None, None, 1e-5, 1e-5, 10, 10, None, None,
                       None) == _lib.WCFD_EINVAL
    m = _lib.WcfdMesh()
    assert lib.wcfd_green_gauss(ctypes.byref(m), None, None, 15, 1, None) == _lib.WCFD_EINVAL
    # a well-formed descriptor except for its slot count (0, 4, 5 or 6 only): the check
    # comes before any pointer is read, so placeholder addresses are never touched
    fake = 4096
    m = _lib.WcfdMesh(n_cells=8, n_faces=12, n_internal=10, n_boundary=2, nnz=28, n_slots=7)
    for f in ('owner', 'slot_face', 'volume', 'slot_nbr'):
        setattr(m, f, fake)
    assert lib.wcfd_green_gauss(ctypes.byref(m), fake, fake, 15, 1, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_cg_ell(ctypes.byref(m), fake, 0, fake, fake, 1e-5, 1e-5, 10, 10, fake, None,
                           None) == _lib.WCFD_EINVAL
    assert lib.wcfd_bicgstab_ell(ctypes.byref(m), fake, 1, fake, fake, 1e-5, 1e-5, 10, 10, fake,
                                 None, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_spmv_ell(ctypes.byref(m), fake, 1, fake, fake, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_convergence_report(ctypes.byref(m), None, None, None, None, None, fake, None,
                                       fake, fake, None) == _lib.WCFD_EINVAL
    m.n_slots = 6
    # no output; a momentum group without its vectors; rUA without a correction
    assert lib.wcfd_convergence_report(ctypes.byref(m), None, None, None, None, None, fake, None,
                                       fake, None, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_convergence_report(ctypes.byref(m), fake, None, fake, None, None, None, None,
                                       fake, fake, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_convergence_report(ctypes.byref(m), None, None, None, fake, None, None, None,
                                       fake, fake, None) == _lib.WCFD_EINVAL
    assert lib.wcfd_convergence_workspace_bytes() == 9 * 1024 * 8


@pytest.mark.skipif(shutil.which('gcc') is None, reason='gcc missing')
def test_struct_layouts_match_header(tmp_path):
    exe = tmp_path / 'layout'
    inc = os.path.join(os.path.dirname(HERE), 'include')
    subprocess.run(['gcc', '-std=c99', '-I', inc, os.path.join(HERE, 'abi_layout.c'), '-o', str(exe)],
                   check=True)
    out = dict(l.rsplit(' ', 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    out = {k: int(v) for k, v in out.items()}
    S = {'wcfd_mesh': _lib.WcfdMesh, 'wcfd_bc': _lib.WcfdBC, 'wcfd_solve_info': _lib.SolveInfo,
         'wcfd_simple_params': _lib.SimpleParams, 'wcfd_simple_state': _lib.SimpleState,
         'wcfd_simple_info': _lib.SimpleInfo, 'wcfd_batch': _lib.WcfdBatch}
    for name, cls in S.items():
        assert ctypes.sizeof(cls) == out[name], name
    for key, val in out.items():
        if '.' in key:
            sname, field = key.split('.')
            assert getattr(S[sname], field).offset == val, key
