"""wcfd_batch_convergence_report_noalloc (static scratch, no allocation)
against wcfd_batch_convergence_report (caller workspace), through the C ABI:
bitwise equal on two calls in a row, the second on a smaller arena, and an
arena wider than the static scratch is refused before any GPU work."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _args(arena):
    from warpcfd_amd.FV.device import ptr
    return (ctypes.byref(arena.desc), ctypes.byref(arena.batch), ptr(arena.mom_values),
            ptr(arena.cv), ptr(arena.mom_rhs), ptr(arena.rua), ptr(arena.cg[9:12]),
            ptr(arena.p_x), ptr(arena.flux))


def _both(arena):
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.FV.device import ptr, stream_ptr
    lib = _lib.load()
    ws = torch.zeros(lib.wcfd_batch_convergence_workspace_bytes(arena.n_blocks), dtype=torch.uint8,
                     device='cuda')
    ref = torch.full((arena.K, 9), -1.0, dtype=torch.float64, device='cuda')
    got = torch.full((arena.K, 9), -2.0, dtype=torch.float64, device='cuda')
    a = _args(arena)
    _lib.check(lib.wcfd_batch_convergence_report(*a, ptr(ws), ptr(ref), stream_ptr()), 'report')
    _lib.check(lib.wcfd_batch_convergence_report_noalloc(*a, ptr(got), stream_ptr()),
               'report_noalloc')
    torch.cuda.synchronize()
    return ref.cpu(), got.cpu()


def test_noalloc_report_equals_workspace_report(gpu):
    import torch
    from test_gpu_batch import KW, cavity_spec, channel_spec
    from warpcfd_amd.batch import CaseArena
    big = CaseArena.for_cases([cavity_spec(16, 400.0), channel_spec(), cavity_spec(12, 800.0)],
                              **KW)
    small = CaseArena.for_cases([cavity_spec(8, 100.0)], **KW)
    assert small.n_blocks < big.n_blocks
    for _ in range(6):
        big.step()
        small.step()
    for arena in (big, small):              # the second call has the smaller input
        print(f'noalloc report: {arena.K} cases, {arena.n_blocks} workgroups', flush=True)
        ref, got = _both(arena)
        assert torch.isfinite(ref).all() and (ref[:, 3] > 0).all()
        assert torch.equal(ref, got), (ref, got)


def test_noalloc_report_rejects_wider_arena(gpu):
    from test_gpu_batch import KW, cavity_spec
    from warpcfd_amd import _lib
    from warpcfd_amd.batch import CaseArena
    from warpcfd_amd.FV.device import ptr, stream_ptr
    import torch
    arena = CaseArena.for_cases([cavity_spec(8, 100.0)], **KW)
    wide = _lib.WcfdBatch.from_buffer_copy(arena.batch)
    wide.n_blocks = _lib.NOALLOC_MAX_BLOCKS + 1       # checked on the host, nothing launched
    out = torch.zeros(9, dtype=torch.float64, device='cuda')
    a = list(_args(arena))
    a[1] = ctypes.byref(wide)
    code = _lib.load().wcfd_batch_convergence_report_noalloc(*a, ptr(out), stream_ptr())
    assert code == _lib.WCFD_ESCRATCH
    assert b'static scratch' in _lib.load().wcfd_error_string(code)
