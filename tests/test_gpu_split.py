"""SplitBatch: cases split over several arenas, each stepped by its own host
thread on its own stream, against one arena holding every case.  A case's
arithmetic depends only on its own rows and the workgroup shape, so the
fields, SIMPLE iteration counts and verdicts must be bitwise the same however
the cases are grouped and however the groups' kernels interleave on the GPU.
Covered: cavities of several sizes and Reynolds numbers, a channel, a case
stopped by its iteration cap, two and three groups."""
import numpy as np
import pytest

from test_gpu_batch import KW, cavity_spec, channel_spec

pytestmark = pytest.mark.gpu


def _specs():
    return [cavity_spec(20, 400.0), cavity_spec(12, 800.0), cavity_spec(16, 250.0, max_outer=60),
            channel_spec(), cavity_spec(24, 100.0), cavity_spec(10, 600.0)]


def test_split_batch_bitwise_equals_one_arena(gpu):
    from warpcfd_amd import _lib
    from warpcfd_amd.batch import CaseArena, SplitBatch
    kw = dict(KW, cg_variant=_lib.CG_SINGLE_REDUCTION)
    one = CaseArena.for_cases(_specs(), **kw)
    assert one.run(max_steps=4000, poll_every=4) < 4000
    ref = [(one.iterations(k), one.converged(k), one.cell_values(k)) for k in range(6)]
    assert ref[2][0] == 60 and not ref[2][1]
    assert all(ref[k][1] for k in (0, 1, 3, 4, 5))
    for n_arenas in (2, 3):
        print(f'{n_arenas} arenas', flush=True)
        sb = SplitBatch(_specs(), n_arenas=n_arenas, **kw)
        assert len(sb.arenas) == n_arenas
        steps = sb.run(max_steps=4000, poll_every=4)
        assert max(steps) < 4000
        assert sorted(sb.done_at) == list(range(6))
        for k in range(6):
            it, conv, cv = ref[k]
            assert sb.iterations(k) == it, (n_arenas, k)
            assert sb.converged(k) == conv, (n_arenas, k)
            assert np.array_equal(sb.cell_values(k), cv), (n_arenas, k)
