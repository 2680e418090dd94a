"""Continuous-batching service (warpcfd_amd.serve) on the GPU: requests
arrive while others run; each served result is bitwise the result of the
same request solved alone and the one-at-a-time server's result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(chunk=128, block=128, atol=0.0, tol=0.0, rtol0=0.1, maxiter=2000, check_every_u=2)


def _stream():
    from warpcfd_amd.serve import Request
    sizes = [('cavity', 16, 300.0), ('channel', (20, 4, 4), 100.0), ('cavity', 12, 800.0),
             ('cavity', 20, 150.0), ('channel', (15, 3, 3), 60.0), ('cavity', 10, 500.0)]
    return [Request(i, 0.01 * i, k, s, re, seed=7 + i) for i, (k, s, re) in enumerate(sizes)]


def test_served_equals_alone_and_sequential(gpu):
    from warpcfd_amd.batch import CaseArena
    from warpcfd_amd.serve import ContinuousBatchServer, Preparer, SequentialServer, summarize
    import time
    prep = Preparer(workers=2, stop=1e-6, stop_u=5e-5, floor_u=1e-7, max_outer=3000)
    # a small budget: at most a few cases resident, so later requests wait for space
    srv = ContinuousBatchServer(hbm_bytes=3 << 20, max_cases=3, keep_fields=True, **KW)
    t0 = time.perf_counter()
    served = srv.serve(_stream(), prep, clock_start=t0)
    s = summarize(served, time.perf_counter() - t0)
    print('served:', s)
    assert s['requests'] == 6 and s['rejected'] == 0 and s['converged'] == 6
    # each request alone in its own arena: bitwise the served fields
    for r in served:
        a = CaseArena.for_cases([r.spec], **KW)
        a.run(max_steps=4000)
        assert a.iterations(0) == r.iterations, r.describe()
        assert np.array_equal(a.cell_values(0), r.fields), r.describe()
    # one at a time on the plain solver: same fields to rounding
    seq = SequentialServer(keep_fields=True, **KW)
    done = seq.serve(_stream(), prep)
    by_id = {r.rid: r for r in served}
    for r in done:
        assert r.converged
        q = by_id[r.rid]
        assert r.iterations == q.iterations and np.array_equal(r.fields, q.fields), r.describe()
    prep.close()


def test_lanes_equal_one_lane(gpu):
    """Two and three lanes (arenas served by their own threads on their own
    streams, requests to the least loaded lane with space): every request's
    fields and SIMPLE iterations bitwise those of the one-lane server, and
    more than one lane takes requests."""
    from warpcfd_amd.serve import ContinuousBatchServer, Preparer
    prep = Preparer(workers=2, stop=1e-6, stop_u=5e-5, floor_u=1e-7, max_outer=3000)
    one = ContinuousBatchServer(hbm_bytes=3 << 20, max_cases=3, keep_fields=True, **KW)
    ref = {r.rid: r for r in one.serve(_stream(), prep)}
    for lanes in (2, 3):
        print(f'{lanes} lanes', flush=True)
        srv = ContinuousBatchServer(hbm_bytes=lanes * (3 << 20), max_cases=3 * lanes,
                                    keep_fields=True, lanes=lanes, **KW)
        got = srv.serve(_stream(), prep)
        assert len(got) == 6 and all(r.error is None and r.converged for r in got)
        for r in got:
            q = ref[r.rid]
            assert r.iterations == q.iterations, r.describe()
            assert np.array_equal(r.fields, q.fields), r.describe()
        assert len({j for _, j, _, _ in srv.log}) > 1
    prep.close()
