"""Continuous-batching service for steady SIMPLE solves.

No counterpart in the reference (one blocking ``IncompressibleSolver.run``
per script).  Requests of uneven mesh sizes arrive over time; each one is
prepared on the host (mesh, device layout, boundary conditions), admitted
into a ``CaseArena`` as soon as a case slot and contiguous arena space are
free (the arena's capacity is set by an HBM budget), iterated together with
every other resident case -- one batched SIMPLE iteration per turn of the
loop -- and released when the device-side monitor reports it converged, so
its space is free for the next request.  Requests that arrive while others
are on the GPU join them at the next iteration boundary.

``SequentialServer`` serves the same stream one case at a time, first come
first served, through the same kernels (an arena holding one case): the
reference point for what continuous batching buys.

Requests are prepared (mesh, renumbered device layout, boundary arrays) in
a pool of worker processes, so host preparation runs in parallel and never
holds the serving loop's interpreter lock.

A request's result does not depend on what else is resident (the arena's
alone-vs-mix property), so both servers return bitwise the same fields.
"""
import concurrent.futures as cf
import math
import multiprocessing as mp
import time

import numpy as np

from .batch import CaseArena, CaseSpec
from .preprocess import DeviceLayout
from .workloads import cavity_model, channel_model


class Request:
    """One solve request: ``kind`` 'cavity' (size = N of an N x N x 1 mesh)
    or 'channel' (size = (nx, ny, nz) hexahedra of the 10 x 2 x 2 box)."""

    def __init__(self, rid, arrival, kind, size, re, seed=0):
        self.rid, self.arrival, self.kind, self.size, self.re, self.seed = \
            rid, arrival, kind, size, re, seed
        self.spec = None
        self.error = None
        self.t_ready = self.t_admit = self.t_done = None
        self.iterations = None
        self.converged = None
        self.fields = None

    @property
    def cells(self):
        if self.kind == 'cavity':
            return self.size * self.size
        return int(np.prod(self.size))

    def describe(self):
        if self.kind == 'cavity':
            return f'cavity {self.size}x{self.size} Re {self.re:.0f}'
        return f'channel {"x".join(map(str, self.size))} Re {self.re:.0f}'


def channel_dims(cells):
    """(nx, ny, nz) of a 10 x 2 x 2 channel with about ``cells`` cubic-ish hexahedra."""
    m = (cells / 5.0) ** (1.0 / 3.0)
    ny = max(2, int(round(m)))
    return (5 * ny, ny, ny)


def poisson_stream(seed=0, rate=20.0, n_requests=32, cavity_n=(32, 512),
                   cavity_re=(100.0, 1000.0), channel_cells=(50_000, 500_000),
                   channel_re=(50.0, 400.0), p_cavity=0.5):
    """Seeded Poisson arrivals (exponential gaps, mean 1/rate) of a mix of
    cavity (N log-uniform) and channel (cell count log-uniform) requests."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        if rng.random() < p_cavity:
            n = int(round(math.exp(rng.uniform(math.log(cavity_n[0]), math.log(cavity_n[1])))))
            out.append(Request(i, t, 'cavity', n, float(rng.uniform(*cavity_re))))
        else:
            c = math.exp(rng.uniform(math.log(channel_cells[0]), math.log(channel_cells[1])))
            out.append(Request(i, t, 'channel', channel_dims(c), float(rng.uniform(*channel_re)),
                               seed=int(rng.integers(1 << 30))))
    return out


def build_model(req, device='cuda'):
    if req.kind == 'cavity':
        return cavity_model(req.size, req.re, device=device)
    return channel_model(req.size, (10.0, 2.0, 2.0), re=req.re, seed=req.seed, device=device)


def prepare_spec(kind, size, re, seed, kw):
    """Host preparation of one request (runs in a worker process): mesh,
    device layout and boundary arrays, as a picklable ``CaseSpec``.  No GPU
    call is made."""
    req = Request(0, 0.0, kind, size, re, seed)
    model = build_model(req, device='cuda')
    L = DeviceLayout(model.mesh, renumber=True)
    return CaseSpec(model, layout=L, name=req.describe(), **kw)


def _warm_worker(hold):
    """Imports what preparation needs; holds the worker briefly so that every
    worker of the pool receives one call."""
    from . import batch, workloads  # noqa: F401
    time.sleep(hold)
    return 0


class Preparer:
    """Prepares requests in ``workers`` processes.  The pool uses the
    forkserver start method, whose server process never touches the GPU;
    create the Preparer before the first GPU call of the program so the
    pool starts from a clean process.  ``workers=0`` prepares inline."""

    def __init__(self, workers=4, stop=1e-6, stop_u=math.inf, floor_u=0.0, max_outer=20000):
        self.kw = dict(stop=stop, stop_u=stop_u, floor_u=floor_u, max_outer=max_outer)
        self.pool = None
        if workers > 0:
            ctx = mp.get_context('forkserver')
            self.pool = cf.ProcessPoolExecutor(max_workers=workers, mp_context=ctx)

    def _finish(self, req, spec):
        req.spec = spec
        req.t_ready = time.perf_counter()
        return req

    def submit(self, req):
        size = req.size if req.kind == 'cavity' else tuple(req.size)
        if self.pool is None:
            f = cf.Future()
            f.set_result(self._finish(req, prepare_spec(req.kind, size, req.re, req.seed,
                                                        self.kw)))
            return f
        inner = self.pool.submit(prepare_spec, req.kind, size, req.re, req.seed, self.kw)
        outer = cf.Future()

        def done(fut):
            try:
                outer.set_result(self._finish(req, fut.result()))
            except BaseException as exc:          # surfaced by outer.result()
                outer.set_exception(exc)
        inner.add_done_callback(done)
        return outer

    def warm(self):
        """Start every worker process (imports done) before the clock starts."""
        if self.pool is not None:
            n = self.pool._max_workers
            list(self.pool.map(_warm_worker, [0.2] * n))

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True)


class ContinuousBatchServer:
    """Serves a request stream with continuous batching.

    ``hbm_bytes`` sets the arena capacity (cells and faces); ``max_cases``
    the number of case slots.  Admission is first come first served, with
    smaller prepared requests allowed to take space the head of the queue
    does not fit in (backfill).  Every other arena keyword (Krylov policy,
    relaxation, chunking) is passed through.

    ``lanes`` > 1 splits the budget and the slots over that many arenas, each
    served by its own host thread on its own stream: a lane admits queued
    requests into its arena, steps it and releases what finished, so one
    lane's long Krylov solves do not hold up the others' iterations, and the
    lanes' kernels overlap on the GPU (as ``batch.SplitBatch``).  A request
    goes to the least loaded lane (resident cells) that has space for it, so
    it must fit one lane's share of the budget (larger ones are rejected).
    Results are bitwise the same in any lane."""

    def __init__(self, hbm_bytes=16 << 30, max_cases=64, keep_fields=False, lanes=1,
                 **arena_kw):
        n = max(1, int(lanes))
        self.arenas = [CaseArena.for_budget(hbm_bytes // n, max(1, max_cases // n), **arena_kw)
                       for _ in range(n)]
        self.arena = self.arenas[0]
        self.keep_fields = keep_fields
        self.log = []
        if n > 1:
            import torch
            self.streams = [torch.cuda.Stream(device=a.device) for a in self.arenas]

    def serve(self, requests, preparer, clock_start=None, max_steps=10_000_000, on_done=None,
              on_step=None, step_every=1000):
        """Serve ``requests`` (``Request`` objects, arrival times in seconds
        after ``clock_start``); returns them in completion order.
        ``on_done(request)`` is called as each one finishes, and
        ``on_step(lane, steps, resident)`` after every ``step_every`` batched
        iterations of a lane (from the lane's thread when there are several
        lanes)."""
        if len(self.arenas) > 1:
            return self._serve_lanes(requests, preparer, clock_start, max_steps, on_done,
                                     on_step, step_every)
        a = self.arena
        t0 = clock_start if clock_start is not None else time.perf_counter()
        pending = sorted(requests, key=lambda r: r.arrival)
        futures = []             # (request, future) in arrival order
        queue = []               # prepared, waiting for space
        resident = {}            # slot -> request
        done = []
        steps = 0
        while (pending or futures or queue or resident) and steps < max_steps:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.pop(0)
                futures.append((r, preparer.submit(r)))
            still = []
            for r, f in futures:
                if f.done():
                    f.result()
                    queue.append(r)
                else:
                    still.append((r, f))
            futures = still
            queue.sort(key=lambda r: r.arrival)
            admitted = []
            for r in queue:
                s = r.spec
                if s.n_cells > a.Cc or s.n_internal > a.Fic or s.n_boundary > a.Fbc:
                    r.error = 'larger than the arena'       # can never be admitted
                    r.t_done = time.perf_counter() - t0
                    admitted.append(r)
                    done.append(r)
                    continue
                if a.fits(s):
                    k = a.admit(r.spec)
                    r.t_admit = time.perf_counter() - t0
                    resident[k] = r
                    admitted.append(r)
            queue = [r for r in queue if r not in admitted]
            if resident:
                a.step()
                steps += 1
                if on_step and steps % step_every == 0:
                    on_step(0, steps, len(resident))
                for k in a.poll():
                    r = resident.pop(k)
                    r.t_done = time.perf_counter() - t0
                    r.iterations = a.iterations(k)
                    r.converged = a.converged(k)
                    if self.keep_fields:
                        r.fields = a.cell_values(k)
                    a.release(k)
                    done.append(r)
                    if on_done:
                        on_done(r)
                    self.log.append((r.t_done, 0, len(resident), a.cells.largest()))
            else:
                # idle: wait for the next arrival or preparation
                nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.0
                time.sleep(min(max(nxt, 0.0), 0.01) if nxt > 0 else 0.001)
        return done

    def _finish(self, a, k, r, t0):
        r.t_done = time.perf_counter() - t0
        r.iterations = a.iterations(k)
        r.converged = a.converged(k)
        if self.keep_fields:
            r.fields = a.cell_values(k)
        a.release(k)

    def _serve_lanes(self, requests, preparer, clock_start, max_steps, on_done, on_step,
                     step_every):
        """serve() over several lanes.  The calling thread takes arrivals and
        moves prepared requests to the shared queue; each lane's thread, on
        its own stream, admits from that queue (under the lock, as admission
        rewrites the arena's descriptors, which only its own stream may
        touch), steps and polls its arena, and releases finished cases.  A
        lane leaves a request to a less loaded lane unless that lane has
        found no space for it."""
        import threading
        import torch
        t0 = clock_start if clock_start is not None else time.perf_counter()
        pending = sorted(requests, key=lambda r: r.arrival)
        futures = []
        lock = threading.Lock()
        queue = []               # prepared, waiting for space (arrival order)
        done = []
        L = len(self.arenas)
        load = [0] * L           # resident cells per lane
        full = {}                # request id -> lanes that found no space for it
        closing = threading.Event()
        abort = threading.Event()        # a lane failed: every lane stops
        errors = []

        def lane(j):
            a = self.arenas[j]
            resident = {}
            steps = 0
            try:
                with torch.cuda.stream(self.streams[j]):
                    while steps < max_steps and not abort.is_set():
                        with lock:
                            for r in list(queue):
                                if not a.fits(r.spec):
                                    full.setdefault(r.rid, set()).add(j)
                                    continue
                                skip = full.get(r.rid, ())
                                if any(load[m] < load[j] and m not in skip
                                       for m in range(L) if m != j):
                                    continue          # a less loaded lane may take it
                                k = a.admit(r.spec)
                                r.t_admit = time.perf_counter() - t0
                                resident[k] = r
                                load[j] += r.spec.n_cells
                                queue.remove(r)
                            idle = not resident
                            last = closing.is_set() and not queue
                        if idle:
                            if last:
                                break
                            time.sleep(0.0005)
                            continue
                        a.step()
                        steps += 1
                        if on_step and steps % step_every == 0:
                            on_step(j, steps, len(resident))
                        for k in a.poll():
                            r = resident.pop(k)
                            self._finish(a, k, r, t0)
                            with lock:
                                load[j] -= r.spec.n_cells
                                done.append(r)
                                self.log.append((r.t_done, j, len(resident), a.cells.largest()))
                            if on_done:
                                on_done(r)
                    self.streams[j].synchronize()
            except BaseException as e:    # re-raised on the calling thread
                errors.append(e)
                abort.set()
                closing.set()

        torch.cuda.synchronize()          # the arenas were built on the default stream
        threads = [threading.Thread(target=lane, args=(j,)) for j in range(L)]
        for t in threads:
            t.start()
        try:
            while (pending or futures) and not errors:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.pop(0)
                    futures.append((r, preparer.submit(r)))
                still = []
                for r, f in futures:
                    if not f.done():
                        still.append((r, f))
                        continue
                    f.result()
                    s = r.spec
                    if all(s.n_cells > a.Cc or s.n_internal > a.Fic or s.n_boundary > a.Fbc
                           for a in self.arenas):
                        r.error = 'larger than the arena'       # can never be admitted
                        r.t_done = time.perf_counter() - t0
                        with lock:
                            done.append(r)
                        continue
                    with lock:
                        queue.append(r)
                        queue.sort(key=lambda q: q.arrival)
                futures = still
                nxt = pending[0].arrival - (time.perf_counter() - t0) if pending else 0.0
                time.sleep(min(max(nxt, 0.0005), 0.002))
        finally:
            closing.set()
            for t in threads:
                t.join()
        if errors:
            raise errors[0]
        return done


class SequentialServer:
    """The same stream, one case at a time (first come first served), each
    alone in an arena sized for it, with the same kernels, Krylov policy and
    stopping rule (``arena_kw``) as the continuous server."""

    def __init__(self, keep_fields=False, **arena_kw):
        self.keep_fields = keep_fields
        self.arena_kw = arena_kw

    def serve(self, requests, preparer, clock_start=None, on_done=None):
        t0 = clock_start if clock_start is not None else time.perf_counter()
        order = sorted(requests, key=lambda r: r.arrival)
        futures = [(r, preparer.submit(r)) for r in order]
        done = []
        for r, f in futures:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            f.result()
            a = CaseArena.for_cases([r.spec], **self.arena_kw)
            r.t_admit = time.perf_counter() - t0
            a.run(max_steps=r.spec.max_outer + 1)
            r.t_done = time.perf_counter() - t0
            r.iterations, r.converged = a.iterations(0), a.converged(0)
            if self.keep_fields:
                r.fields = a.cell_values(0)
            done.append(r)
            del a
            if on_done:
                on_done(r)
        return done


def summarize(done, wall_s):
    ok = [r for r in done if r.error is None]
    lat = np.array([r.t_done - r.arrival for r in ok])
    gpu = np.array([r.t_done - r.t_admit for r in ok])
    return {'requests': len(done), 'rejected': len(done) - len(ok),
            'converged': int(sum(bool(r.converged) for r in ok)),
            'wall_s': round(wall_s, 3), 'solves_per_s': round(len(done) / wall_s, 4),
            'time_to_converged_s': {'p50': round(float(np.percentile(lat, 50)), 3),
                                    'p99': round(float(np.percentile(lat, 99)), 3),
                                    'mean': round(float(lat.mean()), 3)},
            'resident_time_s': {'p50': round(float(np.percentile(gpu, 50)), 3),
                                'p99': round(float(np.percentile(gpu, 99)), 3)},
            'cells_total': int(sum(r.cells for r in done))}
