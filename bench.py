#!/usr/bin/env python3
"""Benchmark: steady SIMPLE iterations per second on one MI355X.

Workload (default): 3D channel, box 10 x 2 x 2, 200 x 40 x 40 hexahedra
(320 000 cells), interior vertices jittered +-10% h (seed 0), uniform inlet
u = 1 (-X), p = 0 outlet (+X), no-slip walls, nu = 0.01 (Re = 200 on the
height), upwind convection, linear + skew face interpolation,
alpha_u = 0.7 / alpha_p = 0.3, the reference's Krylov settings (BiCGSTAB for
momentum, CG for pressure, Jacobi, atol 1e-5, maxiter 500, check every 10).
One step = one SIMPLE outer iteration.  Inputs are resident in HBM before
the timed region.

Prints ONE JSON line.  Besides the contract fields it carries
``roofline`` (the dominant kernel's algorithmic bytes per launch over its
average launch time, from HIP events the runtime stamps with each kernel's
begin and end during a profiled pass of the same number of steps) and
``cpu_baseline``
(the CPU oracle, single thread, on a bounded sample of the same workload).

    python bench.py [--steps K] [--warmup W] [--workload channel|cavity|batch64]

``--workload batch64`` solves the 64-cavity batch (one step = the whole
batch to continuity <= 1e-6) split over four arenas stepped side by side
(``--arenas``) and, for comparison, the same cases one after another
through the plain solver; its value is solves/s.
"""
import argparse
import json
import re
import os
import sys
import time

# Hardware queues HIP gives this process (read when HIP starts, so before
# any import that starts it).  The batch and the served stream step four
# arenas on four streams; with HIP's default of four queues the default
# stream takes one and two of those streams can land on one queue, which
# serialises them (the 64-cavity batch: 39.0 s against 28.1 s with 8
# queues, measured on MI355X).  Six or more busy queues measured slower, so
# the arena count stays at four.  A larger value set by the caller is kept.
try:
    _queues = int(os.environ.get('GPU_MAX_HW_QUEUES', '4'))
except ValueError:
    _queues = 4
if _queues < 8:
    os.environ['GPU_MAX_HW_QUEUES'] = '8'
del _queues

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    'channel': dict(desc='channel 10x2x2, 200x40x40 hex (320000 cells), interior vertices jittered '
                         '+-10% h (seed 0), inlet u=1, outlet p=0, no-slip walls, Re=200 (nu=0.01), '
                         'steady SIMPLE, upwind (implicit convective outflow), alpha_u 0.7 / alpha_p 0.3, '
                         'reference Krylov settings', kind='channel'),
    'cavity': dict(desc='lid-driven cavity 64x64x1 hex (4096 cells), Re=100, steady SIMPLE, upwind, '
                        'alpha_u 0.7 / alpha_p 0.3, reference Krylov settings', kind='cavity'),
}


def build_case(workload, dims=(200, 40, 40), cell_type='hexahedron'):
    """The workload's nodes and cells; ``cell_type`` 'tetra' or 'wedge'
    splits every hexahedron into 6 Kuhn tetrahedra or 2 prisms."""
    from warpcfd_amd.preprocess import channel_mesh, cavity_mesh
    from warpcfd_amd.preprocess.grid import split_hexes
    if workload == 'channel':
        nodes, cells = channel_mesh(*dims, (10.0, 2.0, 2.0), jitter=0.1, seed=0)
    else:
        nodes, cells = cavity_mesh(64)
    if cell_type != 'hexahedron':
        cells = split_hexes(cells, cell_type)
    return nodes, cells


def apply_bcs(b, kind):
    if kind == 'channel':
        b.velocity_BC('-X', u=1.0, v=0.0, w=0.0)
        b.pressure_BC('+X', p=0.0)
        for w in ('-Y', '+Y', '-Z', '+Z'):
            b.no_slip_wall(w)
    else:
        for w in ('-X', '+X', '-Y'):
            b.no_slip_wall(w)
        b.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
        b.slip_wall('+Z'); b.slip_wall('-Z')


def fp64_error(fields32, run64):
    """rel-L2 of the float32 fields against the double build's, per field
    and for the velocity vector; ``run64()`` returns the double build's
    (C, 4) fields after the same iterations."""
    import numpy as np
    a = np.asarray(fields32, np.float64)
    b = np.asarray(run64(), np.float64)
    rel = lambda x, y: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    out = {name: rel(a[:, q], b[:, q]) for q, name in enumerate('uvwp')}
    out['velocity'] = rel(a[:, :3], b[:, :3])
    return out


def calibration_copy(mb):
    """One elementwise kernel of known traffic (reads and writes ``mb`` MiB)
    for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    n = int(mb * (1 << 20) // 4)
    x = torch.ones(n, dtype=torch.float32, device='cuda')
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    del x, y


def algorithmic_bytes(kernel, C, F, Fi, Fb, nnz, slots=6):
    """Compulsory bytes of one launch: every input element read once, every
    output element written once (float32 / int32 = 4 B).  ``slots``: matrix
    slots the Krylov kernels read per row (6 hexahedra, 5 wedges, 4
    tetrahedra); a row is the diagonal and the slots' coefficients (4 B
    each) plus the slots' int16 neighbour offsets.  The face and cell
    kernels read all six slot tables whatever the cell type."""
    rp = 4 * (C + 1)
    mat = 4 * (slots + 1) + 2 * slots          # 40 B for six slots
    table = {
        # slot-major (ELL) matrix: diagonal + 6 coefficients (28 B) + int16
        # neighbour offsets (12 B) per row; q = A (z + beta p) gathered on the
        # fly; reads r, inv, p_in (12 B), writes p_out, q (8 B)
        'cg_1': mat * C + 12 * C + 8 * C,
        # x += a p, r -= a q: reads x p r q inv, writes x r
        'cg_2': 28 * C,
        # fused single-reduction CG iteration: ELL matrix 40 B, gathers of
        # r, w, s, inv (16 B), own p, x (8 B), writes p, s, x, r, w (20 B)
        'cg_1_single': mat * C + 16 * C + 8 * C + 20 * C,
        'bicgstab_1': mat * C + 4 * C + 36 * C + 24 * C + 12 * C,
        'bicgstab_2': mat * C + 4 * C + 24 * C + 24 * C,
        'bicgstab_3': 64 * C + 24 * C,
        # owner/neigh, psi, skew, normal, area, dmag, bc; cell values + grads; face values + flux
        'face_interp': 4 * F + 4 * Fi + 4 * Fi + 12 * Fi + 12 * F + 4 * F + 4 * Fb + 20 * Fb
                       + 16 * C + 48 * C + 16 * F + 4 * F,
        # slots, sides, sn, volume; face values 4 fields; grads out; matrix + rhs; flux; bc
        'assemble_momentum': 24 * C + 6 * C + 72 * C + 4 * C + 16 * F + 48 * C + 4 * nnz + 12 * C
                             + 4 * F + 24 * C + 4 * C + 20 * Fb + 16 * C,
        'assemble_pressure': 24 * C + 6 * C + 72 * C + 4 * C + 4 * F + 12 * C + 4 * nnz + 4 * C
                             + 8 * F + 24 * C + 4 * C + 8 * Fb,
        # internal: owner, neigh, psi, skew, fv(p) out, flux, rua_f, area, dmag, fc out;
        # boundary: owner, bc type + value, dmag, fv out, flux, rua_f, area, fc out;
        # p and grad p per cell
        'face_p_fc': 48 * Fi + 33 * Fb + 16 * C,
        # slot faces, sides, normals, volume; p face values and corrected fluxes
        # (each face once); gradients out; rUA; HbyA in/out; u out; previous u
        'grad_p_correct': 24 * C + 6 * C + 72 * C + 4 * C + 8 * F + 48 * C + 4 * C + 24 * C
                          + 12 * C + 12 * C,
        # volume, diagonal, rUA out, u in / previous u out, HbyA in/out, grad p, u out
        'rua_hbya': 84 * C,
    }
    return table.get(kernel)


def host_info():
    """The host the CPU baselines ran on: usable cores and CPU model."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return {'nproc': cores, 'cpu_model': model}


def _oracle_solver(nodes, cells, kind, cg_variant=1):
    from oracle.mesh import OracleMesh
    from oracle.fv import OracleFVM
    from oracle.simple import OracleSolver
    om = OracleMesh(nodes, cells)
    om.define_boundary_walls()
    of = OracleFVM(om)
    if kind == 'cavity':
        of.set_reference_pressure(0, 0.0)
    apply_bcs(of.boundary, kind)
    of.boundary.finalize()
    return OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=(kind == 'channel'),
                        cg_variant=cg_variant)


def _cpu_run(nodes, cells, kind, warmup, budget_s, max_iters, cg_variant=1):
    """The oracle on one thread: ``warmup`` untimed SIMPLE iterations (the GPU
    run's warm-up), then timed iterations until ``max_iters`` or the budget.
    Returns (iterations, seconds, momentum and pressure Krylov iterations per
    timed step)."""
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(limits=1)
    except Exception:
        pass
    s = _oracle_solver(nodes, cells, kind, cg_variant)
    for _ in range(warmup):
        s.outer_iteration()
    t0 = time.perf_counter()
    n = 0
    while n < max_iters:
        s.outer_iteration()
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    hist = s.history[-n:]
    u_it = sum(h['u_iters'] for h in hist) / n
    p_it = sum(sum(h['p_iters']) for h in hist) / n
    return n, dt, u_it, p_it


def cpu_baseline(nodes, cells, kind, warmup=3, budget_s=15.0, max_iters=20, gpu_window=None,
                 all_cores=True, cg_variant=1, all_cores_budget_s=10.0):
    """The CPU oracle (numpy/scipy restatement of the reference path) on the
    same workload: one thread over the GPU run's iteration window (the same
    untimed warm-up, then as many of its timed iterations as the budget
    allows), and -- the reference's Warp-CPU path and this restatement both run
    one case on one thread -- all host cores as independent copies of the case
    in one process each, whose summed iteration rate is the host's
    throughput."""
    n, dt, u_it, p_it = _cpu_run(nodes, cells, kind, warmup, budget_s, max_iters, cg_variant)
    host = host_info()
    out = {'value': round(n / dt, 5), 'unit': 'iter/s', 'cores': 1, 'kind': 'port',
           'sample': f'SIMPLE iterations {warmup + 1}-{warmup + n} (after {warmup} untimed, as the '
                     f'GPU run) of the same workload by the oracle (numpy/scipy restatement of the '
                     f'reference path), 1 thread, {dt:.1f} s',
           'krylov_iters_per_step': {'momentum': round(u_it, 1), 'pressure': round(p_it, 1)},
           'host': host}
    if gpu_window:
        out['gpu_window'] = gpu_window
    if all_cores and host['nproc'] > 1:
        import concurrent.futures as cf
        import multiprocessing as mp
        P = min(host['nproc'], 16)
        t0 = time.perf_counter()
        with cf.ProcessPoolExecutor(max_workers=P, mp_context=mp.get_context('spawn')) as ex:
            futs = [ex.submit(_cpu_run, nodes, cells, kind, warmup, all_cores_budget_s,
                              max_iters, cg_variant) for _ in range(P)]
            res = [f.result() for f in futs]
        wall = time.perf_counter() - t0
        rate = sum(r[0] / r[1] for r in res)
        out['all_cores'] = {
            'value': round(rate, 5), 'unit': 'iter/s', 'cores': P, 'kind': 'port',
            'sample': f'{P} independent copies of the same case, one process and one thread each, '
                      f'{warmup} untimed then up to {max_iters} timed SIMPLE iterations or '
                      f'{all_cores_budget_s:g} s each ({wall:.1f} s wall); value = the sum of '
                      f'their rates',
            'per_process_iter_s': [round(r[0] / r[1], 5) for r in res]}
    return out


def bench_batch64(args):
    """The 64-cavity batch: every case to continuity <= 1e-6, batched in one
    arena, and the same 64 cases one after another through the plain
    single-case solver.  One step = one solve of the whole batch."""
    print(json.dumps(run_batch64(args, sequential=not args.no_sequential,
                                 profile=not args.no_profile, cpu=not args.no_cpu_baseline,
                                 fp64_check=not args.no_fp64)))


def run_batch64(args, sequential=True, profile=True, cpu=True, fp64_check=False):
    """bench_batch64's measurement; returns the result dict."""
    import numpy as np
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.batch import CaseArena, CaseSpec
    from warpcfd_amd.workloads import batch64_draws, cavity_model
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver

    stop = 1e-6
    stop_u, floor_u = args.stop_u, 1e-7
    pol = dict(atol=0.0, tol=0.0, rtol0=args.rtol0, maxiter=2000, check_every=10,
               check_every_u=args.check_u, maxiter_u=args.maxiter_u,
               cg_variant=_lib.CG_SINGLE_REDUCTION)
    draws = batch64_draws(args.seed)[:args.cases]
    t = time.perf_counter()
    specs = [CaseSpec(cavity_model(n, re), stop=stop, max_outer=args.max_outer,
                      name=f'cavity{n}_re{re:.0f}', stop_u=stop_u, floor_u=floor_u)
             for n, re in draws]
    prep_s = time.perf_counter() - t
    cells = sum(s.n_cells for s in specs)
    print(f'# batch: {len(specs)} cases, {cells} cells, host preprocessing {prep_s:.1f} s',
          file=sys.stderr, flush=True)
    kw = dict(chunk=args.chunk, block=args.block, **pol)
    tuned = None
    if not args.no_tune:
        # the arena's launch shape and CG form, timed on the wide phase of this batch
        from warpcfd_amd.tune import autotune_arena
        tuned = autotune_arena(specs, dict(kw))
        kw.update(cg_variant=tuned['cg_variant'], chunk=tuned['chunk'], block=tuned['block'])
        print(f'# autotuned arena: {kw}', file=sys.stderr, flush=True)
    # warm-up: load every kernel once
    warm = CaseArena.for_cases(specs[:2], **kw)
    for _ in range(3):
        warm.step()
    torch.cuda.synchronize()
    del warm

    split = args.arenas > 1 and not args.graph
    if split:
        # the cases over args.arenas arenas, each stepped by its own host
        # thread on its own stream (bitwise the results of one arena)
        from warpcfd_amd.batch import SplitBatch
        arena = SplitBatch(specs, n_arenas=args.arenas, **kw)
    else:
        arena = CaseArena.for_cases(specs, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done_at = {}
    steps = 0
    cur = None
    n_capture = 0
    if split:
        per_arena = arena.run(max_steps=args.max_outer + 1, poll_every=args.poll_every)
        steps = max(per_arena)
        done_at = dict(arena.done_at)
    while not split and len(done_at) < len(specs) and steps < args.max_outer + 1:
        if args.graph:
            # HIP-graph replay with a Krylov launch budget measured every 100 iterations
            stale = cur is not None and arena._graph['blocks'] != arena.n_blocks
            if cur is None or steps % 100 == 0 or stale:
                need = arena.budget_for(arena.step(info=True), 1.5)
                if (stale or cur is None or need[0] > cur[0] or need[1] > cur[1]
                        or 2 * need[0] < cur[0] or 2 * need[1] < cur[1]):
                    cur = need
                    arena.capture(cur)
                    n_capture += 1
            else:
                arena.replay()
        else:
            arena.step()
        steps += 1
        now = time.perf_counter() - t0
        if steps % args.poll_every == 0 or args.graph:
            for k in arena.poll():
                done_at.setdefault(k, now)
        if steps % 1000 == 0:
            print(f'# batched: {steps} iterations, {len(done_at)} cases done, {now:.1f} s',
                  file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    batched_s = time.perf_counter() - t0
    iters = [arena.iterations(k) for k in range(len(specs))]
    n_conv = sum(arena.converged(k) for k in range(len(specs)))
    lat = np.array([done_at[k] for k in range(len(specs))])
    print(f'# batched: {batched_s:.2f} s, {steps} batch iterations, {n_conv}/{len(specs)} '
          f'converged', file=sys.stderr, flush=True)

    roofline = kernels = None
    if profile:
        prof_arena = CaseArena.for_cases(specs, **kw)
        with _lib.Profiler(capacity=400000) as prof:
            for _ in range(20):
                prof_arena.step()
            torch.cuda.synchronize()
        nnz = sum(s.layout.nnz for s in specs)
        F = sum(s.n_internal + s.n_boundary for s in specs)
        Fi = sum(s.n_internal for s in specs)
        single = prof_arena.cg_variant == _lib.CG_SINGLE_REDUCTION
        roofline, kernels = rooflines_from(
            prof.result, cells, F, Fi, F - Fi, nnz, single,
            note='first 20 batch iterations, all 64 cases active', skipped=prof.skipped)
        del prof_arena

    seq = None
    if sequential:
        seq_t, seq_it, dev, seq_conv = [], [], [], 0
        for k, (n, re) in enumerate(draws):
            model = cavity_model(n, re)
            model.finalize()
            sv = IncompressibleSolver(model, 0.7, 0.3)
            sv.tol, sv.rtol, sv.rtol0 = pol['atol'], pol['tol'], pol['rtol0']
            sv.max_iter, sv.check_every, sv.cg_variant = pol['maxiter'], 10, kw['cg_variant']
            sv.check_every_u = pol['check_every_u']
            sv.max_iter_u = pol['maxiter_u']
            torch.cuda.synchronize()
            t = time.perf_counter()
            it, ok, _ = sv.run_to_convergence(stop, stop_u, floor_u, max_outer=args.max_outer)
            torch.cuda.synchronize()
            seq_t.append(time.perf_counter() - t)
            seq_it.append(it)
            seq_conv += int(ok)
            print(f'# sequential case {k} (N={n}, Re={re:.0f}): {it} iterations, '
                  f'{seq_t[-1]:.2f} s (batched: {iters[k]})', file=sys.stderr, flush=True)
            a = model.cell_values.cpu().numpy().astype(np.float64)
            b = arena.cell_values(k).astype(np.float64)
            dev.append(float(np.linalg.norm(a[:, :2] - b[:, :2]) / np.linalg.norm(a[:, :2])))
            del sv, model
        tot = float(np.sum(seq_t))
        seq = {'total_s': round(tot, 3), 'solves_per_s': round(len(draws) / tot, 4),
               'converged': seq_conv,
               'outer_iterations': {'sum': int(np.sum(seq_it)), 'max': int(np.max(seq_it))},
               'max_rel_l2_velocity_vs_batched': float(np.max(dev))}
        print(f'# sequential: {tot:.2f} s', file=sys.stderr, flush=True)

    fp64 = None
    if fp64_check:
        import copy
        specs64 = []
        for k, sp in enumerate(specs):
            s64 = copy.copy(sp)
            s64.stop, s64.stop_u, s64.floor_u = 0.0, 0.0, 0.0    # run exactly its iterations
            s64.max_outer = iters[k]
            specs64.append(s64)
        t = time.perf_counter()
        a64 = CaseArena.for_cases(specs64, dtype=torch.float64, **kw)
        a64.run(max_steps=max(iters) + 1)
        errs = []
        for k in range(len(specs)):
            a = arena.cell_values(k).astype(np.float64)
            b = a64.cell_values(k).astype(np.float64)
            errs.append(float(np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3])))
        fp64 = {'velocity_rel_l2': {'max': max(errs), 'median': float(np.median(errs))},
                'seconds': round(time.perf_counter() - t, 1),
                'note': 'every case through a double-field arena (libwcfd64.so) for the same '
                        'number of SIMPLE iterations its float32 solve ran'}
        print(f'# fp64 check: {fp64}', file=sys.stderr, flush=True)
        del a64

    cpu_b = None
    if cpu:
        cpu_b = cpu_solve_baseline(draws, stop, pol['rtol0'], stop_u)

    value = len(specs) / batched_s
    out = {
        'metric': 'SIMPLE iter/s + %HBM roofline; served solves/s, p50/p99 at equal residual',
        'value': round(value, 4),
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': 1,
        'warmup': 0,
        'ms_per_step': round(1e3 * batched_s, 1),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': f'{len(specs)} lid-driven cavities in '
                               f'{f"{args.arenas} arenas on {args.arenas} streams" if split else "one arena"}'
                               f', N log-uniform in '
                               f'[32,256], Re uniform in [100,1000] (seed {args.seed}), steady '
                               f'SIMPLE (alpha_u 0.7, alpha_p 0.3) to continuity <= {stop:g} and '
                               f'estimated velocity error <= {stop_u:g}', 'cases': len(specs),
                   'cells': cells, 'chunk': kw['chunk'], 'block': kw['block'],
                   'krylov': {'atol': 0.0, 'tol': 0.0, 'rtol0': pol['rtol0'],
                              'maxiter': pol['maxiter'], 'check_every': 10,
                              'check_every_momentum': pol['check_every_u'],
                              'maxiter_momentum': pol['maxiter_u'] or pol['maxiter'],
                              'pressure_cg': 'single' if kw['cg_variant'] else 'standard'},
                   'stop': {'continuity': stop, 'velocity_error_estimate': stop_u,
                            'velocity_change_floor': floor_u},
                   'reductions': 'f64'},
        'arena_tuning': tuned,
        'batched': {'total_s': round(batched_s, 3), 'batch_iterations': steps,
                    'arenas': ({'count': len(arena.arenas), 'cases': [len(g) for g in arena.groups],
                                'batch_iterations': per_arena,
                                'note': 'cases split by cell count over arenas stepped by '
                                        'separate host threads on separate streams'}
                               if split else None),
                    'hip_graph': ({'captures': n_capture, 'last_budget': cur, 'recheck_every': 100}
                                  if args.graph else None),
                    'converged': n_conv, 'outer_iterations_max': int(max(iters)),
                    'time_to_converged_s': {'p50': round(float(np.percentile(lat, 50)), 3),
                                            'p99': round(float(np.percentile(lat, 99)), 3)}},
        'sequential': seq,
        'speedup_vs_sequential': round(seq['total_s'] / batched_s, 2) if seq else None,
        'error_vs_fp64_build': fp64,
        'host_preprocessing_s': round(prep_s, 1),
        'roofline': roofline,
        'cpu_baseline': cpu_b,
        'kernels': kernels,
    }
    return out


SERVE_MIX = {'cavity_n': (32, 512), 'channel_cells': (50_000, 500_000)}
SERVE_SAMPLE_MIX = {'cavity_n': (32, 96), 'channel_cells': (50_000, 150_000)}


def _long_run_errors(sample, pol, chunk, block):
    """Velocity error of served results against long runs of the same cases
    in the double-field build: every sampled case runs again, all of them
    together in one float64 arena (libwcfd64.so), for a fixed 4 x its served
    iterations (at least served + 3000, no stop), and the served float32 u,
    v, w are compared with that longer float64 solve (rel-L2)."""
    import copy
    import numpy as np
    import torch
    from warpcfd_amd.batch import CaseArena
    specs = []
    for r in sample:
        sp = copy.copy(r.spec)
        sp.stop, sp.stop_u, sp.floor_u = 0.0, 0.0, 0.0     # never stop: run to max_outer
        sp.max_outer = max(4 * r.iterations, r.iterations + 3000)
        specs.append(sp)
    arena = CaseArena.for_cases(specs, chunk=chunk, block=block, dtype=torch.float64, **pol)
    arena.run(max_steps=max(sp.max_outer for sp in specs) + 1)
    out = []
    for k, r in enumerate(sample):
        a = np.asarray(r.fields[:, :3], np.float64)
        b = arena.cell_values(k)[:, :3].astype(np.float64)
        out.append({'request': r.describe(), 'served_iterations': int(r.iterations),
                    'long_run_iterations': int(specs[k].max_outer),
                    'velocity_rel_l2': float(np.linalg.norm(a - b) / np.linalg.norm(b))})
    return out


def run_serve(args, n_requests, mix=SERVE_MIX, sequential=True, error_sample=0, label=''):
    """Serves a Poisson stream (seeded, args.rate per second) of cavity and
    channel requests drawn from ``mix``, each to continuity <= 1e-6 and a
    convergence-monitor velocity-error estimate <= args.stop_u (the batch
    stop; ``--serve-continuity-only`` drops the second test), with continuous
    batching on an arena sized by an HBM budget and, if ``sequential``, one
    at a time through the same kernels; ``error_sample`` served results are
    checked against float64 long runs.  Returns the result dict."""
    import math
    import numpy as np
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.batch import CaseArena
    from warpcfd_amd.serve import (ContinuousBatchServer, Preparer, SequentialServer,
                                   poisson_stream, summarize)
    stop = 1e-6
    stop_u = math.inf if args.serve_continuity_only else args.stop_u
    floor_u = 0.0 if args.serve_continuity_only else 1e-7
    pol = dict(atol=0.0, tol=0.0, rtol0=args.rtol0, maxiter=2000, check_every=10,
               check_every_u=args.check_u, cg_variant=_lib.CG_SINGLE_REDUCTION)
    # preparation workers start before the first GPU call of this process
    prep = Preparer(workers=args.workers, stop=stop, stop_u=stop_u, floor_u=floor_u,
                    max_outer=args.max_outer)
    prep.warm()
    try:
        from warpcfd_amd.batch import CaseSpec
        from warpcfd_amd.workloads import cavity_model
        warm = CaseArena.for_cases([CaseSpec(cavity_model(16, 100.0))], chunk=args.chunk,
                                   block=args.block, **pol)
        for _ in range(3):
            warm.step()
        torch.cuda.synchronize()
        del warm
        stream = poisson_stream(args.seed, rate=args.rate, n_requests=n_requests, **mix)
        cells = sum(r.cells for r in stream)
        print(f'# serve{label}: {len(stream)} requests, {cells} cells, arrivals over '
              f'{stream[-1].arrival:.2f} s', file=sys.stderr, flush=True)
        srv = ContinuousBatchServer(hbm_bytes=int(args.hbm_gb * (1 << 30)), max_cases=args.slots,
                                    chunk=args.chunk, block=args.block, lanes=args.lanes,
                                    keep_fields=error_sample > 0, **pol)

        def report(r):
            print(f'#   done {r.describe()}: {r.iterations} iterations, '
                  f'{r.t_done - r.arrival:.2f} s after arrival', file=sys.stderr, flush=True)

        def progress(lane, steps, resident):
            print(f'#   lane {lane}: {steps} iterations, {resident} resident, '
                  f'{time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)

        t0 = time.perf_counter()
        done = srv.serve(stream, prep, clock_start=t0, on_done=report, on_step=progress,
                         step_every=2000)
        torch.cuda.synchronize()
        served = summarize(done, max(r.t_done for r in done))
        print(f'# continuous batching{label}: {served}', file=sys.stderr, flush=True)
        seq = None
        if sequential:
            stream2 = poisson_stream(args.seed, rate=args.rate, n_requests=n_requests, **mix)
            t0 = time.perf_counter()
            done2 = SequentialServer(chunk=args.chunk, block=args.block, **pol).serve(
                stream2, prep, clock_start=t0, on_done=report)
            torch.cuda.synchronize()
            seq = summarize(done2, max(r.t_done for r in done2))
            print(f'# one at a time{label}: {seq}', file=sys.stderr, flush=True)
        errors = None
        if error_sample:
            ok = [r for r in done if r.error is None and r.converged and r.cells <= 60_000]
            rng = np.random.default_rng(args.seed + 1)
            pick = sorted(rng.choice(len(ok), size=min(error_sample, len(ok)), replace=False))
            errors = _long_run_errors([ok[i] for i in pick], pol, args.chunk, args.block)
            print(f'# velocity error vs float64 long runs{label}: {errors}', file=sys.stderr, flush=True)
    finally:
        prep.close()
    stop_desc = f'continuity <= {stop:g}' + ('' if math.isinf(stop_u) else
                                             f' and estimated velocity error <= {stop_u:g}')
    return {
        'workload': f'served stream: {len(stream)} Poisson arrivals (seed {args.seed}, '
                    f'{args.rate:g}/s) of lid-driven cavities (N in [{mix["cavity_n"][0]},'
                    f'{mix["cavity_n"][1]}], Re in [100,1000]) and 10x2x2 channels '
                    f'({mix["channel_cells"][0] // 1000}k-{mix["channel_cells"][1] // 1000}k cells, '
                    f'Re in [50,400]), steady SIMPLE to {stop_desc}',
        'requests': len(stream), 'cells': cells, 'hbm_budget_gb': args.hbm_gb,
        'slots': args.slots, 'lanes': args.lanes, 'host_workers': args.workers,
        'krylov': {'atol': 0.0, 'tol': 0.0, 'rtol0': args.rtol0, 'maxiter': 2000,
                   'check_every': 10, 'check_every_momentum': args.check_u,
                   'pressure_cg': 'single'},
        'served': served,
        'one_at_a_time': seq,
        'speedup_vs_one_at_a_time': (round(served['solves_per_s'] / seq['solves_per_s'], 2)
                                     if seq else None),
        'velocity_error_vs_float64_long_runs': errors,
        'latency_note': 'time to converged = arrival to convergence, including host '
                        'preparation (mesh, layout; worker processes) and queueing; '
                        'resident time = admission to convergence',
    }


def bench_serve(args):
    """A served stream of ``--requests`` Poisson arrivals: the configured mix
    (cavities N in [32,512], channels 50k-500k cells) or, with ``--serve-mix
    sample``, the reduced mix of the default line's served sample (cavities
    N in [32,96], channels 50k-150k cells).  One step = the whole stream."""
    mix = SERVE_SAMPLE_MIX if args.serve_mix == 'sample' else SERVE_MIX
    res = run_serve(args, args.requests, mix, sequential=not args.no_sequential,
                    error_sample=args.error_sample)
    served = res['served']
    out = {
        'metric': 'SIMPLE iter/s + %HBM roofline; served solves/s, p50/p99 at equal residual',
        'value': served['solves_per_s'],
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': 1,
        'warmup': 0,
        'ms_per_step': round(1e3 * served['wall_s'], 1),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': res.pop('workload'), 'reductions': 'f64'},
        'roofline': None,
        'cpu_baseline': None,
    }
    out.update(res)
    print(json.dumps(out))


KERNELS_NOTE = ('avg_us and gbs average over the launches that did work: a Krylov solve keeps '
                'launching while the host has not yet seen its stop, and those launches exit at '
                'entry; the library profiler retires them when the solve returns and they are '
                'counted apart as exit_at_entry')


def install_launch(spec):
    """GRIDxBLOCK[xFOLD] from the command line: install and describe it."""
    from warpcfd_amd.tune import install
    parts = [int(v) for v in spec.lower().split('x')]
    g, b = parts[:2]
    fold = parts[2] if len(parts) > 2 else 2048
    install(g, b, fold)
    return {'grid': g, 'block': b, 'fold': fold, 'source': 'command line'}


def rooflines_from(result, C, F, Fi, Fb, nnz, cg_single=True, note=None, slots=6, skipped=None):
    """(roofline of the dominant kernel, per-kernel table with each kernel's own
    roofline) from the library profiler's {kernel: (total ms, launches)} over
    the launches that did work; ``skipped`` {kernel: launches that exited at
    entry} is reported beside them."""
    kernels = {k: {'total_ms': round(v[0], 3), 'launches': v[1], 'avg_us': round(1e3 * v[0] / v[1], 3)}
               for k, v in sorted(result.items(), key=lambda kv: -kv[1][0])}
    for k, n in (skipped or {}).items():
        kernels.setdefault(k, {'total_ms': 0.0, 'launches': 0, 'avg_us': 0.0})['exit_at_entry'] = n
    dominant = max(result.items(), key=lambda kv: kv[1][0])[0]
    roofline = None
    for k, v in kernels.items():
        kb = algorithmic_bytes(k + ('_single' if k == 'cg_1' and cg_single else ''),
                               C, F, Fi, Fb, nnz, slots)
        if kb is None or v['avg_us'] <= 0:
            continue
        gbs = kb / (v['avg_us'] * 1e-6) / 1e9
        v['bytes_per_launch'] = kb
        v['gbs'] = round(gbs, 1)
        v['frac'] = round(gbs / HBM_PEAK_GBS, 4)
        if k == dominant:
            roofline = {'bound': 'hbm', 'kernel': k, 'achieved': round(gbs, 1),
                        'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(gbs / HBM_PEAK_GBS, 4),
                        'traffic': None, 'bytes_per_launch': kb, 'avg_launch_us': v['avg_us'],
                        'launches': v['launches'],
                        'exit_at_entry_launches_excluded': v.get('exit_at_entry', 0)}
            if note:
                roofline['note'] = note
    return roofline, kernels


def profile_roofline(run, C, F, Fi, Fb, nnz, cg_single=True, capacity=400000, note=None):
    """Runs ``run()`` under the library's HIP-event profiler; returns
    (roofline dict of the dominant kernel, per-kernel table)."""
    from warpcfd_amd import _lib
    import torch
    with _lib.Profiler(capacity=capacity) as prof:
        run()
        torch.cuda.synchronize()
    return rooflines_from(prof.result, C, F, Fi, Fb, nnz, cg_single, note, skipped=prof.skipped)


def bench_cylinder(args):
    """Transient PISO past a cylinder: box 20 x 8 x 1 with a voxel-carved
    cylinder D = 1 at (5, 4), h = 0.05 (1 273 680 hexahedra, interior nodes
    jittered +-10% h, seed 0), inflow u = 1, outflow p = 0, slip side walls,
    no-slip cylinder, Re = 100, dt = 0.025 (CFL 0.5), two PISO correctors,
    float64 pressure-residual refinement.  One step = one time step."""
    import numpy as np
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.workloads import cylinder_model
    h = args.cyl_h
    t = time.perf_counter()
    model = cylinder_model(h=h, re=100.0, jitter=0.1, seed=0)
    model.finalize()
    prep_s = time.perf_counter() - t
    dt = 0.5 * h / 1.0
    sv = IncompressibleSolver(model, None, None, steady_state=False, implicit_outflow=True)
    sv.set_iterations(1, 2, 1)
    sv.set_time_stepping(args.warmup + args.steps, dt)
    sv.cg_variant = _lib.CG_SINGLE_REDUCTION if args.cg == 'single' else _lib.CG_STANDARD
    sv.p_refine = args.p_refine
    dm = model.dm
    tuned = None
    if args.krylov_launch:
        tuned = install_launch(args.krylov_launch)
    elif not args.no_tune:
        from warpcfd_amd.tune import autotune_krylov
        # the pressure CG runs about 40 iterations per momentum BiCGSTAB
        # iteration in these steps: weigh the two timings accordingly
        tt = autotune_krylov(dm, cg_variant=sv.cg_variant, bicg_weight=0.025,
                             try_persistent=False)
        tuned = {'grid': tt['grid'], 'block': tt['block'], 'fold': tt['fold'],
                 'persistent_rows': tt['persistent']}
    C, F, Fi, Fb, nnz = dm.n_cells, dm.n_faces, dm.n_internal, dm.n_boundary, dm.nnz
    for _ in range(args.warmup):
        sv.outer_iteration(dt=dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        sv.outer_iteration(dt=dt)
        if (i + 1) % 500 == 0:        # long runs (the configured 2 000 steps) report progress
            print(f'# PISO step {i + 1}: {time.perf_counter() - t0:.1f} s', file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    hist = sv.history[-args.steps:]
    fields32 = model.cell_values.numpy()
    roofline = kernels = None
    if not args.no_profile:
        def run():
            for _ in range(min(args.steps, 5)):
                sv.outer_iteration(dt=dt)
        roofline, kernels = profile_roofline(run, C, F, Fi, Fb, nnz, args.cg == 'single')
    fp64 = None
    if not args.no_fp64:
        import numpy as np
        from warpcfd_amd.FV import FVM
        from warpcfd_amd.workloads import cylinder_bcs

        def run64():
            from warpcfd_amd.tune import install
            m64 = FVM(model.mesh, float_dtype=np.float64)
            m64.material.create_incompressible_newtonian_fluid(
                viscosity=model.material.viscosity, density=1.0)
            cylinder_bcs(m64.boundary)
            m64.finalize(layout=model.layout)
            if tuned:
                install(tuned['grid'], tuned['block'], tuned['fold'], m64.lib)
            s64 = IncompressibleSolver(m64, None, None, steady_state=False, implicit_outflow=True)
            s64.set_iterations(1, 2, 1)
            s64.cg_variant, s64.p_refine, s64.tol = sv.cg_variant, sv.p_refine, sv.tol
            for i in range(args.warmup + args.steps):
                s64.outer_iteration(dt=dt)
                if (i + 1) % 500 == 0:
                    print(f'# float64 PISO step {i + 1}', file=sys.stderr, flush=True)
            return m64.cell_values.numpy()
        fp64 = fp64_error(fields32, run64)
        fp64['note'] = (f'the same workload through the double-field build (libwcfd64.so), same '
                        f'Krylov settings, after the same {args.warmup + args.steps} time steps '
                        f'from rest')
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_cylinder_baseline(C, args.p_refine)
    out = {
        'metric': 'SIMPLE iter/s + %HBM roofline; served solves/s, p50/p99 at equal residual',
        'value': round(args.steps / elapsed, 3),
        'unit': 'PISO time steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': f'transient PISO past a cylinder: box 20x8x1, D=1 at (5,4), '
                               f'h={h} ({C} hex cells, interior vertices jittered +-10% h, seed 0), '
                               f'inflow u=1, outflow p=0, slip side walls, no-slip cylinder, '
                               f'Re=100, dt={dt} (CFL 0.5), 2 PISO correctors, float64 '
                               f'pressure-residual refinement ({args.p_refine} sweep)',
                   'cells': C, 'faces': F, 'nnz': nnz, 'reductions': 'f64',
                   'krylov': {'atol': 1e-5, 'maxiter': 500, 'check_every': 10,
                              'pressure_cg': args.cg, 'p_refine': args.p_refine,
                              'cg_stop_test': args.krylov_check}},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'continuity_residual': hist[-1]['continuity'],
        'error_vs_fp64_build': fp64,
        'krylov_iters_per_step': {'momentum': float(np.mean([h_['u_iters'] for h_ in hist])),
                                  'pressure_last_corrector': float(np.mean([h_['p_iters'] for h_ in hist]))},
        'krylov_launch': tuned,
        'host_preprocessing_s': round(prep_s, 1),
        'kernels': kernels,
    }
    print(json.dumps(out))


def cpu_cylinder_baseline(C_full, p_refine, budget_s=30.0):
    """CPU oracle, one thread, on a bounded sample: the same cylinder
    geometry at h = 0.1 (1/8 of the cells), PISO steps until the budget;
    reported per step and scaled by the cell-count ratio to the full mesh
    (iteration counts held equal: a lower bound on the CPU time of a full
    step)."""
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:
        limiter = None
    import numpy as np
    from oracle.mesh import OracleMesh
    from oracle.fv import OracleFVM
    from oracle.simple import OracleSolver
    from warpcfd_amd.preprocess import Mesh, box_and_body_walls, cylinder_mesh
    from warpcfd_amd.workloads import cylinder_bcs
    h = 0.1
    nodes, cells = cylinder_mesh(h, jitter=0.1, seed=0)
    walls = box_and_body_walls(Mesh(nodes, cells), (20.0, 8.0, 1.0))
    om = OracleMesh(nodes, cells)
    om.groups = {k: np.asarray(v) for k, v in walls.items()}
    of = OracleFVM(om, np.float32)
    cylinder_bcs(of.boundary)
    of.boundary.finalize()
    s = OracleSolver(of, None, None, steady_state=False, nu=0.01, implicit_outflow=True,
                     p_refine=p_refine)
    s.set_iterations(1, 2, 1)
    t0 = time.perf_counter()
    n = 0
    while n < 5:
        s.outer_iteration(0.5 * h)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    if limiter is not None and hasattr(limiter, 'unregister'):
        limiter.unregister()
    per_step = dt / n
    scale = C_full / om.num_cells
    return {'value': round(1.0 / (per_step * scale), 5), 'unit': 'PISO time steps/s', 'cores': 1,
            'kind': 'port',
            'sample': f'{n} PISO steps of the same cylinder at h=0.1 ({om.num_cells} cells) by the '
                      f'oracle, 1 thread, {per_step:.2f} s/step; value scaled by the cell ratio '
                      f'{scale:.2f} to the full mesh'}


def cpu_solve_baseline(draws, stop, rtol0, stop_u, budget_s=25.0):
    """CPU oracle, one thread: the smallest case of the batch solved to the
    same continuity target with the same Krylov policy (or as far as the
    time budget allows), reported as solves/s of that case."""
    import numpy as np
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:
        limiter = None
    from oracle.mesh import OracleMesh
    from oracle.fv import OracleFVM
    from oracle.simple import OracleSolver
    from warpcfd_amd.preprocess import cavity_mesh
    n, re = min(draws)
    nodes, cells = cavity_mesh(n)
    om = OracleMesh(nodes, cells)
    om.define_boundary_walls()
    of = OracleFVM(om)
    of.set_reference_pressure(0, 0.0)
    apply_bcs(of.boundary, 'cavity')
    of.boundary.finalize()
    s = OracleSolver(of, 0.7, 0.3, nu=1.0 / re, atol=0.0, tol=0.0, rtol0=rtol0, maxiter=2000,
                     cg_variant=1)
    from warpcfd_amd.FV.implicit_Solvers.incompressible import Monitor
    mon = Monitor(stop, stop_u, 1e-7)      # the same stop rule as the GPU runs
    t0 = time.perf_counter()
    it = 0
    cont = 1.0
    done = False
    prev = of.cell_values[:, :3].astype(np.float64).copy()
    while time.perf_counter() - t0 < budget_s:
        cont = s.outer_iteration()['continuity']
        it += 1
        cur = of.cell_values[:, :3].astype(np.float64)
        d = float(np.linalg.norm(cur - prev) / max(np.linalg.norm(cur), 1e-300))
        prev = cur.copy()
        if mon.update(cont, d):
            done = True
            break
    dt = time.perf_counter() - t0
    if limiter is not None and hasattr(limiter, 'unregister'):
        limiter.unregister()
    return {'value': round((1.0 / dt) if done else 0.0, 5), 'unit': 'solves/s', 'cores': 1,
            'kind': 'port',
            'sample': f'smallest batch case (cavity {n}x{n}, Re {re:.0f}) by the oracle, 1 thread: '
                      f'{it} SIMPLE iterations in {dt:.1f} s, continuity {cont:.2e}, stopped by '
                      f'the same convergence monitor as the GPU runs'
                      + ('' if done else ' (budget reached before convergence; value 0)')}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='channel',
                    choices=sorted(WORKLOADS) + ['batch64', 'serve', 'cylinder'])
    ap.add_argument('--cyl-h', type=float, default=0.05, help='cylinder: cell size')
    ap.add_argument('--cell-type', default='hexahedron', choices=['hexahedron', 'tetra', 'wedge'],
                    help='channel / cavity: split the hexahedra into tetrahedra or wedges')
    ap.add_argument('--channel-dims', default='200x40x40',
                    help='channel: hexahedra along x, y, z (the 10x2x2 box)')
    ap.add_argument('--p-refine', type=int, default=1,
                    help='cylinder: float64 pressure refinement sweeps')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--cases', type=int, default=64)
    ap.add_argument('--chunk', type=int, default=256, help='batch: rows per workgroup')
    ap.add_argument('--block', type=int, default=256, help='batch: threads per workgroup')
    ap.add_argument('--rtol0', type=float, default=0.1,
                    help='batch: Krylov stop relative to the warm-start residual')
    ap.add_argument('--max-outer', type=int, default=40000)
    ap.add_argument('--stop-u', type=float, default=2e-5,
                    help='batch: stop at this estimated velocity error (with continuity <= 1e-6)')
    ap.add_argument('--poll-every', type=int, default=4,
                    help='batch: read the active flags every this many iterations')
    ap.add_argument('--maxiter-u', type=int, default=0,
                    help='batch: momentum-solve iteration cap (0: maxiter)')
    ap.add_argument('--check-u', type=int, default=2,
                    help='batch: check interval of the momentum solve')
    ap.add_argument('--no-sequential', action='store_true')
    ap.add_argument('--arenas', type=int, default=4,
                    help='batch64: split the cases over this many arenas, each stepped by its '
                         'own host thread and stream (4: the hardware queues HIP gives a '
                         'process by default); 1: one arena')
    ap.add_argument('--calibration-copy-mb', type=float, default=0.0,
                    help='run one elementwise copy of this size (rocprofv3 counter calibration)')
    ap.add_argument('--traffic-from', default=None,
                    help='directory with fetch/ and write/ rocprofv3 --pmc runs of this '
                         'workload: fills roofline.traffic')
    ap.add_argument('--traffic-copy-mb', type=float, default=512.0,
                    help='size of the calibration copy in those runs')
    ap.add_argument('--graph', action='store_true',
                    help='batch: replay a HIP-graph-captured iteration (Krylov launch budget '
                         'measured every 100 iterations)')
    ap.add_argument('--requests', type=int, default=200, help='serve: requests in the stream')
    ap.add_argument('--serve-sample-requests', type=int, default=200,
                    help='default line: requests in the served sample')
    ap.add_argument('--serve-mix', default='configured', choices=['configured', 'sample'],
                    help='serve: the configured mix (cavities up to 512^2, channels up to 500k '
                         'cells) or the reduced mix of the default line\'s served sample')
    ap.add_argument('--error-sample', type=int, default=6,
                    help="serve: served results checked against float64 long runs")
    ap.add_argument('--rate', type=float, default=20.0, help='serve: arrivals per second')
    ap.add_argument('--hbm-gb', type=float, default=32.0, help='serve: arena HBM budget')
    ap.add_argument('--slots', type=int, default=64, help='serve: case slots')
    ap.add_argument('--lanes', type=int, default=4,
                    help='serve: arenas (budget and slots split evenly), each served by its own '
                         'host thread and stream; 1: one arena')
    ap.add_argument('--workers', type=int, default=8, help='serve: host preparation threads')
    ap.add_argument('--serve-continuity-only', action='store_true',
                    help='serve: stop at continuity <= 1e-6 alone (without the velocity-error '
                         'estimate of the batch stop)')
    ap.add_argument('--no-extras', action='store_true',
                    help='channel: skip the batch64 and served-sample sub-benchmarks')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-fp64', action='store_true',
                    help='skip the comparison with the double-field build')
    ap.add_argument('--no-profile', action='store_true')
    ap.add_argument('--no-tune', action='store_true')
    ap.add_argument('--cg', default='single', choices=['standard', 'single'],
                    help='pressure CG form: standard (two reductions per iteration) or single '
                         '(Chronopoulos-Gear, one fused kernel per iteration)')
    ap.add_argument('--krylov-launch', default=None,
                    help='GRIDxBLOCK[xFOLD]: fixed Krylov launch shape (grid 0: one row per '
                         'thread; FOLD: inline-fold limit, default 2048) instead of autotuning')
    ap.add_argument('--persistent-rows', type=int, default=0,
                    help='persistent single-reduction CG (one launch per solve, each case on one '
                         'XCD) for solves whose widest case has at most this many rows; 0: off')
    ap.add_argument('--krylov-check', default='inline', choices=['inline', 'kernel'],
                    help='single-case CG stop test: inside the first iteration kernel of each '
                         'check group (default) or in a check kernel after it (bitwise the same)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(cases shard across GPUs with no communication; see the design notes).',
              file=sys.stderr)
        sys.exit(2)
    if args.persistent_rows:
        from warpcfd_amd import _lib
        for lib in (_lib.load(), _lib.load64()):
            _lib.check(lib.wcfd_set_krylov_persistent(args.persistent_rows),
                       'set_krylov_persistent')
    if args.krylov_check != 'inline':
        from warpcfd_amd import _lib
        for lib in (_lib.load(), _lib.load64()):
            _lib.check(lib.wcfd_set_krylov_inline_check(0), 'set_krylov_inline_check')
    if args.workload == 'batch64':
        return bench_batch64(args)
    if args.workload == 'serve':
        return bench_serve(args)
    if args.workload == 'cylinder':
        return bench_cylinder(args)

    import numpy as np
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.preprocess import Mesh, define_boundary_walls
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver

    wl = dict(WORKLOADS[args.workload])
    dims = tuple(int(v) for v in args.channel_dims.lower().split('x'))
    if args.workload == 'channel' and dims != (200, 40, 40):
        C = dims[0] * dims[1] * dims[2]
        wl['desc'] = wl['desc'].replace('200x40x40 hex (320000 cells)', f'{args.channel_dims} hex '
                                        f'({C} cells)')
    if args.cell_type != 'hexahedron':
        per = {'tetra': 6, 'wedge': 2}[args.cell_type]
        name = {'tetra': 'Kuhn tetrahedra', 'wedge': 'prisms'}[args.cell_type]
        wl['desc'] = re.sub(r' hex \((\d+) cells\)',
                            lambda mo: f' hex, each split into {per} {name} '
                                       f'({per * int(mo.group(1))} cells; kernels read the '
                                       f'six-slot layout, unused slots on the null face)',
                            wl['desc'])
    t_prep = time.perf_counter()
    print(f'# building {args.workload} mesh', file=sys.stderr, flush=True)
    nodes, cells = build_case(args.workload, dims, args.cell_type)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
    if wl['kind'] == 'cavity':
        model.set_reference_pressure(0, 0.0)
    apply_bcs(model.boundary, wl['kind'])
    model.finalize()
    prep_s = time.perf_counter() - t_prep
    print(f'# {model.layout.num_cells} cells prepared in {prep_s:.1f} s', file=sys.stderr,
          flush=True)
    solver = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=(wl['kind'] == 'channel'))
    solver.cg_variant = _lib.CG_SINGLE_REDUCTION if args.cg == 'single' else _lib.CG_STANDARD
    dm = model.dm
    tuned = None
    if args.krylov_launch:
        tuned = install_launch(args.krylov_launch)
    elif not args.no_tune:
        from warpcfd_amd.tune import autotune_krylov
        # a SIMPLE iteration of the channel runs about 420 pressure CG
        # iterations and 10 momentum BiCGSTAB iterations: weigh the two
        # timings accordingly
        t = autotune_krylov(dm, cg_variant=solver.cg_variant, bicg_weight=0.025,
                            try_persistent=False)
        tuned = {'grid': t['grid'], 'block': t['block'], 'fold': t['fold'],
                 'persistent_rows': t['persistent'],
                 'cg_ms_per_60_it': round(t['cg_ms'], 3),
                 'candidates': len(t['candidates']),
                 'refined': [(r['grid'], r['block'], r['fold'], r['persistent'],
                              round(r['cg_ms'], 4)) for r in t['refined']]}
    C, F, Fi, Fb, nnz = dm.n_cells, dm.n_faces, dm.n_internal, dm.n_boundary, dm.nnz

    for _ in range(args.warmup):
        solver.outer_iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.outer_iteration()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    hist = solver.history[-args.steps:]
    fields32 = model.cell_values.numpy()

    roofline = None
    kernels = None
    if not args.no_profile:
        with _lib.Profiler(capacity=200000) as prof:
            for _ in range(args.steps):
                solver.outer_iteration()
            torch.cuda.synchronize()
        roofline, kernels = rooflines_from(prof.result, C, F, Fi, Fb, nnz, args.cg == 'single',
                                           slots=model.layout.faces_per_cell,
                                           skipped=prof.skipped)

    if args.calibration_copy_mb:
        calibration_copy(args.calibration_copy_mb)
    if roofline is not None and args.traffic_from:
        from warpcfd_amd.pmc import traffic
        sym = {'cg_1': 'k_cgs_iter' if args.cg == 'single' else 'k_cg_1', 'cg_2': 'k_cg_2',
               'bicgstab_1': 'k_bi_1', 'bicgstab_2': 'k_bi_2', 'bicgstab_3': 'k_bi_3'}
        t = traffic(args.traffic_from, sym.get(roofline['kernel'], roofline['kernel']),
                    args.traffic_copy_mb)
        roofline['traffic'] = t['bytes_per_launch']
        roofline['traffic_detail'] = t
        # the same counters for every kernel with a byte count
        for k, v in (kernels or {}).items():
            if 'bytes_per_launch' in v:
                tk = traffic(args.traffic_from, sym.get(k, 'k_' + k), args.traffic_copy_mb)
                v['traffic'] = tk['bytes_per_launch']

    fp64 = None
    if not args.no_fp64:
        def run64():
            from warpcfd_amd.tune import install
            m64 = FVM(m, float_dtype=np.float64)
            m64.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
            if wl['kind'] == 'cavity':
                m64.set_reference_pressure(0, 0.0)
            apply_bcs(m64.boundary, wl['kind'])
            m64.finalize(layout=model.layout)
            if tuned:
                install(tuned['grid'], tuned['block'], tuned['fold'], m64.lib)
            s64 = IncompressibleSolver(m64, 0.7, 0.3, implicit_outflow=(wl['kind'] == 'channel'))
            s64.cg_variant = solver.cg_variant
            s64.tol = solver.tol           # the same Krylov settings as the float32 run
            for _ in range(args.warmup + args.steps):
                s64.outer_iteration()
            return m64.cell_values.numpy()
        fp64 = fp64_error(fields32, run64)
        fp64['note'] = (f'the same workload through the double-field build (libwcfd64.so), same '
                        f'Krylov settings and launch shape, after the same '
                        f'{args.warmup + args.steps} SIMPLE iterations from rest; each build '
                        f'stops its own Krylov solves at the absolute 1e-5')

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(nodes, cells, wl['kind'], warmup=args.warmup,
                           cg_variant=1 if args.cg == 'single' else 0,
                           gpu_window={'iterations': f'{args.warmup + 1}-{args.warmup + args.steps}',
                                       'krylov_iters_per_step': {
                                           'momentum': float(np.mean([h['u_iters'] for h in hist])),
                                           'pressure': float(np.mean([h['p_iters'] for h in hist]))}})

    persistent = None
    if not args.no_extras and args.cg == 'single':
        def make_solver():
            m2 = FVM(m, float_dtype=np.float32)
            m2.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
            if wl['kind'] == 'cavity':
                m2.set_reference_pressure(0, 0.0)
            apply_bcs(m2.boundary, wl['kind'])
            m2.finalize(layout=model.layout)
            s2 = IncompressibleSolver(m2, 0.7, 0.3, implicit_outflow=(wl['kind'] == 'channel'))
            s2.cg_variant = solver.cg_variant
            s2.tol = solver.tol
            return m2, s2
        try:
            persistent = persistent_cg_line(args, make_solver, C)
        except Exception as exc:             # the main line is printed either way
            persistent = {'error': repr(exc)}
    extras = {}
    if not args.no_extras and args.workload == 'channel' and dims == (200, 40, 40) \
            and args.cell_type == 'hexahedron':
        extras = channel_extras(args)
    out = {
        'metric': 'SIMPLE iter/s + %HBM roofline; served solves/s, p50/p99 at equal residual',
        'value': round(args.steps / elapsed, 3),
        'unit': 'iter/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': wl['desc'], 'cells': C, 'faces': F, 'nnz': nnz,
                   'reductions': 'f64', 'krylov': {'atol': 1e-5, 'maxiter': 500, 'check_every': 10,
                                                   'pressure_cg': args.cg,
                                                   'cg_stop_test': args.krylov_check}},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'continuity_residual': hist[-1]['continuity'],
        'krylov_iters_per_step': {'momentum': float(np.mean([h['u_iters'] for h in hist])),
                                  'pressure': float(np.mean([h['p_iters'] for h in hist]))},
        'error_vs_fp64_build': fp64,
        'krylov_launch': tuned,
        'persistent_cg': persistent,
        'host_preprocessing_s': round(prep_s, 1),
        'kernels': kernels,
        'kernels_note': KERNELS_NOTE,
    }
    out.update(extras)
    print(json.dumps(out))


def persistent_cg_line(args, make_solver, C):
    """The same steps with the persistent pressure CG (one cooperative launch
    per solve, the case spread over every XCD, bitwise the per-iteration
    kernels at the installed launch shape), from rest on a fresh model: SIMPLE
    iter/s over the same warm-up and timed window, and the launch's own
    roofline on the bytes it must move.  Its threads keep the matrix, the
    Jacobi values at the columns, x and p in registers, so a launch reads
    those once (52 B per row: 7 coefficients, 6 int16 offsets, inv, x in and
    out) and an iteration moves w, s and r in and out (24 B per row)."""
    import torch
    from warpcfd_amd import _lib
    lib = _lib.load()
    _lib.check(lib.wcfd_set_krylov_persistent(int(C)), 'set_krylov_persistent')
    try:
        model, sv = make_solver()
        for _ in range(args.warmup):
            sv.outer_iteration()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sv.outer_iteration()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        n0 = len(sv.history)
        with _lib.Profiler(capacity=200000) as prof:
            for _ in range(args.steps):
                sv.outer_iteration()
            torch.cuda.synchronize()
        ran = lib.wcfd_get_krylov_persistent() > 0
    finally:
        _lib.check(lib.wcfd_set_krylov_persistent(0), 'set_krylov_persistent')
    its = sum(h['p_iters'] for h in sv.history[n0:])
    ms, launches = prof.result.get('cg_xcd', (0.0, 0))
    out = {'value': round(args.steps / dt, 3), 'unit': 'iter/s',
           'ms_per_step': round(1e3 * dt / args.steps, 3), 'persistent_kernel_ran': ran,
           'note': 'same workload, launch shape and Krylov settings as the main line, from rest; '
                   'every pressure solve one cooperative launch over every XCD, bitwise the '
                   'per-iteration kernels (tests/test_gpu_persistent.py)'}
    if launches and its:
        by = C * (52.0 * launches + 24.0 * its) / launches
        gbs = by / (1e-3 * ms / launches) / 1e9
        out.update({'launches': launches, 'cg_iterations': its,
                    'us_per_launch': round(1e3 * ms / launches, 2),
                    'us_per_cg_iteration': round(1e3 * ms / its, 3),
                    'roofline': {'bound': 'hbm', 'kernel': 'cg_xcd', 'achieved': round(gbs, 1),
                                 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                                 'frac': round(gbs / HBM_PEAK_GBS, 4),
                                 'bytes_per_launch': round(by),
                                 'note': 'the bytes the launch must move (52 B per row once, '
                                         '24 B per row per iteration); the barrier between '
                                         'iterations, not HBM, bounds it'}})
    return out


def run_channel_to_stop(stop=1e-6, max_iters=600, float32_tail=200):
    """The 320k channel run to its stated stop, continuity <= 1e-6, with the
    reference Krylov settings: the double-field build until the stop, then
    the float32 build for the same iterations (its distance from the double
    run there) and ``float32_tail`` more, whose continuity levels off above
    the stop: the corrected flux takes p_N - p_O of pressures near 0.3 that
    differ by about 1.5e-3 from cell to cell, which float32 rounds to about
    1e-5 of the difference."""
    import numpy as np
    import torch
    from warpcfd_amd import _lib
    from warpcfd_amd.workloads import channel_model
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    out = {'workload': 'channel 200x40x40 (320000 cells), as the main line, to continuity <= 1e-6',
           'krylov': {'atol': 1e-5, 'maxiter': 500, 'check_every': 10, 'pressure_cg': 'single'}}
    fields = {}
    n_stop = None
    for dt in (np.float64, np.float32):
        model = channel_model(float_dtype=dt)
        model.finalize()
        sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
        sv.cg_variant = _lib.CG_SINGLE_REDUCTION
        n_max = max_iters if n_stop is None else n_stop + float32_tail
        cont = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(1, n_max + 1):
            cont.append(sv.outer_iteration()['continuity'])
            if n_stop is None and cont[-1] <= stop:
                break
            if n_stop is not None and it == n_stop:
                torch.cuda.synchronize()
                t_eq = time.perf_counter() - t0
                fields[dt] = model.cell_values.cpu().numpy().astype(np.float64)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        name = 'float64_build' if dt is np.float64 else 'float32'
        if n_stop is None:
            n_stop = it if cont[-1] <= stop else None
            fields[dt] = model.cell_values.cpu().numpy().astype(np.float64)
            out[name] = {'reached_stop': n_stop is not None, 'iterations': it,
                         'seconds': round(secs, 3), 'iter_per_s': round(it / secs, 1),
                         'continuity': cont[-1]}
            if n_stop is None:
                break
        else:
            out[name] = {'iterations_run': n_max, 'seconds_to_same_iterations': round(t_eq, 3),
                         'continuity_at_same_iterations': cont[n_stop - 1],
                         'continuity_min': min(cont), 'continuity_last': cont[-1],
                         'reached_stop': min(cont) <= stop}
        del sv, model
        torch.cuda.empty_cache()
    if np.float32 in fields and np.float64 in fields:
        a, b = fields[np.float32], fields[np.float64]
        out['float32_vs_float64_at_the_stop'] = {
            'velocity_rel_l2': float(np.linalg.norm(a[:, :3] - b[:, :3]) / np.linalg.norm(b[:, :3])),
            'p_rel_l2': float(np.linalg.norm(a[:, 3] - b[:, 3]) / np.linalg.norm(b[:, 3]))}
    return out


def channel_extras(args):
    """The other half of the metric, measured in the default run beside the
    channel line: the 64-cavity batch solved batched (the configured batch,
    timed as ``--workload batch64`` times it, without its one-after-another
    leg), and a served sample -- ``--serve-sample-requests`` (200) Poisson
    requests of a reduced mix (cavities N in [32,96], channels 50k-150k
    cells) to the batch stop, continuous batching against one at a time on
    the same stream, with a velocity-error check of 12 results against long
    runs.  The configured mix (N up to 512) runs for minutes and is measured
    by ``--workload serve``."""
    out = {}
    t = time.perf_counter()
    try:
        out['channel_to_stop'] = run_channel_to_stop()
    except Exception as exc:
        out['channel_to_stop'] = {'error': repr(exc)}
    print(f'# channel to its stop: {time.perf_counter() - t:.1f} s', file=sys.stderr, flush=True)
    t = time.perf_counter()
    try:
        b = run_batch64(args, sequential=False, profile=False, cpu=False)
        out['batch64'] = {'workload': b['config']['workload'], 'cases': b['config']['cases'],
                          'cells': b['config']['cells'], 'solves_per_s': b['value'],
                          'batched': b['batched'], 'arena_tuning': {
                              k: b['arena_tuning'][k] for k in ('cg_variant', 'chunk', 'block')}
                          if b['arena_tuning'] else None}
    except Exception as exc:                # the channel line is printed either way
        out['batch64'] = {'error': repr(exc)}
    print(f'# batch64 sub-benchmark: {time.perf_counter() - t:.1f} s', file=sys.stderr, flush=True)
    t = time.perf_counter()
    try:
        out['served_sample'] = run_serve(args, args.serve_sample_requests, SERVE_SAMPLE_MIX,
                                         sequential=True, error_sample=12, label=' sample')
    except Exception as exc:
        out['served_sample'] = {'error': repr(exc)}
    print(f'# served sample: {time.perf_counter() - t:.1f} s', file=sys.stderr, flush=True)
    return out


if __name__ == '__main__':
    main()
