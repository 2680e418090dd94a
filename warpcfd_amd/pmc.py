"""Per-dispatch figures of a kernel from rocprofv3 runs.

HBM traffic: two profiled runs of the same program, one with ``--pmc
FETCH_SIZE`` and one with ``--pmc WRITE_SIZE`` (rocprofv3 collects counters
per dispatch, in KiB; ``TCC`` channels summed), give the bytes each dispatch
moved between the L2 and memory.  The program also runs a plain device copy
of known size (the calibration copy: ``bench.py --calibration-copy-mb``);
the ratio of its known bytes to its counter reading scales every other
reading.

Durations: a ``--kernel-trace`` run gives each dispatch's begin and end.

A Krylov solve keeps launching until the host has seen its stop, and those
launches exit at entry; both summaries count them apart (a dispatch far
below the kernel's typical bytes or duration), as the library profiler does.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR/fetch -o run -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d DIR/write -o run -- python3 bench.py ...
    python -m warpcfd_amd.pmc traffic DIR k_cgs_iter 512
    rocprofv3 --kernel-trace --stats --output-format csv -d TDIR -o run -- python3 bench.py ...
    python -m warpcfd_amd.pmc trace TDIR k_cgs_iter
    python -m warpcfd_amd.pmc overlap TDIR

Overlap (kernels from several streams running at once): the sum of every
dispatch's duration against the union of the intervals in which at least one
kernel ran; their ratio is the mean number of kernels running whenever one
is.
"""
import csv
import glob
import os
import sys

# the calibration op is torch.neg(x, out=y) on float32 tensors: one
# elementwise kernel that reads and writes ``copy_bytes`` each
COPY_KERNEL_HINTS = ('neg_kernel',)


def _rows(run_dir):
    paths = glob.glob(os.path.join(run_dir, '**', '*counter_collection.csv'), recursive=True)
    if not paths:
        raise FileNotFoundError(f'no *counter_collection.csv under {run_dir}')
    out = []
    for p in paths:
        with open(p) as f:
            out.extend(csv.DictReader(f))
    return out


def per_dispatch(run_dir, counter):
    """{dispatch id: (kernel name, counter value in KiB)} from one run."""
    acc = {}
    for r in _rows(run_dir):
        name = r.get('Counter_Name', '')
        if name != counter:
            continue
        d = r.get('Dispatch_Id') or r.get('Correlation_Id')
        k = r.get('Kernel_Name', '')
        v = float(r.get('Counter_Value', 0.0))
        prev = acc.get(d, (k, 0.0))
        acc[d] = (k, prev[1] + v)
    return acc


# a dispatch reading or writing less than this share of the kernel's largest
# dispatch exited at entry
IDLE_BYTES_FRAC = 0.05
# a dispatch shorter than this share of the kernel's 90th-percentile duration
# exited at entry (about 1.6 us against 7 us and more for an iteration)
IDLE_TIME_FRAC = 0.4


def kernel_mean_bytes(run_dir, counter, kernel_substr):
    """Mean bytes per working dispatch of the kernels whose name contains
    ``kernel_substr``, the number of dispatches averaged and the number left
    out as exits at entry."""
    vals = [v for k, v in per_dispatch(run_dir, counter).values() if kernel_substr in k]
    if not vals:
        return None, 0, 0
    cut = IDLE_BYTES_FRAC * max(vals)
    work = [v for v in vals if v >= cut]
    return 1024.0 * sum(work) / len(work), len(work), len(vals) - len(work)


def calibration_factor(run_dir, counter, copy_bytes):
    """Known bytes of the calibration copy over the counter's reading of its
    largest copy dispatch (one direction: the copy reads and writes
    ``copy_bytes`` each)."""
    best = 0.0
    for k, v in per_dispatch(run_dir, counter).values():
        if any(h in k for h in COPY_KERNEL_HINTS) and 'wcfd' not in k:
            best = max(best, 1024.0 * v)
    if best <= 0.0:
        return None
    return copy_bytes / best


def traffic(pmc_dir, kernel_substr, copy_mb):
    """HBM bytes per launch of ``kernel_substr``: FETCH_SIZE + WRITE_SIZE,
    each scaled by the calibration copy.  ``pmc_dir`` holds ``fetch/`` and
    ``write/`` run directories.  Returns a dict (None entries if missing)."""
    copy_bytes = copy_mb * (1 << 20)
    out = {'kernel': kernel_substr}
    total = 0.0
    for sub, counter in (('fetch', 'FETCH_SIZE'), ('write', 'WRITE_SIZE')):
        d = os.path.join(pmc_dir, sub)
        raw, n, idle = kernel_mean_bytes(d, counter, kernel_substr)
        f = calibration_factor(d, counter, copy_bytes)
        out[counter] = {'raw_bytes': raw, 'dispatches': n, 'exit_at_entry': idle, 'scale': f,
                        'bytes': raw * f if (raw is not None and f) else None}
        if out[counter]['bytes'] is None:
            total = None
        elif total is not None:
            total += out[counter]['bytes']
    out['bytes_per_launch'] = total
    return out


def trace_summary(trace_dir, kernel_substr):
    """Durations of the dispatches whose kernel name contains
    ``kernel_substr`` in a ``--kernel-trace`` run: all of them (what
    ``--stats`` averages) and the working ones (exits at entry left out),
    per kernel instance and together."""
    paths = glob.glob(os.path.join(trace_dir, '**', '*kernel_trace.csv'), recursive=True)
    if not paths:
        raise FileNotFoundError(f'no *kernel_trace.csv under {trace_dir}')
    per = {}
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                k = r.get('Kernel_Name', '')
                if kernel_substr in k:
                    t = (int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-3
                    per.setdefault(k, []).append(t)
    if not per:
        return None
    every = sorted(t for ts in per.values() for t in ts)
    cut = IDLE_TIME_FRAC * every[min(len(every) - 1, int(0.9 * len(every)))]

    def stats(ts):
        work = [t for t in ts if t >= cut]
        return {'dispatches': len(ts), 'avg_us': round(sum(ts) / len(ts), 3),
                'working': len(work),
                'working_avg_us': round(sum(work) / len(work), 3) if work else None,
                'exit_at_entry': len(ts) - len(work)}
    out = {'kernel': kernel_substr, 'exit_at_entry_below_us': round(cut, 3),
           'instances': {k: stats(ts) for k, ts in sorted(per.items())}}
    out.update(stats(every))
    return out


def overlap_summary(trace_dir):
    """Dispatch count, traced span, summed kernel durations and the union of
    the busy intervals (seconds) of a ``--kernel-trace`` run, and their
    ratio: the mean number of kernels running whenever at least one runs."""
    paths = glob.glob(os.path.join(trace_dir, '**', '*kernel_trace.csv'), recursive=True)
    if not paths:
        raise FileNotFoundError(f'no *kernel_trace.csv under {trace_dir}')
    iv = []
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                iv.append((int(r['Start_Timestamp']), int(r['End_Timestamp'])))
    iv.sort()
    busy = sum(e - b for b, e in iv)
    union, cur_b, cur_e = 0, iv[0][0], iv[0][1]
    for b, e in iv[1:]:
        if b > cur_e:
            union += cur_e - cur_b
            cur_b, cur_e = b, e
        elif e > cur_e:
            cur_e = e
    union += cur_e - cur_b
    return {'dispatches': len(iv), 'span_s': round((max(e for _, e in iv) - iv[0][0]) * 1e-9, 3),
            'kernel_time_s': round(busy * 1e-9, 3), 'busy_union_s': round(union * 1e-9, 3),
            'mean_concurrent': round(busy / union, 3) if union else None}


if __name__ == '__main__':
    import json
    if sys.argv[1] == 'trace':
        print(json.dumps(trace_summary(sys.argv[2], sys.argv[3]), indent=1))
    elif sys.argv[1] == 'overlap':
        print(json.dumps(overlap_summary(sys.argv[2]), indent=1))
    else:
        a = sys.argv[2:] if sys.argv[1] == 'traffic' else sys.argv[1:]
        print(json.dumps(traffic(a[0], a[1], float(a[2])), indent=1))
