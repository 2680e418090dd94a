"""HBM traffic of a kernel from rocprofv3 counter runs.

Two profiled runs of the same program, one with ``--pmc FETCH_SIZE`` and one
with ``--pmc WRITE_SIZE`` (rocprofv3 collects counters per dispatch, in KiB;
``TCC`` channels summed), give the bytes each dispatch moved between the L2
and memory.  The program also runs a plain device copy of known size (the
calibration copy: ``bench.py --calibration-copy-mb``); the ratio of its
known bytes to its counter reading scales every other reading.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR/fetch -o run -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d DIR/write -o run -- python3 bench.py ...
    python -m warpcfd_amd.pmc DIR k_cgs_iter 512
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


def kernel_mean_bytes(run_dir, counter, kernel_substr):
    """Mean bytes per dispatch of the kernels whose name contains
    ``kernel_substr``, and the number of dispatches averaged."""
    vals = [v for k, v in per_dispatch(run_dir, counter).values() if kernel_substr in k]
    if not vals:
        return None, 0
    return 1024.0 * sum(vals) / len(vals), len(vals)


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
        raw, n = kernel_mean_bytes(d, counter, kernel_substr)
        f = calibration_factor(d, counter, copy_bytes)
        out[counter] = {'raw_bytes': raw, 'dispatches': n, 'scale': f,
                        'bytes': raw * f if (raw is not None and f) else None}
        if out[counter]['bytes'] is None:
            total = None
        elif total is not None:
            total += out[counter]['bytes']
    out['bytes_per_launch'] = total
    return out


if __name__ == '__main__':
    import json
    print(json.dumps(traffic(sys.argv[1], sys.argv[2], float(sys.argv[3])), indent=1))
