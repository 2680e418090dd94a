"""rocprofv3 counter parsing and calibration (warpcfd_amd.pmc) on
synthetic counter files (no GPU)."""
import csv
import os

from warpcfd_amd import pmc


def _write(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, 'run_counter_collection.csv'), 'w', newline='') as f:
        w = csv.DictWriter(f, fieldnames=['Dispatch_Id', 'Kernel_Name', 'Counter_Name',
                                          'Counter_Value'])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_traffic_scaled_by_calibration(tmp_path):
    mb = 64
    # calibration reads 64 MiB; the counter reports it as half (scale 2), and
    # the kernel's two dispatches report 1000 and 3000 KiB split over channels
    fetch = [{'Dispatch_Id': 1, 'Kernel_Name': 'at::native::neg_kernel_cuda', 'Counter_Name':
              'FETCH_SIZE', 'Counter_Value': mb * 1024 / 2},
             {'Dispatch_Id': 2, 'Kernel_Name': 'void wcfd::k_cgs_iter<x>', 'Counter_Name':
              'FETCH_SIZE', 'Counter_Value': 600.0},
             {'Dispatch_Id': 2, 'Kernel_Name': 'void wcfd::k_cgs_iter<x>', 'Counter_Name':
              'FETCH_SIZE', 'Counter_Value': 400.0},
             {'Dispatch_Id': 3, 'Kernel_Name': 'void wcfd::k_cgs_iter<x>', 'Counter_Name':
              'FETCH_SIZE', 'Counter_Value': 3000.0},
             {'Dispatch_Id': 4, 'Kernel_Name': 'void wcfd::k_check', 'Counter_Name':
              'FETCH_SIZE', 'Counter_Value': 7.0}]
    write = [{'Dispatch_Id': 1, 'Kernel_Name': 'at::native::neg_kernel_cuda', 'Counter_Name':
              'WRITE_SIZE', 'Counter_Value': mb * 1024},
             {'Dispatch_Id': 2, 'Kernel_Name': 'void wcfd::k_cgs_iter<x>', 'Counter_Name':
              'WRITE_SIZE', 'Counter_Value': 100.0}]
    _write(str(tmp_path / 'fetch'), fetch)
    _write(str(tmp_path / 'write'), write)
    t = pmc.traffic(str(tmp_path), 'k_cgs_iter', mb)
    assert t['FETCH_SIZE']['dispatches'] == 2
    assert abs(t['FETCH_SIZE']['scale'] - 2.0) < 1e-12
    assert abs(t['FETCH_SIZE']['bytes'] - 2.0 * 1024 * 2000.0) < 1e-6
    assert abs(t['WRITE_SIZE']['scale'] - 1.0) < 1e-12
    assert abs(t['bytes_per_launch'] - (2.0 * 1024 * 2000.0 + 1024 * 100.0)) < 1e-6


def test_traffic_leaves_out_exits_at_entry(tmp_path):
    """Dispatches far below the kernel's largest (a launch that found its
    solve stopped) are counted apart, not averaged in."""
    mb = 64
    rows = [{'Dispatch_Id': 1, 'Kernel_Name': 'at::native::neg_kernel_cuda', 'Counter_Name':
             'FETCH_SIZE', 'Counter_Value': mb * 1024}]
    for d, v in ((2, 5000.0), (3, 5200.0), (4, 12.0), (5, 4800.0), (6, 3.0)):
        rows.append({'Dispatch_Id': d, 'Kernel_Name': 'void wcfd::k_cgs_iter2<x, false, false>',
                     'Counter_Name': 'FETCH_SIZE', 'Counter_Value': v})
    _write(str(tmp_path / 'fetch'), rows)
    mean, n, idle = pmc.kernel_mean_bytes(str(tmp_path / 'fetch'), 'FETCH_SIZE', 'k_cgs_iter')
    assert (n, idle) == (3, 2)
    assert abs(mean - 1024 * 5000.0) < 1e-6


def test_trace_summary_splits_exits_at_entry(tmp_path):
    d = tmp_path / 'trace'
    os.makedirs(d)
    durs = {'void wcfd::k_cgs_iter2<E, false, false>': [8.0, 8.5, 9.0, 1.6, 1.5, 8.2, 8.4, 8.8],
            'void wcfd::k_cgs_iter2<E, false, true>': [9.0, 1.7],
            'void wcfd::k_bi_1<E>': [14.0]}
    with open(d / 'run_kernel_trace.csv', 'w', newline='') as f:
        w = csv.DictWriter(f, fieldnames=['Kernel_Name', 'Start_Timestamp', 'End_Timestamp'])
        w.writeheader()
        t = 1000
        for k, ts in durs.items():
            for us in ts:
                w.writerow({'Kernel_Name': k, 'Start_Timestamp': t,
                            'End_Timestamp': t + int(round(us * 1000))})
                t += 20000
    s = pmc.trace_summary(str(d), 'k_cgs_iter')
    assert s['dispatches'] == 10 and s['working'] == 7 and s['exit_at_entry'] == 3
    assert abs(s['working_avg_us'] - (8.0 + 8.5 + 9.0 + 8.2 + 8.4 + 8.8 + 9.0) / 7) < 1e-3
    inst = s['instances']['void wcfd::k_cgs_iter2<E, false, true>']
    assert (inst['working'], inst['exit_at_entry']) == (1, 1)
    assert abs(s['avg_us'] - sum(durs['void wcfd::k_cgs_iter2<E, false, false>']
                                 + durs['void wcfd::k_cgs_iter2<E, false, true>']) / 10) < 1e-3


def test_overlap_summary_counts_concurrent_kernels(tmp_path):
    """Three dispatches, [0, 10], [5, 15] and [20, 30] ns: 30 ns of kernel
    time in a 25 ns union of busy intervals, 1.2 kernels running on average
    whenever one runs; a second trace file in another folder is merged."""
    import csv
    from warpcfd_amd.pmc import overlap_summary
    for sub, ivs in (('a', [(0, 10), (20, 30)]), ('b', [(5, 15)])):
        d = tmp_path / sub
        d.mkdir()
        with open(d / 'run_kernel_trace.csv', 'w', newline='') as f:
            w = csv.writer(f)
            w.writerow(['Kernel_Name', 'Start_Timestamp', 'End_Timestamp'])
            for b, e in ivs:
                w.writerow(['k', b, e])
    s = overlap_summary(str(tmp_path))
    assert s['dispatches'] == 3
    assert s['mean_concurrent'] == 1.2
    # serial kernels: the ratio is exactly 1
    d = tmp_path / 'c'
    d.mkdir()
    with open(d / 'run_kernel_trace.csv', 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(['Kernel_Name', 'Start_Timestamp', 'End_Timestamp'])
        for b, e in [(100, 110), (120, 125)]:
            w.writerow(['k', b, e])
    assert overlap_summary(str(d))['mean_concurrent'] == 1.0
