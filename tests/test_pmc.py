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
