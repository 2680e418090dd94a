"""Compile libwcfd.so for gfx950 with hipcc (in-tree, next to this file).

    python -m warpcfd_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ['csrc/fv_kernels.hip', 'csrc/krylov.hip', 'csrc/simple.hip', 'csrc/terms.hip',
           'csrc/profile.cpp']
OUT = os.path.join(HERE, 'libwcfd.so')
ARCH = os.environ.get('WCFD_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def build_library(verbose=False):
    """Build libwcfd.so; always recompiles.  Packed-FP32 instructions are
    disabled for device code (they are not needed by these kernels)."""
    tmp = OUT + '.tmp'
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-O3', '-fPIC', '-shared', '-std=c++17',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-o', tmp] + [os.path.join(HERE, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libwcfd.so')
    if verbose:
        sys.stdout.write(r.stdout)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build_library(verbose=True))
