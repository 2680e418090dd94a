"""Compile libwcfd.so (float fields) and libwcfd64.so (double fields) for
gfx950 with hipcc, in-tree next to this file.

    python -m warpcfd_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ['csrc/fv_kernels.hip', 'csrc/krylov.hip', 'csrc/simple.hip', 'csrc/terms.hip',
           'csrc/profile.cpp', 'csrc/mesh.cpp']
OUT = os.path.join(HERE, 'libwcfd.so')
OUT64 = os.path.join(HERE, 'libwcfd64.so')
ARCH = os.environ.get('WCFD_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def _command(out, defines):
    return ([hipcc(), f'--offload-arch={ARCH}', '-O3', '-fPIC', '-shared', '-std=c++17',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops'] + defines
            + ['-o', out] + [os.path.join(HERE, s) for s in SOURCES])


def build_library(verbose=False):
    """Build both libraries (two hipcc processes at once); always recompiles.
    Packed-FP32 instructions are disabled for device code (they are not
    needed by these kernels).  Returns the two paths."""
    jobs = [(OUT, []), (OUT64, ['-DWCFD_REAL=double'])]
    procs = [(out, subprocess.Popen(_command(out + '.tmp', d), stdout=subprocess.PIPE,
                                    stderr=subprocess.STDOUT, text=True)) for out, d in jobs]
    failed = []
    for out, p in procs:
        log, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(log)
            failed.append(os.path.basename(out))
        elif verbose:
            sys.stdout.write(log)
    if failed:
        raise RuntimeError(f'hipcc failed building {", ".join(failed)}')
    for out, _ in jobs:
        os.replace(out + '.tmp', out)
    return OUT, OUT64


if __name__ == '__main__':
    print(build_library(verbose=True))
