"""Build the HIP shared library ``libshwfs.so`` in-tree for gfx950.

Usage: ``python -m specula_amd.build`` (also called by ``__graft_entry__.build()``).
hipcc cross-compiles for gfx950 without a GPU present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'lib', 'libshwfs.so')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')

SOURCES = ['shwfs.hip', 'ccd.hip', 'gemm.hip']

# Packed-FP32 VALU ops are disabled: hipcc still vectorises float code into them at -O3
# otherwise, and some of their forms have produced silently wrong results on this hardware.
HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    return os.path.join(rocm, 'bin', 'hipcc')


def build(verbose=True):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', INCLUDE, '-I', CSRC] + srcs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-q' not in sys.argv)
