"""Builds liblensfdtd.so (HIP, gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps the reference's float32 operation order (no fused
multiply-add), -Xclang -target-feature -Xclang -packed-fp32-ops keeps the compiler from
emitting packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
SRC = os.path.join(PKG, 'csrc', 'lensfdtd.hip')
LIB = os.path.join(PKG, 'liblensfdtd.so')
ARCH = os.environ.get('LENSFDTD_ARCH', 'gfx950')


def build(verbose=False):
    cmd = ['hipcc', f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-shared',
           '-ffp-contract=off', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-o', LIB + '.tmp', SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f'hipcc failed ({r.returncode})')
    os.replace(LIB + '.tmp', LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
