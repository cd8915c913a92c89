"""Builds liblensfdtd.so (HIP, gfx950) in-tree with hipcc.

Translation units: csrc/lensfdtd.hip (step kernels, monitors, batch engine, C ABI),
csrc/ops.hip (operator-level CPML updates on caller-owned device arrays),
csrc/reduce.hip (device steady-state reductions),
csrc/tb2d_pass.hip and csrc/tb2d_mon.hip (plain and monitored 2D blocked passes, once per
polarisation each) and
csrc/yee_tb3.hip (the temporally blocked 3D pass, once per depth / loss variant); they
compile in parallel into lens_sensor_amd/_build/ and link into one library.

Flags: -ffp-contract=off keeps the reference's float32 operation order (no fused
multiply-add), -Xclang -target-feature -Xclang -packed-fp32-ops keeps the compiler from
emitting packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
CSRC = os.path.join(PKG, 'csrc')
SRC = os.path.join(CSRC, 'lensfdtd.hip')
SRC_TB3 = os.path.join(CSRC, 'yee_tb3.hip')
SRC_TB2M = os.path.join(CSRC, 'tb2d_mon.hip')
SRC_TB2P = os.path.join(CSRC, 'tb2d_pass.hip')
SRC_OPS = os.path.join(CSRC, 'ops.hip')
SRC_RED = os.path.join(CSRC, 'reduce.hip')
LIB = os.path.join(PKG, 'liblensfdtd.so')
BUILD_DIR = os.path.join(PKG, '_build')
ARCH = os.environ.get('LENSFDTD_ARCH', 'gfx950')
FLAGS = [f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
TB3_VARIANTS = [(1, 0), (1, 1), (2, 0), (2, 1), (3, 0), (3, 1)]       # (depth, loss factors present)


def _jobs():
    jobs = [(['hipcc'] + FLAGS + ['-c', SRC, '-o', os.path.join(BUILD_DIR, 'lensfdtd.o')],
             os.path.join(BUILD_DIR, 'lensfdtd.o'))]
    obj = os.path.join(BUILD_DIR, 'ops.o')
    jobs.append((['hipcc'] + FLAGS + ['-c', SRC_OPS, '-o', obj], obj))
    obj = os.path.join(BUILD_DIR, 'reduce.o')
    jobs.append((['hipcc'] + FLAGS + ['-c', SRC_RED, '-o', obj], obj))
    for tm in (0, 1):
        obj = os.path.join(BUILD_DIR, f'tb2d_mon_{tm}.o')
        jobs.append((['hipcc'] + FLAGS + [f'-DLF_TB2M_TM={tm}', '-c', SRC_TB2M, '-o', obj], obj))
        obj = os.path.join(BUILD_DIR, f'tb2d_pass_{tm}.o')
        jobs.append((['hipcc'] + FLAGS + [f'-DLF_TB2P_TM={tm}', '-c', SRC_TB2P, '-o', obj], obj))
    for T, L in TB3_VARIANTS:
        obj = os.path.join(BUILD_DIR, f'yee_tb3_{T}_{L}.o')
        jobs.append((['hipcc'] + FLAGS + [f'-DLF_TB3_T={T}', f'-DLF_TB3_LOSS={L}', '-c', SRC_TB3,
                                          '-o', obj], obj))
    return jobs


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = _jobs()
    procs = [(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True),
              cmd) for cmd, _ in jobs]
    failed = []
    for p, cmd in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(' '.join(cmd) + '\n' + out)
        raise RuntimeError(f'hipcc failed on {len(failed)} translation unit(s)')
    objs = [obj for _, obj in jobs]
    link = ['hipcc', f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', LIB + '.tmp'] + objs
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f'hipcc link failed ({r.returncode})')
    os.replace(LIB + '.tmp', LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
