"""The wavefront plan of the temporally blocked 2D passes, checked on the CPU.

tb2d_plan.h is host code; tests/plan_check.cpp (built here with g++) runs the plan for
the grids the drivers and tests use, at depths 2-8 and x chunks 3-128, and checks that
every owned cell is advanced by exactly one wavefront, that the slab-free bodies only
meet planes and lanes inside the grid and outside the CPML slabs, and that a rectangle
without excitation code meets no source or drive plane.  The device tests
(test_gpu_temporal*.py) then check the values the plan's kernels compute.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    from lens_sensor_amd.planar_beads import sweep
    rows = []
    for p in sweep.sweep_instances(64, seed=2024)[:16]:                # bench sweep grids
        rows.append((p['nx'], p['ny'], 24, p['nx'] - 24, 24, p['ny'] - 24, 26, 0))
    rows += [(8220, 1536, 24, 8196, 24, 1512, 26, 0),                 # stitch window 0
             (8220, 1536, 0, 8196, 24, 1512, -1, 2048),               # inner windows (drive)
             (8192, 1536, 0, 8168, 24, 1512, -1, 2048)]               # last window
    for nx, ny in [(41, 66), (160, 250), (37, 70), (150, 240), (300, 200), (20, 20), (9, 300)]:
        rows += [(nx, ny, 5, nx - 3, 4, ny - 6, 7, 0), (nx, ny, 0, nx - 3, 4, ny - 6, -1, 6),
                 (nx, ny, 0, nx, 0, ny, 3, 0), (nx, ny, 5, nx - 3, 4, ny - 6, nx // 2, 6)]
    return rows


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs g++')
def test_tb_plan_covers_every_cell_once(tmp_path):
    exe = str(tmp_path / 'plan_check')
    subprocess.run(['g++', '-O2', '-std=c++17', os.path.join(HERE, 'plan_check.cpp'), '-o', exe],
                   check=True)
    inp = '\n'.join(' '.join(map(str, r)) for r in _cases()) + '\n'
    r = subprocess.run([exe], input=inp, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert ' 0 problems' in r.stdout
