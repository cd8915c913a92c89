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


def _strip_range(c, first, last):
    """Slab bounds of one profile over [first, last]: positions < lo or >= hi carry psi
    (c != 0), as the library's strip_range finds them."""
    span = last - first + 1
    nlo = 0
    while nlo < span and c[first + nlo] != 0:
        nlo += 1
    nhi = 0
    while nhi < span and c[last - nhi] != 0:
        nhi += 1
    if nlo + nhi >= span:
        return len(c), len(c)
    return first + nlo, last + 1 - nhi


def _slabs(N, npml, dx, dt):
    """The plan's slab bounds of one axis: the tighter of the H-node strip (nodes
    0..N-2) and the E-node strip (nodes 1..N-2), as tb_plan_in combines them."""
    from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
    e, h = cpml_profile_1d(N, npml, npml, dx, dt, lam_m=850e-9)
    (hlo, hhi), (elo, ehi) = _strip_range(h[2], 0, N - 2), _strip_range(e[2], 1, N - 2)
    return max(hlo, elo), min(hhi, ehi)


def _cases():
    import numpy as np
    from lens_sensor_amd.planar_beads import sweep
    rows = []
    dx = 25e-9
    dt = 0.5 * dx / (299792458.0 * np.sqrt(2))
    for p in sweep.sweep_instances(64, seed=2024):                     # bench sweep grids
        xlo, xhi = _slabs(p['nx'], 24, dx, dt)
        ylo, yhi = _slabs(p['ny'], 24, dx, dt)
        rows.append((p['nx'], p['ny'], xlo, xhi, ylo, yhi, 26, 0))
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
