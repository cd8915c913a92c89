"""Coefficient arrays assembled on the device from their compact form (lf_build_coeff)
against the host's full arrays (needs an MI355X): the assembly only moves float32
values, so the arrays must agree BIT FOR BIT, on the stitch workload's window shape
(8220 x 1536, several beads, lossy aqueous and beads) and on small TE / TM grids."""
import numpy as np
import pytest

from lens_sensor_amd.planar_beads import fdtd2d_core as d2
from lens_sensor_amd.planar_beads import stitch_bench as sb
from test_window_coeffs import CASES, STACK, _full

pytestmark = pytest.mark.gpu


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


@pytest.mark.parametrize('pol', ['p', 's'])
def test_device_assembly_of_a_stitch_window(pol):
    H = _hip()
    dx = 25e-9
    dt = 0.5 * dx / (d2.C0 * np.sqrt(2))
    Ny = 1536
    old_k, old_tb = d2.KAPPA, d2.TEAR_BUFFER
    d2.KAPPA = {'bead': 0.002}
    d2.TEAR_BUFFER = (Ny * 0.025 - 2.0 - d2.AIR_BUFFER * 1e6) * 1e-6
    stack = [dict(name='Aqueous', t_um=3.0, n=complex(1.331, 0.001), kappa=0.001, mat=None),
             dict(name='Mucin', t_um=0.5, n=complex(1.34, 0.0), kappa=0.0, mat=None)]
    try:
        args = (8220, Ny, dx, 1.49, 1.59, 2e-6, 0.0, 0.0, 6144 * 0.025, False)
        kw = dict(n_cornea=1.376, stack=stack, beads=sb.beads_for(2024, 8192, 6144, 3),
                  length_um=(8192 + 2 * 6144) * 0.025, input_gap_um=0.5)
        want, _ = _full(pol, args, kw, dt, dx)
        packed, ch = d2.window_coeffs(pol, d2._material_parts(*args, **kw), dt, dx)
    finally:
        d2.KAPPA, d2.TEAR_BUFFER = old_k, old_tb
    print(f'[device coefficients] {pol} 8220x1536', flush=True)
    e = H('tm2d' if pol == 'p' else 'te2d', (8220, Ny), 'fp64acc')
    e.build_coeffs(np.float32(0.4) if ch is None else ch, packed)
    for name, arr in want.items():
        if arr is None:
            continue
        got = e.read_coeff(name)
        np.testing.assert_array_equal(got.view(np.uint32), arr.view(np.uint32), err_msg=name)
    e.close()


@pytest.mark.parametrize('case', range(len(CASES)))
def test_device_assembly_of_small_grids(case):
    H = _hip()
    pol, kappa, blist, x0, stack = CASES[case]
    dx = 50e-9
    dt = 0.5 * dx / (d2.C0 * np.sqrt(2))
    old_k, old_tb = d2.KAPPA, d2.TEAR_BUFFER
    d2.KAPPA = dict(kappa)
    d2.TEAR_BUFFER = 5e-6
    try:
        beads = [dict(x_um=x, d_um=2 * r, y_um=y, n=1.59) for x, r, y in blist] or None
        args = (110, 172, dx, 1.49, 1.59, 1e-6, 2.0, 0.5 if beads is None else 0.0, x0,
                beads is not None)
        kw = dict(n_cornea=1.376, stack=stack, beads=beads, length_um=5.0, input_gap_um=0.7)
        want, _ = _full(pol, args, kw, dt, dx)
        packed, ch = d2.window_coeffs(pol, d2._material_parts(*args, **kw), dt, dx)
    finally:
        d2.KAPPA, d2.TEAR_BUFFER = old_k, old_tb
    e = H('tm2d' if pol == 'p' else 'te2d', (110, 172), 'ref32')
    e.build_coeffs(np.float32(0.4) if ch is None else ch, packed)
    for name, arr in want.items():
        if arr is not None:
            np.testing.assert_array_equal(e.read_coeff(name).view(np.uint32), arr.view(np.uint32),
                                          err_msg=name)
    e.close()
