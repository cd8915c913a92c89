"""Compact per-cell coefficients of a 2D grid (fdtd2d_core.window_coeffs) against the full
arrays of the reference's host path.

The full path is build_materials + _coeffs (TE) or _tm_coeffs + _apply_bead_tm + _tm_loss
(TM), pinned bit for bit against the reference's own arrays by test_setup_vectors.py.  The
compact form holds one row per layer-column class and the explicit rows the beads touch;
expanded, it must equal the full arrays BIT FOR BIT, on grids with several beads, loss in
the layers and in the beads, and beads near the top and bottom rows.
"""
import numpy as np
import pytest

from lens_sensor_amd.planar_beads import fdtd2d_core as d2

STACK = [dict(name='Aqueous', t_um=1.0, n=complex(1.331, 0.001), kappa=0.001, mat=None),
         dict(name='Mucin', t_um=0.4, n=complex(1.34, 0.0), kappa=0.0, mat=None)]


def _full(pol, args, kw, dt, dx):
    bp = {}
    eps, sig = d2.build_materials(*args, tm_patch=bp, **kw)
    if pol == 'p':
        cax, cay, ch = d2._tm_coeffs(eps, dt, dx)
        d2._apply_bead_tm(bp, cax, cay, dt, dx)
        dax, day = d2._tm_loss(eps, sig, cax, cay, dt)
        return dict(cax=cax, cay=cay, dax=dax, day=day), ch
    ceE, ceH = d2._coeffs(eps, sig, dt, dx)
    return dict(ceE=ceE, ceH=ceH), None


CASES = [
    # (pol, kappa, beads (x_um, r_um, y_um), x_start_um, stack)
    ('p', {'bead': 0.002}, [(1.2, 0.3, -0.3), (3.0, 0.5, -0.5), (4.1, 0.2, -0.9)], 0.0, STACK),
    ('s', {'bead': 0.002}, [(1.2, 0.3, -0.3), (3.0, 0.5, -0.5)], 0.0, STACK),
    ('p', {}, [(2.0, 0.4, -0.4)], 0.0, None),                    # lossless: no D arrays
    ('p', {'aq': 0.001}, [], 0.0, None),                           # no beads: tables only
    ('s', {}, [(2.5, 0.6, 1.2)], 1.0, None),                       # bead in the core, shifted window
    ('p', {'bead': 0.003}, [(0.5, 0.3, -4.95), (4.4, 0.3, 2.95)], 0.0, STACK),   # near the edges
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_compact_coefficients_expand_to_the_full_arrays(case):
    pol, kappa, blist, x0, stack = CASES[case]
    dx = 50e-9
    dt = 0.5 * dx / (d2.C0 * np.sqrt(2))
    old_k, old_tb = d2.KAPPA, d2.TEAR_BUFFER
    d2.KAPPA = dict(kappa)
    d2.TEAR_BUFFER = 5e-6
    try:
        Nx, Ny = 110, 172
        beads = [dict(x_um=x, d_um=2 * r, y_um=y, n=1.59) for x, r, y in blist] or None
        args = (Nx, Ny, dx, 1.49, 1.59, 1e-6, 2.0, 0.5 if beads is None else 0.0, x0,
                beads is not None)
        kw = dict(n_cornea=1.376, stack=stack, beads=beads, length_um=5.0, input_gap_um=0.7)
        want, ch_w = _full(pol, args, kw, dt, dx)
        mp = d2._material_parts(*args, **kw)
        packed, ch = d2.window_coeffs(pol, mp, dt, dx)
        assert ch == ch_w
        for name, arr in want.items():
            if arr is None:
                assert name not in packed
                continue
            got = d2.expand_coeffs(packed[name])
            assert got.dtype == np.float32 and got.shape == arr.shape, name
            np.testing.assert_array_equal(got.view(np.uint32), arr.view(np.uint32), err_msg=name)
        assert set(packed) == {k for k, v in want.items() if v is not None}
    finally:
        d2.KAPPA, d2.TEAR_BUFFER = old_k, old_tb


def test_compact_form_catches_a_planted_row_offset():
    """A band written one row off (a plausible indexing slip) no longer expands to the full
    array: the comparison above would fail."""
    dx = 50e-9
    dt = 0.5 * dx / (d2.C0 * np.sqrt(2))
    old_k, old_tb = d2.KAPPA, d2.TEAR_BUFFER
    d2.KAPPA = {'bead': 0.002}
    d2.TEAR_BUFFER = 5e-6
    try:
        args = (90, 172, dx, 1.49, 1.59, 1e-6, 2.0, 0.0, 0.0, True)
        kw = dict(n_cornea=1.376, stack=STACK, beads=[dict(x_um=2.0, d_um=0.8, y_um=-0.4, n=1.59)],
                  length_um=4.5)
        want, _ = _full('p', args, kw, dt, dx)
        packed, _ = d2.window_coeffs('p', d2._material_parts(*args, **kw), dt, dx)
        xc, tab, j0, band = packed['cax']
        assert not np.array_equal(d2.expand_coeffs((xc, tab, j0 + 1, band)), want['cax'])
    finally:
        d2.KAPPA, d2.TEAR_BUFFER = old_k, old_tb
