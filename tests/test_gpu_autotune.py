"""Autotuner of the temporal blocking (needs an MI355X).

lf_autotune times every candidate depth / x chunk on the live engine and restores every
field and psi slab afterwards, so an engine tuned in the middle of a run continues
BIT FOR BIT like an untuned one (the blocked passes agree with each other exactly).
"""
import numpy as np
import pytest

from test_gpu_parity import _setup, _compare
from test_gpu_temporal import _table

pytestmark = pytest.mark.gpu

CHOICES = {'tm2d': ({2, 4, 6, 8}, {32, 64, 128}), 'te2d': ({2, 4, 6, 8}, {32, 64, 128}),
           'yee3d': ({2, 3}, {32, 48, 64, 96})}


@pytest.mark.parametrize('kind,dims', [('tm2d', (300, 200)), ('te2d', (260, 180)),
                                       ('yee3d', (60, 70, 150))])
def test_autotune_keeps_the_state(kind, dims):
    from lens_sensor_amd.engine.hip import HipEngine
    print(f'[autotune] {kind} {dims}', flush=True)
    a, b = HipEngine(kind, dims, 'fp64acc'), HipEngine(kind, dims, 'fp64acc')
    ma = _setup(a, kind, dims, np.random.default_rng(6), source=True, pec=False)
    _setup(b, kind, dims, np.random.default_rng(6), source=True, pec=False)
    tab = _table(48, 0, 48)
    a.run(tab[:12])
    b.run(tab[:12])
    depth, lx, lead, ms = a.autotune(tab[12], reps=2)
    depths, lxs = CHOICES[kind]
    assert depth in depths and lx in lxs and ms > 0
    assert lead in ((0,) if kind == 'yee3d' else (0, 1, 2, 3, 4))
    b.set_temporal(depth, lx)
    b.set_tb_lead(lead)
    a.run(tab[12:])
    b.run(tab[12:])
    _compare(kind, a, b, ma)
    a.close()
    b.close()


@pytest.mark.parametrize('kind', ['tm2d', 'te2d'])
def test_lead_class_is_bit_exact(kind):
    """Every launch order of the pass classes, and the serial order on one stream (4),
    gives the depth-1 result bit for bit (source and stitch drive inside the grid, so all
    four classes have work)."""
    from lens_sensor_amd.engine.hip import HipEngine
    dims = (300, 200)
    ref = HipEngine(kind, dims, 'fp64acc')
    mr = _setup(ref, kind, dims, np.random.default_rng(12), source=True, drive=True, pec=False)
    ref.set_temporal(1, 64)
    tab = _table(40, 0, 40)
    ref.run(tab)
    for lead in range(5):
        print(f'[lead class] {kind} lead={lead}', flush=True)
        e = HipEngine(kind, dims, 'fp64acc')
        _setup(e, kind, dims, np.random.default_rng(12), source=True, drive=True, pec=False)
        e.set_temporal(8, 32)
        e.set_tb_lead(lead)
        e.run(tab)
        _compare(kind, e, ref, mr)
        e.close()
    ref.close()


def test_lead_class_setter():
    from lens_sensor_amd.engine.hip import HipEngine
    e = HipEngine('tm2d', (20, 20), 'ref32')
    for bad in (-1, 5):
        with pytest.raises(RuntimeError):
            e.set_tb_lead(bad)
    e.close()
    e3 = HipEngine('yee3d', (12, 12, 12), 'ref32')
    with pytest.raises(RuntimeError):
        e3.set_tb_lead(2)
    e3.set_tb_lead(0)
    e3.close()


def test_tuned_3d_stitch_windows_equal_untuned(in_tmp):
    """run_3d_stitched tunes the blocking of each window shape (source window, driven
    windows) once; the tuned run gives the untuned run's results bit for bit."""
    from lens_sensor_amd.common import backend
    from lens_sensor_amd.planar_3d import fdtd3d_core as f3d
    case = dict(label='tune3d', wg_n=1.491, window_w_um=4.8, slide_um=2.4, lx_um=7.2,
                air_um=1.0, tear_um=2.0, lz_um=2.0, dx_nm=50.0, t_wg_um=1.0, t_aq_um=1.0,
                t_mu_um=0.5, n_snapshots=1, vcsel_waist_um=1.0, vcsel_waist_z_um=1.0,
                bead=dict(x_um=3.6, d_um=0.5, n=1.59), engine_mode='fp64acc')
    out = {}
    old = backend.AUTOTUNE
    try:
        for on in (True, False):
            backend.AUTOTUNE = on
            print(f'[3d stitch] autotune={on}', flush=True)
            out[on] = f3d.run_3d_stitched(**case)
    finally:
        backend.AUTOTUNE = old
    a, b = out[True], out[False]
    assert a['steps_total'] // a['n_windows'] >= backend.AUTOTUNE_MIN_STEPS
    assert a['transmission'] == b['transmission']
    np.testing.assert_array_equal(np.asarray(a['Iavg']), np.asarray(b['Iavg']))
    np.testing.assert_array_equal(np.asarray(a['vols_array']), np.asarray(b['vols_array']))
