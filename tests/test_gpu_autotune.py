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
    depth, lx, ms = a.autotune(tab[12], reps=2)
    depths, lxs = CHOICES[kind]
    assert depth in depths and lx in lxs and ms > 0
    b.set_temporal(depth, lx)
    a.run(tab[12:])
    b.run(tab[12:])
    _compare(kind, a, b, ma)
    a.close()
    b.close()
