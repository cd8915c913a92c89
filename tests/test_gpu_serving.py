"""Serving loop on the GPU: requests served together (2D through the batch engine, 3D
overlapped on their own streams) give the same result as each request alone."""
import numpy as np
import pytest

from lens_sensor_amd import serving

pytestmark = pytest.mark.gpu

REQS = [dict(id=0, arrival=0.0, kind='2d', dims=(700, 470), pol='s', bead_x_frac=0.5, bead_r_um=0.3,
             kappa_aq=0.001),
        dict(id=1, arrival=0.0, kind='2d', dims=(900, 480), pol='p', bead_x_frac=0.4, bead_r_um=0.5,
             kappa_aq=0.002),
        dict(id=2, arrival=0.0, kind='2d', dims=(520, 600), pol='p', bead_x_frac=0.6, bead_r_um=0.2,
             kappa_aq=0.0),
        dict(id=3, arrival=0.0, kind='3d', dims=(96, 128, 64), bead_x_frac=0.5, bead_r_um=0.3)]


def test_results_identical_alone_and_in_the_mix():
    mix, summ = serving.Server(round_steps=384).run(REQS, realtime=False)
    assert summ['requests'] == len(REQS)
    by_id = {r['id']: r for r in mix}
    for q in REQS:
        one, _ = serving.Server(round_steps=384).run([q], realtime=False)
        a, b = one[0], by_id[q['id']]
        assert a['steps'] == b['steps'] and a['converged'] == b['converged']
        assert np.array_equal([a['transmission'], a['change'], a['phasor_norm']],
                              [b['transmission'], b['change'], b['phasor_norm']], equal_nan=True)
