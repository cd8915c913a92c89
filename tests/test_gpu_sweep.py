"""Batched sweep vs the same instances solved one after another with one step per launch:
every solve must give the identical transmission and CW phasor (needs an MI355X)."""
import pytest

from lens_sensor_amd.planar_beads import sweep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('mode', ['fp64acc', 'ref32'])
def test_batched_sweep_equals_sequential(mode):
    ps = sweep.sweep_instances(6, seed=11, nx_range=(400, 900), ny_range=(520, 640))
    rb, tb, _ = sweep.run_sweep(ps, batched=True, mode=mode, steps_cap=900, depth=6, lx=32)
    rs, ts, _ = sweep.run_sweep(ps, batched=False, mode=mode, steps_cap=900)
    for a, b in zip(rb, rs):
        assert a['transmission'] == b['transmission']
        assert a['ecw_norm'] == b['ecw_norm']
        assert a['S_in'] == b['S_in'] and a['S_out'] == b['S_out']
