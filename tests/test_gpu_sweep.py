"""Batched sweep vs the same instances solved one after another with one step per launch
(bit-identical), and vs the CPU oracle solving each instance (needs an MI355X).

The oracle leg runs the same Solve set-up (materials, coefficients, CPML, source, CW
window, flux planes) on OracleEngine.  In fp64acc the source phase comes from FP64
tables on both sides and the device's float32 field arithmetic is the oracle's, so the
transmission, fluxes and phasor norm agree to the fixture tolerance (1e-5 relative)."""
import numpy as np
import pytest

from lens_sensor_amd.planar_beads import sweep
from oracle.engine import OracleEngine

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


def test_batched_sweep_matches_the_oracle():
    # 50 nm cells keep the oracle leg to a few seconds while the wave still crosses the
    # domain (transmissions 0.55-0.89 over the 4 instances)
    ps = sweep.sweep_instances(4, seed=21, dx_nm=50.0, nx_range=(200, 300), ny_range=(260, 290))
    ro, _, _ = sweep.run_sweep(ps, batched=False, mode='fp64acc', steps_cap=1400,
                               engine_factory=OracleEngine)
    assert {p['pol'] for p in ps} == {'s', 'p'}      # TE and TM instances both checked
    for depth in (6, 8):
        rb, _, _ = sweep.run_sweep(ps, batched=True, mode='fp64acc', steps_cap=1400, depth=depth,
                                   lx=32)
        for p, a, b in zip(ps, rb, ro):
            assert a['steps'] == b['steps']
            for k in ('transmission', 'S_in', 'S_out', 'ecw_norm'):
                assert np.isclose(a[k], b[k], rtol=1e-5, atol=0.0), (depth, p, k, a[k], b[k])
