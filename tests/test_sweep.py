"""Sweep driver: instance generation and one solve per instance on the CPU oracle."""
import numpy as np

from lens_sensor_amd.planar_beads import sweep
from oracle.engine import OracleEngine


def test_sweep_parameters_follow_the_stated_ranges():
    ps = sweep.sweep_instances(64, seed=7)
    assert len(ps) == 64
    assert all(2048 <= p['nx'] <= 16384 and 512 <= p['ny'] <= 2048 for p in ps)
    assert {p['pol'] for p in ps} == {'s', 'p'}
    assert all(0.2 <= p['bead_r_um'] <= 0.6 and 0.0 <= p['kappa_aq'] <= 0.003 for p in ps)
    assert all(2.4 <= p['t_aq_um'] <= 3.6 and 0.4 <= p['t_mu_um'] <= 0.6 for p in ps)
    assert all(-p['t_aq_um'] + p['bead_r_um'] - 1e-9 <= p['bead_y_um'] <= -p['bead_r_um'] for p in ps)
    assert ps == sweep.sweep_instances(64, seed=7)          # seeded


def test_small_sweep_on_oracle_gives_finite_transmission():
    ps = sweep.sweep_instances(2, seed=3, nx_range=(160, 200), ny_range=(300, 330))
    res, _, _ = sweep.run_sweep(ps, batched=False, steps_cap=260,
                                engine_factory=lambda k, d, m: OracleEngine(k, d, m))
    for r in res:
        assert np.isfinite(r['transmission']) and r['S_in'] != 0.0
        assert r['steps'] == 260
