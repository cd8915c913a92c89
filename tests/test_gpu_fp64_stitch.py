"""Error of the benchmark's stitch path against the FP64 oracle (needs an MI355X).

The cropped copy of the 2D TM stitch workload (lens_sensor_amd.planar_beads.
stitch_bench.CROP: two windows of 1024 x 512 cells, 20000 steps per window, the lossy
stack, beads, CPML, phasor handoff and monitors of the full run) runs on the device at
the benchmark's blocking setting (depth 8, x chunk 128, fp64acc) and is compared with
the f64 oracle's result for the same case, written in the build container by
tests/golden/make_fp64_stitch.py (every 4th cell of the assembled phasors).

Bound: relative L2 <= 1e-4 for the CW phasors of Ey and node-averaged Ex, and for the
transmission.  The measured values are printed (pytest -s) for the record.
"""
import contextlib
import io
import os

import numpy as np
import pytest

from lens_sensor_amd.common import backend
from lens_sensor_amd.planar_beads import stitch_bench as sb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-4


def test_stitch_path_within_1e4_of_the_fp64_oracle(monkeypatch):
    from lens_sensor_amd.engine.hip import HipEngine
    ref = np.load(os.path.join(HERE, 'golden', 'fp64_tm_stitch_crop.npz'))

    def blocked(kind, dims, mode):
        e = HipEngine(kind, dims, mode)
        e.set_temporal(8, 128)
        return e
    monkeypatch.setattr(backend, 'AUTOTUNE', False)
    with backend.use_engine(blocked), contextlib.redirect_stdout(io.StringIO()):
        r = sb.run(engine_mode='fp64acc', return_efull=True, **sb.CROP)
    sx, sy = sb.CROP_SUB
    E = np.asarray(r['Efull'])
    assert E.shape == tuple(ref['efull_shape'])
    e_rel = np.linalg.norm(E[::sx, ::sy] - ref['efull']) / np.linalg.norm(ref['efull'])
    x = np.asarray(r['Ex_cw_amp'])[::sx, ::sy]
    x_rel = np.linalg.norm(x - ref['ex_amp']) / np.linalg.norm(ref['ex_amp'])
    t_rel = abs(r['transmission'] - float(ref['transmission'])) / abs(float(ref['transmission']))
    print(f'\nfp64 error of the stitch path: Ey phasor {e_rel:.2e}, node-Ex phasor {x_rel:.2e}, '
          f'transmission {t_rel:.2e}')
    assert e_rel <= TOL and x_rel <= TOL and t_rel <= TOL, (e_rel, x_rel, t_rel)
