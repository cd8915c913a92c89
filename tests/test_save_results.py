"""Result files: the package's drivers, called with exactly the keyword sets the
reference CLI passes (run_simulation._run_2d / _run_3d, run_simulation.py:290-431), run on
the CPU oracle and write their results through save_results / save_results_3d; the file
names and every NPZ key (dtype kind, number of dimensions) must be those the reference's
own writers produced for the same call, recorded in tests/golden/ref_save_keys.json by
tests/golden/make_save_keys.py.  PKL summaries must hold every key of the reference's."""
import contextlib
import io
import json
import os
import pickle

import numpy as np
import pytest

from lens_sensor_amd.common import backend
from lens_sensor_amd.planar_beads import fdtd2d_core as d2
from lens_sensor_amd.planar_3d import fdtd3d_core as f3d
from oracle.engine import OracleEngine

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, 'golden', 'ref_save_keys.json')))


def _written(out_dir):
    files = {}
    for name in sorted(os.listdir(out_dir)):
        path = os.path.join(out_dir, name)
        if name.endswith('.npz'):
            with np.load(path, allow_pickle=False) as z:
                files[name] = {k: [z[k].dtype.kind, int(z[k].ndim)] for k in z.files}
        elif name.endswith('.pkl'):
            with open(path, 'rb') as f:              # written by this test
                files[name] = sorted(pickle.load(f).keys())
    return files


@pytest.mark.parametrize('case', sorted(GOLD['cases']))
def test_result_files_have_the_reference_keys(case, in_tmp):
    ref = GOLD['cases'][case]
    kw = dict(ref['kwargs'])
    out_dir = str(in_tmp / 'results_out')
    with backend.use_engine(OracleEngine), contextlib.redirect_stdout(io.StringIO()):
        if case.startswith('2d'):
            r = d2.run_beads(**kw)
            r['layer_names'] = GOLD['names']
            r['n_lip'] = 1.48
            d2.save_results(r, out_dir=out_dir)
        elif 'pulse_nfreq' in kw:                     # the 3D Fourier pulse stitch
            r = f3d.run_3d_stitched_pulse(**kw)
            f3d.save_results_3d(r, out_dir=out_dir, prefix='planar3d', material_names=GOLD['names'])
        else:
            kw['pec_faces'] = tuple(kw['pec_faces'])
            r = f3d.run_3d(**kw)
            f3d.save_results_3d(r, out_dir=out_dir, prefix='planar3d', material_names=GOLD['names'])
    got = _written(out_dir)
    assert sorted(got) == sorted(ref['files']), (sorted(got), sorted(ref['files']))
    for name, want in ref['files'].items():
        if name.endswith('.npz'):
            assert got[name] == want, (name, {k: (got[name].get(k), v) for k, v in want.items()
                                              if got[name].get(k) != v},
                                       sorted(set(got[name]) - set(want)))
        else:
            missing = sorted(set(want) - set(got[name]))
            assert not missing, (name, missing)
