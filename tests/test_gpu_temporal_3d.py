"""3D temporally blocked pass (several Yee steps per launch) against the oracle and the
one-step-per-launch HIP path (needs an MI355X).

The blocked pass visits cells in another order and recomputes tile halos, but performs
exactly the float32 operations of yee_h_kernel + yee_e_kernel per cell and step, so
fields, psi and every monitor agree BIT FOR BIT with the oracle when no float32 source
phase is involved, and with depth 1 on the device otherwise.  The grids are large
enough that every tile class occurs (slab-free tiles and tiles meeting z / y / x slabs,
edges and their combinations); small x chunks put chunk seams inside the slabs.
"""
import numpy as np
import pytest

from oracle.engine import OracleEngine
from test_gpu_parity import _setup, _compare
from test_gpu_temporal import _table

pytestmark = pytest.mark.gpu


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


@pytest.mark.parametrize('dims,depth,lx', [((19, 23, 70), 2, 5), ((19, 23, 70), 3, 64),
                                            ((30, 34, 150), 3, 7), ((30, 34, 150), 2, 16),
                                            ((26, 41, 140), 2, 3), ((26, 41, 140), 3, 64),
                                            ((40, 70, 200), 3, 9), ((40, 70, 200), 2, 32)])
def test_yee_blocked_bit_exact_vs_oracle(dims, depth, lx):
    H = _hip()
    print(f'[3d blocked] dims={dims} depth={depth} lx={lx}', flush=True)
    ho, oo = H('yee3d', dims, 'ref32'), OracleEngine('yee3d', dims, 'ref32')
    ho.set_temporal(depth, lx)
    mh = _setup(ho, 'yee3d', dims, np.random.default_rng(23))
    _setup(oo, 'yee3d', dims, np.random.default_rng(23))
    tab = _table(37, 3, 33)                 # monitors on the first 3 and last 4 steps
    ho.set_timing(True)
    ho.run(tab)
    oo.run(tab)
    ms, launches, d = ho.tb_stats()
    assert d == depth and launches == 30 // depth
    _compare('yee3d', ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('mode', ['ref32', 'fp64acc'])
def test_yee_blocked_source_and_drive_equal_depth1(mode):
    H = _hip()
    dims = (30, 34, 150)
    print(f'[3d blocked source/drive] mode={mode}', flush=True)
    a, b = H('yee3d', dims, mode), H('yee3d', dims, mode)
    a.set_temporal(3, 6)
    b.set_temporal(1, 64)
    ma = _setup(a, 'yee3d', dims, np.random.default_rng(9), source=True, drive=True, pec=False)
    _setup(b, 'yee3d', dims, np.random.default_rng(9), source=True, drive=True, pec=False)
    tab = _table(26, 0, 24)
    a.run(tab)
    b.run(tab)
    _compare('yee3d', a, b, ma)
    a.close()
    b.close()


def test_yee_blocked_lossless_equal_oracle():
    H = _hip()
    dims = (28, 30, 140)
    print('[3d blocked lossless]', flush=True)
    ho, oo = H('yee3d', dims, 'ref32'), OracleEngine('yee3d', dims, 'ref32')
    ho.set_temporal(3, 9)
    mh = _setup(ho, 'yee3d', dims, np.random.default_rng(2), lossy=False)
    _setup(oo, 'yee3d', dims, np.random.default_rng(2), lossy=False)
    tab = _table(20, 0, 18)
    ho.run(tab)
    oo.run(tab)
    _compare('yee3d', ho, oo, mh)
    ho.close()


def test_yee_temporal_setter_rejects_bad_depth():
    H = _hip()
    e = H('yee3d', (12, 12, 12), 'ref32')
    with pytest.raises(RuntimeError):
        e.set_temporal(4, 8)
    e.close()
