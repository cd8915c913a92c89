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


@pytest.mark.parametrize('mode,dims', [('ref32', (40, 70, 200)), ('fp64acc', (40, 70, 200)),
                                       ('fp64acc', (72, 90, 150))])
def test_yee_blocked_fused_boxes_and_pass_end_monitors(mode, dims):
    """Detector boxes on every step are summed inside the pass; the intensity and phasor
    monitors of a pass's last step run after it.  Fields, intensity and phasors equal
    depth 1 bit for bit, the FP64 box sums to 1e-12 (another summation order).  The second
    box lies in the slab-free region, away from the source and drive planes."""
    H = _hip()
    print(f'[3d blocked fused boxes] mode={mode} dims={dims}', flush=True)
    a, b = H('yee3d', dims, mode), H('yee3d', dims, mode)
    a.set_temporal(3, 16)
    b.set_temporal(1, 64)
    ma = _setup(a, 'yee3d', dims, np.random.default_rng(5), source=True, drive=True, pec=False)
    _setup(b, 'yee3d', dims, np.random.default_rng(5), source=True, drive=True, pec=False)
    box2 = ((dims[0] // 2 - 8, dims[0] // 2 + 9), (30, 47), (10, -10))
    ma['box2'] = a.add_box(box2)
    b.add_box(box2)
    n = 30
    tab = _table(n, 0, n)
    mask = np.zeros(n, np.uint32)
    mask[:] |= (1 << ma['box']) | (1 << ma['box2'])    # boxes every step
    mask[2::3] |= 1 << ma['sq_s']                  # intensity on every pass's last step
    mask[5::6] |= (1 << ma['dft_pl']) | (1 << ma['dft_q'])
    tab['mask'] = mask
    a.set_timing(True)
    a.run(tab)
    b.run(tab)
    ms, launches, d = a.tb_stats()
    assert launches == n // 3
    _compare('yee3d', a, b, ma)
    a.close()
    b.close()
