"""Single-pass 3D Yee step (one launch per step for H and E) against the oracle and the
two-kernel path (needs an MI355X).

The single pass visits cells in another order but performs exactly the float32
operations of the H and E kernels, so fields, psi-driven updates, loss, PEC, the stitch
drive and every monitor agree BIT FOR BIT with the oracle without a float32 source phase,
and with the two-kernel device path with one.  The grids are large enough that every
wavefront class occurs: slab-free interior waves, waves touching z / y / x slabs or edges
and their combinations; small x chunks put chunk seams inside the slabs.
"""
import numpy as np
import pytest

from oracle.engine import OracleEngine
from test_gpu_parity import _setup, _compare, _run

pytestmark = pytest.mark.gpu


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


@pytest.mark.parametrize('dims,lx,rows', [((19, 23, 70), 5, 4), ((30, 34, 150), 7, 2),
                                          ((30, 34, 150), 16, 4), ((26, 41, 140), 3, 4)])
def test_single_pass_bit_exact_vs_oracle(dims, lx, rows):
    H = _hip()
    ho, oo = H('yee3d', dims, 'ref32'), OracleEngine('yee3d', dims, 'ref32')
    ho.set_variant(0, lx=lx)
    ho.set_tile(rows)
    mh = _setup(ho, 'yee3d', dims, np.random.default_rng(17))
    _setup(oo, 'yee3d', dims, np.random.default_rng(17))
    _run(ho, 14)
    _run(oo, 14)
    _compare('yee3d', ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('mode', ['ref32', 'fp64acc'])
def test_single_pass_source_and_drive_equal_two_kernel(mode):
    H = _hip()
    dims = (30, 34, 150)
    a, b = H('yee3d', dims, mode), H('yee3d', dims, mode)
    a.set_variant(0, lx=6)
    a.set_tile(2)
    b.set_variant(1)
    ma = _setup(a, 'yee3d', dims, np.random.default_rng(9), source=True, drive=True, pec=False)
    _setup(b, 'yee3d', dims, np.random.default_rng(9), source=True, drive=True, pec=False)
    _run(a, 16)
    _run(b, 16)
    _compare('yee3d', a, b, ma)
    a.close()
    b.close()


def test_single_pass_lossless():
    H = _hip()
    dims = (28, 30, 140)
    ho, oo = H('yee3d', dims, 'ref32'), OracleEngine('yee3d', dims, 'ref32')
    ho.set_variant(0, lx=8)
    mh = _setup(ho, 'yee3d', dims, np.random.default_rng(2), lossy=False)
    _setup(oo, 'yee3d', dims, np.random.default_rng(2), lossy=False)
    _run(ho, 10)
    _run(oo, 10)
    _compare('yee3d', ho, oo, mh)
    ho.close()
