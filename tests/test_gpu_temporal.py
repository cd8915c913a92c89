"""Temporally blocked passes (several Yee steps per launch) against the oracle and against
the one-step-per-launch HIP path (needs an MI355X).

Blocking changes only the order in which cells are visited, never the arithmetic of a
cell, so fields, psi and every monitor must agree BIT FOR BIT: with the oracle when no
float32 source phase is involved, and with depth 1 on the device otherwise (both use
the device sinf).  Small x chunks (lx) put chunk seams, CPML slabs, PEC edges, the source
planes and the stitch overlap inside the halo regions.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, fill_phasors, E_COMPS, H_COMPS
from oracle.engine import OracleEngine
from test_gpu_parity import _setup, _compare

pytestmark = pytest.mark.gpu


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


def _table(nsteps, quiet_from=0, quiet_to=None, seed=3):
    """Step table; steps outside [quiet_from, quiet_to) carry random monitor masks."""
    tab = new_steps(nsteps)
    t = np.arange(nsteps) * 1e-16
    env = 1 - np.exp(-(t / 4e-15) ** 2)
    fill_phasors(tab, 2 * np.pi * 3.5e14, 1e-16, t, env)
    r = np.random.default_rng(seed)
    mask = r.integers(0, 2 ** 12, nsteps).astype(np.uint32) | 1
    quiet_to = nsteps if quiet_to is None else quiet_to
    mask[quiet_from:quiet_to] = 0
    tab['mask'] = mask
    return tab


@pytest.mark.parametrize('dims,depth,lx', [((41, 66), 2, 5), ((41, 66), 4, 5), ((41, 66), 4, 64),
                                            ((41, 66), 6, 9), ((41, 66), 8, 3), ((41, 66), 8, 64),
                                            ((160, 250), 4, 16), ((160, 250), 8, 24)])
def test_tm_blocked_bit_exact_vs_oracle(dims, depth, lx):
    """The larger grid has wavefronts clear of every CPML slab (the psi-free variant)."""
    H = _hip()
    ho, oo = H('tm2d', dims, 'ref32'), OracleEngine('tm2d', dims, 'ref32')
    ho.set_temporal(depth, lx)
    mh = _setup(ho, 'tm2d', dims, np.random.default_rng(21))
    _setup(oo, 'tm2d', dims, np.random.default_rng(21))
    tab = _table(37, 3, 33)                 # monitors on the first 3 and last 4 steps
    ho.set_timing(True)
    ho.run(tab)
    oo.run(tab)
    ms, launches, d = ho.tb_stats()
    assert d == depth and launches == 30 // depth
    _compare('tm2d', ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('mode', ['ref32', 'fp64acc'])
@pytest.mark.parametrize('depth,lx', [(4, 7), (8, 16)])
def test_tm_blocked_source_and_drive_equal_depth1(mode, depth, lx):
    H = _hip()
    dims = (140, 150)
    a, b = H('tm2d', dims, mode), H('tm2d', dims, mode)
    a.set_temporal(depth, lx)
    b.set_temporal(1, 64)
    ma = _setup(a, 'tm2d', dims, np.random.default_rng(8), source=True, drive=True, pec=False)
    _setup(b, 'tm2d', dims, np.random.default_rng(8), source=True, drive=True, pec=False)
    tab = _table(50, 0, 48)
    a.run(tab)
    b.run(tab)
    _compare('tm2d', a, b, ma)
    a.close()
    b.close()


def test_tm_blocked_lossless_no_cpml_equal_oracle():
    """No loss arrays, PEC only on y: the lossless branch and slab-free columns."""
    H = _hip()
    dims = (90, 130)
    ho, oo = H('tm2d', dims, 'ref32'), OracleEngine('tm2d', dims, 'ref32')
    ho.set_temporal(6, 11)
    mh = _setup(ho, 'tm2d', dims, np.random.default_rng(4), lossy=False)
    _setup(oo, 'tm2d', dims, np.random.default_rng(4), lossy=False)
    tab = _table(25, 0, 24)
    ho.run(tab)
    oo.run(tab)
    _compare('tm2d', ho, oo, mh)
    ho.close()


def test_temporal_setter_rejects_bad_depth():
    H = _hip()
    e = H('tm2d', (20, 20), 'ref32')
    with pytest.raises(RuntimeError):
        e.set_temporal(3, 8)
    e.close()


@pytest.mark.parametrize('dims,depth,lx', [((37, 70), 2, 5), ((37, 70), 4, 9), ((37, 70), 8, 64),
                                            ((170, 240), 4, 16), ((170, 240), 6, 32),
                                            ((170, 240), 8, 24)])
def test_te_blocked_bit_exact_vs_oracle(dims, depth, lx):
    H = _hip()
    ho, oo = H('te2d', dims, 'ref32'), OracleEngine('te2d', dims, 'ref32')
    ho.set_temporal(depth, lx)
    mh = _setup(ho, 'te2d', dims, np.random.default_rng(31))
    _setup(oo, 'te2d', dims, np.random.default_rng(31))
    tab = _table(37, 3, 33)
    ho.set_timing(True)
    ho.run(tab)
    oo.run(tab)
    ms, launches, d = ho.tb_stats()
    assert d == depth and launches == 30 // depth
    _compare('te2d', ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('mode', ['ref32', 'fp64acc'])
def test_te_blocked_source_and_drive_equal_depth1(mode):
    H = _hip()
    dims = (150, 140)
    a, b = H('te2d', dims, mode), H('te2d', dims, mode)
    a.set_temporal(6, 12)
    b.set_temporal(1, 64)
    ma = _setup(a, 'te2d', dims, np.random.default_rng(12), source=True, drive=True, pec=False)
    _setup(b, 'te2d', dims, np.random.default_rng(12), source=True, drive=True, pec=False)
    tab = _table(50, 0, 48)
    a.run(tab)
    b.run(tab)
    _compare('te2d', a, b, ma)
    a.close()
    b.close()


@pytest.mark.parametrize('kind', ['tm2d', 'te2d'])
@pytest.mark.parametrize('source,drive', [(True, False), (False, True), (True, True)])
def test_blocked_excitation_variants_equal_depth1(kind, source, drive):
    """The passes take a body with only the source code, only the drive code, or both,
    by what the context carries; each equals depth 1 bit for bit."""
    H = _hip()
    dims = (150, 140)
    print(f'[excitation variant] {kind} source={source} drive={drive}', flush=True)
    a, b = H(kind, dims, 'fp64acc'), H(kind, dims, 'fp64acc')
    a.set_temporal(8, 16)
    b.set_temporal(1, 64)
    ma = _setup(a, kind, dims, np.random.default_rng(5), source=source, drive=drive, pec=False)
    _setup(b, kind, dims, np.random.default_rng(5), source=source, drive=drive, pec=False)
    tab = _table(42, 0, 40)
    a.run(tab)
    b.run(tab)
    _compare(kind, a, b, ma)
    a.close()
    b.close()
