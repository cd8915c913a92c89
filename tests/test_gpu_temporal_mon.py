"""Monitored 2D blocked passes: the CW-window monitors summed inside the pass (needs an
MI355X).

Volume phasors (Ey, Ex, node Ey, node Ex; TE Ez) and the intensity sums ride with their
plane through the levels and receive exactly the additions of the one-step monitor
kernels in the same order; flux planes are staged per level and added in step order
after the pass.  So fields, psi and every accumulator agree BIT FOR BIT with depth 1
(and with the oracle where depth 1 does).  Step masks switch monitors on and off inside
a pass, the window starts in the middle of a pass, and small x chunks put chunk seams
next to the flux planes.
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


def _fusable(kind, mons, which):
    """Monitor subset the monitored pass sums (at most two phasors, one intensity sum)."""
    if kind == 'tm2d':
        sets = [('dft_p', 'exn', 'ex2', 'flux'), ('eyn', 'dft_c', 'sq', 'flux')]
    else:
        sets = [('dft_p', 'sq', 'flux'), ('dft_c', 'flux')]
    return {k: mons[k] for k in sets[which]}


def _mon_table(n, mons, window_from, seed=4):
    """Steps before window_from are monitor-free; from there on every step carries a
    random non-empty subset of `mons`."""
    tab = _table(n, 0, n)
    r = np.random.default_rng(seed)
    ids = np.array(sorted(mons.values()), np.uint32)
    mask = np.zeros(n, np.uint32)
    for t in range(window_from, n):
        pick = r.integers(0, 2, len(ids)).astype(bool)
        pick[r.integers(0, len(ids))] = True
        mask[t] = np.bitwise_or.reduce(np.uint32(1) << ids[pick])
    tab['mask'] = mask
    return tab


@pytest.mark.parametrize('kind,dims,dm,lx,which', [
    ('tm2d', (41, 66), 2, 5, 0), ('tm2d', (41, 66), 3, 64, 1), ('tm2d', (41, 66), 4, 7, 0),
    ('tm2d', (160, 250), 3, 16, 0), ('tm2d', (160, 250), 4, 64, 1), ('tm2d', (160, 250), 2, 9, 1),
    ('te2d', (37, 70), 3, 5, 0), ('te2d', (150, 240), 4, 16, 0), ('te2d', (150, 240), 2, 64, 1)])
def test_monitored_pass_bit_exact_vs_oracle(kind, dims, dm, lx, which):
    H = _hip()
    print(f'[monitored pass] {kind} dims={dims} depth={dm} lx={lx} set={which}', flush=True)
    ho, oo = H(kind, dims, 'ref32'), OracleEngine(kind, dims, 'ref32')
    mh = _setup(ho, kind, dims, np.random.default_rng(31))
    _setup(oo, kind, dims, np.random.default_rng(31))
    ho.set_temporal(4, lx)
    ho.set_temporal_mon(dm)
    sub = _fusable(kind, mh, which)
    tab = _mon_table(41, sub, 9)
    ho.run(tab)
    oo.run(tab)
    _compare(kind, ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('kind,mode,dm', [('tm2d', 'ref32', 3), ('tm2d', 'fp64acc', 4),
                                          ('tm2d', 'fp64acc', 2), ('te2d', 'ref32', 4),
                                          ('te2d', 'fp64acc', 3)])
def test_monitored_pass_source_drive_equal_depth1(kind, mode, dm):
    """Source and drive planes inside the marches; fp64acc sums in complex128."""
    H = _hip()
    dims = (150, 240)
    print(f'[monitored pass source/drive] {kind} {mode} depth={dm}', flush=True)
    a, b = H(kind, dims, mode), H(kind, dims, mode)
    ma = _setup(a, kind, dims, np.random.default_rng(8), source=True, drive=True, pec=False)
    _setup(b, kind, dims, np.random.default_rng(8), source=True, drive=True, pec=False)
    a.set_temporal(6 if kind == 'tm2d' else 4, 32)
    a.set_temporal_mon(dm)
    b.set_temporal(1, 64)
    b.set_temporal_mon(1)
    sub = _fusable(kind, ma, 0)
    tab = _mon_table(60, sub, 17, seed=9)
    a.run(tab)
    b.run(tab)
    _compare(kind, a, b, ma)
    a.close()
    b.close()


def test_monitored_pass_skipped_with_boxes():
    """A detector box on a step keeps that step on the one-step path; results equal depth 1."""
    H = _hip()
    dims = (160, 250)
    a, b = H('tm2d', dims, 'ref32'), H('tm2d', dims, 'ref32')
    ma = _setup(a, 'tm2d', dims, np.random.default_rng(2))
    _setup(b, 'tm2d', dims, np.random.default_rng(2))
    a.set_temporal_mon(3)
    b.set_temporal(1, 64)
    b.set_temporal_mon(1)
    sub = dict(_fusable('tm2d', ma, 0), box=ma['box'])
    tab = _mon_table(30, sub, 5)
    a.run(tab)
    b.run(tab)
    _compare('tm2d', a, b, ma)
    a.close()
    b.close()


def test_monitored_depth_setter():
    H = _hip()
    e = H('tm2d', (20, 20), 'ref32')
    for bad in (0, 5):
        with pytest.raises(RuntimeError):
            e.set_temporal_mon(bad)
    e.close()
    e3 = H('yee3d', (12, 12, 12), 'ref32')
    with pytest.raises(RuntimeError):
        e3.set_temporal_mon(2)
    e3.close()


@pytest.mark.parametrize('kind,dm,every', [('tm2d', 4, False), ('tm2d', 3, False), ('te2d', 4, False),
                                            ('tm2d', 4, True), ('tm2d', 2, True)])
def test_monitored_pass_complex128_variant_equal_depth1(kind, dm, every):
    """All phasors in complex128 (the stitch runs' fp64acc setup): the pass takes the kernel
    variant without accumulator-format tests; with every monitor on every step (the TM
    stitch CW window: Ey and node-Ex phasors, Ex^2) the variant with the layout fixed at
    compile time.  Accumulators equal depth 1 bit for bit."""
    H = _hip()
    dims = (150, 240)
    print(f'[monitored pass c128] {kind} depth={dm} every={every}', flush=True)
    engs = []
    for depth in (dm, 1):
        e = H(kind, dims, 'fp64acc')
        _setup(e, kind, dims, np.random.default_rng(13), source=True, drive=True, pec=False)
        comp = 'Ey' if kind == 'tm2d' else 'Ez'
        mons = dict(d0=e.add_dft(comp, 0, dims[0], acc='c128', phasor='p'),
                    flux=e.add_flux_plane(dims[0] // 3))
        if kind == 'tm2d':
            mons['d1'] = e.add_dft('ex_node', 0, dims[0], acc='c128', phasor='p')
            mons['sq'] = e.add_sq('tm_ex2')
        e.set_temporal(8 if kind == 'tm2d' else 4, 32)
        e.set_temporal_mon(depth)
        engs.append((e, mons))
    tab = _mon_table(50, engs[0][1], 13, seed=21)
    if every:
        tab['mask'][13:] = np.uint32(sum(1 << int(v) for v in engs[0][1].values()))
    for e, _ in engs:
        e.run(tab)
    (a, ma), (b, _) = engs
    _compare(kind, a, b, ma)
    a.close()
    b.close()
