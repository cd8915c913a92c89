"""Time-domain TF/SF stitch primitives on the HIP engine against the oracle (needs an
MI355X).

The reference's time-domain stitch (fdtd2d_core.py:1586-1863, fdtd3d_core.py:1785-2084)
adds, inside every step: records of E at the start of the step and of H after the H
update at one x plane, TF/SF corrections F +- c inc after the H and after the E update,
and per-step sums of the core intensity (2D TM: node Ey^2 only) and of the instantaneous
Poynting flux through a plane; the host tests max |F| to stop.  All of it is float32
elementwise arithmetic (plus FP64 sums), so the device agrees with the oracle BIT FOR BIT
on random fields, coefficients and series, in 2D TE, 2D TM (lossy) and 3D (CPML with its
PEC faces, and the sponge), for both precision modes.  Rows of the series and records
come from each step's `row` entry; steps whose row lies outside a series are refused.
Cases with the source run in fp64acc only: in ref32 the device's float32 sinf of the
source phase may differ from NumPy's float32 sin by 2 ulp.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import comp_shape, new_steps, E_COMPS, H_COMPS
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
from oracle.engine import OracleEngine

pytestmark = pytest.mark.gpu
ROWS = 16


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


def _plane(kind, dims, comp):
    return comp_shape(kind, dims, comp)[1:]


# per kind: TF/SF corrections (comp, plane, sign, coef, ranges as fractions of the plane:
# 'all' or 'inner'), records (comp, plane, phase), the driven component
TFSF = {
    'te2d': [('Hy', 11, -1, None, 'all'), ('Ez', 12, -1, 'ceH', 'inner')],
    'tm2d': [('Hz', 11, +1, None, 'all'), ('Ey', 12, +1, 'cay', 'all')],
    'yee3d': [('Hy', 11, -1, None, ('all', 'all')), ('Hz', 11, +1, None, ('all', 'all')),
              ('Ey', 12, +1, 'cey', ('all', 'inner')), ('Ez', 12, -1, 'cez', ('inner', 'all'))],
}
RECS = {'te2d': [('Ez', 30, 0), ('Hy', 29, 1)], 'tm2d': [('Ey', 30, 0), ('Hz', 29, 1)],
        'yee3d': [('Ey', 20, 0), ('Ez', 20, 0), ('Hy', 19, 1), ('Hz', 19, 1)]}
DIMS = {'te2d': (48, 40), 'tm2d': (48, 40), 'yee3d': (30, 22, 26)}


def _rng_of(kind, dims, comp, spec):
    shp = _plane(kind, dims, comp)
    specs = (spec,) if kind != 'yee3d' else spec
    return [(0, n) if s == 'all' else (1, n - 1) for s, n in zip(specs, shp)]


def _setup(eng, kind, seed, absorber, source):
    dims = DIMS[kind]
    rng = np.random.default_rng(seed)
    if kind == 'te2d':
        coef = dict(ceE=rng.uniform(0.8, 1.0, dims), ceH=rng.uniform(0.05, 0.2, dims))
    elif kind == 'tm2d':
        coef = dict(cax=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ex')),
                    cay=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ey')),
                    dax=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ex')),
                    day=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ey')))
    else:
        coef = {}
        for c, n in (('Ex', 'x'), ('Ey', 'y'), ('Ez', 'z')):
            coef['ce' + n] = rng.uniform(0.05, 0.2, comp_shape(kind, dims, c))
    eng.set_coeffs(np.float32(0.4), **{k: v.astype(np.float32) for k, v in coef.items()})
    dx, dt = 50e-9, 0.5 * 50e-9 / (3e8 * np.sqrt(3))
    for ax, N in enumerate(dims):
        npml = 0 if absorber == 'sponge' else 8
        e, h = cpml_profile_1d(N, npml, npml, dx, dt, lam_m=850e-9)
        eng.set_cpml(ax, e, h, absorber != 'sponge', absorber != 'sponge')
    if absorber == 'sponge':
        g = np.exp(-0.3 * ((6 - np.arange(6)) / 6) ** 3).astype(np.float32)
        eng.set_sponge(6, g, 63)
    eng.finalize()
    for c in E_COMPS[kind] + H_COMPS[kind]:
        eng.write_field(c, rng.standard_normal(comp_shape(kind, dims, c)).astype(np.float32))
    drv = {'te2d': 'Ez', 'tm2d': 'Ey', 'yee3d': 'Ez'}[kind]
    if source:
        plane = _plane(kind, dims, drv)
        eng.set_source(drv, 9, rng.uniform(0.1, 1, plane), rng.uniform(-3, 3, plane),
                       rng.uniform(-3, 3, plane))
    ids = []
    for comp, il, sign, cf, spec in TFSF[kind]:
        ids.append(eng.add_tfsf(comp, il, sign, cf, _rng_of(kind, dims, comp, spec), ROWS))
        eng.load_tfsf(ids[-1], rng.standard_normal((ROWS,) + _plane(kind, dims, comp))
                      .astype(np.float32))
    mons = {('rec', c, il): eng.add_rec(c, il, ph, ROWS) for c, il, ph in RECS[kind]}
    mons['sx_a'] = eng.add_sx(10)
    mons['sx_b'] = eng.add_sx(dims[0] - 10)
    if kind == 'yee3d':
        mons['box'] = eng.add_box(((25, 26), (4, 15), (6, -6)))
    else:
        mons['box'] = eng.add_box(((25, 26), (10, 30)), form=1 if kind == 'tm2d' else 0)
    return mons


def _table(n, mask, rows, src):
    tab = new_steps(n)
    t = np.arange(n) * 1e-16
    tab['env'] = 0.7 if src else 0.0
    tab['e32'] = np.float32(0.7 if src else 0.0)
    wt = 2.2e15 * t
    tab['wt32'] = wt.astype(np.float32)
    tab['ssin'] = tab['env'] * np.sin(wt)
    tab['scos'] = tab['env'] * np.cos(wt)
    tab['mask'] = mask
    tab['row'] = rows
    return tab


CASES = [('te2d', 'cpml', False), ('te2d', 'cpml', True), ('tm2d', 'cpml', False),
         ('tm2d', 'cpml', True), ('yee3d', 'cpml', False), ('yee3d', 'sponge', False),
         ('yee3d', 'cpml', True)]


@pytest.mark.parametrize('mode', ['ref32', 'fp64acc'])
@pytest.mark.parametrize('kind,absorber,source', CASES)
def test_td_primitives_bit_exact_vs_oracle(kind, absorber, source, mode):
    if source and mode == 'ref32':
        pytest.skip('ref32 source phase: the device sinf and NumPy float32 sin may differ by '
                    '2 ulp; the source cases run in fp64acc (phase from FP64 host tables)')
    print(f'[td] {kind} {absorber} source={source} {mode}', flush=True)
    dims = DIMS[kind]
    ho, oo = _hip()(kind, dims, mode), OracleEngine(kind, dims, mode)
    mh = _setup(ho, kind, 5, absorber, source)
    _setup(oo, kind, 5, absorber, source)
    mask = sum(1 << m for m in mh.values())
    rows = (np.arange(23) * 7) % ROWS                    # rows out of order, reused
    tab = _table(23, mask, rows, source)
    ho.run(tab)
    oo.run(tab)
    for c in E_COMPS[kind] + H_COMPS[kind]:
        np.testing.assert_array_equal(ho.read_field(c), oo.read_field(c), err_msg=f'{kind} {c}')
        assert ho.field_maxabs(c) == oo.field_maxabs(c)
    for key, mid in mh.items():
        a, b = ho.read_monitor(mid), oo.read_monitor(mid)
        if key[0] == 'rec':
            np.testing.assert_array_equal(a, b, err_msg=str(key))
        else:
            assert abs(a - b) <= 1e-12 * abs(b), (key, a, b)
    ho.close()


def test_td_rows_and_setup_are_checked():
    H = _hip()
    e = H('te2d', (40, 30), 'ref32')
    e.set_coeffs(np.float32(0.4), ceE=np.ones((40, 30), np.float32),
                 ceH=np.full((40, 30), 0.1, np.float32))
    for ax in (0, 1):
        ek, hk = cpml_profile_1d(40 if ax == 0 else 30, 5, 5, 50e-9, 1e-16, lam_m=850e-9)
        e.set_cpml(ax, ek, hk, 1, 1)
    e.finalize()
    tid = e.add_tfsf('Hy', 11, -1, None, [(0, 30)], 4)
    mons = e.add_rec('Ez', 20, 0, 4)
    with pytest.raises(RuntimeError):                    # row outside the series
        e.run(_table(2, 0, [0, 4], False))
    with pytest.raises(RuntimeError):                    # row outside the record
        e.run(_table(1, 1 << mons, [5], False))
    with pytest.raises(RuntimeError):                    # E corrections need coefficients
        e.add_tfsf('Ez', 12, -1, None, [(1, 29)], 4)
    with pytest.raises(RuntimeError):                    # range outside the plane
        e.add_tfsf('Hy', 11, -1, None, [(0, 31)], 4)
    with pytest.raises(RuntimeError):                    # more rows than the series holds
        e.load_tfsf(tid, np.zeros((5, 30), np.float32))
    from lens_sensor_amd.engine.hip import HipBatch
    b = HipBatch()
    with pytest.raises(RuntimeError):                    # runs alone, not in a batch
        b.add(e)
    b.close()
    e.close()
