"""Secondary absorbers on the HIP engine against the oracle (needs an MI355X).

The graded sponge (fdtd3d_core.py:397-414) and the first-order Mur faces
(boundaries3d.py:99-140; 2D inline fdtd2d_core.py:805-833) act after each step's E update
and source, in the reference's face order.  Their arithmetic is float32 products and
a + m (b - c) per cell, so the device agrees with the oracle BIT FOR BIT on random fields
(fp64acc: the source phase comes from FP64 host tables on both sides).  The cases cover
the PEC orderings the reference has: PEC faces applied after the Mur faces (run_3d's
mirror faces, the 2D PML edges between the x and the y Mur edges) and a seam CPML whose
PEC acts inside the step, before the Mur update (3D stitch windows).  A context in
blocked mode with absorbers runs one step per launch and equals depth 1.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import comp_shape, E_COMPS, H_COMPS
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
from oracle.engine import OracleEngine
from test_gpu_temporal import _table

pytestmark = pytest.mark.gpu


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


def _plane_shape(kind, dims, comp, axis):
    sh = comp_shape(kind, dims, comp)
    return tuple(n for d, n in enumerate(sh) if d != axis)


def _setup(eng, kind, dims, seed, npml, pec, absorb):
    """npml[ax] = (lo, hi) CPML cells, pec[ax] = (lo, hi) PEC ends, absorb = callable(eng, rng)
    adding the sponge / Mur faces."""
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
            coef['da' + n] = rng.uniform(0.9, 1.0, comp_shape(kind, dims, c))
    eng.set_coeffs(np.float32(0.4), **{k: v.astype(np.float32) for k, v in coef.items()})
    dx, dt = 50e-9, 0.5 * 50e-9 / (3e8 * np.sqrt(3))
    for ax, N in enumerate(dims):
        e, h = cpml_profile_1d(N, npml[ax][0], npml[ax][1], dx, dt, lam_m=850e-9)
        eng.set_cpml(ax, e, h, *pec[ax])
    absorb(eng, np.random.default_rng(seed + 1))
    eng.finalize()
    for c in E_COMPS[kind] + H_COMPS[kind]:
        eng.write_field(c, rng.standard_normal(comp_shape(kind, dims, c)).astype(np.float32))
    comp = {'te2d': 'Ez', 'tm2d': 'Ey', 'yee3d': 'Ez'}[kind]
    plane = comp_shape(kind, dims, comp)[1:]
    eng.set_source(comp, 9, rng.uniform(0.1, 1, plane), rng.uniform(-3, 3, plane),
                   rng.uniform(-3, 3, plane))
    mons = {'dft': eng.add_dft(comp, 0, dims[0], acc='c128')}
    if kind == 'yee3d':
        mons['box'] = eng.add_box(((3, 9), (2, 11), (4, -4)))
    else:
        mons['box'] = eng.add_box(((4, 20), (10, 30)))
    return mons


def _sponge(n, faces):
    def add(eng, rng):
        g = np.exp(-0.3 * ((n - np.arange(n)) / n) ** 3).astype(np.float32)
        eng.set_sponge(n, g, faces)
    return add


def _mur3d(faces, flags, dims):
    tang = {0: ('Ey', 'Ez'), 1: ('Ex', 'Ez'), 2: ('Ex', 'Ey')}

    def add(eng, rng):
        for ax, hi in faces:
            for c in tang[ax]:
                m = (np.float32(rng.uniform(-0.45, -0.35)) if ax == 0
                     else rng.uniform(-0.45, -0.35, _plane_shape('yee3d', dims, c, ax)).astype(np.float32))
                eng.add_mur(c, ax, hi, m, flags)
    return add


def _mur2d(kind, dims, x_lo, x_hi, y):
    cx, cy = ('Ez', 'Ez') if kind == 'te2d' else ('Ey', 'Ex')

    def add(eng, rng):
        ny = comp_shape(kind, dims, cy)[0]
        if x_lo:
            eng.add_mur(cx, 0, 0, np.float32(rng.uniform(-0.45, -0.35)), 3)
        if x_hi:
            eng.add_mur(cx, 0, 1, np.float32(rng.uniform(-0.45, -0.35)), 3)
        if y:
            eng.add_mur(cy, 1, 0, rng.uniform(-0.45, -0.35, ny).astype(np.float32), 1)
            eng.add_mur(cy, 1, 1, rng.uniform(-0.45, -0.35, ny).astype(np.float32), 1)
    return add


D3 = (30, 24, 40)
CASES = {
    # run_3d with the sponge on every face / on some faces with PEC mirrors on others
    'sponge_all': ('yee3d', D3, [(0, 0)] * 3, [(0, 0)] * 3, _sponge(6, 63)),
    'sponge_pec': ('yee3d', D3, [(0, 0)] * 3, [(0, 1), (1, 0), (0, 0)],
                   _sponge(5, 1 | 8 | 16 | 32)),
    # stitch window: seam CPML (with PEC) on xmax, sponge on the other faces
    'sponge_seam': ('yee3d', D3, [(0, 6), (0, 0), (0, 0)], [(0, 1), (0, 0), (0, 0)],
                    _sponge(5, 1 | 4 | 8 | 16 | 32)),
    # run_3d with Mur on every face; Mur with a PEC mirror on ymin and zmax (PEC after)
    'mur_all': ('yee3d', D3, [(0, 0)] * 3, [(0, 0)] * 3,
                _mur3d([(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)], 2, D3)),
    'mur_pec': ('yee3d', D3, [(0, 0)] * 3, [(0, 0), (1, 0), (0, 1)],
                _mur3d([(0, 0), (0, 1), (1, 1), (2, 0)], 2, D3)),
    # stitch window: seam CPML (PEC inside the step) on xmax, Mur on the rest, no PEC after
    'mur_seam': ('yee3d', D3, [(0, 6), (0, 0), (0, 0)], [(0, 1), (0, 0), (0, 0)],
                 _mur3d([(0, 0), (1, 0), (1, 1), (2, 0), (2, 1)], 0, D3)),
    # 2D full-domain Mur (x ends, then y edges)
    'te_mur': ('te2d', (60, 44), [(0, 0)] * 2, [(0, 0)] * 2, _mur2d('te2d', (60, 44), 1, 1, 1)),
    'tm_mur': ('tm2d', (60, 44), [(0, 0)] * 2, [(0, 0)] * 2, _mur2d('tm2d', (60, 44), 1, 1, 1)),
    # 2D stitch windows: seam CPML with PEC on x high, Mur on x low (window 0) and on y;
    # and a PEC y edge beside the Mur x ends (PEC between the x and the y updates)
    'te_mur_seam': ('te2d', (60, 44), [(0, 6), (0, 0)], [(0, 1), (0, 0)],
                    _mur2d('te2d', (60, 44), 1, 0, 1)),
    'tm_mur_seam': ('tm2d', (60, 44), [(0, 6), (0, 0)], [(0, 1), (0, 0)],
                    _mur2d('tm2d', (60, 44), 1, 0, 1)),
    'te_mur_ypec': ('te2d', (60, 44), [(0, 0), (0, 5)], [(0, 0), (0, 1)],
                    _mur2d('te2d', (60, 44), 1, 1, 0)),
}


@pytest.mark.parametrize('name', list(CASES))
def test_absorber_bit_exact_vs_oracle(name):
    kind, dims, npml, pec, absorb = CASES[name]
    print(f'[absorbers] {name} {kind} {dims}', flush=True)
    ho, oo = _hip()(kind, dims, 'fp64acc'), OracleEngine(kind, dims, 'fp64acc')
    mh = _setup(ho, kind, dims, 11, npml, pec, absorb)
    _setup(oo, kind, dims, 11, npml, pec, absorb)
    tab = _table(23, 0, 23)
    tab['mask'] = (1 << mh['dft']) | (1 << mh['box'])
    ho.run(tab)
    oo.run(tab)
    for c in E_COMPS[kind] + H_COMPS[kind]:
        np.testing.assert_array_equal(ho.read_field(c), oo.read_field(c), err_msg=f'{name} {c}')
    np.testing.assert_array_equal(ho.read_monitor(mh['dft']), oo.read_monitor(mh['dft']))
    a, b = ho.read_monitor(mh['box']), oo.read_monitor(mh['box'])
    assert abs(a - b) <= 1e-12 * abs(b), (name, a, b)
    ho.close()


@pytest.mark.parametrize('name,depth', [('sponge_seam', 3), ('mur_seam', 2), ('tm_mur', 8),
                                        ('te_mur_seam', 4)])
def test_absorbers_run_one_step_per_launch_in_blocked_mode(name, depth):
    kind, dims, npml, pec, absorb = CASES[name]
    print(f'[absorbers blocked mode] {name} depth {depth}', flush=True)
    a, b = _hip()(kind, dims, 'fp64acc'), _hip()(kind, dims, 'fp64acc')
    a.set_temporal(depth, 16)
    _setup(a, kind, dims, 4, npml, pec, absorb)
    _setup(b, kind, dims, 4, npml, pec, absorb)
    tab = _table(24, 0, 24)
    a.run(tab)
    b.run(tab)
    for c in E_COMPS[kind] + H_COMPS[kind]:
        np.testing.assert_array_equal(a.read_field(c), b.read_field(c), err_msg=f'{name} {c}')
    a.close()
    b.close()


def test_absorber_setters_reject_bad_input():
    H = _hip()
    e = H('te2d', (20, 20), 'ref32')
    with pytest.raises(RuntimeError):               # profiles of every axis come first
        e.add_mur('Ez', 0, 0, np.float32(-0.4), 1)
    e.close()
    e = H('yee3d', (12, 12, 12), 'ref32')
    with pytest.raises(RuntimeError):               # thicker than the grid
        e.set_sponge(20, np.ones(20, np.float32), 63)
    e.close()
    e = H('te2d', (20, 20), 'ref32')
    with pytest.raises(RuntimeError):               # the sponge is 3D only
        e.set_sponge(4, np.ones(4, np.float32), 15)
    e.close()
