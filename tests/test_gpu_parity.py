"""HIP step engine vs the CPU oracle engine on identical inputs (needs an MI355X).

Without a source term the device kernels perform exactly the float32 operations of the
reference (boundaries.py:186-268, boundaries3d.py:240-362 in the same order, no FMA
contraction), so fields, psi-driven updates, TM/3D loss, PEC, the stitch overlap
imposition (complex products in NumPy's fused multiply-add form) and every
float32/complex64 monitor must agree BIT FOR BIT.  Detector box
sums are float64 reductions in a different order (relative 1e-12).  The float32 source
phase uses the device sinf against NumPy's float32 sin (<= 2 ulp apart), hence a 1e-6
relative-L2 bar for runs with a source.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, fill_phasors, comp_shape, E_COMPS, H_COMPS
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
from oracle.engine import OracleEngine

pytestmark = pytest.mark.gpu

DIMS = {'te2d': (37, 70), 'tm2d': (41, 66), 'yee3d': (19, 23, 70)}


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


def _setup(eng, kind, dims, rng, lossy=True, source=False, drive=False, pec=True):
    coef = {}
    if kind == 'te2d':
        coef = dict(ceE=rng.uniform(0.8, 1.0, dims), ceH=rng.uniform(0.05, 0.2, dims))
    elif kind == 'tm2d':
        coef = dict(cax=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ex')),
                    cay=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ey')))
        if lossy:
            coef.update(dax=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ex')),
                        day=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ey')))
    else:
        for c, n in (('Ex', 'x'), ('Ey', 'y'), ('Ez', 'z')):
            coef['ce' + n] = rng.uniform(0.05, 0.2, comp_shape(kind, dims, c))
            if lossy:
                coef['da' + n] = rng.uniform(0.9, 1.0, comp_shape(kind, dims, c))
    coef = {k: v.astype(np.float32) for k, v in coef.items()}
    eng.set_coeffs(np.float32(0.4), **coef)
    dx, dt = 50e-9, 0.5 * 50e-9 / (3e8 * np.sqrt(3))
    npml = [(5, 3), (4, 6), (3, 3)]
    for ax, N in enumerate(dims):
        lo, hi = npml[ax]
        e, h = cpml_profile_1d(N, lo, hi, dx, dt, lam_m=850e-9)
        eng.set_cpml(ax, e, h, pec and ax != 1, pec)
    eng.finalize()
    for c in E_COMPS[kind] + H_COMPS[kind]:
        eng.write_field(c, rng.standard_normal(comp_shape(kind, dims, c)).astype(np.float32))
    comp = {'te2d': 'Ez', 'tm2d': 'Ey', 'yee3d': 'Ez'}[kind]
    plane = comp_shape(kind, dims, comp)[1:]
    if source:
        eng.set_source(comp, 7, rng.uniform(0.1, 1, plane), rng.uniform(-3, 3, plane),
                       rng.uniform(-3, 3, plane))
    if drive:
        O = 6
        w = (0.5 * (1 + np.cos(np.pi * np.arange(O) / (O - 1)))).astype(np.float32)
        comps = ['Ex', 'Ey', 'Ez'] if kind == 'yee3d' else [comp]
        D = {c: (rng.standard_normal((O,) + comp_shape(kind, dims, c)[1:])
                 + 1j * rng.standard_normal((O,) + comp_shape(kind, dims, c)[1:])).astype(np.complex64)
             for c in comps}
        eng.set_drive(O, w, D, 'c64' if kind == 'yee3d' else 'c128')
    mons = {}
    if kind == 'yee3d':
        mons['box'] = eng.add_box(((3, 9), (2, 11), (4, -4)))
        mons['sq_s'] = eng.add_sq('3d_s', dims[0])
        mons['sq_p'] = eng.add_sq('3d_p', dims[0] - 3)
        mons['dft_p'] = eng.add_dft('Ez', 0, dims[0], acc='c64p')
        mons['dft_q'] = eng.add_dft('Ey', 2, 11, acc='c64', phasor='q')
        mons['dft_pl'] = eng.add_dft('Ez', 0, dims[0], kplane=dims[2] // 2, acc='c64p')
    else:
        mons['box'] = eng.add_box(((4, 20), (10, 30)))
        mons['sq'] = eng.add_sq('te' if kind == 'te2d' else 'tm')
        mons['dft_p'] = eng.add_dft(comp, 0, dims[0], acc='c64p')
        mons['dft_c'] = eng.add_dft(comp, 3, 15, acc='c64', phasor='p')
        if kind == 'tm2d':
            mons['ex2'] = eng.add_sq('tm_ex2')
            mons['exn'] = eng.add_dft('ex_node', 0, dims[0], acc='c64p')
            mons['eyn'] = eng.add_dft('ey_node', 0, dims[0], acc='c64')
    mons['flux'] = eng.add_flux_plane(dims[0] // 2)
    return mons


def _run(eng, nsteps, rng_seed=3):
    tab = new_steps(nsteps)
    t = np.arange(nsteps) * 1e-16
    env = 1 - np.exp(-(t / 4e-15) ** 2)
    fill_phasors(tab, 2 * np.pi * 3.5e14, 1e-16, t, env)
    r = np.random.default_rng(rng_seed)
    tab['mask'] = r.integers(0, 2 ** 12, nsteps).astype(np.uint32) | 1
    eng.run(tab)


def _compare(kind, ho, oo, mons, exact=True, tol=1e-6):
    for c in E_COMPS[kind] + H_COMPS[kind]:
        a, b = ho.read_field(c), oo.read_field(c)
        if exact:
            np.testing.assert_array_equal(a, b, err_msg=f'{kind} field {c}')
        else:
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert rel <= tol, (kind, c, rel)
    for name, mid in mons.items():
        a, b = ho.read_monitor(mid), oo.read_monitor(mid)
        if name.startswith('box'):
            btol = 1e-12 if exact else tol
            assert abs(a - b) <= btol * abs(b) + 1e-300, (kind, name, a, b)
        elif name.startswith('flux'):
            for (ae, ah), (be, bh) in zip(a, b):
                if exact:
                    np.testing.assert_array_equal(ae, be)
                    np.testing.assert_array_equal(ah, bh)
                else:
                    assert np.linalg.norm(ae - be) <= tol * np.linalg.norm(be)
        elif exact:
            np.testing.assert_array_equal(a, b, err_msg=f'{kind} monitor {name}')
        else:
            assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b), (kind, name)


@pytest.mark.parametrize('kind,variant', [('te2d', 0), ('tm2d', 0), ('tm2d', 1), ('yee3d', 0),
                                          ('yee3d', 1)])
def test_kernels_bit_exact_without_source(kind, variant):
    H = _hip()
    dims = DIMS[kind]
    ho, oo = H(kind, dims, 'ref32'), OracleEngine(kind, dims, 'ref32')
    ho.set_variant(variant, lx=5)
    mh = _setup(ho, kind, dims, np.random.default_rng(11))
    mo = _setup(oo, kind, dims, np.random.default_rng(11))
    _run(ho, 25)
    _run(oo, 25)
    _compare(kind, ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('kind', ['te2d', 'tm2d', 'yee3d'])
def test_stitch_drive_bit_exact(kind):
    H = _hip()
    dims = DIMS[kind]
    ho, oo = H(kind, dims, 'ref32'), OracleEngine(kind, dims, 'ref32')
    mh = _setup(ho, kind, dims, np.random.default_rng(5), drive=True, pec=False)
    _setup(oo, kind, dims, np.random.default_rng(5), drive=True, pec=False)
    _run(ho, 12)
    _run(oo, 12)
    _compare(kind, ho, oo, mh)
    ho.close()


@pytest.mark.parametrize('kind', ['te2d', 'tm2d', 'yee3d'])
def test_source_ref32_and_fp64_phase(kind):
    H = _hip()
    dims = DIMS[kind]
    for mode, tol in (('ref32', 1e-6), ('fp64acc', 1e-6)):
        ho, oo = H(kind, dims, mode), OracleEngine(kind, dims, mode)
        mh = _setup(ho, kind, dims, np.random.default_rng(7), source=True)
        _setup(oo, kind, dims, np.random.default_rng(7), source=True)
        _run(ho, 30)
        _run(oo, 30)
        _compare(kind, ho, oo, mh, exact=False, tol=tol)
        ho.close()
