"""Error of the device path against the FP64 oracle (needs an MI355X).

The float32 device stepping with FP64 phase and accumulators ('fp64acc', the mode of the
long runs) is compared with the oracle running every array in float64 ('f64') on the
same physical set-up: CPML, lossy coefficients, a CW Huygens source, 300 steps.  The
relative L2 error of the fields and of the CW phasor must stay <= 1e-4; the measured
values are printed (pytest -s) for the record.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, fill_phasors, comp_shape, E_COMPS, H_COMPS
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
from oracle.engine import OracleEngine

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _setup(eng, kind, dims, rng):
    dx = 25e-9
    dt = 0.5 * dx / (3e8 * np.sqrt(3 if kind == 'yee3d' else 2))
    if kind == 'te2d':
        coef = dict(ceE=rng.uniform(0.995, 1.0, dims), ceH=rng.uniform(0.2, 0.35, dims))
    elif kind == 'tm2d':
        coef = dict(cax=rng.uniform(0.2, 0.35, comp_shape(kind, dims, 'Ex')),
                    cay=rng.uniform(0.2, 0.35, comp_shape(kind, dims, 'Ey')),
                    dax=rng.uniform(0.995, 1.0, comp_shape(kind, dims, 'Ex')),
                    day=rng.uniform(0.995, 1.0, comp_shape(kind, dims, 'Ey')))
    else:
        coef = {'ce' + n: rng.uniform(0.15, 0.25, comp_shape(kind, dims, c))
                for c, n in (('Ex', 'x'), ('Ey', 'y'), ('Ez', 'z'))}
    eng.set_coeffs(np.float32(0.3 if kind != 'yee3d' else 0.25),
                   **{k: v.astype(np.float32) for k, v in coef.items()})
    for ax, N in enumerate(dims):
        e, h = cpml_profile_1d(N, 8, 8, dx, dt, lam_m=850e-9)
        eng.set_cpml(ax, e, h, True, True)
    eng.finalize()
    comp = {'te2d': 'Ez', 'tm2d': 'Ey', 'yee3d': 'Ez'}[kind]
    plane = comp_shape(kind, dims, comp)[1:]
    c = [np.linspace(-1, 1, n) for n in plane]
    amp = np.exp(-sum(np.meshgrid(*[x * x * 6 for x in c], indexing='ij'))).astype(np.float32)
    eng.set_source(comp, 11, amp, np.zeros(plane, np.float32), np.full(plane, 0.9, np.float32))
    mid = eng.add_dft(comp, 0, dims[0], acc='c128')
    return dt, mid


@pytest.mark.parametrize('kind,dims', [('te2d', (160, 120)), ('tm2d', (150, 130)),
                                       ('yee3d', (48, 40, 40))])
def test_device_error_against_fp64_oracle(kind, dims):
    from lens_sensor_amd.engine.hip import HipEngine
    dev, ref = HipEngine(kind, dims, 'fp64acc'), OracleEngine(kind, dims, 'f64')
    dt, md = _setup(dev, kind, dims, np.random.default_rng(3))
    _, mr = _setup(ref, kind, dims, np.random.default_rng(3))
    n = 300
    tab = new_steps(n)
    t = np.arange(n) * dt
    fill_phasors(tab, 2 * np.pi * 3.527e14, dt, t, 1 - np.exp(-(t / (25 * dt)) ** 2))
    tab['mask'][n - 96:] = 1
    dev.run(tab)
    ref.run(tab)
    errs = {}
    for c in E_COMPS[kind] + H_COMPS[kind]:
        a, b = dev.read_field(c).astype(np.float64), ref.read_field(c)
        errs[c] = np.linalg.norm(a - b) / np.linalg.norm(b)
    pa, pb = dev.read_monitor(md), ref.read_monitor(mr)
    errs['cw_phasor'] = np.linalg.norm(pa - pb) / np.linalg.norm(pb)
    print(kind, {k: f'{v:.2e}' for k, v in errs.items()})
    assert max(errs.values()) <= TOL, errs
    dev.close()
