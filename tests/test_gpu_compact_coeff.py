"""Blocked 2D passes read the coefficient arrays that lf_build_coeff assembles in their
compact form (one table row per layer-column class, explicit band rows) straight from
that form; a context whose arrays were uploaded plainly reads them through one table row
per plane.  Both views hold the same float32 values, so one window stepped with its
coefficients built from a compact form and the same window with the assembled arrays
uploaded must agree BIT FOR BIT in every field (needs an MI355X).

The compact forms are random (classes drawn per plane, random table and band values, the
band at a random offset), so a wrong class row, band offset or plane index changes the
fields.  Grids and x chunks put slab-free chunks, x and y slabs, PEC edges and the
band's edges inside the blocked passes."""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import comp_shape, E_COMPS, H_COMPS
from test_gpu_parity import _setup
from test_gpu_temporal import _table

pytestmark = pytest.mark.gpu

_SHAPE = {'ceE': 'node', 'ceH': 'node', 'cax': 'Ex', 'cay': 'Ey', 'dax': 'Ex', 'day': 'Ey'}
_RANGE = {'ceE': (0.8, 1.0), 'ceH': (0.05, 0.2), 'cax': (0.05, 0.2), 'cay': (0.05, 0.2),
          'dax': (0.9, 1.0), 'day': (0.9, 1.0)}


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine
    return HipEngine


def _packed(kind, dims, rng, lossy, rows):
    names = ['ceE', 'ceH'] if kind == 'te2d' else ['cax', 'cay'] + (['dax', 'day'] if lossy else [])
    out = {}
    for name in names:
        ex, _, ez = (lambda s: (s[0], 1, s[1]))(comp_shape(kind, dims, _SHAPE[name]))
        lo, hi = _RANGE[name]
        ncls = 5
        xc = rng.integers(0, ncls, ex).astype(np.uint8)
        table = rng.uniform(lo, hi, (ncls, ez)).astype(np.float32)
        r = min(rows, ez)
        j0 = int(rng.integers(0, ez - r + 1))
        band = rng.uniform(lo, hi, (ex, r)).astype(np.float32)
        out[name] = (xc, table, j0, band)
    return out


@pytest.mark.parametrize('kind,dims,depth,lx,lossy,rows,src,drv', [
    ('tm2d', (41, 66), 8, 5, True, 9, True, False),
    ('tm2d', (300, 200), 8, 64, True, 40, False, True),
    ('tm2d', (300, 200), 6, 128, False, 0, True, False),
    ('te2d', (41, 66), 4, 7, False, 12, True, False),
    ('te2d', (300, 200), 8, 64, False, 64, False, True),
])
def test_compact_form_steps_like_plain_arrays(kind, dims, depth, lx, lossy, rows, src, drv):
    H = _hip()
    print(f'[compact coeff] {kind} {dims} depth {depth} lx {lx} band {rows}', flush=True)
    packed = _packed(kind, dims, np.random.default_rng(7), lossy, rows)
    a, b = H(kind, dims, 'ref32'), H(kind, dims, 'ref32')
    for e in (a, b):
        e.set_temporal(depth, lx)
        _setup(e, kind, dims, np.random.default_rng(31), lossy=lossy, source=src, drive=drv)
    a.build_coeffs(np.float32(0.4), packed)
    plain = {name: a.read_coeff(name) for name in packed}
    b.set_coeffs(np.float32(0.4), **plain)
    for tab in (_table(40, 0, 40), _table(37, 3, 33)):   # monitor-free; monitors at the ends
        a.run(tab)
        b.run(tab)
    for c in E_COMPS[kind] + H_COMPS[kind]:
        fa, fb = a.read_field(c), b.read_field(c)
        assert np.isfinite(fa).all()
        np.testing.assert_array_equal(fa.view(np.uint32), fb.view(np.uint32), err_msg=c)
    a.close()
    b.close()
