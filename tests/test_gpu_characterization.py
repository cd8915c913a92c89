"""The solver drivers on the HIP engine (MI355X) reproduce the reference's outputs.

Same cases and fixtures as tests/test_characterization.py (written by the reference
itself, tests/golden/make_golden.py: full domain, stitch, absorbers and the Fourier pulse
stitch), now with every time step executed by the gfx950 kernels in the
reference-faithful 'ref32' mode.  Tolerance: the device computes
the float32 source phase with sinf instead of NumPy's float32 sin (<= 2 ulp), every
other float32 operation is identical; over a few hundred steps this stays below
1e-5 relative on scalars and relative L2 on arrays (the reference's own tolerance is
1e-6 between NumPy versions).
"""
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu

CASES = [('2d', n) for n in rc.CASES_2D] + [('3d', n) for n in rc.CASES_3D]


@pytest.mark.parametrize('dim,name', CASES)
def test_drivers_on_hip_match_reference(dim, name, in_tmp):
    from lens_sensor_amd.engine.hip import HipEngine
    print(f'[gpu characterization] {dim}/{name}', flush=True)
    got = rc.run_case(dim, name, HipEngine)
    want = rc.load_fixture(dim, name)
    rep = rc.compare(got, want, rtol_scalar=1e-5, rtol_array=1e-5, atol_frac=1e-4)
    print({k: f'{v[0]:.2e}' for k, v in rep.items()}, flush=True)
    bad = {k: v[0] for k, v in rep.items() if not v[1]}
    assert not bad, f'{dim}/{name} deviates: {bad}'
