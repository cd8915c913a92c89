"""Device steady-state reduction (needs an MI355X): sum |A/da - B/db|^2 and sum |A/da|^2
of two phasor windows, reduced in FP64 on the device, against NumPy float64 on the same
data: complex128 and complex64 windows, sizes that leave partial workgroups, and the
convergence numbers of a served request and of the drivers' CW test."""
import numpy as np
import pytest

from oracle.engine import OracleEngine

pytestmark = pytest.mark.gpu


def _ref(a, b, da, db):
    a = a.astype(np.complex128) / da
    b = b.astype(np.complex128) / db
    d = a - b
    return np.sum(np.abs(d) ** 2), np.sum(np.abs(a) ** 2)


@pytest.mark.parametrize('acc', ['c128', 'c64'])
@pytest.mark.parametrize('dims', [(37, 70), (301, 517)])
def test_monitor_change_matches_numpy(acc, dims):
    from lens_sensor_amd.engine.hip import HipEngine
    from test_gpu_parity import _setup
    mode = 'fp64acc' if acc == 'c128' else 'ref32'
    h = HipEngine('tm2d', dims, mode)
    o = OracleEngine('tm2d', dims, mode)
    _setup(h, 'tm2d', dims, np.random.default_rng(1))
    _setup(o, 'tm2d', dims, np.random.default_rng(1))
    ma = h.add_dft('Ey', 0, dims[0], acc=acc)
    mb = h.add_dft('Ey', 0, dims[0], acc=acc)
    oa = o.add_dft('Ey', 0, dims[0], acc=acc)
    ob = o.add_dft('Ey', 0, dims[0], acc=acc)
    rng = np.random.default_rng(7)
    dt = np.complex128 if acc == 'c128' else np.complex64
    shp = h.read_monitor(ma).shape
    A = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(dt)
    B = (A + 1e-3 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))).astype(dt)
    for e, x, y in ((h, ma, mb), (o, oa, ob)):
        e.write_monitor(x, A)
        e.write_monitor(y, B)
    for da, db in ((1.0, 1.0), (96.0, 97.0)):
        num, den = h.monitor_change(ma, mb, da, db)
        rn, rd = _ref(A, B, da, db)
        assert abs(num - rn) <= 1e-12 * rn and abs(den - rd) <= 1e-12 * rd, (num, rn, den, rd)
        on, od = o.monitor_change(oa, ob, da, db)
        assert abs(on - rn) <= 1e-12 * rn and abs(od - rd) <= 1e-12 * rd
    # deterministic: the same data gives the same bits
    assert h.monitor_change(ma, mb, 96.0, 97.0) == h.monitor_change(ma, mb, 96.0, 97.0)
    h.close()


@pytest.mark.parametrize('n', [1, 255, 4096, 4097, 1 << 20])
def test_norm_diff_on_device_arrays(n):
    import ctypes
    import torch
    from lens_sensor_amd.engine import hip
    rng = np.random.default_rng(n)
    A = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)
    B = (A + 0.01 * rng.standard_normal(n)).astype(np.complex128)
    ta = torch.from_numpy(A.view(np.float64)).to('cuda:0')
    tb = torch.from_numpy(B.view(np.float64)).to('cuda:0')
    out = torch.zeros(2, dtype=torch.float64, device='cuda:0')
    st = torch.cuda.current_stream().cuda_stream
    hip.op_call('lf_norm_diff', ta.data_ptr(), tb.data_ptr(), ctypes.c_longlong(n), 2, 1.0, 3.0,
                out.data_ptr(), st)
    torch.cuda.synchronize()
    num, den = out.cpu().numpy()
    rn, rd = _ref(A, B, 1.0, 3.0)
    assert abs(num - rn) <= 1e-12 * rn and abs(den - rd) <= 1e-12 * rd, (num, rn, den, rd)
