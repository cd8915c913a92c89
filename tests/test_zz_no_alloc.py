"""Allocation-free steady-state reduction (needs an MI355X): lf_norm_diff_noalloc, whose
partials live in the library's static per-device array, called through the C ABI on
device arrays, gives the same bits as lf_norm_diff (library-allocated partials) on two
calls in a row, the second smaller than the first (the array keeps the first call's
partials), and the float64 NumPy sums to 1e-12.  Sizes past the array's bound are
rejected on the host before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref(a, b, da, db):
    a = a.astype(np.complex128) / da
    b = b.astype(np.complex128) / db
    return np.sum(np.abs(a - b) ** 2), np.sum(np.abs(a) ** 2)


def _pair(n, seed, dtype):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)
    B = (A + 0.01 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(dtype)
    return A, B


def _call(name, ta, tb, n, acc, da, db, out, stream):
    from lens_sensor_amd.engine import hip
    lib = hip.load()
    rc = getattr(lib, name)(ta.data_ptr(), tb.data_ptr(), ctypes.c_longlong(n), acc, da, db,
                            out.data_ptr(), stream)
    assert rc == 0, (name, lib.lf_last_error())


@pytest.mark.parametrize('acc,dtype,ftype', [(2, np.complex128, np.float64),
                                             (1, np.complex64, np.float32)])
def test_noalloc_matches_the_allocating_entry_point(acc, dtype, ftype):
    import torch
    st = torch.cuda.current_stream().cuda_stream
    # second call smaller: partials of the first call stay in the static array
    for n, seed, da, db in ((3 * 4096 * 16 + 77, 5, 96.0, 97.0), (4097, 6, 1.0, 3.0)):
        A, B = _pair(n, seed, dtype)
        ta = torch.from_numpy(A.view(ftype)).to('cuda:0')
        tb = torch.from_numpy(B.view(ftype)).to('cuda:0')
        o_alloc = torch.zeros(2, dtype=torch.float64, device='cuda:0')
        o_static = torch.full((2,), float('nan'), dtype=torch.float64, device='cuda:0')
        _call('lf_norm_diff', ta, tb, n, acc, da, db, o_alloc, st)
        _call('lf_norm_diff_noalloc', ta, tb, n, acc, da, db, o_static, st)
        torch.cuda.synchronize()
        ra, rs = o_alloc.cpu().numpy(), o_static.cpu().numpy()
        assert ra.tobytes() == rs.tobytes(), (n, ra, rs)
        rn, rd = _ref(A, B, da, db)
        assert abs(rs[0] - rn) <= 1e-12 * rn and abs(rs[1] - rd) <= 1e-12 * rd, (n, rs, rn, rd)


def test_noalloc_rejects_sizes_past_its_bound():
    import torch
    from lens_sensor_amd.engine import hip
    lib = hip.load()
    t = torch.zeros(4, dtype=torch.float64, device='cuda:0')
    out = torch.zeros(2, dtype=torch.float64, device='cuda:0')
    st = torch.cuda.current_stream().cuda_stream
    # rejected on the host (size check before any launch): the arrays are never read
    rc = lib.lf_norm_diff_noalloc(t.data_ptr(), t.data_ptr(), ctypes.c_longlong((1 << 28) + 1), 2,
                                  1.0, 1.0, out.data_ptr(), st)
    assert rc != 0
    assert b'static scratch' in lib.lf_last_error()
