"""Batch engine: many 2D instances (mixed TE / TM, different sizes, sources, stitch
drives, monitors at different times, different step counts) advanced together must end
BIT-IDENTICAL to each instance run alone on the device, and must agree with the CPU
oracle run of the same instance: bit for bit without a source, within 1e-6 relative L2
with one (the device sinf against NumPy's float32 sin in ref32).  Needs an MI355X."""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, fill_phasors
from oracle.engine import OracleEngine
from test_gpu_parity import _setup, _compare

pytestmark = pytest.mark.gpu

SPECS = [('tm2d', (41, 66), dict()), ('te2d', (37, 70), dict(source=True)),
         ('tm2d', (170, 250), dict(source=True)), ('te2d', (160, 230), dict(drive=True, pec=False)),
         ('tm2d', (300, 140), dict(drive=True, pec=False)), ('te2d', (90, 400), dict())]


def _hip():
    from lens_sensor_amd.engine.hip import HipEngine, HipBatch
    return HipEngine, HipBatch


def _build(H, mode, seed0):
    engs, mons = [], []
    for n, (kind, dims, kw) in enumerate(SPECS):
        e = H(kind, dims, mode)
        mons.append(_setup(e, kind, dims, np.random.default_rng(seed0 + n), **kw))
        engs.append(e)
    return engs, mons


def _table(nsteps):
    tab = new_steps(nsteps)
    t = np.arange(nsteps) * 1e-16
    fill_phasors(tab, 2 * np.pi * 3.5e14, 1e-16, t, 1 - np.exp(-(t / 4e-15) ** 2))
    return tab


# depth 8 included: a depth-8 batch pass once launched the depth-6 kernels over a depth-8
# wave plan, which left cells of the output buffers unwritten
@pytest.mark.parametrize('mode,depth', [('ref32', 6), ('fp64acc', 4), ('ref32', 2), ('fp64acc', 6),
                                        ('ref32', 8), ('fp64acc', 8)])
def test_batch_equals_instances_alone_and_the_oracle(mode, depth):
    H, B = _hip()
    nsteps = 61
    tab = _table(nsteps)
    rng = np.random.default_rng(4)
    masks = np.zeros((len(SPECS), nsteps), np.uint32)
    nact = np.array([61, 50, 61, 33, 47, 58], np.int32)
    for i in range(len(SPECS)):
        a = int(rng.integers(0, nsteps - 10))
        masks[i, a:a + 5] = rng.integers(1, 2 ** 8, 5).astype(np.uint32)   # a monitor window
        masks[i, nact[i] - 2:nact[i]] = 0xff
    batch_eng, mons = _build(H, mode, 100)
    alone_eng, _ = _build(H, mode, 100)
    oracle_eng, _ = _build(OracleEngine, mode, 100)
    b = B(batch_eng, depth=depth, lx=16)
    b.set_timing(True)
    b.run(tab, masks=masks, nact=nact)
    ms, passes = b.stats()
    assert passes > 0
    for i, e in enumerate(alone_eng):
        e.set_temporal(1, 64)
        tt = tab[:nact[i]].copy()
        tt['mask'] = masks[i, :nact[i]]
        e.run(tt)
        oracle_eng[i].run(tt)
    for i, (kind, dims, kw) in enumerate(SPECS):
        _compare(kind, batch_eng[i], alone_eng[i], mons[i])
        if kw.get('source'):
            _compare(kind, batch_eng[i], oracle_eng[i], mons[i], exact=False, tol=1e-6)
        else:
            _compare(kind, batch_eng[i], oracle_eng[i], mons[i])
    b.close()
    for e in batch_eng + alone_eng:
        e.close()


def test_batch_rejects_unknown_depths():
    H, B = _hip()
    b = B([], depth=6)
    for bad in (3, 5, 7, 10):
        with pytest.raises(RuntimeError):
            b.set_temporal(bad, 64)
    b.close()
