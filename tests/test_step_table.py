"""The step table (lens_sensor_amd.engine.spec.fill_phasors) is filled with whole-array
NumPy expressions; the reference evaluates the same expressions once per step on Python
scalars (planar_beads/fdtd2d_core.py:839, common/flux.py:32-33, stitch drive
fdtd2d_core.py:1426, fdtd3d_core.py:1627-1639, vcsel.py:154-155).  Every field of every
row must be the scalar value bit for bit, on tables of the drivers' time steps and
frequencies, with CW and pulse envelopes, at offsets deep into a run.
"""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, fill_phasors, cw_envelope, pulse_envelope


def _scalar_rows(tab, omega, dt, t_phys, env, src_omega):
    """The reference's per-step expressions, one Python scalar at a time."""
    out = new_steps(len(t_phys))
    for i, t in enumerate(np.asarray(t_phys, dtype=np.float64).tolist()):
        e = float(env[i])
        wt = src_omega * t
        out['env'][i] = e
        out['e32'][i] = np.float32(e)
        out['wt32'][i] = np.float32(wt)
        out['ssin'][i] = e * np.sin(wt)
        out['scos'][i] = e * np.cos(wt)
        p = np.exp(-1j * omega * t)
        out['pr'][i], out['pi'][i] = p.real, p.imag
        p32 = np.complex64(p)
        out['pr32'][i], out['pi32'][i] = p32.real, p32.imag
        h = np.exp(-1j * omega * (t - 0.5 * dt))
        out['hr'][i], out['hi'][i] = h.real, h.imag
        h32 = np.complex64(h)
        out['hr32'][i], out['hi32'][i] = h32.real, h32.imag
        d = np.exp(1j * omega * t)
        out['dr'][i], out['di'][i] = d.real, d.imag
        d32 = np.complex64(d)
        out['dr32'][i], out['di32'][i] = d32.real, d32.imag
        out['qr32'][i] = np.float32(np.cos(omega * t))
        out['qi32'][i] = -np.float32(np.sin(omega * t))
    return out


@pytest.mark.parametrize('dt,omega,start,pulse', [
    (2.3584e-17, 2.2161e15, 0, False),          # 2D, dx 10 nm class, 850 nm
    (5.8959e-17, 2.2161e15, 180000, True),      # 2D 25 nm, deep into a run, pulse
    (2.8890e-17, 2.2161e15 * 1.028, 7000, False),   # 3D 30 nm, another comb line
])
def test_table_matches_the_per_step_scalars(dt, omega, start, pulse):
    n = 6000
    t = np.arange(start, start + n) * dt
    f0 = omega / (2 * np.pi)
    env = pulse_envelope(t, f0, 30.0) if pulse else cw_envelope(t, 4 / omega)
    src = omega * 1.0003
    got = fill_phasors(new_steps(n), omega, dt, t, env, src_omega=src)
    want = _scalar_rows(None, omega, dt, t, env, src)
    for f in got.dtype.names:
        if f in ('mask', 'row'):
            continue
        a = got[f].view(np.uint64 if got[f].dtype.itemsize == 8 else np.uint32)
        b = want[f].view(np.uint64 if want[f].dtype.itemsize == 8 else np.uint32)
        assert np.array_equal(a, b), (f, int((a != b).sum()))
