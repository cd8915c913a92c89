"""The Fourier pulse synthesis (lens_sensor_amd.common.fourier_pulse; reference
planar_beads/fdtd2d_core.py:1891-1974, planar_3d/fdtd3d_core.py:2102-2179) against what
the mathematics fixes, independent of the reference's own expressions:

* the comb's S(w_k) equals the Fourier integral of the source pulse
  s(t) = exp(-((t - tc)/sp)^2) sin(w0 t), taken by quadrature;
* a structure that passes the source unchanged (CW phasor -i, the response to sin(w t))
  synthesises the pulse itself, and a pure delay tau (phasor -i e^{-i w tau}) the delayed
  pulse: a dropped factor i, a sign slip in an exponent, a wrong comb step or a missing
  1/pi each fail these;
* the x-blocked accumulation equals the whole-array expression bit for bit, in both the
  complex128 (2D) and the complex64 (3D) arithmetic.

The whole-run results are pinned separately by the reference-written fixtures
(tests/test_characterization.py, cases *_pulse*).
"""
import numpy as np
import pytest

from lens_sensor_amd.common.fourier_pulse import PulseComb, add_frequency


def _pulse(comb, t):
    return np.exp(-((t - comb.tc) / comb.sp) ** 2) * np.sin(comb.w0 * t)


@pytest.mark.parametrize('cycles', [5.0, 30.0])
def test_comb_spectrum_is_the_fourier_integral_of_the_pulse(cycles):
    comb = PulseComb(850.0, cycles, 2e-13)
    t = np.linspace(-comb.tc, 3 * comb.tc, 800001)      # the tails are exp(-36)
    s = _pulse(comb, t)
    for k in (0, comb.centre, len(comb) // 3, len(comb) - 1):
        f = s * np.exp(-1j * comb.w[k] * t)
        num = np.sum(0.5 * (f[1:] + f[:-1])) * (t[1] - t[0])
        assert abs(comb.S[k] - num) <= 1e-6 * abs(comb.S).max(), k
    # the band: the outermost lines lie where |S| has fallen below exp(-(5.26/2)^2) = 9.91e-4
    # of its peak, the next ones inside above it
    rel = np.abs(comb.S) / np.abs(comb.S).max()
    assert np.all(rel[[0, -1]] <= 9.92e-4) and np.all(rel[[1, -2]] > 9.90e-4)


@pytest.mark.parametrize('wide', [True, False])
@pytest.mark.parametrize('delay_fs', [0.0, 120.0])
def test_synthesis_of_a_transparent_and_a_delaying_structure(wide, delay_fs):
    comb = PulseComb(850.0, 8.0, 1.5e-13)
    tau = delay_fs * 1e-15
    t_j = np.linspace(0.0, comb.t_max, 61)
    nx = 70
    frames = np.zeros((len(t_j), nx, 3), dtype=np.float32)
    for wk, Sk in zip(comb.w, comb.S):
        E = np.full((nx, 3), -1j * np.exp(-1j * wk * tau), dtype=np.complex64)
        add_frequency(frames, E, Sk * comb.dw / np.pi, wk, t_j, wide=wide, block=16)
    want = _pulse(comb, t_j - tau)
    tol = 3e-3 if wide else 1e-2
    for x in (0, nx - 1):
        np.testing.assert_allclose(frames[:, x, 1], want, atol=tol)
    assert np.abs(want).max() > 0.9


def test_comb_step_and_aliasing_flag():
    comb = PulseComb(850.0, 30.0, 4e-12)
    assert np.isclose(comb.dw, 2 * np.pi / (1.6 * comb.t_max))
    assert not comb.aliased
    assert np.isclose(comb.w[comb.centre], comb.w0)
    few = PulseComb(850.0, 30.0, 4e-12, nfreq=3)
    assert len(few) == 3 and few.aliased
    assert np.isclose(few.w[-1] - few.w0, 5.26 / few.sp)
    fixed = PulseComb(850.0, 30.0, 4e-12, tmax_fs=900.0)
    assert np.isclose(fixed.t_max, 900e-15)


@pytest.mark.parametrize('wide', [True, False])
def test_blocked_accumulation_equals_the_whole_array_expression(wide):
    rng = np.random.default_rng(3)
    E = (rng.standard_normal((203, 37)) + 1j * rng.standard_normal((203, 37))).astype(np.complex64)
    coeff, w = complex(0.3, -1.7) * 1e-3, 2.2e15
    t_j = np.linspace(0.0, 4e-13, 7)
    got = np.zeros((7, 203, 37), dtype=np.float32)
    add_frequency(got, E, coeff, w, t_j, wide=wide)
    want = np.zeros_like(got)
    if wide:
        c = (1j * E) * np.complex128(coeff)
        for j, t in enumerate(t_j):
            want[j] += np.real(c * np.exp(1j * w * t)).astype(np.float32)
    else:
        c = (1j * E) * np.complex64(coeff)
        for j, t in enumerate(t_j):
            want[j] += np.real(c * np.complex64(np.exp(1j * w * t))).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
