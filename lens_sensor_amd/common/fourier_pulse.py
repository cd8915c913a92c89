"""Transient fields from CW runs by Fourier synthesis (reference
planar_beads/fdtd2d_core.py:1866-2017, planar_3d/fdtd3d_core.py:2087-2215).

The pulsed source of the full-domain runs, s(t) = exp(-((t - tc)/sp)^2) sin(w0 t) with
sp = cycles/f0 and tc = 3 sp, has the spectrum

    S(w) = [G(w - w0) - G(w + w0)] / (2i),   G(v) = sqrt(pi) sp exp(-(v sp)^2/4) exp(-i v tc).

A CW run at w answers sin(w t) = Re(-i e^{iwt}) with the phasor E_w, so its transfer
function is H_w = i E_w and the pulsed response is

    E(r, t) = (1/pi) Re sum_k H_k(r) S(w_k) e^{i w_k t} dw

over a comb w_k = w0 + k dw that covers |w - w0| <= 5.26/sp (there |S| has fallen to
1e-3 of its peak).  The comb step is 2 pi / (1.6 t_max), so the synthesis does not wrap
round before t_max, unless the caller fixes the number of frequencies.

The frames are summed one frequency at a time into a memory-mapped float32 array, so
the memory held does not grow with the number of frequencies.
"""
import numpy as np

from lens_sensor_amd.common.physics import C0

BAND = 5.26            # |S(w0 +- BAND/sp)| = 1e-3 |S(w0)|
REPEAT = 1.6           # comb repetition period / t_max


class PulseComb:
    """The frequency comb of a Gaussian pulse (reference fdtd2d_core.py:1891-1909,
    fdtd3d_core.py:2102-2118).

    lambda_nm: centre wavelength; cycles: pulse length in periods; t_transit: time the
    pulse needs to cross the structure (s); tmax_fs: explicit synthesis span; nfreq: a
    fixed number of comb lines (the step then follows from the band, and a period shorter
    than t_max aliases, which ``aliased`` reports)."""

    def __init__(self, lambda_nm, cycles, t_transit, tmax_fs=None, nfreq=None):
        self.f0 = C0 / (lambda_nm * 1e-9)
        self.w0 = 2 * np.pi * self.f0
        self.sp = cycles / self.f0
        self.tc = 3 * self.sp
        self.t_max = (tmax_fs * 1e-15) if tmax_fs else (self.tc + 3 * self.sp + t_transit)
        half = BAND / self.sp
        if nfreq:
            self.dw = 2 * half / max(int(nfreq) - 1, 1)
        else:
            self.dw = 2 * np.pi / (REPEAT * self.t_max)
        self.aliased = bool(nfreq) and 2 * np.pi / self.dw < self.t_max
        m = int(np.ceil(half / self.dw))
        w = self.w0 + np.arange(-m, m + 1) * self.dw
        self.w = w[w > 0]
        self.lam_nm = 2 * np.pi * C0 / self.w * 1e9
        self.S = (self._G(self.w - self.w0) - self._G(self.w + self.w0)) / 2j
        self.centre = int(np.argmin(np.abs(self.w - self.w0)))

    def _G(self, v):
        return np.sqrt(np.pi) * self.sp * np.exp(-(v * self.sp) ** 2 / 4.0) * np.exp(-1j * v * self.tc)

    def __len__(self):
        return len(self.w)

    def weights(self):
        """|S_k|^2, the weights of the spectrally averaged transmission."""
        return np.abs(self.S) ** 2

    def weighted(self, values):
        w = self.weights()
        return float(np.sum(w * np.asarray(values, dtype=np.float64)) / np.sum(w))


def add_frequency(frames, E, coeff, w, times, wide=True, block=64):
    """frames[j] += Re(i E coeff e^{i w t_j}) for every frame j, one x block at a time.

    coeff = S_k dw / pi.  ``wide``: the product in complex128 (the 2D stitch,
    fdtd2d_core.py:1972-1974); else in complex64 as the 3D stitch does
    (fdtd3d_core.py:2175-2179).  E's first axis is x; frames has shape (n_t,) + E.shape."""
    ct = np.complex128 if wide else np.complex64
    c = np.complex128(coeff) if wide else np.complex64(coeff)
    rot = [ct(np.exp(1j * w * t)) for t in times]
    for x0 in range(0, E.shape[0], block):
        h = (1j * E[x0:x0 + block]).astype(ct, copy=False) * c
        for j, z in enumerate(rot):
            frames[j, x0:x0 + block] += np.real(h * z).astype(np.float32)
