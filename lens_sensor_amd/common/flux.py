"""CW Poynting flux through x-planes from running DFTs of E and H.

    S_x = 1/2 Re( sum_k s_k E_k conj(H_k) )  over the whole cross-section
          2D TE: -Ez Hy*    2D TM: +Ey Hz*    3D: +Ey Hz* - Ez Hy*

Follows common/flux.py of the reference (PlaneFluxDFT :19-57, guard_in_plane :63-74,
pairs_2d/pairs_3d :77-86).  The running sums themselves live on the device (the step
engine accumulates E(t) e^{-iwt} and the x-averaged H(t) e^{-iw(t-dt/2)} per plane);
this module keeps the host bookkeeping and the final reduction, which is done with the
same NumPy expression as the reference on the downloaded accumulators.
"""
import numpy as np

SRC_MARGIN = 5


# Field pairs per polarisation: (sign, E component, H component).  H lives at i +- 1/2
# and is averaged onto the E plane il from H[il-1] and H[il].
PAIRS = {
    'te2d': ((-1.0, 'Ez', 'Hy'),),
    'tm2d': ((+1.0, 'Ey', 'Hz'),),
    'yee3d': ((+1.0, 'Ey', 'Hz'), (-1.0, 'Ez', 'Hy')),
}


def flux_from_sums(sums, n):
    """Net flux in +x from per-pair accumulators [(sign, accE, accH), ...] after n samples."""
    if not n:
        return float('nan')
    scale = (2.0 / n) ** 2
    total = 0.0
    for sign, aE, aH in sums:
        total += sign * 0.5 * float(np.real(np.sum(aE * np.conj(aH), dtype=np.complex128)))
    return total * scale


class PlaneFluxBook:
    """Host-side registry of flux planes (key -> local plane index) and their engine
    monitor ids; ``flux(key)`` downloads the sums and reduces them."""

    def __init__(self, engine, kind):
        self.engine = engine
        self.kind = kind
        self.mon = {}
        self.samples = {}

    def add_plane(self, key, il, carry=None):
        """Register plane ``key`` at local index ``il``.  ``carry`` (a FluxAccumulator)
        continues sums started in an earlier window: the reference keeps one running
        sum per key across stitch windows (fdtd3d_core.py:1640-1641), so a plane that
        two windows assemble keeps accumulating into the same sum."""
        self.mon[key] = self.engine.add_flux_plane(int(il))
        self.samples[key] = 0
        if carry is not None and carry.has(key):
            sums, n = carry.data[key]
            self.engine.write_monitor(self.mon[key], [(aE, aH) for _s, aE, aH in sums])
            self.samples[key] = n

    def keys(self):
        return list(self.mon)

    def note_samples(self, keys, count):
        for k in keys:
            self.samples[k] += count

    def reset(self):
        for k, m in self.mon.items():
            self.engine.zero_monitor(m)
            self.samples[k] = 0

    def has(self, key):
        return self.samples.get(key, 0) > 0

    def sums(self, key):
        acc = self.engine.read_flux(self.mon[key])
        return [(sign, aE, aH) for (sign, _e, _h), (aE, aH) in zip(PAIRS[self.kind], acc)]

    def flux(self, key):
        n = self.samples.get(key, 0)
        if not n:
            return float('nan')
        return flux_from_sums(self.sums(key), n)


class FluxAccumulator:
    """Host-side flux totals gathered over several engines (stitch windows): keeps the
    downloaded sums per key once a window has finished."""

    def __init__(self):
        self.data = {}

    def take(self, book, keys):
        for k in keys:
            if book.has(k):
                self.data[k] = (book.sums(k), book.samples[k])

    def has(self, key):
        return key in self.data and self.data[key][1] > 0

    def flux(self, key):
        if not self.has(key):
            return float('nan')
        sums, n = self.data[key]
        return flux_from_sums(sums, n)


def guard_in_plane(i_in, src_ix, n_max, margin=SRC_MARGIN):
    """Keep the input detector plane at least ``margin`` cells downstream of the source."""
    lo = int(src_ix) + int(margin)
    if i_in >= lo:
        return int(i_in)
    new = min(lo, int(n_max) - 2)
    print(f'  [Detector] input plane was at/before the source (ix={i_in} <= source {src_ix}) '
          f'-> moved to ix={new}')
    return new
