"""VCSEL Gaussian beam source and the one-sided (Huygens pair) soft source.

Behaviour of common/vcsel.py of the reference:
  defaults / air path      common/vcsel.py:43-83
  Gaussian beam w(z), R(z) common/vcsel.py:86-99
  profile_1d / profile_2d  common/vcsel.py:102-123
  numerical_kx_dx          common/vcsel.py:126-133 (Yee dispersion along x)
  HuygensSource            common/vcsel.py:136-175

The source here is host-side data only: amplitude and the two phase offsets of the
planes ix and ix+1.  Injection runs inside the device E-update (the step engine
receives the planes once and per-step scalars per chunk).
"""
import math

import numpy as np

C0 = 299792458.0

VCSEL_DEFAULTS = dict(lambda_nm=850.0, waist_um=3.0, distance_um=0.0, power_mw=20.0, tilt_deg=0.0)
MIN_AIR_UM = 1.0
MIN_AIR_CELLS = 8


def rayleigh_um(waist_um, lambda_nm, n=1.0):
    return math.pi * n * waist_um ** 2 / (lambda_nm * 1e-3)


def divergence_deg(waist_um, lambda_nm, n=1.0):
    return 2.0 * math.degrees(math.atan((lambda_nm * 1e-3) / (math.pi * n * waist_um)))


def air_path_um(distance_um, dx_um):
    """Air path source plane -> facet [um]: at least MIN_AIR_UM or MIN_AIR_CELLS cells."""
    s_min = max(MIN_AIR_UM, MIN_AIR_CELLS * float(dx_um))
    return max(float(distance_um or 0.0), s_min)


def z_at_source(distance_um, air_um):
    """Propagation coordinate of the source plane relative to the waist [um]."""
    return float(distance_um or 0.0) - float(air_um)


def beam_params(z_um, waist_um, lambda_nm, n=1.0):
    zr = rayleigh_um(waist_um, lambda_nm, n)
    w = waist_um * math.sqrt(1.0 + (z_um / zr) ** 2)
    R = math.inf if abs(z_um) < 1e-12 else z_um + zr ** 2 / z_um
    return w, R


def _phase_1d(u_um, R_um, lambda_nm, n, tilt_deg):
    k = 2 * math.pi * n / (lambda_nm * 1e-3)
    ph = k * math.sin(math.radians(tilt_deg)) * u_um
    if math.isfinite(R_um):
        ph = ph - k * u_um ** 2 / (2.0 * R_um)
    return ph


def profile_1d(y_um, yc_um, z_um, waist_um, lambda_nm, n=1.0, tilt_deg=0.0):
    """Line source (2D): amplitude and phase [rad] on the points y_um."""
    y = np.asarray(y_um, dtype=np.float64) - yc_um
    w, R = beam_params(z_um, waist_um, lambda_nm, n)
    amp = math.sqrt(waist_um / w) * np.exp(-(y / w) ** 2)
    return amp, _phase_1d(y, R, lambda_nm, n, tilt_deg)


def profile_2d(y_um, z_um_grid, yc_um, zc_um, zprop_um, waist_y_um, waist_z_um, lambda_nm,
               n=1.0, tilt_deg=0.0):
    """Area source (3D): amplitude/phase on y_um[:, None] x z_um_grid[None, :]."""
    y = np.asarray(y_um, dtype=np.float64)[:, None] - yc_um
    zz = np.asarray(z_um_grid, dtype=np.float64)[None, :] - zc_um
    wy, Ry = beam_params(zprop_um, waist_y_um, lambda_nm, n)
    wz, Rz = beam_params(zprop_um, waist_z_um, lambda_nm, n)
    amp = math.sqrt((waist_y_um / wy) * (waist_z_um / wz)) * np.exp(-(y / wy) ** 2 - (zz / wz) ** 2)
    ph = _phase_1d(y, Ry, lambda_nm, n, tilt_deg) + np.zeros_like(zz)
    if math.isfinite(Rz):
        ph = ph - (2 * math.pi * n / (lambda_nm * 1e-3)) * zz ** 2 / (2.0 * Rz)
    return amp, ph


def numerical_kx_dx(lambda_nm, dx_m, dt_m_per_c, n=1.0, tilt_deg=0.0, dims=2):
    """Phase advance per cell for the Huygens pair from the on-axis Yee dispersion
    sin(k dx/2) = (n dx/(c dt)) sin(w dt/2); times cos(tilt)."""
    half_wdt = math.pi * dt_m_per_c / (lambda_nm * 1e-9)
    arg = min(1.0, n * dx_m / dt_m_per_c * math.sin(half_wdt))
    return 2.0 * math.asin(arg) * math.cos(math.radians(tilt_deg))


class HuygensSource:
    """One-sided soft source on planes ix (phase ph1) and ix+1 (phase ph2, sign -1).
    The amplitude is divided by 2|sin(k dx)| so the forward wave has unit scale."""

    def __init__(self, ix, amp, phase, omega, kx_dx):
        self.ix = int(ix)
        g = 2.0 * abs(math.sin(kx_dx)) or 1.0
        self.amp = np.asarray(np.asarray(amp) / g, dtype=np.float32)
        self.ph1 = np.asarray(np.asarray(phase), dtype=np.float32)
        self.ph2 = np.asarray(np.asarray(phase) + kx_dx, dtype=np.float32)
        self.omega = float(omega)
        self.kx_dx = float(kx_dx)

    @property
    def plane_up(self):
        """Flux plane before the pair (sees only the reflected wave)."""
        return self.ix - 1

    @property
    def plane_dn(self):
        """Flux plane after the pair (forward minus reflected)."""
        return self.ix + 3

    def to_ey(self):
        """Same source on the TM Ey grid (average of neighbouring y nodes, float32)."""
        s = HuygensSource.__new__(HuygensSource)
        s.ix, s.omega, s.kx_dx = self.ix, self.omega, self.kx_dx
        s.amp = 0.5 * (self.amp[:-1] + self.amp[1:])
        s.ph1 = 0.5 * (self.ph1[:-1] + self.ph1[1:])
        s.ph2 = 0.5 * (self.ph2[:-1] + self.ph2[1:])
        return s


def describe(waist_um, lambda_nm, distance_um, air_um, power_mw, waist_z_um=None):
    z = z_at_source(distance_um, air_um)
    w, R = beam_params(z, waist_um, lambda_nm)
    wz = '' if waist_z_um is None or waist_z_um == waist_um else f'/{waist_z_um:g}'
    return (f'VCSEL {lambda_nm:g} nm: w0={waist_um:g}{wz} um, waist {distance_um or 0:g} um before '
            f'facet, source plane {air_um:.2f} um before facet (w={w:.2f} um), '
            f'P={power_mw:g} mW, forward-only (Huygens)')
