"""2D CPML (Roden-Gedney, CFS) profiles and boundary descriptors.

Profiles follow planar_beads/boundaries.py of the reference:
  cfs_a_max            boundaries.py:95-101   a_max = 2 pi eps0 f0 / 10
  _cpml_profiles       boundaries.py:104-151  sigma = smax d^m, smax = -(m+1) ln R0/(2 eta0 npml dx),
                                              kappa = 1 + (kmax-1) d^m, a = a_max (1-d),
                                              b = exp(-(sigma/kappa + a) dt/eps0),
                                              c = sigma (b-1)/(sigma kappa + kappa^2 a)
  CPML_TE / CPML_TM    boundaries.py:164-274  npml per end, PEC on the PML ends
  make_boundary        boundaries.py:277-289

Unlike the reference the psi auxiliaries are not host/CuPy arrays here: the step engine
owns them and stores them only in the PML slabs (c != 0), where psi can become non-zero.
The descriptor carries the 1-D profiles (float32, as in the reference) and the PEC ends.
"""
import numpy as np

from lens_sensor_amd.common.physics import C0, EPS0


def cfs_a_max(lam_m=None):
    return 2.0 * np.pi * EPS0 * (C0 / (lam_m or 850e-9)) / 10.0


def cpml_profile_1d(N, npml_lo, npml_hi, dx, dt, m=3, kappa_max=1.0, a_max=None, R0=1e-6,
                    lam_m=None):
    """(kappa, b, c) float32 on the N integer nodes and the N-1 half nodes of one axis."""
    if npml_lo <= 0 and npml_hi <= 0:
        one_e, zero_e = np.ones(N, np.float32), np.zeros(N, np.float32)
        one_h, zero_h = np.ones(N - 1, np.float32), np.zeros(N - 1, np.float32)
        return (one_e, one_e.copy(), zero_e), (one_h, one_h.copy(), zero_h)
    eta0 = np.sqrt(4e-7 * np.pi / EPS0)
    if a_max is None:
        a_max = cfs_a_max(lam_m)

    def sigma_max(npml):
        return (-(m + 1) * np.log(R0) / (2.0 * eta0 * npml * dx)) if npml > 0 else 0.0

    s_lo, s_hi = sigma_max(npml_lo), sigma_max(npml_hi)

    def at(pos):
        depth = np.zeros_like(pos, dtype=np.float64)
        smax = np.zeros_like(pos, dtype=np.float64)
        if npml_lo > 0:
            sel = pos < npml_lo
            depth[sel] = (npml_lo - pos[sel]) / npml_lo
            smax[sel] = s_lo
        if npml_hi > 0:
            sel = pos > (N - 1 - npml_hi)
            depth[sel] = (pos[sel] - (N - 1 - npml_hi)) / npml_hi
            smax[sel] = s_hi
        depth = np.clip(depth, 0.0, 1.0)
        sigma = smax * depth ** m
        kappa = 1.0 + (kappa_max - 1.0) * depth ** m
        a = np.where(depth > 0.0, a_max * (1.0 - depth), 0.0)
        b = np.exp(-(sigma / kappa + a) * dt / EPS0)
        denom = sigma * kappa + kappa * kappa * a
        c = np.zeros_like(sigma)
        nz = denom > 0.0
        c[nz] = sigma[nz] * (b[nz] - 1.0) / denom[nz]
        return kappa.astype(np.float32), b.astype(np.float32), c.astype(np.float32)

    return at(np.arange(N, dtype=np.float64)), at(np.arange(N - 1, dtype=np.float64) + 0.5)


def _pml_pair(v, default):
    if v is None:
        v = default
    if isinstance(v, (tuple, list)):
        return int(v[0]), int(v[1])
    return int(v), int(v)


class CPML2D:
    """CPML descriptor for a 2D domain (TE or TM): profiles per axis and PEC ends.
    ``prof[axis] = (e, h)`` with e = (kappa, b, c) on nodes and h on half nodes."""

    def __init__(self, pol, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, **kw):
        self.pol = pol
        self.Nx, self.Ny = Nx, Ny
        self.npx_lo, self.npx_hi = _pml_pair(npml_x, npml)
        self.npy_lo, self.npy_hi = _pml_pair(npml_y, npml)
        self.prof = (cpml_profile_1d(Nx, self.npx_lo, self.npx_hi, dx, dt, **kw),
                     cpml_profile_1d(Ny, self.npy_lo, self.npy_hi, dx, dt, **kw))

    @property
    def pec(self):
        """PEC closes every end that carries a PML (boundaries.py:207-215, :260-268)."""
        return ((self.npx_lo > 0, self.npx_hi > 0), (self.npy_lo > 0, self.npy_hi > 0))

    def attach(self, engine):
        for ax in (0, 1):
            engine.set_cpml(ax, self.prof[ax][0], self.prof[ax][1], *self.pec[ax])


def make_boundary(kind, pol, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, lam_m=None):
    kind = (kind or 'cpml').lower()
    if kind == 'mur1':
        raise NotImplementedError("boundary 'mur1' is not part of the device path; use 'cpml'")
    if kind != 'cpml':
        raise ValueError(f"unknown boundary '{kind}' (cpml)")
    te = str(pol).lower() in ('s', 'te')
    return CPML2D('s' if te else 'p', Nx, Ny, dx, dt, npml=npml, npml_x=npml_x, npml_y=npml_y,
                  lam_m=lam_m)
