"""2D CPML (Roden-Gedney, CFS) profiles and boundary descriptors.

Profiles follow planar_beads/boundaries.py of the reference:
  cfs_a_max            boundaries.py:95-101   a_max = 2 pi eps0 f0 / 10
  _cpml_profiles       boundaries.py:104-151  sigma = smax d^m, smax = -(m+1) ln R0/(2 eta0 npml dx),
                                              kappa = 1 + (kmax-1) d^m, a = a_max (1-d),
                                              b = exp(-(sigma/kappa + a) dt/eps0),
                                              c = sigma (b-1)/(sigma kappa + kappa^2 a)
  CPML_TE / CPML_TM    boundaries.py:164-274  npml per end, PEC on the PML ends;
                                              update_H / update_E / terminate on device
                                              arrays through fused2d (HIP kernels)
  make_boundary        boundaries.py:277-289
  mur1_coeff           boundaries.py:22-29    m = (c dt - dx)/(c dt + dx), c = c0/n
  guided_neff          boundaries.py:32-53    n_eff of the fundamental mode of eps(y)
  mur_speeds / MurSet  boundaries.py:56-93    local speeds of the Mur edges (x ends: n_eff,
                                              y edges: n of the edge cells)
  attach_mur_2d                               the inline first-order Mur of
                                              fdtd2d_core.py:805-810, 831-833 (stitch
                                              :1422-1433, :1458-1469) as engine faces

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


class _CPMLOps(CPML2D):
    """Operator-level interface of the reference's CPML objects (boundaries.py:164-274) on
    device arrays.  The 1-D profiles and the full-domain psi arrays are created on the
    device of the first arrays the object updates (torch tensors, float32, the
    reference's attribute names and shapes); the updates run the HIP kernels of
    fused2d.  Host (NumPy) arrays are refused: there is no CPU path."""

    _PSI = ()

    def __init__(self, pol, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, **kw):
        super().__init__(pol, Nx, Ny, dx, dt, npml=npml, npml_x=npml_x, npml_y=npml_y, **kw)
        self._dev = None

    def _device_state(self, like):
        if self._dev is not None and self._dev == like.device:
            return
        import torch
        dev = like.device
        names = (('kex', 'bex', 'cex'), ('khx', 'bhx', 'chx'), ('key', 'bey', 'cey'), ('khy', 'bhy', 'chy'))
        profs = (self.prof[0][0], self.prof[0][1], self.prof[1][0], self.prof[1][1])
        for nm, pr in zip(names, profs):
            for n, a in zip(nm, pr):
                setattr(self, n, torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev))
        for n, shape in self._PSI:
            setattr(self, n, torch.zeros(shape, dtype=torch.float32, device=dev))
        self._dev = dev

    def _op2d(self, like, psi0, psi1):
        from lens_sensor_amd.engine.hip import OpProfiles2D
        return OpProfiles2D(*[getattr(self, n).data_ptr() for n in
                              ('bex', 'cex', 'kex', 'bhx', 'chx', 'khx',
                               'bey', 'cey', 'key', 'bhy', 'chy', 'khy')],
                            psi0.data_ptr(), psi1.data_ptr())

    def _need(self, what, *arrays):
        from lens_sensor_amd.planar_beads import fused2d
        if not fused2d.available(*arrays):
            raise TypeError(f'{type(self).__name__}.{what} needs float32, contiguous device '
                            'tensors on the HIP device (there is no host path)')
        self._device_state(arrays[0])


class CPML_TE(_CPMLOps):
    """TE CPML (Ez, Hx, Hy), reference boundaries.py:164-221: psi over the full domain,
    PEC on the E ends that carry a PML."""

    def __init__(self, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, **kw):
        self._PSI = (('psi_Hxy', (Nx, Ny - 1)), ('psi_Hyx', (Nx - 1, Ny)),
                     ('psi_Ezx', (Nx - 2, Ny - 2)), ('psi_Ezy', (Nx - 2, Ny - 2)))
        super().__init__('s', Nx, Ny, dx, dt, npml=npml, npml_x=npml_x, npml_y=npml_y, **kw)

    def update_H(self, Ez, Hx, Hy, Ch):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('update_H', Ez, Hx, Hy)
        return fused2d.te_update_H(self, Ez, Hx, Hy, Ch)

    def update_E(self, Ez, Hx, Hy, Ce_E, Ce_H):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('update_E', Ez, Hx, Hy, Ce_E, Ce_H)
        return fused2d.te_update_E(self, Ez, Hx, Hy, Ce_E, Ce_H)

    def terminate(self, Ez):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('terminate', Ez)
        return fused2d.te_terminate(self, Ez)


class CPML_TM(_CPMLOps):
    """TM CPML (Hz, Ex, Ey), reference boundaries.py:224-274: PEC sets the tangential E
    (Ey on the x ends, Ex on the y ends) that carry a PML."""

    def __init__(self, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, **kw):
        self._PSI = (('psi_Hzx', (Nx - 1, Ny - 1)), ('psi_Hzy', (Nx - 1, Ny - 1)),
                     ('psi_Exy', (Nx - 1, Ny - 2)), ('psi_Eyx', (Nx - 2, Ny - 1)))
        super().__init__('p', Nx, Ny, dx, dt, npml=npml, npml_x=npml_x, npml_y=npml_y, **kw)

    def update_H(self, Ex, Ey, Hz, Ch_tm):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('update_H', Ex, Ey, Hz)
        return fused2d.tm_update_H(self, Ex, Ey, Hz, Ch_tm)

    def update_E(self, Ex, Ey, Hz, Ca_ex, Ca_ey):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('update_E', Ex, Ey, Hz, Ca_ex, Ca_ey)
        return fused2d.tm_update_E(self, Ex, Ey, Hz, Ca_ex, Ca_ey)

    def terminate(self, Ex, Ey):
        from lens_sensor_amd.planar_beads import fused2d
        self._need('terminate', Ex, Ey)
        return fused2d.tm_terminate(self, Ex, Ey)


def make_boundary(kind, pol, Nx, Ny, dx, dt, npml=10, npml_x=None, npml_y=None, lam_m=None):
    """CPML object for kind 'cpml' (reference boundaries.py:277-289): CPML_TE for 's'/'te',
    CPML_TM for 'p'/'tm'.  It attaches to a step engine (attach) and updates device
    arrays (update_H / update_E / terminate)."""
    kind = (kind or 'cpml').lower()
    if kind == 'mur1':
        raise NotImplementedError("boundary 'mur1' is not part of the device path; use 'cpml'")
    if kind != 'cpml':
        raise ValueError(f"unknown boundary '{kind}' (cpml)")
    te = str(pol).lower() in ('s', 'te')
    return (CPML_TE if te else CPML_TM)(Nx, Ny, dx, dt, npml=npml, npml_x=npml_x, npml_y=npml_y,
                                         lam_m=lam_m)


# ----------------------------------------------------------------------------- Mur (first order)
def mur1_coeff(dx, dt, n=1.0):
    """Mur-1 coefficient at phase velocity c0/n (n scalar or array), float32 like the
    reference's (boundaries.py:22-29)."""
    c = C0 / np.asarray(n, dtype=np.float64)
    m = (c * dt - dx) / (c * dt + dx)
    return np.float32(m) if np.ndim(m) == 0 else m.astype(np.float32)


def guided_neff(eps_col, dx, lam_m):
    """n_eff of the fundamental mode of a 1-D profile eps(y): the largest eigenvalue of the
    finite-difference operator E'' + k0^2 eps E near k0^2 max(eps) (shift-invert Lanczos,
    scipy), sqrt(max eps) when the profile guides nothing (boundaries.py:32-53)."""
    e = np.asarray(eps_col, dtype=np.float64)
    k0 = 2 * np.pi / lam_m
    try:
        import scipy.sparse as sps
        from scipy.sparse.linalg import eigsh
        N = len(e)
        diag = -2.0 / dx ** 2 + k0 ** 2 * e
        off = np.full(N - 1, 1.0 / dx ** 2)
        A = sps.diags([off, diag, off], [-1, 0, 1], format='csc')
        beta2 = eigsh(A, k=1, sigma=k0 ** 2 * float(e.max()), which='LM',
                      return_eigenvectors=False)[0]
        ne = float(np.sqrt(max(beta2, 0.0)) / k0)
        if np.isfinite(ne) and 1.0 <= ne <= np.sqrt(e.max()) + 1e-6:
            return ne
    except Exception:
        pass
    return float(np.sqrt(e.max()))


def mur_speeds(eps_r, dx, lam_m):
    """Local indices of the Mur edges from eps_r (Nx, Ny): the x ends take the guided
    mode's n_eff of their column, the y edges the n of their cells (boundaries.py:56-65)."""
    e = np.asarray(eps_r, dtype=np.float64)
    return dict(xl=guided_neff(e[1, :], dx, lam_m), xr=guided_neff(e[-2, :], dx, lam_m),
                b=np.sqrt(e[:, 0]), t=np.sqrt(e[:, -1]))


class MurSet:
    """Mur-1 coefficients per edge: scalars on the x ends, vectors along the y edges
    (b/t on the Nx nodes, bh/th on the Nx-1 half cells of TM Ex), boundaries.py:68-93."""

    def __init__(self, dx, dt, speeds=None):
        sp = speeds or {}
        self.xl = mur1_coeff(dx, dt, sp.get('xl', 1.0))
        self.xr = mur1_coeff(dx, dt, sp.get('xr', 1.0))
        if 'b' in sp:
            nb, nt = np.asarray(sp['b']), np.asarray(sp['t'])
            self.b, self.t = mur1_coeff(dx, dt, nb), mur1_coeff(dx, dt, nt)
            self.bh = mur1_coeff(dx, dt, 0.5 * (nb[:-1] + nb[1:]))
            self.th = mur1_coeff(dx, dt, 0.5 * (nt[:-1] + nt[1:]))
        else:
            self.b = self.t = self.bh = self.th = mur1_coeff(dx, dt)
        self.speeds = sp

    def describe(self):
        sp = self.speeds
        if not sp:
            return 'Mur (vacuum c0)'
        d = 'Mur with local velocity: x ends n_eff=%.4f / %.4f' % (sp.get('xl', 1), sp.get('xr', 1))
        if 'b' in sp:
            d += ', y edges n=%.3f..%.3f / %.3f..%.3f' % (np.min(sp['b']), np.max(sp['b']),
                                                         np.min(sp['t']), np.max(sp['t']))
        return d


def attach_mur_2d(engine, pol, ms, x_lo=True, x_hi=True, y=True):
    """The reference's inline first-order Mur of a 2D run as engine faces, in its order:
    the x ends (TE Ez, TM Ey) with F0 = the current boundary value and the PEC of the
    PML edges after them, then the y edges (TE Ez, TM Ex) (fdtd2d_core.py:805-810,
    831-833; stitch :1422-1433, :1458-1469).  Call after the CPML profiles are attached."""
    te = str(pol).lower() in ('s', 'te')
    cx, cy = ('Ez', 'Ez') if te else ('Ey', 'Ex')
    yb, yt = (ms.b, ms.t) if te else (ms.bh, ms.th)
    if x_lo:
        engine.add_mur(cx, 0, 0, ms.xl, 3)
    if x_hi:
        engine.add_mur(cx, 0, 1, ms.xr, 3)
    if y:
        engine.add_mur(cy, 1, 0, yb, 1)
        engine.add_mur(cy, 1, 1, yt, 1)
