"""3D CPML descriptor (Roden-Gedney, CFS) for the Yee step on the device.

Reference: planar_3d/boundaries3d.py
  _cpml_bc_1d      :150-194  same profile formulas as the 2D CPML (shared helper here)
  _PsiStrip        :201-251  psi stored only in the PML slabs of its stretch axis
  CPML_3D          :262-362  12 psi components; PEC (tangential E = 0) on the CPML faces;
                             step / step_H / step_E on device arrays (lf_op3d_*)
  make_boundary_3d :365-377
  mur1_coeff, mur_speeds_3d, _face_n, _TANG, Mur1_3D   :24-140   first-order Mur on the
                   absorbing faces (local speeds: x faces n_eff, y/z faces n of the cells)
  sponge (fdtd3d_core.py:397-414)  graded exp(-alpha u^3) damping of all six fields

The engine stores psi per slab (lo/hi part of each axis where c != 0) in HBM; this
descriptor only carries the 1-D profiles and the set of faces.
"""
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d

ALL_FACES = ('xmin', 'xmax', 'ymin', 'ymax', 'zmin', 'zmax')
_AXIS_FACES = (('xmin', 'xmax'), ('ymin', 'ymax'), ('zmin', 'zmax'))


class CPML3D:
    """Profiles per axis; PEC on the CPML faces and on ``pec_faces`` (mirror faces without
    a PML, reference fdtd3d_core.py:417-425)."""

    def __init__(self, Nx, Ny, Nz, dx, dt, faces, npml=10, pec_faces=(), **kw):
        self.faces = tuple(faces)
        self.pec_faces = tuple(pec_faces)
        self.npml = int(npml)
        self.prof = []
        for ax, N in enumerate((Nx, Ny, Nz)):
            lo = self.npml if _AXIS_FACES[ax][0] in self.faces else 0
            hi = self.npml if _AXIS_FACES[ax][1] in self.faces else 0
            self.prof.append(cpml_profile_1d(N, lo, hi, dx, dt, **kw))

    def pec(self, ax):
        return tuple(f in self.faces or f in self.pec_faces for f in _AXIS_FACES[ax])

    def attach(self, engine):
        for ax in range(3):
            engine.set_cpml(ax, self.prof[ax][0], self.prof[ax][1], *self.pec(ax))


class CPML_3D(CPML3D):
    """The reference's CPML_3D object (boundaries3d.py:262-362) on device arrays.

    step / step_H / step_E update float32, contiguous device tensors (torch on the HIP
    device) in the reference's shapes, in place, on the current stream, through the
    operator-level kernels of liblensfdtd.so (lf_op3d_*).  The 12 psi slabs live in the
    library object, in the PML slabs only, as _PsiStrip keeps them (:201-251).  It
    attaches to a step engine like the descriptor it extends.  Host arrays are refused:
    there is no CPU path."""

    def __init__(self, Nx, Ny, Nz, dx, dt, faces, npml=10, **kw):
        super().__init__(Nx, Ny, Nz, dx, dt, faces, npml=npml, **kw)
        self.shape = (int(Nx), int(Ny), int(Nz))
        self._h = None

    def _op(self, *arrays):
        import ctypes
        import numpy as np
        from lens_sensor_amd.engine import hip
        from lens_sensor_amd.planar_beads import fused2d
        if not fused2d.available(*arrays):
            raise TypeError('CPML_3D needs float32, contiguous device tensors on the HIP device '
                            '(there is no host path)')
        if self._h is None:
            h = ctypes.c_void_p()
            hip.op_call('lf_op3d_create', *self.shape, ctypes.byref(h))
            fp = ctypes.POINTER(ctypes.c_float)
            keep = []
            for ax in range(3):
                e, hh = self.prof[ax]
                arrs = [np.ascontiguousarray(a, np.float32) for a in (*e, *hh)]
                keep += arrs
                hip.op_call('lf_op3d_set_cpml', h, ax, *[a.ctypes.data_as(fp) for a in arrs],
                            *[int(f) for f in self.pec(ax)])
            hip.op_call('lf_op3d_finalize', h)
            self._h = h
        return self._h

    def step(self, Ex, Ey, Ez, Hx, Hy, Hz, ce_x, ce_y, ce_z, Ch, da=None):
        """One Yee step with CPML: H, then E (boundaries3d.py:305-309)."""
        self.step_H(Ex, Ey, Ez, Hx, Hy, Hz, Ch)
        self.step_E(Ex, Ey, Ez, Hx, Hy, Hz, ce_x, ce_y, ce_z, da=da)

    def step_H(self, Ex, Ey, Ez, Hx, Hy, Hz, Ch):
        """H^{n+1/2} from E^n with the psi corrections in the slabs (:311-330)."""
        import numpy as np
        from lens_sensor_amd.engine import hip
        self._check(Ex, Ey, Ez, Hx, Hy, Hz)
        h = self._op(Ex, Ey, Ez, Hx, Hy, Hz)
        hip.op_call('lf_op3d_step_H', h, *[a.data_ptr() for a in (Ex, Ey, Ez, Hx, Hy, Hz)],
                    float(np.float32(Ch)), self._stream(Ex))

    def step_E(self, Ex, Ey, Ez, Hx, Hy, Hz, ce_x, ce_y, ce_z, da=None):
        """Optional loss E <- D E, interior E update with psi, PEC on the CPML faces
        (:332-362)."""
        from lens_sensor_amd.engine import hip
        self._check(Ex, Ey, Ez, Hx, Hy, Hz)
        extra = [ce_x, ce_y, ce_z] + (list(da) if da is not None else [])
        h = self._op(Ex, Ey, Ez, Hx, Hy, Hz, *extra)
        for c, f in zip((ce_x, ce_y, ce_z) + (tuple(da) if da is not None else ()),
                        (Ex, Ey, Ez, Ex, Ey, Ez)):
            if tuple(c.shape) != tuple(f.shape):
                raise ValueError('coefficient and field shapes differ')
        dp = [a.data_ptr() for a in da] if da is not None else [None, None, None]
        hip.op_call('lf_op3d_step_E', h, *[a.data_ptr() for a in (Ex, Ey, Ez, Hx, Hy, Hz)],
                    ce_x.data_ptr(), ce_y.data_ptr(), ce_z.data_ptr(), *dp, self._stream(Ex))

    def _check(self, Ex, Ey, Ez, Hx, Hy, Hz):
        X, Y, Z = self.shape
        want = ((X - 1, Y, Z), (X, Y - 1, Z), (X, Y, Z - 1), (X, Y - 1, Z - 1), (X - 1, Y, Z - 1),
                (X - 1, Y - 1, Z))
        for a, w, n in zip((Ex, Ey, Ez, Hx, Hy, Hz), want, ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')):
            if tuple(a.shape) != w:
                raise ValueError(f'{n} has shape {tuple(a.shape)}, expected {w}')

    @staticmethod
    def _stream(a):
        import torch
        return torch.cuda.current_stream(a.device).cuda_stream

    def close(self):
        if self._h is not None:
            from lens_sensor_amd.engine import hip
            hip.load().lf_op3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_boundary_3d(kind, dx, dt, faces, shape=None, npml=10, speeds=None, lam_m=None,
                     pec_faces=()):
    """Boundary object (reference boundaries3d.py:365-377): CPML_3D for 'cpml', Mur1_3D for
    'mur1', None for 'sponge' (the sponge is applied by the driver, Sponge3D)."""
    kind = (kind or 'cpml').lower()
    if kind == 'sponge':
        return None
    if kind == 'mur1':
        return Mur1_3D(dx, dt, faces, speeds=speeds)
    if kind == 'cpml':
        if shape is None:
            raise ValueError('cpml (3D) requires shape=(Nx,Ny,Nz)')
        return CPML_3D(shape[0], shape[1], shape[2], dx, dt, faces, npml=npml, lam_m=lam_m,
                       pec_faces=pec_faces)
    raise ValueError(f"unknown boundary '{kind}' (sponge|cpml|mur1)")


# ----------------------------------------------------------------------------- Mur-1 (3D)
def mur1_coeff(dx, dt, n=1.0):
    """Mur-1 coefficient at phase velocity c0/n (boundaries3d.py:24-28), float64."""
    from lens_sensor_amd.common.physics import C0
    import numpy as np
    c = C0 / np.asarray(n, dtype=np.float64)
    m = (c * dt - dx) / (c * dt + dx)
    return float(m) if np.ndim(m) == 0 else m


def mur_speeds_3d(eps, dx, lam_m):
    """Local indices per face from eps (Nx, Ny, Nz): x faces the n_eff of the guided mode
    of the y profile at the z centre, y/z faces the n of their cells (:31-44)."""
    import numpy as np
    from lens_sensor_amd.planar_beads.boundaries import guided_neff
    e = np.asarray(eps)
    kz = e.shape[2] // 2
    return {'xmin': guided_neff(e[1, :, kz], dx, lam_m),
            'xmax': guided_neff(e[-2, :, kz], dx, lam_m),
            'ymin': np.sqrt(e[:, 0, :].astype(np.float64)),
            'ymax': np.sqrt(e[:, -1, :].astype(np.float64)),
            'zmin': np.sqrt(e[:, :, 0].astype(np.float64)),
            'zmax': np.sqrt(e[:, :, -1].astype(np.float64))}


def _face_n(speeds, face, shape):
    """n on a component's boundary plane: the face's array cut or extended (last row /
    column repeated) to the plane shape; the half-cell offsets of the Yee components are
    ignored, as in the reference (:47-66)."""
    import numpy as np
    n = (speeds or {}).get(face, 1.0)
    if np.ndim(n) == 0:
        return float(n)
    n = np.asarray(n)
    a, b = shape
    out = np.empty((a, b))
    ra, rb = min(a, n.shape[0]), min(b, n.shape[1])
    out[:ra, :rb] = n[:ra, :rb]
    if a > n.shape[0]:
        out[n.shape[0]:, :] = out[n.shape[0] - 1:n.shape[0], :]
    if b > n.shape[1]:
        out[:, n.shape[1]:] = out[:, n.shape[1] - 1:n.shape[1]]
    return out


# tangential E components of each face: (component, normal axis, high end), in the
# reference's order (:78-85)
_TANG = {'xmin': (('Ey', 0, 0), ('Ez', 0, 0)), 'xmax': (('Ey', 0, 1), ('Ez', 0, 1)),
         'ymin': (('Ex', 1, 0), ('Ez', 1, 0)), 'ymax': (('Ex', 1, 1), ('Ez', 1, 1)),
         'zmin': (('Ex', 2, 0), ('Ey', 2, 0)), 'zmax': (('Ex', 2, 1), ('Ey', 2, 1))}


def _comp_plane_shape(comp, axis, shape):
    Nx, Ny, Nz = shape
    ext = {'Ex': (Nx - 1, Ny, Nz), 'Ey': (Nx, Ny - 1, Nz), 'Ez': (Nx, Ny, Nz - 1)}[comp]
    return tuple(e for d, e in enumerate(ext) if d != axis)


class Mur1_3D:
    """First-order Mur on the absorbing faces (boundaries3d.py:99-140): the tangential E
    planes of each face take F0 = F1_prev + m (F1_new - F0_prev) after every step, with m
    from the local speed of the face (set_speeds, mur_speeds_3d).  attach() hands the
    faces to a step engine in the reference's order; the engine captures the planes
    before each step and updates them after the source."""

    def __init__(self, dx, dt, faces, speeds=None):
        self.dx, self.dt = dx, dt
        self.faces = tuple(faces)
        self.set_speeds(speeds)

    def set_speeds(self, speeds):
        self.speeds = speeds

    def attach(self, engine, shape, pec_faces=()):
        """Zero-width profiles on every axis (PEC on pec_faces), then the Mur faces."""
        zero = CPML3D(*shape, self.dx, self.dt, (), npml=0, pec_faces=pec_faces)
        zero.attach(engine)
        self.attach_faces(engine, shape)

    def attach_faces(self, engine, shape, pec_after=True):
        """The Mur faces in the reference's order.  F0_prev is the plane before the step
        (flag bit 0 clear).  pec_after: the run applies the PEC faces after the Mur update
        (run_3d's mirror faces, fdtd3d_core.py:1137-1138); in the stitch windows the seam
        CPML's PEC comes inside the step, before it (:1619-1634)."""
        for face in ALL_FACES:
            if face not in self.faces:
                continue
            for comp, axis, hi in _TANG[face]:
                n = _face_n(self.speeds, face, _comp_plane_shape(comp, axis, shape))
                engine.add_mur(comp, axis, hi, mur1_coeff(self.dx, self.dt, n),
                               2 if pec_after else 0)


class Sponge3D:
    """Graded sponge of n_sp cells on the faces it absorbs: per step every field in the
    outer slabs is multiplied by exp(-alpha u^3), u = depth (n_sp - d)/n_sp, float32
    (fdtd3d_core.py:397-414)."""

    def __init__(self, n_sp, alpha, faces):
        import numpy as np
        self.n_sp = int(n_sp)
        u = (self.n_sp - np.arange(self.n_sp, dtype=np.float64)) / self.n_sp
        self.g = np.exp(-alpha * u ** 3).astype(np.float32)
        self.faces = tuple(faces)

    @property
    def mask(self):
        return sum(1 << ALL_FACES.index(f) for f in self.faces)

    def attach(self, engine, shape, dx, dt, pec_faces=()):
        zero = CPML3D(*shape, dx, dt, (), npml=0, pec_faces=pec_faces)
        zero.attach(engine)
        self.attach_faces(engine)

    def attach_faces(self, engine):
        if self.faces:
            engine.set_sponge(self.n_sp, self.g, self.mask)
