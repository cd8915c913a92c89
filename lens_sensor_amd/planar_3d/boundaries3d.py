"""3D CPML descriptor (Roden-Gedney, CFS) for the Yee step on the device.

Reference: planar_3d/boundaries3d.py
  _cpml_bc_1d      :150-194  same profile formulas as the 2D CPML (shared helper here)
  _PsiStrip        :201-251  psi stored only in the PML slabs of its stretch axis
  CPML_3D          :262-362  12 psi components; PEC (tangential E = 0) on the CPML faces;
                             step / step_H / step_E on device arrays (lf_op3d_*)
  make_boundary_3d :365-377

The engine stores psi per slab (lo/hi part of each axis where c != 0) in HBM; this
descriptor only carries the 1-D profiles and the set of faces.
"""
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d

ALL_FACES = ('xmin', 'xmax', 'ymin', 'ymax', 'zmin', 'zmax')
_AXIS_FACES = (('xmin', 'xmax'), ('ymin', 'ymax'), ('zmin', 'zmax'))


class CPML3D:
    def __init__(self, Nx, Ny, Nz, dx, dt, faces, npml=10, **kw):
        self.faces = tuple(faces)
        self.npml = int(npml)
        self.prof = []
        for ax, N in enumerate((Nx, Ny, Nz)):
            lo = self.npml if _AXIS_FACES[ax][0] in self.faces else 0
            hi = self.npml if _AXIS_FACES[ax][1] in self.faces else 0
            self.prof.append(cpml_profile_1d(N, lo, hi, dx, dt, **kw))

    def pec(self, ax):
        return tuple(f in self.faces for f in _AXIS_FACES[ax])

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


def make_boundary_3d(kind, dx, dt, faces, shape=None, npml=10, lam_m=None):
    """Boundary object for kind 'cpml' (reference boundaries3d.py:365-377)."""
    kind = (kind or 'cpml').lower()
    if kind == 'cpml':
        if shape is None:
            raise ValueError('cpml (3D) requires shape=(Nx,Ny,Nz)')
        return CPML_3D(shape[0], shape[1], shape[2], dx, dt, faces, npml=npml, lam_m=lam_m)
    if kind in ('sponge', 'mur1'):
        raise NotImplementedError(f"boundary '{kind}' is not part of the device path; use 'cpml'")
    raise ValueError(f"unknown boundary '{kind}' (cpml)")
