"""3D CPML descriptor (Roden-Gedney, CFS) for the Yee step on the device.

Reference: planar_3d/boundaries3d.py
  _cpml_bc_1d      :150-194  same profile formulas as the 2D CPML (shared helper here)
  _PsiStrip        :201-251  psi stored only in the PML slabs of its stretch axis
  CPML_3D          :262-362  12 psi components; PEC (tangential E = 0) on the CPML faces
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


def make_boundary_3d(kind, dx, dt, faces, shape=None, npml=10, lam_m=None):
    kind = (kind or 'cpml').lower()
    if kind == 'cpml':
        if shape is None:
            raise ValueError('cpml (3D) requires shape=(Nx,Ny,Nz)')
        return CPML3D(shape[0], shape[1], shape[2], dx, dt, faces, npml=npml, lam_m=lam_m)
    if kind in ('sponge', 'mur1'):
        raise NotImplementedError(f"boundary '{kind}' is not part of the device path; use 'cpml'")
    raise ValueError(f"unknown boundary '{kind}' (cpml)")
