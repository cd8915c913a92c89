"""Step-engine contract shared by the drivers and every engine implementation.

An engine owns the Yee state of one domain (fields, CPML psi slabs, per-cell update
coefficients, source planes, monitor accumulators) and advances it in chunks of steps.
The host never touches fields inside a chunk: everything the reference's time loop does
per step (planar_beads/fdtd2d_core.py:781-873, planar_3d/fdtd3d_core.py:1105-1197,
stitch loops fdtd2d_core.py:1403-1474 and fdtd3d_core.py:1616-1656) is expressed as

  * the Yee update itself (kind 'te2d', 'tm2d' or 'yee3d'),
  * per-step scalars in a step table (envelope, phases, DFT phasors),
  * a bit mask per step that selects which registered monitors accumulate.

The step table is a packed C struct (``lf_step`` in include/lensfdtd.h); the host fills
it with the same NumPy expressions the reference evaluates per step, so every engine
sees bit-identical scalars.

Modes:
  'ref32'   float32 fields with the reference's float32 arithmetic: float32 source phase,
            complex64 DFT/flux accumulators (the golden-master behaviour).
  'fp64acc' float32 fields, source phase reduced in FP64 on the host, FP64 DFT and flux
            accumulators (long runs; 1e-4 relative L2 against FP64).
  'f64'     everything in float64 (the CPU oracle's FP64 reference only).
"""
import numpy as np

KINDS = ('te2d', 'tm2d', 'yee3d')
MODES = ('ref32', 'fp64acc', 'f64')

STEP_DTYPE = np.dtype([
    ('env', '<f8'),                 # source envelope (float64)
    ('ssin', '<f8'), ('scos', '<f8'),  # env*sin(wt), env*cos(wt)   (FP64 phase source)
    ('pr', '<f8'), ('pi', '<f8'),   # exp(-i w t)
    ('hr', '<f8'), ('hi', '<f8'),   # exp(-i w (t - dt/2))        (H flux phasor)
    ('dr', '<f8'), ('di', '<f8'),   # exp(+i w t)                  (stitch drive)
    ('wt32', '<f4'), ('e32', '<f4'),  # float32(w t), float32(env)  (float32 source)
    ('pr32', '<f4'), ('pi32', '<f4'),  # complex64(exp(-i w t))
    ('hr32', '<f4'), ('hi32', '<f4'),  # complex64(exp(-i w (t - dt/2)))
    ('qr32', '<f4'), ('qi32', '<f4'),  # (float32(cos w t), -float32(sin w t))
    ('dr32', '<f4'), ('di32', '<f4'),  # complex64(exp(+i w t))
    ('mask', '<u4'), ('row', '<u4'),   # monitor bits; row of the plane records / TF/SF series
], align=False)
assert STEP_DTYPE.itemsize == 120

# Field component shapes (C order, x slowest) as in the reference:
#   2D TE  Ez (Nx,Ny)   Hx (Nx,Ny-1)  Hy (Nx-1,Ny)          fdtd2d_core.py:702-704
#   2D TM  Ex (Nx-1,Ny) Ey (Nx,Ny-1)  Hz (Nx-1,Ny-1)        fdtd2d_core.py:706-708
#   3D     Ex (Nx-1,Ny,Nz) Ey (Nx,Ny-1,Nz) Ez (Nx,Ny,Nz-1)
#          Hx (Nx,Ny-1,Nz-1) Hy (Nx-1,Ny,Nz-1) Hz (Nx-1,Ny-1,Nz) fdtd3d_core.py:957-962
_OFF = {
    'te2d': {'Ez': (0, 0), 'Hx': (0, 1), 'Hy': (1, 0)},
    'tm2d': {'Ex': (1, 0), 'Ey': (0, 1), 'Hz': (1, 1)},
    'yee3d': {'Ex': (1, 0, 0), 'Ey': (0, 1, 0), 'Ez': (0, 0, 1),
              'Hx': (0, 1, 1), 'Hy': (1, 0, 1), 'Hz': (1, 1, 0)},
}

E_COMPS = {'te2d': ('Ez',), 'tm2d': ('Ex', 'Ey'), 'yee3d': ('Ex', 'Ey', 'Ez')}
H_COMPS = {'te2d': ('Hx', 'Hy'), 'tm2d': ('Hz',), 'yee3d': ('Hx', 'Hy', 'Hz')}
COEF_NAMES = {
    'te2d': ('ceE', 'ceH'),
    'tm2d': ('cax', 'cay', 'dax', 'day'),
    'yee3d': ('cex', 'cey', 'cez', 'dax', 'day', 'daz'),
}
# coefficient -> component whose grid it lives on
COEF_GRID = {'ceE': 'node', 'ceH': 'node', 'cax': 'Ex', 'cay': 'Ey', 'dax': 'Ex',
             'day': 'Ey', 'cex': 'Ex', 'cey': 'Ey', 'cez': 'Ez', 'daz': 'Ez'}


def comp_shape(kind, dims, comp):
    if comp == 'node':
        return tuple(dims)
    off = _OFF[kind][comp]
    return tuple(n - o for n, o in zip(dims, off))


def new_steps(n):
    return np.zeros(int(n), dtype=STEP_DTYPE)


def fill_phasors(tab, omega, dt, t_phys, env, src_omega=None):
    """Per-step scalars with the reference's NumPy expressions.

    t_phys, env: float64 arrays (one entry per step); omega drives the DFT/flux phasors,
    src_omega (default omega) the source phase.  Each line names the reference
    expression it reproduces."""
    t_phys = np.asarray(t_phys, dtype=np.float64)
    env = np.asarray(env, dtype=np.float64)
    wt = (omega if src_omega is None else src_omega) * t_phys   # vcsel.py:154 wt = omega*t
    tab['env'] = env
    tab['e32'] = env.astype(np.float32)                   # env*f32 array: weak scalar -> f32
    tab['wt32'] = wt.astype(np.float32)                   # vcsel.py:155 np.float32(wt)
    tab['ssin'] = env * np.sin(wt)
    tab['scos'] = env * np.cos(wt)
    # whole-table NumPy expressions; element for element the same values as the
    # reference's per-step scalar expressions (tests/test_step_table.py compares them)
    p = np.exp(-1j * omega * t_phys)                      # fdtd2d_core.py:839
    tab['pr'], tab['pi'] = p.real, p.imag
    p32 = p.astype(np.complex64)                          # flux.py:32, stitch :1438
    tab['pr32'], tab['pi32'] = p32.real, p32.imag
    h = np.exp(-1j * omega * (t_phys - 0.5 * dt))         # flux.py:33
    tab['hr'], tab['hi'] = h.real, h.imag
    h32 = h.astype(np.complex64)
    tab['hr32'], tab['hi32'] = h32.real, h32.imag
    d = np.exp(1j * omega * t_phys)                       # stitch drive, fdtd2d_core.py:1426
    tab['dr'], tab['di'] = d.real, d.imag
    d32 = d.astype(np.complex64)                          # fdtd3d_core.py:1627
    tab['dr32'], tab['di32'] = d32.real, d32.imag
    tab['qr32'] = np.cos(omega * t_phys).astype(np.float32)       # fdtd3d_core.py:1638
    tab['qi32'] = -np.sin(omega * t_phys).astype(np.float32)      # fdtd3d_core.py:1639 (acc_im -= ...)
    return tab


def cw_envelope(t_phys, sigma_t):
    """Ramped CW envelope 1 - exp(-(t/(2 sigma_t))^2), as a Python float per step
    (fdtd2d_core.py:787, fdtd3d_core.py:1126)."""
    return np.array([float(1 - np.exp(-((t / (2 * sigma_t)) ** 2)))
                     for t in np.asarray(t_phys, dtype=np.float64).tolist()])


def pulse_envelope(t_phys, f0, pulse_cycles):
    sp = pulse_cycles / f0
    return np.array([float(np.exp(-((t - 3 * sp) / sp) ** 2))
                     for t in np.asarray(t_phys, dtype=np.float64).tolist()])


def resolve_box(shape, box):
    """Python slice semantics of a box (tuple of (start, stop) pairs, negative allowed)
    against one component's shape -> list of (lo, hi)."""
    out = []
    for n, (a, b) in zip(shape, box):
        lo, hi, _ = slice(a, b).indices(n)
        out.append((lo, max(lo, hi)))
    return out
