"""Operator-level API on device arrays (needs an MI355X): the reference's CPML objects
and fused2d functions (CPML_TE / CPML_TM .update_H / .update_E / .terminate,
boundaries.py:164-274; CPML_3D.step, boundaries3d.py:262-362), driven from a caller's
own time loop over torch device tensors, against the CPU oracle stepping the same
state.  Without a source the float32 operations are the reference's, in its order, so
fields must agree BIT FOR BIT; the loss factor E <- D E is the caller's multiply
(fdtd2d_core.py:796-797), here a torch multiply."""
import numpy as np
import pytest

from lens_sensor_amd.engine.spec import new_steps, comp_shape, E_COMPS, H_COMPS
from lens_sensor_amd.planar_beads.boundaries import make_boundary
from lens_sensor_amd.planar_3d.boundaries3d import make_boundary_3d
from oracle.engine import OracleEngine

pytestmark = pytest.mark.gpu

DX = 50e-9
DT = 0.5 * DX / (3e8 * np.sqrt(3))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to('cuda:0')


def _fields(kind, dims, rng):
    return {c: rng.standard_normal(comp_shape(kind, dims, c)).astype(np.float32)
            for c in E_COMPS[kind] + H_COMPS[kind]}


def _oracle(kind, dims, ch, coef, bnd, fields, nsteps):
    o = OracleEngine(kind, dims, 'ref32')
    o.set_coeffs(np.float32(ch), **coef)
    if kind == 'yee3d':
        for ax in range(3):
            o.set_cpml(ax, bnd.prof[ax][0], bnd.prof[ax][1], *bnd.pec(ax))
    else:
        for ax in range(2):
            o.set_cpml(ax, bnd.prof[ax][0], bnd.prof[ax][1], *bnd.pec[ax])
    o.finalize()
    for c, a in fields.items():
        o.write_field(c, a)
    o.run(new_steps(nsteps))
    return {c: o.read_field(c) for c in fields}


@pytest.mark.parametrize('pol', ['s', 'p'])
def test_cpml_objects_2d_bit_exact_vs_oracle(pol):
    import torch
    kind = 'te2d' if pol == 's' else 'tm2d'
    dims = (37, 70) if pol == 's' else (41, 66)
    Nx, Ny = dims
    rng = np.random.default_rng(31)
    ch = np.float32(0.4)
    if pol == 's':
        coef = dict(ceE=rng.uniform(0.8, 1.0, dims), ceH=rng.uniform(0.05, 0.2, dims))
    else:
        coef = dict(cax=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ex')),
                    cay=rng.uniform(0.05, 0.2, comp_shape(kind, dims, 'Ey')),
                    dax=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ex')),
                    day=rng.uniform(0.9, 1.0, comp_shape(kind, dims, 'Ey')))
    coef = {k: v.astype(np.float32) for k, v in coef.items()}
    bnd = make_boundary('cpml', pol, Nx, Ny, DX, DT, npml_x=(5, 3), npml_y=(4, 6), lam_m=850e-9)
    f0 = _fields(kind, dims, rng)
    nsteps = 17
    ref = _oracle(kind, dims, ch, coef, bnd, f0, nsteps)
    F = {c: _dev(a) for c, a in f0.items()}
    C = {k: _dev(v) for k, v in coef.items()}
    for _ in range(nsteps):
        if pol == 's':
            bnd.update_H(F['Ez'], F['Hx'], F['Hy'], ch)
            bnd.update_E(F['Ez'], F['Hx'], F['Hy'], C['ceE'], C['ceH'])
            bnd.terminate(F['Ez'])
        else:
            bnd.update_H(F['Ex'], F['Ey'], F['Hz'], ch)
            F['Ex'].mul_(C['dax'])
            F['Ey'].mul_(C['day'])
            bnd.update_E(F['Ex'], F['Ey'], F['Hz'], C['cax'], C['cay'])
            bnd.terminate(F['Ex'], F['Ey'])
    torch.cuda.synchronize()
    for c in f0:
        np.testing.assert_array_equal(F[c].cpu().numpy(), ref[c], err_msg=f'{kind} {c}')


def test_cpml_3d_step_bit_exact_vs_oracle():
    import torch
    kind, dims = 'yee3d', (19, 23, 70)
    rng = np.random.default_rng(32)
    ch = np.float32(0.3)
    coef = {}
    for c, n in (('Ex', 'x'), ('Ey', 'y'), ('Ez', 'z')):
        coef['ce' + n] = rng.uniform(0.05, 0.2, comp_shape(kind, dims, c)).astype(np.float32)
        coef['da' + n] = rng.uniform(0.9, 1.0, comp_shape(kind, dims, c)).astype(np.float32)
    faces = ('xmin', 'xmax', 'ymin', 'zmin', 'zmax')          # ymax open: no PML, no PEC
    bnd = make_boundary_3d('cpml', DX, DT, faces, shape=dims, npml=4, lam_m=850e-9)
    f0 = _fields(kind, dims, rng)
    nsteps = 9
    ref = _oracle(kind, dims, ch, coef, bnd, f0, nsteps)
    F = {c: _dev(a) for c, a in f0.items()}
    C = {k: _dev(v) for k, v in coef.items()}
    for _ in range(nsteps):
        bnd.step(F['Ex'], F['Ey'], F['Ez'], F['Hx'], F['Hy'], F['Hz'], C['cex'], C['cey'], C['cez'],
                 ch, da=(C['dax'], C['day'], C['daz']))
    torch.cuda.synchronize()
    for c in f0:
        np.testing.assert_array_equal(F[c].cpu().numpy(), ref[c], err_msg=f'3d {c}')
    bnd.close()
