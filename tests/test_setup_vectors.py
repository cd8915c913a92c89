"""Material and coefficient host code vs the reference's arrays (bit for bit).

tests/golden/ref_setup.npz holds the reference's outputs of build_materials, _coeffs,
_tm_coeffs + _apply_bead_tm + _tm_loss (planar_beads/fdtd2d_core.py:122-349, 445-455)
and build_eps3d, _split_eps3, _inv_loss, _ce_arrays, _stagger_da
(planar_3d/fdtd3d_core.py:119-378, 585-613) on small lossy grids with a bead, a lipid
layer, the layered cornea and an entry facet.
"""
import os

import numpy as np

from lens_sensor_amd.planar_beads import fdtd2d_core as d2
from lens_sensor_amd.planar_3d import fdtd3d_core as f3d
from lens_sensor_amd.common.physics import C0, cornea_default

FIX = os.path.join(os.path.dirname(__file__), 'golden', 'ref_setup.npz')


def _ref():
    with np.load(FIX) as z:
        return {k: z[k] for k in z.files}


def test_2d_materials_and_coefficients_bit_exact():
    ref = _ref()
    dx = 50e-9
    dt = 0.5 * dx / (C0 * np.sqrt(2))
    bp = {}
    old_k, old_t = d2.KAPPA, d2.TEAR_BUFFER
    d2.KAPPA = {'aq': 0.0015, 'bead': 0.002}
    d2.LAM_LOSS = 850e-9
    d2.TEAR_BUFFER = 1.5e-6
    try:
        eps, sig = d2.build_materials(60, 70, dx, 1.4835, 1.5764, 1.0e-6, 1.3, 0.8, 0.0, True,
                                      n_aq=1.3313, n_mucin=1.3331, n_cornea=cornea_default(850.0),
                                      t_aq_um=1.0, t_mu_um=0.4, t_lip_um=0.05, n_lip=1.4729,
                                      input_gap_um=0.6, length_um=3.0, tm_patch=bp)
    finally:
        d2.KAPPA, d2.TEAR_BUFFER = old_k, old_t
    assert np.array_equal(eps, ref['te2d_eps'])
    assert np.array_equal(sig, ref['te2d_sig'])
    ceE, ceH = d2._coeffs(eps, sig, dt, dx)
    assert np.array_equal(ceE, ref['te2d_ceE']) and np.array_equal(ceH, ref['te2d_ceH'])
    cax, cay, _ = d2._tm_coeffs(eps, dt, dx)
    d2._apply_bead_tm(bp, cax, cay, dt, dx)
    dax, day = d2._tm_loss(eps, sig, cax, cay, dt)
    for k, v in dict(cax=cax, cay=cay, dax=dax, day=day).items():
        assert np.array_equal(v, ref['tm2d_' + k]), k


def test_3d_materials_and_coefficients_bit_exact():
    ref = _ref()
    dxu = 0.05
    dx3 = dxu * 1e-6
    dt3 = 0.5 * dx3 / (C0 * np.sqrt(3))
    tp = {}
    old = f3d.KAPPA
    f3d.KAPPA = {'aq': 0.0015, 'bead': 0.002}
    try:
        e3 = f3d.build_eps3d(30, 44, 26, dxu, -1.2, -0.65, 1.4835, 0.8, 0.05, 0.6, 0.3,
                             1.4729, 1.3313, 1.3331, 1.3704,
                             bead=dict(x_um=0.8, d_um=0.5, n=1.5764, z_um=0.05, y_um=-0.26),
                             x_wg_start_um=0.3, tensor_patch=tp)
    finally:
        f3d.KAPPA = old
    epsr, epsi = f3d.split_eps(e3)
    assert np.array_equal(epsr, ref['y3d_epsr']) and np.array_equal(epsi, ref['y3d_epsi'])
    inv, da = f3d.inv_loss(epsr, epsi, dt3, dx3, 850e-9)
    ces = f3d.ce_arrays(inv, tp, dt3, dx3, dxu, da=da)
    das = f3d.stagger_da(da)
    for k, v in zip(('cex', 'cey', 'cez', 'dax', 'day', 'daz'), ces + das):
        assert np.array_equal(v, ref['y3d_' + k]), k
