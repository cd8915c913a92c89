"""Writes tests/golden/ref_*.npz: input/output vectors of the reference solver.

Runs the reference's NumPy path (planar_beads/fdtd2d_core.py, planar_3d/fdtd3d_core.py,
imported from /root/reference in this container only) on small cases and stores the
results as fixtures.  Why an import is needed: the golden master committed in the
reference (tests/test_characterization.py:37-64) was recorded before the VCSEL
source-in-air geometry (common/vcsel.py:71-83, fdtd2d_core.py:631-634,
fdtd3d_core.py:533-537) and is not reproduced by the current reference code; the
three golden cases are re-run here together with TM, lossy and stitch cases that no
published value covers.

Usage (writes into this directory):  python tests/golden/make_golden.py [2d_<case> 3d_<case> ...]
The fixtures are data only (inputs, reduced outputs and small arrays); no reference
source is stored.
"""
import builtins
import contextlib
import io
import os
import shutil
import sys
import tempfile

import numpy as np

REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference():
    for p in (REF, os.path.join(REF, 'planar_3d'), os.path.join(REF, 'planar_beads')):
        if p not in sys.path:
            sys.path.insert(0, p)
    with contextlib.redirect_stdout(io.StringIO()):
        import fdtd2d_core as d2
        import fdtd3d_core as f3d
    return d2, f3d


def _red(a):
    a = np.asarray(a, dtype=np.float64)
    return np.array([a.sum(), np.abs(a).max(), a.mean()])


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


CASES_2D = {
    # the reference's own characterization case (tests/test_characterization.py:102-116)
    'te_full_char': dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=0.5, dx_nm=100.0,
                         save_frames=True, n_snapshots=2, wg_thickness_um=1.0, lambda_nm=850.0,
                         method='full', length_um=8.0, bead_x_um=4.0),
    'tm_full_lossy': dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=0.6, dx_nm=100.0,
                          save_frames=True, n_snapshots=3, wg_thickness_um=1.0, lambda_nm=850.0,
                          method='full', length_um=7.0, bead_x_um=3.5, polarization='p',
                          kappa={'aq': 0.001, 'bead': 0.002}),
    'te_stitch': dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=0.5, dx_nm=100.0,
                      save_frames=True, n_snapshots=2, wg_thickness_um=1.0, lambda_nm=850.0,
                      method='stitch', length_um=9.0, bead_x_um=5.0, window_w_um=4.0,
                      slide_um=2.5),
    'tm_stitch_lossy': dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=0.6, dx_nm=100.0,
                            save_frames=True, n_snapshots=2, wg_thickness_um=1.0,
                            lambda_nm=850.0, method='stitch', length_um=9.0, bead_x_um=5.5,
                            window_w_um=4.0, slide_um=2.5, polarization='p',
                            kappa={'aq': 0.001, 'bead': 0.002}),
}
# the same cases with the first-order Mur outer boundary (fdtd2d_core.py:805-833, stitch
# :1422-1469)
for _n in ('te_full_char', 'tm_full_lossy', 'te_stitch', 'tm_stitch_lossy'):
    CASES_2D[_n + '_mur1'] = dict(CASES_2D[_n], boundary='mur1')
# the Fourier pulse stitch (fdtd2d_core.py:1866-2017): three comb lines of a 10-period
# pulse; TE with every index fixed at lambda0, lossy TM with a layer stack whose material
# layers follow n(lambda_k)
PULSE = dict(source_type='pulse', pulse_cycles=10.0, pulse_nfreq=3)
STACK_PULSE = [dict(name='Aqueous', t_um=1.0, n=complex(1.329, 0.001), kappa=0.001, mat='aqueous'),
               dict(name='Mucin', t_um=0.5, n=complex(1.34, 0.0), kappa=0.0, mat=None)]
CASES_2D['te_stitch_pulse'] = dict(CASES_2D['te_stitch'], pulse_dispersion=False, **PULSE)
CASES_2D['tm_stitch_pulse_stack'] = dict(CASES_2D['tm_stitch_lossy'], stack=STACK_PULSE, **PULSE)
# the time-domain TF/SF stitch (fdtd2d_core.py:1586-1863): windows of 4 um every 2 um
# (the TF/SF plane needs an overlap of npml + 6 cells), TE and lossy TM
TD = dict(source_type='pulse', pulse_mode='td', pulse_cycles=10.0, slide_um=2.0)
CASES_2D['te_stitch_td'] = dict(CASES_2D['te_stitch'], **TD)
CASES_2D['tm_stitch_td_lossy'] = dict(CASES_2D['tm_stitch_lossy'], **TD)

CASES_3D = {
    'te_full_char': dict(fn='run_3d', label='golden3d', wg_n=1.491, lx_um=4.0, air_um=1.0,
                         tear_um=2.0, lz_um=2.0, dx_nm=100.0, t_wg_um=1.0, t_aq_um=1.0,
                         t_mu_um=0.5, lam_nm=850.0, n_snapshots=2, steps_factor=1.0,
                         vcsel_waist_um=1.0, vcsel_waist_z_um=1.0,
                         bead=dict(x_um=2.0, d_um=0.5, n=1.59), n_sponge=8),
    'te_stitch_char': dict(fn='run_3d_stitched', label='goldenstitch', wg_n=1.491,
                           window_w_um=2.0, slide_um=1.2, lx_um=4.0, air_um=1.0, tear_um=2.0,
                           lz_um=2.0, dx_nm=100.0, t_wg_um=1.0, t_aq_um=1.0, t_mu_um=0.5,
                           lam_nm=850.0, n_snapshots=2, vcsel_waist_um=1.0,
                           vcsel_waist_z_um=1.0, bead=dict(x_um=2.0, d_um=0.5, n=1.59),
                           n_sponge=8),
    'tm_full_lossy': dict(fn='run_3d', label='tm3d', wg_n=1.491, lx_um=4.0, air_um=0.8,
                          tear_um=2.0, lz_um=2.4, dx_nm=100.0, t_wg_um=1.0, t_aq_um=1.0,
                          t_mu_um=0.5, lam_nm=850.0, n_snapshots=2, steps_factor=1.0,
                          vcsel_waist_um=1.0, vcsel_waist_z_um=1.0, polarization='p',
                          bead=dict(x_um=3.0, d_um=0.6, n=1.59, z_um=0.1), n_sponge=8,
                          kappa={'aq': 0.001, 'bead': 0.002}),
}
# the graded sponge and the first-order Mur as outer absorbers (fdtd3d_core.py:397-414,
# 1112-1135; boundaries3d.py:99-140), full domain and stitch windows
for _n in ('te_full_char', 'tm_full_lossy', 'te_stitch_char'):
    for _b in ('sponge', 'mur1'):
        CASES_3D[f'{_n}_{_b}'] = dict(CASES_3D[_n], boundary=_b)
# the 3D Fourier pulse stitch (fdtd3d_core.py:2087-2215): three comb lines, the core,
# bead, aqueous and cornea indices following n(lambda_k)
CASES_3D['te_stitch_char_pulse'] = dict(
    CASES_3D['te_stitch_char'], fn='run_3d_stitched_pulse', pulse_cycles=10.0, pulse_nfreq=3,
    disp_map={'wg_n': 'pmma', 'bead': 'polystyrol', 'n_aq': 'aqueous', 'n_co': 'cornea_layered'})
# the 3D time-domain TF/SF stitch (fdtd3d_core.py:1785-2084): CPML on every face, and the
# sponge on every face for the lossy TM case
CASES_3D['te_stitch_char_td'] = dict(
    CASES_3D['te_stitch_char'], fn='run_3d_stitched_td', pulse_cycles=10.0, window_w_um=3.0,
    slide_um=1.2, n_sponge=10)
CASES_3D['tm_stitch_td_sponge'] = dict(
    CASES_3D['tm_full_lossy'], fn='run_3d_stitched_td', label='tm3dtd', pulse_cycles=10.0,
    window_w_um=3.0, slide_um=1.0, n_sponge=8, boundary='sponge')
for _k in ('steps_factor',):
    CASES_3D['tm_stitch_td_sponge'].pop(_k, None)
# windows of at least 54 cells on every axis let the TF/SF plane (i0 = n_sp + 2 = 11) lie
# behind a 10-cell CPML: the CPML path of the 3D time-domain stitch
CASES_3D['te_stitch_td_cpml'] = dict(
    CASES_3D['te_stitch_char'], fn='run_3d_stitched_td', label='tdcpml', pulse_cycles=10.0,
    lx_um=8.0, window_w_um=5.5, slide_um=1.5, tear_um=3.5, lz_um=5.4, n_sponge=None,
    bead=dict(x_um=4.0, d_um=0.5, n=1.59))


def run_2d(d2, name, case):
    case = dict(case)
    kappa = case.pop('kappa', {})
    d2.KAPPA = dict(kappa)
    try:
        r = _quiet(d2.run_beads, **case)
    finally:
        d2.KAPPA = {}
    out = dict(transmission=r['transmission'], P_in=r['P_in'], P_out=r['P_out'])
    frames = r.get('frames') or []
    mm = r.get('frames_memmap')
    if mm and os.path.exists(mm):
        out['frames'] = np.array(np.load(mm, mmap_mode='r')[:len(frames)], dtype=np.float32)
    if r.get('Iavg') is not None:
        out['Iavg'] = np.asarray(r['Iavg'], dtype=np.float32)
    if r.get('Ex_cw_amp') is not None:
        out['Ex_cw_amp'] = np.asarray(r['Ex_cw_amp'], dtype=np.complex64)
    if r.get('spectrum') is not None:
        out['spectrum_T'] = np.asarray(r['spectrum']['T'], dtype=np.float64)
        out['spectrum_lam'] = np.asarray(r['spectrum']['lam_nm'], dtype=np.float64)
    cwm = r.get('cw_frames_memmap')
    if cwm and os.path.exists(cwm):
        out['cw_frames'] = np.array(np.load(cwm, mmap_mode='r'), dtype=np.float32)
    for k in ('cw_rel_change', 'steps_total', 'p_in_si', 'coupling_efficiency',
              'facet_reflection', 'transmission_intensity'):
        if r.get(k) is not None:
            out[k] = float(r[k])
    return out


def run_3d(f3d, name, case):
    case = dict(case)
    kappa = case.pop('kappa', {})
    fn = getattr(f3d, case.pop('fn'))
    f3d.KAPPA = dict(kappa)
    try:
        r = _quiet(fn, **case)
    finally:
        f3d.KAPPA = {}
    out = dict(transmission=r['transmission'], P_in=r['P_in'], P_out=r['P_out'],
               Iavg=np.asarray(r['Iavg'], dtype=np.float32),
               vols=np.asarray(r['vols_array'], dtype=np.float32))
    if r.get('spectrum') is not None:
        out['spectrum_T'] = np.asarray(r['spectrum']['T'], dtype=np.float64)
        out['spectrum_lam'] = np.asarray(r['spectrum']['lam_nm'], dtype=np.float64)
    if r.get('cw_vol') is not None:
        out['cw_vol'] = np.asarray(r['cw_vol'], dtype=np.float32)
    for k in ('cw_rel_change', 'steps_total', 'p_in_si', 'coupling_efficiency',
              'facet_reflection', 'transmission_intensity'):
        if r.get(k) is not None:
            out[k] = float(r[k])
    return out


def setup_vectors(d2, f3d):
    """Coefficient builders on a small grid: pins the material/coefficient host code
    (fdtd2d_core.py:122-349, :445-455; fdtd3d_core.py:119-378, :593-613) bit for bit."""
    out = {}
    dx = 50e-9
    dt = 0.5 * dx / (d2.C0 * np.sqrt(2))
    bp = {}
    d2.KAPPA = {'aq': 0.0015, 'bead': 0.002}
    d2.LAM_LOSS = 850e-9
    tear_old = d2.TEAR_BUFFER
    d2.TEAR_BUFFER = 1.5e-6
    try:
        eps, sig = _quiet(d2.build_materials, 60, 70, dx, 1.4835, 1.5764, 1.0e-6, 1.3, 0.8,
                          0.0, True, n_aq=1.3313, n_mucin=1.3331, n_cornea=d2.cornea_default(850.0),
                          t_aq_um=1.0, t_mu_um=0.4, t_lip_um=0.05, n_lip=1.4729,
                          input_gap_um=0.6, length_um=3.0, tm_patch=bp)
    finally:
        d2.KAPPA = {}
        d2.TEAR_BUFFER = tear_old
    out['te2d_eps'], out['te2d_sig'] = np.asarray(eps), np.asarray(sig)
    ceE, ceH = d2._coeffs(eps, sig, dt, dx)
    out['te2d_ceE'], out['te2d_ceH'] = np.asarray(ceE), np.asarray(ceH)
    cax, cay, ch = d2._tm_coeffs(eps, dt, dx)
    d2._apply_bead_tm(bp, cax, cay, dt, dx)
    dax, day = d2._tm_loss(eps, sig, cax, cay, dt)
    out.update(tm2d_cax=np.asarray(cax), tm2d_cay=np.asarray(cay), tm2d_dax=np.asarray(dax),
               tm2d_day=np.asarray(day))
    # 3D
    dxu = 0.05
    dx3 = dxu * 1e-6
    dt3 = 0.5 * dx3 / (f3d.C0 * np.sqrt(3))
    tp = {}
    f3d.KAPPA = {'aq': 0.0015, 'bead': 0.002}
    try:
        e3 = _quiet(f3d.build_eps3d, 30, 44, 26, dxu, -1.2, -0.65, 1.4835, 0.8, 0.05, 0.6, 0.3,
                    1.4729, 1.3313, 1.3331, 1.3704,
                    bead=dict(x_um=0.8, d_um=0.5, n=1.5764, z_um=0.05, y_um=-0.26),
                    x_wg_start_um=0.3, tensor_patch=tp)
    finally:
        f3d.KAPPA = {}
    eps3, eim3 = f3d._split_eps3(e3)
    inv, da = f3d._inv_loss(eps3, eim3, dt3, dx3, 850e-9)
    cex, cey, cez = f3d._ce_arrays(inv, tp, dt3, dx3, dxu, da=da)
    dax3, day3, daz3 = f3d._stagger_da(da)
    out.update(y3d_epsr=np.asarray(eps3), y3d_epsi=np.asarray(eim3), y3d_cex=np.asarray(cex),
               y3d_cey=np.asarray(cey), y3d_cez=np.asarray(cez), y3d_dax=np.asarray(dax3),
               y3d_day=np.asarray(day3), y3d_daz=np.asarray(daz3))
    return out


def main(only=None):
    """only: names (e.g. 2d_te_full_char_mur1) to write; None writes every fixture."""
    d2, f3d = _import_reference()
    prev = os.getcwd()
    tmp = tempfile.mkdtemp(prefix='lens_fixtures_')
    try:
        os.chdir(tmp)
        if only:
            for key in only:
                dim, name = key.split('_', 1)
                res = (run_2d(d2, name, CASES_2D[name]) if dim == '2d'
                       else run_3d(f3d, name, CASES_3D[name]))
                np.savez_compressed(os.path.join(HERE, f'ref_{key}.npz'), **res)
                print(key, res['transmission'], res['P_in'], res['P_out'], flush=True)
            return
        for name, case in CASES_2D.items():
            res = run_2d(d2, name, case)
            np.savez_compressed(os.path.join(HERE, f'ref_2d_{name}.npz'), **res)
            print('2d', name, res['transmission'], res['P_in'], res['P_out'], flush=True)
        for name, case in CASES_3D.items():
            res = run_3d(f3d, name, case)
            np.savez_compressed(os.path.join(HERE, f'ref_3d_{name}.npz'), **res)
            print('3d', name, res['transmission'], res['P_in'], res['P_out'], flush=True)
        np.savez_compressed(os.path.join(HERE, 'ref_setup.npz'), **setup_vectors(d2, f3d))
        print('setup vectors written', flush=True)
    finally:
        os.chdir(prev)
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == '__main__':
    main(sys.argv[1:] or None)
