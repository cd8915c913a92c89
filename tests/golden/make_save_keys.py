"""Writes tests/golden/ref_save_keys.json: the files and keys the reference's result
writers produce for the keyword sets its CLI passes.

Runs, in this container only, the reference's solvers as run_simulation._run_2d / _run_3d
call them (run_simulation.py:290-431: the full keyword set, the post-run result entries,
then save_results / save_results_3d) on small cases, and records for every written file
its name, its NPZ keys with dtype kind and number of dimensions, or its PKL keys.  Why an
import is needed: the result formats are pinned by no published value or fixture of the
reference; tests/test_save_results.py checks the package's writers against this list.
The fixture is data only (keyword inputs and key lists); no reference source is stored.

Usage (writes into this directory):  python tests/golden/make_save_keys.py
"""
import contextlib
import io
import json
import os
import pickle
import shutil
import sys
import tempfile

import numpy as np

REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

NAMES = ['WG', 'Lipid', 'Aqueous', 'Mucin', 'Cornea']

# the keyword set of run_simulation._run_2d (run_simulation.py:295-321), small case
KW_2D = dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=0.5, dx_nm=100.0, save_frames=True,
             n_snapshots=2, wg_thickness_um=1.0, lambda_nm=850.0, vcsel_waist=1.0, vcsel_tilt=0.0,
             vcsel_offset=0.0, source_type='cw', method='full', window_w_um=4.0, slide_um=2.5,
             place_bead=True, length_um=8.0, t_aqueous_um=1.0, t_mucin_um=0.5, t_lipid_um=0.0,
             input_gap_um=0.0, bead_x_um=4.0, bead_y_um=None, wg_n=None, bead_n=None,
             n_aqueous=None, n_mucin=None, n_cornea=None, n_lipid=None, polarization='s',
             curved=False, boundary='cpml', cad=None, pulse_cycles=30.0, pulse_tmax_fs=None,
             pulse_nfreq=None, pulse_dispersion=True, n_aq_tear_offset=None, pulse_mode='fourier',
             tear_buffer_um=None, stack=None)
CASES_2D = {
    '2d_te_full': dict(KW_2D),
    '2d_tm_stitch': dict(KW_2D, method='stitch', polarization='p', length_um=9.0, bead_x_um=5.5),
    # the Fourier pulse stitch (run_beads -> run_beads_stitched_pulse)
    '2d_te_stitch_pulse': dict(KW_2D, method='stitch', source_type='pulse', length_um=9.0,
                               bead_x_um=5.0, pulse_cycles=10.0, pulse_nfreq=3),
}
# the keyword set of run_simulation._run_3d for run_3d (run_simulation.py:360-429)
KW_3D = dict(label='Bead-polystyrol_d0.5_full', wg_n=1.491, t_wg_um=1.0, t_lip_um=0.0,
             t_aq_um=1.0, t_mu_um=0.5, n_lip=1.48, n_aq=1.336, n_mu=1.342, n_co=1.376,
             lam_nm=850.0, lx_um=4.0, air_um=1.0, tear_um=2.0, lz_um=2.0, dx_nm=100.0,
             vcsel_waist_um=1.0, vcsel_waist_z_um=1.0, source_type='cw', n_snapshots=2,
             bead=dict(x_um=2.0, d_um=0.5, n=1.59), vol_dtype='float32', wg_width_um=None,
             n_clad_side=1.0, vcsel_tilt_deg=0.0, vcsel_offset_y_um=0.0, vcsel_offset_z_um=0.0,
             input_gap_um=0.0, polarization='s', stack=None, steps_factor=1.0, calibrate_steps=0,
             save_vector=False, check_resources=False, allow_large=False, boundary='cpml',
             pec_faces=[], end_facet_um=0.0, cad=None, pulse_cycles=30.0)
# run_simulation._run_3d's call of run_3d_stitched_pulse (run_simulation.py:365-413): the
# keywords common to the 3D drivers without source_type, plus the pulse and window ones
_COMMON_3D = ('label', 'wg_n', 't_wg_um', 't_lip_um', 't_aq_um', 't_mu_um', 'n_lip', 'n_aq',
              'n_mu', 'n_co', 'lam_nm', 'lx_um', 'air_um', 'tear_um', 'lz_um', 'dx_nm',
              'vcsel_waist_um', 'vcsel_waist_z_um', 'n_snapshots', 'bead', 'vol_dtype',
              'wg_width_um', 'n_clad_side', 'vcsel_tilt_deg', 'vcsel_offset_y_um',
              'vcsel_offset_z_um', 'input_gap_um', 'polarization', 'stack')
KW_3D_PULSE = dict({k: KW_3D[k] for k in _COMMON_3D}, label='Bead-polystyrol_d0.5_stitch',
                   window_w_um=2.0, slide_um=1.2, pulse_cycles=10.0, pulse_tmax_fs=None,
                   pulse_nfreq=3, pulse_dispersion=True, cad=None, boundary='cpml',
                   disp_map={'wg_n': 'pmma', 'n_mu': 'mucin', 'n_co': 'cornea_layered',
                             'n_lip': 'lipid', 'bead': 'polystyrol', 'n_aq': 'aqueous'})
CASES_3D = {'3d_te_full': dict(KW_3D), '3d_te_stitch_pulse': dict(KW_3D_PULSE)}


def _import_reference():
    for p in (REF, os.path.join(REF, 'planar_3d'), os.path.join(REF, 'planar_beads')):
        if p not in sys.path:
            sys.path.insert(0, p)
    with contextlib.redirect_stdout(io.StringIO()):
        import fdtd2d_core as d2
        import fdtd3d_core as f3d
    return d2, f3d


def describe(out_dir):
    files = {}
    for name in sorted(os.listdir(out_dir)):
        path = os.path.join(out_dir, name)
        if name.endswith('.npz'):
            with np.load(path, allow_pickle=False) as z:
                files[name] = {k: [z[k].dtype.kind, int(z[k].ndim)] for k in z.files}
        elif name.endswith('.pkl'):
            with open(path, 'rb') as f:      # written by the reference in this very run
                files[name] = sorted(pickle.load(f).keys())
    return files


def main():
    d2, f3d = _import_reference()
    prev = os.getcwd()
    out = {}
    for name, kw in list(CASES_2D.items()) + list(CASES_3D.items()):
        tmp = tempfile.mkdtemp(prefix='lens_save_')
        try:
            os.chdir(tmp)
            res_dir = os.path.join(tmp, 'results_out')
            with contextlib.redirect_stdout(io.StringIO()):
                if name.startswith('2d'):
                    r = d2.run_beads(**kw)
                    r['layer_names'] = NAMES
                    r['n_lip'] = 1.48
                    d2.save_results(r, out_dir=res_dir)
                elif 'pulse_nfreq' in kw:
                    r = f3d.run_3d_stitched_pulse(**kw)
                    f3d.save_results_3d(r, out_dir=res_dir, prefix='planar3d', material_names=NAMES)
                else:
                    k = dict(kw)
                    k['pec_faces'] = tuple(k['pec_faces'])
                    r = f3d.run_3d(**k)
                    f3d.save_results_3d(r, out_dir=res_dir, prefix='planar3d', material_names=NAMES)
            out[name] = dict(kwargs=kw, files=describe(res_dir))
            print(name, sorted(out[name]['files']), flush=True)
        finally:
            os.chdir(prev)
            shutil.rmtree(tmp, ignore_errors=True)
    with open(os.path.join(HERE, 'ref_save_keys.json'), 'w') as f:
        json.dump(dict(names=NAMES, cases=out), f, indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
