"""Result files of the solvers: a PKL summary and NPZ frame/volume files.

The file names and the NPZ / PKL keys are the reference's (save_results,
planar_beads/fdtd2d_core.py:2055-2240; save_results_3d, planar_3d/fdtd3d_core.py:2218-2423),
so that its analyzer and viewers (fdtd_analyzer.py:420-586) read what the drivers of
this package write.  CAD geometry is not produced by the device path, so its optional
keys are never written; the Fourier pulse stitch adds its spectrum keys.

2D:  <out>/beads_<scenario>_<method>.pkl         every scalar / small entry of the result
     <out>/beads_<scenario>_<method>_frames.npz   Ez frames (x stride keeps <= 5000 columns,
                                                   y stride <= 2000 rows), meta rows,
                                                   stack, materials, bead, flags, <|E|^2>,
                                                   TM node-Ex phasor, VCSEL scalars
3D:  <out>/<prefix>_<scenario>.pkl, <prefix>_<scenario>_frames.npz (z mid-plane slices),
     <prefix>_<scenario>_vol3d.npz (snapshot volumes, <|E|^2>, geometry)
"""
import os
import pickle

import numpy as np

from lens_sensor_amd.common.layer_stack import legacy_stack, stack_to_npz

# VCSEL scalars stored when present and finite (common/vcsel.py:189-204)
VCSEL_KEYS = ('laser_power_mw', 'p_src_si', 'coupling_efficiency', 'facet_reflection',
              'vcsel_distance_um', 'vcsel_air_um', 'vcsel_waist_um', 'vcsel_waist_z_um',
              'x_facet_um')
DEFAULT_NAMES = ('WG', 'Lipid', 'Aqueous', 'Mucin', 'Cornea')
_LEGACY_POS = {'lipid': 1, 'aqueous': 2, 'mucin': 3}        # material -> names position


def vcsel_npz(result):
    out = {}
    for k in VCSEL_KEYS:
        v = result.get(k)
        try:
            if v is not None and np.isfinite(float(v)):
                out[k] = np.array([float(v)])
        except (TypeError, ValueError):
            pass
    return out


def _arr1(v):
    return np.array([v])


def _view_flags(result, cw_only, has_cw):
    return dict(has_transient=_arr1(0 if cw_only else 1), has_cw=_arr1(1 if has_cw else 0),
                cw_only=_arr1(1 if cw_only else 0),
                view_primary=_arr1(str(result.get('view_primary', 'transient'))))


def _spectrum_npz(result, with_tmax):
    """Fourier pulse stitch: the spectrum T(lambda_k), the comb weights and the pulse
    parameters (reference fdtd2d_core.py:2229-2236, fdtd3d_core.py:2301-2308)."""
    sp = result.get('spectrum')
    if not sp:
        return {}
    pu = result.get('pulse') or {}
    out = dict(spec_lam_nm=np.asarray(sp['lam_nm'], dtype=np.float64),
               spec_T=np.asarray(sp['T'], dtype=np.float64),
               spec_weight=np.asarray(sp['weight'], dtype=np.float64),
               pulse_cycles=_arr1(pu.get('cycles', 0.0)))
    if with_tmax:
        out['pulse_tmax_fs'] = _arr1(pu.get('tmax_fs', 0.0))
    out['pulse_dispersion'] = _arr1(int(bool(pu.get('dispersion', True))))
    return out


def _first_stride(n, limit, choices):
    for s in choices:
        if n // s <= limit:
            return s
    return choices[-1]


def _pickle_summary(path, result, exclude):
    with open(path, 'wb') as f:
        pickle.dump({k: v for k, v in result.items() if k not in exclude}, f)


def _named_stack(stack, names):
    """Legacy layers take their display names from the material-name list."""
    if len(names) <= 3:
        return [dict(Ly) for Ly in stack]
    return [dict(Ly, name=str(names[_LEGACY_POS[Ly['mat']]])) if Ly.get('mat') in _LEGACY_POS
            else dict(Ly) for Ly in stack]


def _materials_2d(result, defaults):
    """(names, indices, stack keys) of a 2D result; sets result['stack'] and
    ['stack_custom'] (fdtd2d_core.py:2055-2082)."""
    names = list(result.get('layer_names') or DEFAULT_NAMES)
    stack = result.get('stack')
    custom = bool(result.get('stack_custom', False))
    if stack is None:
        t_lip, t_aq, t_mu, n_aq = result.get('layers') or (0.0, defaults['t_aq'], defaults['t_mu'],
                                                            defaults['n_aq'])
        stack = legacy_stack(t_lip * 1e6, result.get('n_lip', defaults['n_lip']), t_aq * 1e6,
                             result.get('n_aq', defaults['n_aq']), t_mu * 1e6,
                             result.get('n_mucin', defaults['n_mu']))
        custom = False
    if custom:
        mat_names = np.array([str(names[0])] + [str(Ly['name']) for Ly in stack] + [str(names[-1])])
        idx = np.array([float(np.real(result['wg_n']))] + [complex(Ly['n']).real for Ly in stack]
                       + [float(result['n_cornea'])], dtype=np.float64)
    else:
        stack = _named_stack(stack, names)
        mat_names = np.array(names)
        idx = np.array([result['wg_n'], result.get('n_lip', defaults['n_lip']), result['n_aq'],
                        result['n_mucin'], result['n_cornea']], dtype=np.float64)
    result['stack'] = stack
    result['stack_custom'] = custom
    return mat_names, idx, stack_to_npz(stack)


def _frame_meta(frames):
    return np.array([(fr['x_start'], fr['x_end'], fr['y_start'], fr['y_end'], fr['slide_i'],
                      fr['time'], fr.get('step', 0)) for fr in frames], dtype=np.float64)


def save_results(result, out_dir='results', defaults=None):
    """2D result files (reference fdtd2d_core.py:2085-2240).  ``defaults``: the legacy
    layer values (t_aq, t_mu, n_aq, n_mu, n_lip) for result dicts without a stack."""
    defaults = defaults or dict(t_aq=3.5e-6, t_mu=0.5e-6, n_aq=1.336, n_mu=1.342, n_lip=1.48)
    os.makedirs(out_dir, exist_ok=True)
    tag = f"{result['scenario']}_{result.get('method', 'full')}"
    mat_names, mat_idx, stack_kw = _materials_2d(result, defaults)
    _pickle_summary(f'{out_dir}/beads_{tag}.pkl', result, ('frames', 'cad_region', 'Iavg', 'Ex_cw_amp'))
    print(f'  Saved: {out_dir}/beads_{tag}.pkl')
    frames = result.get('frames')
    if not frames:
        return
    n_fr = len(frames)
    mm_path = result.get('frames_memmap')
    if mm_path and os.path.exists(mm_path):        # frames streamed to a disk memmap
        stack_src = np.load(mm_path, mmap_mode='r')[:n_fr]
        shape = stack_src.shape[1:]
    else:
        stack_src = None
        shape = frames[0]['Ez'].shape
    sy = _first_stride(shape[1], 2000, (1, 2, 4, 8))
    sx = _first_stride(shape[0], 5000, (1, 2, 3, 4, 6, 8, 12, 16, 32))
    if stack_src is not None:
        ez = stack_src[:, ::sx, ::sy]
    else:
        ez = np.empty((n_fr, len(range(0, shape[0], sx)), len(range(0, shape[1], sy))), np.float32)
        for i, fr in enumerate(frames):
            ez[i] = fr['Ez'][::sx, ::sy]
    meta = _frame_meta(frames)
    cw_kw = {}
    cw_path, cw_frames = result.get('cw_frames_memmap'), result.get('cw_frames') or []
    if cw_path and os.path.exists(cw_path) and cw_frames:
        cw_kw = dict(Ez_cw=np.load(cw_path, mmap_mode='r')[:len(cw_frames), ::sx, ::sy],
                     meta_cw=_frame_meta(cw_frames))
    cw_only = bool(result.get('cw_only', False)) or result.get('method') == 'stitch'
    extra = {}
    if result.get('Iavg') is not None:             # exact time average <|E|^2>
        extra.update(Iavg=np.asarray(result['Iavg'])[::sx, ::sy].astype(np.float32),
                     iavg_full=_arr1(1))
    if result.get('Ex_cw_amp') is not None:        # TM: complex CW amplitude of node Ex
        extra['Ex_cw_amp'] = np.asarray(result['Ex_cw_amp'])[::sx, ::sy].astype(np.complex64)
    if result.get('p_in_si') is not None:
        extra.update(p_in_si=_arr1(float(result['p_in_si'])),
                     p_in_unit=_arr1(str(result.get('p_in_unit', 'W/m'))))
    extra.update(vcsel_npz(result))
    kw = dict(Ez=ez, meta=meta, scenario=result['scenario'],
              layers=np.array(result['layers'], dtype=np.float64),
              material_names=mat_names, material_indices=mat_idx, **stack_kw,
              bead_n=_arr1(result.get('bead_n', 0.0)),
              bead_material=_arr1(str(result.get('bead_material', ''))),
              bead_x_um=_arr1(result.get('bead_x_um', 0.0)),
              bead_diameter_um=_arr1(result.get('bead_diameter_um', 0.0)),
              t_wg_um=_arr1(result.get('wg_thickness_um', 5.0)),
              x_wg_start_um=_arr1(result.get('x_wg_start_um', 0.0) or 0.0),
              x_wg_end_um=_arr1(result.get('x_wg_end_um', 0.0) or 0.0),
              polarization=_arr1(str(result.get('polarization', 's'))),
              lam_nm=_arr1(result.get('lambda_nm', 850.0)))
    if result.get('bead_y_um') is not None:
        kw['bead_y_um'] = _arr1(result['bead_y_um'])
    kw.update(_spectrum_npz(result, with_tmax=True))
    npz = f'{out_dir}/beads_{tag}_frames.npz'
    np.savez_compressed(npz, **kw, **cw_kw, **_view_flags(result, cw_only, cw_only or bool(cw_kw)),
                        **extra)
    print(f'  Saved: {npz} (stride={sx}, {ez.nbytes / 1e6:.1f} MB)' + ('  +CW set' if cw_kw else '')
          + ('  [cw_only]' if cw_only else ''))


def save_results_3d(result, out_dir='results', prefix='fdtd3d', material_names=None):
    """3D result files (reference fdtd3d_core.py:2218-2423)."""
    os.makedirs(out_dir, exist_ok=True)
    tag = result['scenario']
    names = list(material_names or DEFAULT_NAMES)
    mat_idx = np.array([result['wg_n'], result['n_lip'], result['n_aq'], result['n_mu'], result['n_co']],
                       dtype=np.float64)
    stack = result.get('stack')
    if stack is None:
        t_lip, t_aq, t_mu, _n = result.get('layers') or (0.0, 3.5e-6, 0.5e-6, result.get('n_aq', 1.336))
        stack = legacy_stack(t_lip * 1e6, result.get('n_lip', 1.48), t_aq * 1e6,
                             result.get('n_aq', 1.336), t_mu * 1e6, result.get('n_mu', 1.342))
    if result.get('stack_custom'):
        names = [str(names[0])] + [str(Ly['name']) for Ly in stack] + [str(names[-1])]
        mat_idx = np.array([float(np.real(result['wg_n']))] + [complex(Ly['n']).real for Ly in stack]
                           + [float(result['n_co'])], dtype=np.float64)
    else:
        stack = _named_stack(stack, names)
    result['stack'] = stack
    stack_kw = stack_to_npz(stack)
    big = ('vols', 'vols_array', 'vols_ex', 'vols_ey', 'Iavg', 'cw_vol', 'cad_region')
    try:
        _pickle_summary(f'{out_dir}/{prefix}_{tag}.pkl', result, big)
        print(f'  Saved: {out_dir}/{prefix}_{tag}.pkl')
    except Exception as e:                        # the NPZ files are written regardless
        print(f'  [WARN] PKL summary not saved ({e}); continuing with NPZ.')
    vols = result.get('vols_array')
    if vols is None or vols.shape[0] == 0:
        return
    steps = np.array(result.get('snap_steps', []), dtype=np.int64)
    times = np.array(result.get('snap_times', []), dtype=np.float64)
    kmid = vols.shape[3] // 2
    prim = vols
    if (result.get('vector') and result.get('vols_ey') is not None
            and str(result.get('polarization', 's')).lower() in ('p', 'tm')):
        prim = result['vols_ey']                  # TM drives Ey
    ez_xy = prim[:, :, :, kmid].astype(np.float32)
    x_end, y0 = result['lx_um'], result['y0_um']
    y1 = y0 + result['ly_um']
    meta = np.array([(0.0, x_end, y0, y1, 0, t, s) for t, s in zip(times, steps)], dtype=np.float64)
    cw_vol, cw_only = result.get('cw_vol'), bool(result.get('cw_only', False))
    cw_fr, cw_v = {}, {}
    if cw_vol is not None and not cw_only:
        cw_xy = np.asarray(cw_vol[:, :, :, cw_vol.shape[3] // 2]).astype(np.float32)
        f0 = result.get('f0') or (2.99792458e8 / (result.get('lam_nm', 850.0) * 1e-9))
        ncw = cw_xy.shape[0]
        cw_fr = dict(Ez_cw=cw_xy, meta_cw=np.array([(0.0, x_end, y0, y1, 0, (k / ncw) / f0, k)
                                                    for k in range(ncw)], dtype=np.float64))
        cw_v = dict(Ez_cw=cw_vol)
    flags = _view_flags(result, cw_only, cw_only or cw_vol is not None)
    flags.update(_spectrum_npz(result, with_tmax=False))      # into both NPZ files
    bead = result.get('bead') or {}
    bd = bead.get('d_um', 0.0)
    bead_kw = dict(bead_x_um=_arr1(bead.get('x_um', 0.0)), bead_diameter_um=_arr1(bd),
                   bead_y_um=_arr1(bead.get('y_um', -bd / 2.0)), bead_z_um=_arr1(bead.get('z_um', 0.0)),
                   bead_n=_arr1(bead.get('n', 0.0)))
    power = ({'p_in_si': _arr1(float(result['p_in_si'])),
              'p_in_unit': _arr1(str(result.get('p_in_unit', 'W')))}
             if result.get('p_in_si') is not None else {})
    power.update(vcsel_npz(result))
    iavg = {}
    ia = result.get('Iavg')
    if ia is not None and np.ndim(ia) == 3:
        ia = np.asarray(ia)
        iavg = dict(Iavg=ia[:, :, ia.shape[2] // 2].astype(np.float32), iavg_full=_arr1(1))
    common = dict(layers=np.array(result['layers'], dtype=np.float64), material_names=np.array(names),
                  material_indices=mat_idx, **bead_kw,
                  x_wg_start_um=_arr1(result.get('x_wg_start_um', 0.0) or 0.0),
                  end_facet_um=_arr1(result.get('end_facet_um', 0.0) or 0.0),
                  x_src_um=_arr1(result.get('x_src_um', 0.0) or 0.0),
                  lam_nm=_arr1(result.get('lam_nm', 850.0)),
                  polarization=_arr1(str(result.get('polarization', 's'))), **stack_kw)
    fr_path = f'{out_dir}/{prefix}_{tag}_frames.npz'
    np.savez_compressed(fr_path, Ez=ez_xy, meta=meta, scenario=tag, **iavg, **power, **common,
                        **cw_fr, **flags)
    print(f'  Saved: {fr_path} (z mid-plane)' + ('  +CW' if cw_fr else '')
          + ('  [cw_only]' if cw_only else ''))
    ia_full = result['Iavg']
    ia_full = ia_full if getattr(ia_full, 'dtype', None) == np.float32 else np.asarray(ia_full, np.float32)
    vol_kw = dict(**power, Ez=vols, Iavg=ia_full, iavg_full=_arr1(1), steps=steps, times=times,
                  dx_um=result['dx_nm'] / 1000.0, x0_um=0.0, y0_um=result['y0_um'],
                  z0_um=result['z0_um'], scenario=tag, dt=result['dt'],
                  t_wg_um=_arr1(result.get('t_wg_um', 0.0)),
                  wg_width_um=_arr1(bead.get('wg_width_um', result.get('wg_width_um', 0.0)) or 0.0),
                  lx_um=_arr1(result.get('lx_um', 0.0)), **common, **cw_v, **flags)
    if result.get('vector') and result.get('vols_ex') is not None:
        vol_kw.update(Ex=result['vols_ex'], Ey=result['vols_ey'], vector=_arr1(1))
    vol_path = f'{out_dir}/{prefix}_{tag}_vol3d.npz'
    try:
        np.savez_compressed(vol_path, **vol_kw)
    except (MemoryError, OSError) as e:           # compression needs extra memory
        print(f'  [WARN] compressed volume NPZ failed ({e}); writing it uncompressed')
        np.savez(vol_path, **vol_kw)
    print(f'  Saved: {vol_path} ({vols.shape[0]} volumes {vols.shape[1:]})')
