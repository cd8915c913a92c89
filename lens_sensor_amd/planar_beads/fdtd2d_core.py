"""2D Yee FDTD solver (planar waveguide + bead at the waveguide/tear interface) on the
MI355X step engine.

Entry points keep the reference's signatures (planar_beads/fdtd2d_core.py):
  run_beads           :531-927   full domain (method='full'), dispatch to stitch
  run_beads_stitched  :1197-1583 domain decomposition with phasor handoff
Materials and coefficients (host NumPy, as in the reference):
  build_materials :122-259, _coeffs :262-270, _tm_loss :283-294, _apply_bead_tm :297-349,
  _tm_coeffs :445-455, node colocation :458-486
TE = polarisation 's' (Ez, Hx, Hy); TM = 'p' (Hz, Ex, Ey).

The time loop is executed by the step engine (HIP library); per step the host only
fills table entries (envelope, phasors) with the reference's expressions.
"""
import contextlib
import io
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_REPO_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _REPO_ROOT not in sys.path:
    sys.path.insert(0, _REPO_ROOT)

from lens_sensor_amd.common import backend                                    # noqa: E402
from lens_sensor_amd.common.physics import (C0, EPS0, MU0, N_AIR, n_at,       # noqa: E402
                                            cornea_default, loss_sigma, core_source_mask)
from lens_sensor_amd.common.subpixel import circle_fraction, blend, cell_average_column  # noqa: E402
from lens_sensor_amd.common.flux import PlaneFluxBook, FluxAccumulator, guard_in_plane  # noqa: E402
from lens_sensor_amd.common import vcsel                                      # noqa: E402
from lens_sensor_amd.common.td_stitch import TD_CHECK, TD_ROWS, td_next_event, signal_span  # noqa: E402
from lens_sensor_amd.common.layer_stack import (stack_total_um, legacy_stack,  # noqa: E402
                                                legacy_layers_tuple, stack_at_lambda,
                                                describe as describe_stack)
from lens_sensor_amd.engine.spec import new_steps, fill_phasors, cw_envelope, pulse_envelope  # noqa: E402
from lens_sensor_amd.planar_beads.boundaries import (make_boundary, MurSet, mur_speeds,  # noqa: E402
                                                    attach_mur_2d)

BEAD_SUBPIXEL = True
N_POLYSTYR, N_PMMA, N_AQ, N_MUCIN, N_CORNEA, N_LIPID = 1.590, 1.491, 1.336, 1.342, 1.376, 1.480
L_DEMO = 1.0e-3
AIR_BUFFER = 5.0e-6
TEAR_BUFFER = 10.0e-6
KAPPA = {}
SRC_CORE_ONLY = False
LASER_POWER_MW = vcsel.VCSEL_DEFAULTS['power_mw']
LAM_LOSS = 850e-9
T_MU = 0.5e-6
T_AQ = 3.5e-6
BEAD_X_UM = 500.0
DET_LEN = 30e-6
N_G_FACTOR = 1.10
CW_TOL = 0.01
CW_MAX_EXTEND = 2.0
NPML = 10


def _bead_list(place_bead, bead_x_um, bead_d_um, bead_y_um, bead_n, beads):
    if beads is not None:
        return [b for b in beads if b.get('d_um', 0) > 0]
    if not place_bead or not bead_d_um or bead_d_um <= 0:
        return []
    return [dict(x_um=bead_x_um, d_um=bead_d_um, y_um=bead_y_um, n=bead_n)]


# ----------------------------------------------------------------------------- materials
def _material_parts(Nx, Ny, dx, wg_n, bead_n, wg_top, bead_x_um, bead_d_um, x_start_um=0.0,
                    place_bead=True, n_aq=N_AQ, n_mucin=N_MUCIN, n_cornea=N_CORNEA, t_aq_um=None,
                    t_mu_um=None, t_lip_um=0.0, n_lip=N_LIPID, input_gap_um=0.0, length_um=None,
                    bead_y_um=None, curved=False, r_bend_m=8.3e-3, tm_patch=None, stack=None,
                    beads=None):
    """What build_materials fills the grid from: the two layer columns (outside / inside
    the core's x range), the core's x range, the beads' boxes, the curvature metric.
    Reference fdtd2d_core.py:122-259."""
    t_aq = (t_aq_um * 1e-6) if t_aq_um is not None else T_AQ
    t_mu = (t_mu_um * 1e-6) if t_mu_um is not None else T_MU
    t_lip = (t_lip_um or 0.0) * 1e-6
    L = (length_um * 1e-6) if length_um is not None else L_DEMO
    x_wg_start = (input_gap_um or 0.0) * 1e-6
    x_wg_end = L
    kp = dict(KAPPA or {})
    lossy = any(v > 0 for v in kp.values())
    if stack is not None:
        stack = list(stack)
        lossy = lossy or any(complex(Ly['n']).imag > 0 for Ly in stack)
    blist = _bead_list(place_bead, bead_x_um, bead_d_um, bead_y_um, bead_n, beads)
    if lossy:
        wg_n = complex(wg_n, kp.get('wg', 0.0))
        n_aq = complex(n_aq, kp.get('aq', 0.0))
        n_mucin = complex(n_mucin, kp.get('mu', 0.0))
        n_lip = complex(n_lip, kp.get('lip', 0.0))
        k_co = kp.get('co', 0.0)
    else:
        k_co = 0.0
    edt = np.complex128 if lossy else np.float64
    ys = -TEAR_BUFFER + np.arange(Ny) * dx
    y_lip_bot = -t_lip
    y_aq_bot = y_lip_bot - t_aq
    y_mu_bot = y_aq_bot - t_mu
    bounds = None
    if stack is not None:
        bounds, yt = [], 0.0
        for Ly in stack:
            yb = yt - float(Ly['t_um']) * 1e-6
            nL = complex(Ly['n'])
            bounds.append((yt, yb, nL if lossy else nL.real))
            yt = yb
        y_mu_bot = yt

    def eps_y(y, core):
        e = np.full(y.shape, N_AIR ** 2, dtype=edt)
        below = y < y_mu_bot
        if hasattr(n_cornea, 'at_depth'):
            e[below] = ((n_cornea.at_depth((y_mu_bot - y[below]) * 1e6) + 1j * k_co) ** 2 if lossy
                        else n_cornea.at_depth((y_mu_bot - y[below]) * 1e6) ** 2)
        else:
            e[below] = (float(n_cornea) + 1j * k_co) ** 2 if lossy else n_cornea ** 2
        if bounds is not None:
            for yt_, yb_, n_ in bounds:
                e[(y >= yb_) & (y < yt_)] = n_ ** 2
        else:
            e[(y >= y_mu_bot) & (y < y_aq_bot)] = n_mucin ** 2
            e[(y >= y_aq_bot) & (y < y_lip_bot)] = n_aq ** 2
            if t_lip > 1e-12:
                e[(y >= y_lip_bot) & (y < 0.0)] = n_lip ** 2
        if core:
            e[(y >= 0.0) & (y < wg_top)] = wg_n ** 2
        return e

    col = cell_average_column(ys, dx, lambda y: eps_y(y, False))
    col_core = cell_average_column(ys, dx, lambda y: eps_y(y, True))
    xs_abs = x_start_um * 1e-6 + np.arange(Nx) * dx
    in_wg_x = (xs_abs >= x_wg_start) & (xs_abs <= x_wg_end)
    layer = None
    if not curved:
        layer = dict(
            in_wg_x=in_wg_x,
            h_out=cell_average_column(ys[:-1], dx, lambda y: eps_y(y, False).real,
                                      harmonic=True, shift=0.5 * dx),
            h_core=cell_average_column(ys[:-1], dx, lambda y: eps_y(y, True).real,
                                       harmonic=True, shift=0.5 * dx))
        if tm_patch is not None:
            tm_patch['layer'] = layer
    x0 = x_start_um * 1e-6
    xs = x0 + np.arange(Nx) * dx
    boxes = []
    for bd in blist:
        r = bd['d_um'] * 1e-6 / 2.0
        xc = bd['x_um'] * 1e-6
        bn = bd['n']
        bn_c = complex(bn, kp.get('bead', 0.0)) if lossy else bn
        if xc + r < xs[0] or xc - r > xs[-1]:
            continue
        yc = (bd['y_um'] * 1e-6) if bd.get('y_um') is not None else -r
        ix0 = max(0, int((xc - r - x0) / dx) - 2)
        ix1 = min(Nx, int((xc + r - x0) / dx) + 2)
        jy0 = max(0, int((yc - r - ys[0]) / dx) - 2)
        jy1 = min(Ny, int((yc + r - ys[0]) / dx) + 3)
        boxes.append(dict(r=r, xc=xc, yc=yc, bn=bn, bn_c=bn_c, ix0=ix0, ix1=ix1, jy0=jy0, jy1=jy1))
    metric = None
    if curved:
        u = ys - wg_top / 2.0
        metric = (1.0 + u / r_bend_m).astype(np.float32) ** 2
    return dict(Nx=Nx, Ny=Ny, dx=dx, lossy=lossy, col=col, col_core=col_core, in_wg_x=in_wg_x,
                xs=xs, ys=ys, boxes=boxes, metric=metric, curved=curved, edt=edt, layer=layer)


def _material_rows(mp, j0, j1, tm_patch=None):
    """eps_r, sigma (float32) of the grid rows [j0, j1), all x: the layer columns, then
    the beads blended by area fraction (every bead's box lies inside the rows), the
    curvature metric, then the complex index split (fdtd2d_core.py:219-259)."""
    Nx = mp['Nx']
    col, col_core = mp['col'], mp['col_core']
    eps = np.full((Nx, j1 - j0), N_AIR ** 2, dtype=np.complex128 if mp['lossy'] else np.float32)
    eps[:, :] = col[None, j0:j1]
    if mp['in_wg_x'].any():
        eps[mp['in_wg_x'], :] = col_core[None, j0:j1]
    xs, ys, dx = mp['xs'], mp['ys'], mp['dx']
    for b in mp['boxes']:
        ix0, ix1, jy0, jy1 = b['ix0'], b['ix1'], b['jy0'], b['jy1']
        if BEAD_SUBPIXEL:
            if ix1 > ix0 and jy1 > jy0:
                if not (j0 <= jy0 and jy1 <= j1):
                    raise ValueError('bead box outside the material rows')
                blk = eps[ix0:ix1, jy0 - j0:jy1 - j0]
                if tm_patch is not None and not mp['curved']:
                    tm_patch.setdefault('beads', []).append(dict(
                        eps_bg=blk.real.astype(np.float64), ix0=ix0, jy0=jy0,
                        xs=xs[ix0:ix1], ys=ys[jy0:jy1], xc=b['xc'], yc=b['yc'], r=b['r'], n=b['bn']))
                blend(blk, circle_fraction(xs[ix0:ix1], ys[jy0:jy1], b['xc'], b['yc'], b['r'], dx),
                      b['bn_c'])
        else:
            r, xc, yc = b['r'], b['xc'], b['yc']
            yr = ys[j0:j1]
            for i in range(ix0, ix1):
                ddx = xs[i] - xc
                if abs(ddx) > r:
                    continue
                ddy = np.sqrt(max(0.0, r * r - ddx * ddx))
                eps[i, (yr >= yc - ddy) & (yr <= yc + ddy)] = b['bn_c'] ** 2
    if mp['metric'] is not None:
        eps = eps * mp['metric'][None, j0:j1]
    if mp['lossy']:
        sig = loss_sigma(eps.imag, LAM_LOSS).astype(np.float32)
        eps = eps.real.astype(np.float32)
    else:
        sig = np.zeros((Nx, j1 - j0), dtype=np.float32)
    return eps, sig


def build_materials(Nx, Ny, dx, wg_n, bead_n, wg_top, bead_x_um, bead_d_um, x_start_um=0.0,
                    place_bead=True, n_aq=N_AQ, n_mucin=N_MUCIN, n_cornea=N_CORNEA, t_aq_um=None,
                    t_mu_um=None, t_lip_um=0.0, n_lip=N_LIPID, input_gap_um=0.0, length_um=None,
                    bead_y_um=None, curved=False, r_bend_m=8.3e-3, tm_patch=None, stack=None,
                    beads=None):
    """eps_r (Nx,Ny) float32 and sigma (Nx,Ny) float32 of the planar stack with the core on
    [input_gap, L] x [0, t_wg] and disc beads blended by area fraction.
    Reference fdtd2d_core.py:122-259.  ``beads`` (list of {x_um, d_um, y_um, n}) replaces
    the single bead for several beads per window."""
    mp = _material_parts(Nx, Ny, dx, wg_n, bead_n, wg_top, bead_x_um, bead_d_um, x_start_um,
                         place_bead, n_aq, n_mucin, n_cornea, t_aq_um, t_mu_um, t_lip_um, n_lip,
                         input_gap_um, length_um, bead_y_um, curved, r_bend_m, tm_patch, stack,
                         beads)
    return _material_rows(mp, 0, Ny, tm_patch)


def _coeffs(eps_r, sig, dt, dx):
    """TE: Ce_E = (1-a)/(1+a), Ce_H = dt/(eps0 eps dx)/(1+a), a = sigma dt/(2 eps0 eps)
    (fdtd2d_core.py:262-270)."""
    Ce_n = (1 - sig * dt / (2 * EPS0 * eps_r))
    Ce_d = (1 + sig * dt / (2 * EPS0 * eps_r))
    Ce_E = (Ce_n / Ce_d).astype(np.float32)
    Ce_H = (dt / (EPS0 * eps_r * dx) / Ce_d).astype(np.float32)
    return Ce_E, Ce_H


def _tm_coeffs(eps_r, dt, dx):
    """TM: Ca at the staggered Ex/Ey positions from the averaged node eps
    (fdtd2d_core.py:445-455)."""
    eps_ex = 0.5 * (eps_r[:-1, :] + eps_r[1:, :])
    eps_ey = 0.5 * (eps_r[:, :-1] + eps_r[:, 1:])
    Ca_ex = (dt / (EPS0 * eps_ex * dx)).astype(np.float32)
    Ca_ey = (dt / (EPS0 * eps_ey * dx)).astype(np.float32)
    return Ca_ex, Ca_ey, np.float32(dt / (MU0 * dx))


def _tm_loss(eps_r, sig, Ca_ex, Ca_ey, dt, lossy=None):
    """Semi-implicit TM loss: Ca /= (1+a) in place, returns D = (1-a)/(1+a) per component
    or (None, None) (fdtd2d_core.py:283-294).  ``lossy`` overrides the test of the given
    sigma (a part of a grid takes the whole grid's decision)."""
    if lossy is None:
        lossy = sig is not None and bool(np.any(sig > 0))
    if not lossy:
        return None, None
    s_ex = 0.5 * (sig[:-1, :] + sig[1:, :])
    e_ex = 0.5 * (eps_r[:-1, :] + eps_r[1:, :])
    s_ey = 0.5 * (sig[:, :-1] + sig[:, 1:])
    e_ey = 0.5 * (eps_r[:, :-1] + eps_r[:, 1:])
    a_ex = s_ex * dt / (2 * EPS0 * e_ex)
    a_ey = s_ey * dt / (2 * EPS0 * e_ey)
    Ca_ex /= (1 + a_ex).astype(np.float32)
    Ca_ey /= (1 + a_ey).astype(np.float32)
    return ((1 - a_ex) / (1 + a_ex)).astype(np.float32), ((1 - a_ey) / (1 + a_ey)).astype(np.float32)


def _apply_bead_tm(bp, Ca_ex, Ca_ey, dt, dx, row0=0):
    """Layer-harmonic Ey coefficient and the anisotropic (Farjadpour) bead-edge tensor for
    Ex/Ey: (1/eps)_ii = n_i^2 <1/eps> + (1 - n_i^2)/<eps>  (fdtd2d_core.py:297-349).
    ``row0``: the arrays hold the grid rows from row0 on (a band of rows)."""
    if not bp:
        return
    c0 = dt / (EPS0 * dx)
    lp = bp.get('layer')
    h = None
    if lp is not None:
        h = np.where(lp['in_wg_x'][:, None], lp['h_core'][None, row0:row0 + Ca_ey.shape[1]],
                     lp['h_out'][None, row0:row0 + Ca_ey.shape[1]])
        Ca_ey[...] = (c0 / h).astype(np.float32)
    for bd in bp.get('beads', []):
        e_bg, ix0, jy0 = bd['eps_bg'], bd['ix0'], bd['jy0'] - row0
        xs, ys, xc, yc, r, eb = bd['xs'], bd['ys'], bd['xc'], bd['yc'], bd['r'], bd['n'] ** 2
        c = dt / (EPS0 * dx)

        def tensor(xp_, yp_, ebg_a, ebg_h, comp):
            f = circle_fraction(xp_, yp_, xc, yc, r, dx)
            X = xp_[:, None] - xc
            Y = yp_[None, :] - yc
            d = np.maximum(np.sqrt(X * X + Y * Y), 1e-30)
            nc2 = (X / d) ** 2 if comp == 'x' else (Y / d) ** 2
            mean = (1 - f) * ebg_a + f * eb
            hmean = (1 - f) / ebg_h + f / eb
            return f, nc2 * hmean + (1 - nc2) / mean

        f, inv = tensor(xs[:-1] + 0.5 * dx, ys, 0.5 * (e_bg[:-1] + e_bg[1:]),
                        1.0 / (0.5 * (1.0 / e_bg[:-1] + 1.0 / e_bg[1:])), 'x')
        sl = (slice(ix0, ix0 + inv.shape[0]), slice(jy0, jy0 + inv.shape[1]))
        blk = Ca_ex[sl].astype(np.float64)
        m = f > 0
        blk[m] = c * inv[m]
        Ca_ex[sl] = blk.astype(np.float32)
        eh = 1.0 / (0.5 * (1.0 / e_bg[:, :-1] + 1.0 / e_bg[:, 1:]))
        if h is not None:
            hb = h[ix0:ix0 + eh.shape[0], jy0:jy0 + eh.shape[1]]
            eh[:hb.shape[0], :hb.shape[1]] = hb
        f, inv = tensor(xs, ys[:-1] + 0.5 * dx, 0.5 * (e_bg[:, :-1] + e_bg[:, 1:]), eh, 'y')
        sl = (slice(ix0, ix0 + inv.shape[0]), slice(jy0, jy0 + inv.shape[1]))
        blk = Ca_ey[sl].astype(np.float64)
        m = f > 0
        blk[m] = c * inv[m]
        Ca_ey[sl] = blk.astype(np.float32)


def window_coeffs(pol, mp, dt, dx):
    """Per-cell update coefficients of a 2D grid in compact form, for the engine to
    assemble the full arrays on the device.  Outside the rows the beads touch, a cell's
    coefficient depends only on its row and on which layer columns (outside / inside the
    core's x range) its x neighbours hold, so each array is: a class per x (the columns
    it reads), one row of values per class, and the explicit values of the rows the beads
    touch (one more row each side for the staggered neighbours).  Every value comes from
    the same NumPy expressions as the full arrays of build_materials + _coeffs (TE) or
    _tm_coeffs + _apply_bead_tm + _tm_loss (TM), applied to sub-grids; those are
    element-wise in float32, so expanding the compact form gives the full arrays bit for
    bit.  Returns {name: (xclass uint8, table float32 (classes, ny_c), j0, band (nx_c, rows))}
    and the TM Ch."""
    Nx, Ny = mp['Nx'], mp['Ny']
    core = mp['in_wg_x'].astype(np.uint8)
    boxes = [b for b in mp['boxes'] if b['ix1'] > b['ix0'] and b['jy1'] > b['jy0']]
    if boxes:
        b0 = max(0, min(b['jy0'] for b in boxes) - 1)
        b1 = min(Ny, max(b['jy1'] for b in boxes) + 1)
    else:
        b0 = b1 = 0
    bp = {'layer': mp['layer']} if mp['layer'] is not None else {}
    eps_b, sig_b = _material_rows(mp, b0, b1, tm_patch=bp)
    # the two layer columns (outside / inside the core), full height
    eps_c, sig_c = _material_rows(dict(mp, Nx=2, in_wg_x=np.array([False, True]), boxes=[]), 0, Ny)
    used = sorted(set(core.tolist()))
    out_rows = np.ones(Ny, bool)
    out_rows[b0:b1] = False
    lossy = bool(np.any(sig_b > 0)) or any(bool(np.any(sig_c[c][out_rows] > 0)) for c in used)
    res = {}
    if str(pol).lower() in ('s', 'te'):
        tE, tH = _coeffs(eps_c, sig_c, dt, dx)
        bE, bH = _coeffs(eps_b, sig_b, dt, dx)
        res['ceE'] = (core, tE, b0, bE)
        res['ceH'] = (core, tH, b0, bH)
        return res, None
    # TM.  Ex between columns i and i+1: four classes; Ey inside column i: two
    cls_x = (2 * core[:-1] + core[1:]).astype(np.uint8)
    tab = {k: [] for k in ('cax', 'dax', 'cay', 'day')}
    for a, b in ((0, 0), (0, 1), (1, 0), (1, 1)):
        e2, s2 = eps_c[[a, b]], sig_c[[a, b]]
        cax, cay, ch = _tm_coeffs(e2, dt, dx)
        dax, _ = _tm_loss(e2, s2, cax, cay, dt, lossy=lossy)
        tab['cax'].append(cax[0])
        if dax is not None:
            tab['dax'].append(dax[0])
    lp = bp.get('layer')
    for c in (0, 1):
        e1, s1 = eps_c[[c]], sig_c[[c]]
        cax, cay, _ = _tm_coeffs(e1, dt, dx)
        if lp is not None:
            _apply_bead_tm({'layer': dict(lp, in_wg_x=np.array([c == 1]))}, cax, cay, dt, dx)
        _, day = _tm_loss(e1, s1, cax, cay, dt, lossy=lossy)
        tab['cay'].append(cay[0])
        if day is not None:
            tab['day'].append(day[0])
    cax_b, cay_b, ch = _tm_coeffs(eps_b, dt, dx) if b1 > b0 else (None, None, ch)
    if b1 > b0:
        _apply_bead_tm(bp, cax_b, cay_b, dt, dx, row0=b0)
        dax_b, day_b = _tm_loss(eps_b, sig_b, cax_b, cay_b, dt, lossy=lossy)
    else:
        cax_b = np.zeros((Nx - 1, 0), np.float32)
        cay_b = np.zeros((Nx, 0), np.float32)
        dax_b, day_b = (cax_b, cay_b) if lossy else (None, None)
    res['cax'] = (cls_x, np.array(tab['cax'], np.float32), b0, cax_b)
    res['cay'] = (core, np.array(tab['cay'], np.float32), b0, cay_b)
    if lossy:
        res['dax'] = (cls_x, np.array(tab['dax'], np.float32), b0, dax_b)
        res['day'] = (core, np.array(tab['day'], np.float32), b0, day_b)
    return res, ch


def expand_coeffs(packed):
    """The full array of one compact coefficient (window_coeffs)."""
    xc, table, j0, band = packed
    full = table[xc.astype(np.intp)]
    full[:, j0:j0 + band.shape[1]] = band
    return full


def _tm_node_field(Ey):
    """Ey (Nx,Ny-1) -> node grid (Nx,Ny): neighbour mean inside, copies at the edges."""
    out = np.zeros((Ey.shape[0], Ey.shape[1] + 1), dtype=np.float32)
    out[:, 1:-1] = 0.5 * (Ey[:, :-1] + Ey[:, 1:])
    out[:, 0] = Ey[:, 0]
    out[:, -1] = Ey[:, -1]
    return out


def _n_source_2d(src_x_um, iy_src, y0_m, dx, wg_n, T_WG, n_below, gap_um):
    if gap_um and src_x_um < gap_um:
        return float(N_AIR)
    y = y0_m + iy_src * dx
    if 0.0 <= y < T_WG:
        return float(np.real(wg_n))
    return float(N_AIR) if y >= T_WG else float(np.real(n_below))


def _vcsel_geom_2d(src_ix, dx, distance_um):
    air = vcsel.air_path_um(distance_um, dx * 1e6)
    return src_ix * dx * 1e6 + air, air, vcsel.z_at_source(distance_um, air)


def _vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, vcsel_offset, vcsel_waist, lambda_nm, tilt,
                     n_src, z_um, f0, y0_m=None, y_center_m=None):
    """Gaussian line source as a one-sided Huygens pair in column src_ix
    (fdtd2d_core.py:378-397); returns (iy_src, TE source, TM Ey source)."""
    if y0_m is None:
        y0_m = -TEAR_BUFFER
    yc = (T_WG / 2 if y_center_m is None else y_center_m) + vcsel_offset * 1e-6
    iy_src = int(round((yc - y0_m) / dx))
    y_um = (y0_m + np.arange(Ny) * dx) * 1e6
    amp, ph = vcsel.profile_1d(y_um, yc * 1e6, z_um, vcsel_waist, lambda_nm, n_src, tilt)
    if SRC_CORE_ONLY and y_center_m is None:
        amp = amp * core_source_mask(y_um, T_WG * 1e6)
    kdx = vcsel.numerical_kx_dx(lambda_nm, dx, C0 * dt, n_src, tilt)
    src = vcsel.HuygensSource(src_ix, amp, ph, 2 * np.pi * f0, kdx)
    return iy_src, src, src.to_ey()


def _result_dict(label, P_in, P_out, Tr, frames, total_time, bead_d_um, wg_mat, bead_mat, wg_n,
                 bead_n, wg_th, lam, waist, tilt, off, src, method, place_bead=True, n_aq=N_AQ,
                 n_mucin=N_MUCIN, n_cornea=N_CORNEA, frames_memmap=None, t_aq_um=None,
                 t_mu_um=None, t_lip_um=0.0, bead_x_um=None, x_wg_start_um=0.0,
                 x_wg_end_um=None, polarization='s', stack=None, n_lip=None):
    """Result dictionary with the reference's keys (fdtd2d_core.py:2020-2048)."""
    t_aq = (t_aq_um * 1e-6) if t_aq_um is not None else T_AQ
    t_mu = (t_mu_um * 1e-6) if t_mu_um is not None else T_MU
    t_lip = (t_lip_um or 0.0) * 1e-6
    if stack is None:
        kp = dict(KAPPA or {})
        nl = float(n_lip) if n_lip is not None else N_LIPID
        stack_out = legacy_stack(t_lip * 1e6, complex(nl, kp.get('lip', 0.0)),
                                 t_aq * 1e6, complex(float(n_aq), kp.get('aq', 0.0)),
                                 t_mu * 1e6, complex(float(n_mucin), kp.get('mu', 0.0)))
        layers = (t_lip, t_aq, t_mu, n_aq)
    else:
        stack_out = [dict(Ly) for Ly in stack]
        layers = legacy_layers_tuple(stack)
    return dict(scenario=label, P_in=P_in, P_out=P_out, transmission=Tr, frames=frames,
                frames_memmap=frames_memmap, stack=stack_out, stack_custom=stack is not None,
                total_time=total_time, layers=layers, n_aq=n_aq, n_mucin=n_mucin,
                n_cornea=n_cornea, wg_material=wg_mat, bead_material=bead_mat, wg_n=wg_n,
                bead_n=bead_n, bead_diameter_um=(bead_d_um if place_bead else 0.0),
                bead_x_um=(bead_x_um if bead_x_um is not None else BEAD_X_UM),
                x_wg_start_um=x_wg_start_um, x_wg_end_um=x_wg_end_um, wg_thickness_um=wg_th,
                lambda_nm=lam, vcsel_waist=waist, vcsel_tilt=tilt, vcsel_offset=off,
                source_type=src, method=method, polarization=polarization)


def _emit_power(fl, dx):
    if not (fl.has('src_up') and fl.has('src_dn')):
        return None, None
    up, dn = fl.flux('src_up'), fl.flux('src_dn')
    return (dn - up) * dx, -up * dx


def _vcsel_result(res, p_emit, p_refl, distance_um, air_um, x_facet_um, vcsel_waist,
                  p_in=None, energy=False):
    """VCSEL figures of a result (reference fdtd2d_core.py:415-435); energy=True: p_emit,
    p_refl and p_in are pulse energies (J/m), so no source power is stored."""
    res['laser_power_mw'] = LASER_POWER_MW
    res['vcsel_distance_um'] = float(distance_um or 0.0)
    res['vcsel_air_um'] = float(air_um)
    res['vcsel_waist_um'] = float(vcsel_waist)
    res['x_facet_um'] = float(x_facet_um)
    res['p_src_si'] = None if energy else p_emit
    if p_emit and p_emit > 0:
        if p_in is None:
            p_in = res.get('p_in_si')
        res['coupling_efficiency'] = (p_in / p_emit) if p_in is not None else float('nan')
        res['facet_reflection'] = (p_refl / p_emit) if p_refl is not None else float('nan')


def _engine_mode(mode):
    return mode or backend.DEFAULT_MODE


def _safe(label):
    return ''.join(c if (c.isalnum() or c in '-_.') else '_' for c in label)


# ----------------------------------------------------------------------------- run_beads
def run_beads(wg_mat, bead_mat, bead_d_um, dx_nm=20.0, save_frames=True, n_snapshots=12,
              wg_thickness_um=5.0, lambda_nm=850.0, vcsel_waist=3.0, vcsel_tilt=0.0,
              vcsel_offset=0.0, source_type='cw', method='full', window_w_um=350.0,
              slide_um=150.0, place_bead=True, t_aqueous_um=None, t_mucin_um=None,
              t_lipid_um=0.0, input_gap_um=0.0, length_um=None, bead_x_um=None, bead_y_um=None,
              wg_n=None, bead_n=None, n_aqueous=None, n_mucin=None, n_cornea=None, n_lipid=None,
              polarization='s', curved=False, boundary='cpml', cad=None, pulse_cycles=30.0,
              pulse_tmax_fs=None, pulse_nfreq=None, pulse_dispersion=True, pulse_mode='fourier',
              n_aq_tear_offset=None, tear_buffer_um=None, stack=None,
              npml=None, engine_mode=None, beads=None, steps_override=None):
    """2D full-domain run (reference fdtd2d_core.py:531-927); method='stitch' dispatches
    to run_beads_stitched.  Extra keywords of this build: ``npml`` (CPML cells, reference
    fixed 10), ``engine_mode`` ('ref32' | 'fp64acc'), ``beads`` (several beads),
    ``steps_override`` (exact step count, no CW extension)."""
    global TEAR_BUFFER, LAM_LOSS
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    if method == 'sliding':
        raise NotImplementedError("method 'sliding' is not part of the device path")
    npml_v = NPML if npml is None else int(npml)
    if stack is not None and tear_buffer_um is None:
        need = stack_total_um(stack) + 1.5 + npml_v * dx_nm * 1e-3
        if need * 1e-6 > TEAR_BUFFER + 1e-12:
            tear_buffer_um = need
    if tear_buffer_um is not None:
        old = TEAR_BUFFER
        TEAR_BUFFER = float(tear_buffer_um) * 1e-6
        try:
            kw = dict(locals())
            args = {k: kw[k] for k in _RUN_BEADS_ARGS}
            args['tear_buffer_um'] = None
            args['stack'] = stack
            return _run_beads_body(**args)
        finally:
            TEAR_BUFFER = old
    kw = dict(locals())
    return _run_beads_body(**{k: kw[k] for k in _RUN_BEADS_ARGS})


def _run_beads_body(wg_mat, bead_mat, bead_d_um, dx_nm=20.0, save_frames=True, n_snapshots=12,
                    wg_thickness_um=5.0, lambda_nm=850.0, vcsel_waist=3.0, vcsel_tilt=0.0,
                    vcsel_offset=0.0, source_type='cw', method='full', window_w_um=350.0,
                    slide_um=150.0, place_bead=True, t_aqueous_um=None, t_mucin_um=None,
                    t_lipid_um=0.0, input_gap_um=0.0, length_um=None, bead_x_um=None,
                    bead_y_um=None, wg_n=None, bead_n=None, n_aqueous=None, n_mucin=None,
                    n_cornea=None, n_lipid=None, polarization='s', curved=False, boundary='cpml',
                    cad=None, pulse_cycles=30.0, pulse_tmax_fs=None, pulse_nfreq=None,
                    pulse_dispersion=True, pulse_mode='fourier', n_aq_tear_offset=None,
                    tear_buffer_um=None, stack=None, npml=None, engine_mode=None, beads=None,
                    steps_override=None):
    global LAM_LOSS
    if method == 'stitch' and str(source_type).lower() == 'pulse' and pulse_mode == 'td':
        return run_beads_stitched_td(
            wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots, wg_thickness_um,
            lambda_nm, vcsel_waist, vcsel_tilt, vcsel_offset, window_w_um, slide_um, place_bead,
            pulse_cycles=pulse_cycles, pulse_tmax_fs=pulse_tmax_fs, cad=cad,
            t_aqueous_um=t_aqueous_um, t_mucin_um=t_mucin_um, t_lipid_um=t_lipid_um,
            input_gap_um=input_gap_um, length_um=length_um, bead_x_um=bead_x_um,
            bead_y_um=bead_y_um, wg_n=wg_n, bead_n=bead_n, n_aqueous=n_aqueous, n_mucin=n_mucin,
            n_cornea=n_cornea, n_lipid=n_lipid, polarization=polarization, curved=curved,
            boundary=boundary, stack=stack, npml=npml, engine_mode=engine_mode, beads=beads)
    if method == 'stitch' and str(source_type).lower() == 'pulse':
        return run_beads_stitched_pulse(
            wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots, wg_thickness_um,
            lambda_nm, vcsel_waist, vcsel_tilt, vcsel_offset, window_w_um, slide_um, place_bead,
            t_aqueous_um=t_aqueous_um, t_mucin_um=t_mucin_um, t_lipid_um=t_lipid_um,
            input_gap_um=input_gap_um, length_um=length_um, bead_x_um=bead_x_um,
            bead_y_um=bead_y_um, wg_n=wg_n, bead_n=bead_n, n_aqueous=n_aqueous, n_mucin=n_mucin,
            n_cornea=n_cornea, n_lipid=n_lipid, polarization=polarization, curved=curved,
            boundary=boundary, cad=cad, pulse_cycles=pulse_cycles, pulse_tmax_fs=pulse_tmax_fs,
            pulse_nfreq=pulse_nfreq, pulse_dispersion=pulse_dispersion,
            n_aq_tear_offset=n_aq_tear_offset, stack=stack, npml=npml, engine_mode=engine_mode,
            beads=beads)
    if method == 'stitch':
        return run_beads_stitched(wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots,
                                  wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt,
                                  vcsel_offset, source_type, window_w_um, slide_um, place_bead,
                                  t_aqueous_um, t_mucin_um=t_mucin_um, t_lipid_um=t_lipid_um,
                                  input_gap_um=input_gap_um, length_um=length_um,
                                  bead_x_um=bead_x_um, bead_y_um=bead_y_um, wg_n=wg_n,
                                  bead_n=bead_n, n_aqueous=n_aqueous, n_mucin=n_mucin,
                                  n_cornea=n_cornea, n_lipid=n_lipid, polarization=polarization,
                                  curved=curved, boundary=boundary, stack=stack, npml=npml,
                                  engine_mode=engine_mode, beads=beads)
    pol = 'p' if str(polarization).lower() in ('p', 'tm') else 's'
    kind = 'tm2d' if pol == 'p' else 'te2d'
    t_aq_um = t_aqueous_um if t_aqueous_um is not None else T_AQ * 1e6
    t_total = time.time()
    wg_n = wg_n if wg_n is not None else n_at(wg_mat, lambda_nm)
    bead_n = bead_n if bead_n is not None else n_at(bead_mat, lambda_nm)
    n_aq_l = n_aqueous if n_aqueous is not None else n_at('aqueous', lambda_nm)
    n_mu_l = n_mucin if n_mucin is not None else n_at('mucin', lambda_nm)
    n_co_l = n_cornea if n_cornea is not None else cornea_default(lambda_nm)
    n_lip_l = n_lipid if n_lipid is not None else n_at('lipid', lambda_nm)
    L = (length_um * 1e-6) if length_um is not None else L_DEMO
    L_um = L * 1e6
    bx_um = bead_x_um if bead_x_um is not None else L_um / 2.0
    T_WG = wg_thickness_um * 1e-6
    LAM = lambda_nm * 1e-9
    LAM_LOSS = LAM
    label = f'WG-{wg_mat}_Bead-{bead_mat}_d{bead_d_um:g}_L{lambda_nm:g}' + ('' if place_bead else '_ref')
    dx = dx_nm * 1e-9
    dt = 0.5 * dx / (C0 * np.sqrt(2))
    NP = NPML if npml is None else int(npml)
    bkind = str(boundary).lower()
    if bkind not in ('cpml', 'mur1'):
        raise ValueError(f"unknown boundary '{boundary}' (cpml|mur1)")
    use_cpml = bkind == 'cpml'
    src_ix = (NP + 2) if use_cpml else 5            # fdtd2d_core.py:628-629
    _vd = float(input_gap_um or 0.0)
    x_facet_um, _air, _zsrc = _vcsel_geom_2d(src_ix, dx, _vd)
    input_gap_um = x_facet_um
    y0_m = -TEAR_BUFFER
    H_total = T_WG + AIR_BUFFER + TEAR_BUFFER
    Nx, Ny = int(round(L / dx)), int(round(H_total / dx))
    print(f'\n===== Bead-Demo [full]: {label} =====  grid {Nx}x{Ny} dx={dx_nm}nm  pol {pol}  '
          f'engine {backend.engine_name()}')
    # materials and coefficients (fdtd2d_core.py:664-672): computed in compact form on
    # the host and assembled on the device (window_coeffs, lf_build_coeff)
    mp = _material_parts(Nx, Ny, dx, wg_n, bead_n, T_WG, bx_um, bead_d_um, 0.0, place_bead,
                         n_aq=n_aq_l, n_mucin=n_mu_l, n_cornea=n_co_l, t_aq_um=t_aq_um,
                         t_mu_um=t_mucin_um, t_lip_um=t_lipid_um, n_lip=n_lip_l,
                         input_gap_um=input_gap_um, length_um=length_um, curved=curved,
                         bead_y_um=bead_y_um, stack=stack, beads=beads)
    Ch = np.float32(dt / (MU0 * dx))
    eng = backend.make_engine(kind, (Nx, Ny), _engine_mode(engine_mode))
    packed, ch_tm = window_coeffs(pol, mp, dt, dx)
    eng.build_coeffs(ch_tm if pol == 'p' else Ch, packed)
    del packed
    # Mur with the local speeds (x ends: n_eff, y edges: n of the edge cells),
    # fdtd2d_core.py:675-678
    ms = None if use_cpml else MurSet(dx, dt, mur_speeds(_material_rows(mp, 0, Ny)[0], dx, LAM))
    f0 = C0 / LAM
    sigma_t = 4 / (2 * np.pi * f0)
    iy0 = int(round((T_WG / 2 + vcsel_offset * 1e-6 - y0_m) / dx))
    n_src = _n_source_2d(src_ix * dx * 1e6, iy0, y0_m, dx, wg_n, T_WG,
                         n_aq_l if stack is None else (float(complex(stack[0]['n']).real) if stack
                                                       else float(n_co_l)), input_gap_um)
    iy_src, vs, vs_ey = _vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, vcsel_offset, vcsel_waist,
                                         lambda_nm, vcsel_tilt, n_src, _zsrc, f0, y0_m=y0_m)
    bnd = make_boundary('cpml', pol, Nx, Ny, dx, dt, npml=NP if use_cpml else 0, lam_m=LAM)
    bnd.attach(eng)
    if ms is not None:                               # inline Mur-1 (fdtd2d_core.py:805-833)
        attach_mur_2d(eng, pol, ms)
        print('  ' + ms.describe())
    eng.finalize()
    src = vs_ey if pol == 'p' else vs
    field_comp = 'Ey' if pol == 'p' else 'Ez'
    eng.set_source(field_comp, src.ix, src.amp, src.ph1, src.ph2)
    steps_total = int((6 * sigma_t + 2 * L * N_G_FACTOR * wg_n / C0) / dt)
    if steps_override:
        steps_total = int(steps_override)
    steps_max = int(CW_MAX_EXTEND * steps_total)
    snap_every = max(1, steps_total // n_snapshots)
    iy_lo = int(round((0.0 + TEAR_BUFFER) / dx))
    iy_hi = int(round((T_WG + TEAR_BUFFER) / dx))
    hw = max(1, int(round(min(DET_LEN / 2, 0.05 * L_um * 1e-6) / dx)))

    def xwin(xu):
        c = int(round(xu * 1e-6 / dx))
        return max(0, c - hw), min(Nx, c + hw)
    ixin0, ixin1 = xwin(0.1 * L_um)
    ixout0, ixout1 = xwin(0.9 * L_um)
    ci = (ixin0 + ixin1) // 2
    sh = guard_in_plane(ci, max(src_ix, int(round(x_facet_um / (dx * 1e6)))), Nx) - ci
    ixin0, ixin1 = ixin0 + sh, min(Nx, ixin1 + sh)
    omega = 2 * np.pi * f0
    steps_per_T = max(4, int(round((1.0 / f0) / dt)))
    n_acc_cw = 8 * steps_per_T
    conv = source_type != 'pulse' and not steps_override
    cw_rel_change = float('nan')
    n_avg0 = steps_total // 2
    wide = eng.wide_acc
    acc_mode = 'c128' if wide else 'c64p'
    dft_src = 'ey_node' if pol == 'p' else 'Ez'
    mon_pin = eng.add_box(((ixin0, ixin1), (iy_lo, iy_hi)))
    mon_pout = eng.add_box(((ixout0, ixout1), (iy_lo, iy_hi)))
    mon_cw = eng.add_dft(dft_src, 0, Nx, acc=acc_mode)
    mon_cwex = eng.add_dft('ex_node', 0, Nx, acc=acc_mode) if pol == 'p' else None
    mon_prev = eng.add_dft(dft_src, 0, Nx, acc=acc_mode) if conv else None
    mon_I = eng.add_sq('tm' if pol == 'p' else 'te')
    book = PlaneFluxBook(eng, kind)
    for key, il in (('in', max(1, (ixin0 + ixin1) // 2)), ('out', max(1, (ixout0 + ixout1) // 2)),
                    ('src_up', vs.plane_up), ('src_dn', vs.plane_dn)):
        book.add_plane(key, il)
    flux_bits = sum(1 << book.mon[k] for k in book.keys())
    frames = []
    mm_path = None
    ez_mm = None
    k_fr = 0
    if save_frames:
        os.makedirs('results', exist_ok=True)
        mm_path = os.path.join('results', f'_stream_beads_{_safe(label)}_full.npy')
        ez_mm = np.lib.format.open_memmap(mm_path, mode='w+', dtype=np.float32,
                                          shape=(max(1, n_snapshots + 2), Nx, Ny))
    nacc_cw = nacc_prev = nacc_I = 0
    t_start = time.time()

    def masks(n0, n1):
        nonlocal nacc_cw, nacc_prev, nacc_I
        ns = np.arange(n0, n1)
        mk = np.full(len(ns), (1 << mon_pin) | (1 << mon_pout), dtype=np.uint32)
        in_cw = ns >= steps_total - n_acc_cw
        bits = (1 << mon_cw) | flux_bits | ((1 << mon_cwex) if mon_cwex is not None else 0)
        mk[in_cw] |= np.uint32(bits)
        cnt = int(in_cw.sum())
        nacc_cw += cnt
        book.note_samples(book.keys(), cnt)
        if conv:
            in_prev = (~in_cw) & (ns >= steps_total - 2 * n_acc_cw)
            mk[in_prev] |= np.uint32(1 << mon_prev)
            nacc_prev += int(in_prev.sum())
        in_I = ns >= n_avg0
        mk[in_I] |= np.uint32(1 << mon_I)
        nacc_I += int(in_I.sum())
        return mk

    n = 0
    while n < steps_total:
        stop = steps_total
        if ez_mm is not None and k_fr < ez_mm.shape[0]:
            stop = min(stop, ((n // snap_every) + 1) * snap_every)
        tab = new_steps(stop - n)
        t_phys = np.arange(n, stop) * dt
        env = (pulse_envelope(t_phys, f0, pulse_cycles) if source_type == 'pulse'
               else cw_envelope(t_phys, sigma_t))
        fill_phasors(tab, omega, dt, t_phys, env, src_omega=src.omega)
        tab['mask'] = masks(n, stop)
        eng.run(tab)
        n = stop
        if ez_mm is not None and n % snap_every == 0 and k_fr < ez_mm.shape[0]:
            f = eng.read_field(field_comp)
            ez_mm[k_fr] = (_tm_node_field(f) if pol == 'p' else f).astype(np.float32)
            frames.append({'x_start': 0.0, 'x_end': L_um, 'y_start': y0_m * 1e6,
                           'y_end': (y0_m + H_total) * 1e6, 'slide_i': 0, 'step': n,
                           'time': time.time() - t_start})
            k_fr += 1
        if conv and n == steps_total and nacc_prev > 0 and nacc_cw > 0:
            if wide:                             # FP64 windows: reduced on the device
                num, den = eng.monitor_change(mon_cw, mon_prev, nacc_cw, nacc_prev)
                cw_rel_change = float(np.sqrt(num) / max(float(np.sqrt(den)), 1e-30))
            else:                                # the reference's complex64 arithmetic
                a = (eng.read_monitor(mon_cw) / nacc_cw).astype(np.complex64)
                b = (eng.read_monitor(mon_prev) / nacc_prev).astype(np.complex64)
                cw_rel_change = float(np.linalg.norm(a - b) / max(float(np.linalg.norm(a)), 1e-30))
            if cw_rel_change > CW_TOL and steps_total + n_acc_cw <= steps_max:
                print(f'  [CW] not yet in steady state (change {100 * cw_rel_change:.2f}%) '
                      f'-> +{n_acc_cw} steps')
                mon_cw, mon_prev = mon_prev, mon_cw
                nacc_prev, nacc_cw = nacc_cw, 0
                eng.zero_monitor(mon_cw)
                if mon_cwex is not None:
                    eng.zero_monitor(mon_cwex)
                book.reset()
                steps_total += n_acc_cw
            else:
                print(f'  [CW] phasor change over the last 2 DFT windows: {100 * cw_rel_change:.3f}%')
    if ez_mm is not None:
        ez_mm.flush()
        del ez_mm
    Ecw = ((2.0 / nacc_cw) * eng.read_monitor(mon_cw)) if nacc_cw > 0 else None
    Ex_cw = ((2.0 / nacc_cw) * eng.read_monitor(mon_cwex)) if (mon_cwex is not None and nacc_cw) else None
    # steady-state frame set: Re(E_cw e^{i w t_k}) at n_snapshots phases of one period
    # (reference fdtd2d_core.py:876-890)
    cw_path, cw_frames = None, []
    if save_frames and Ecw is not None:
        nfr_cw = max(1, n_snapshots)
        cw_path = os.path.join('results', f'_stream_beads_{_safe(label)}_full_cw.npy')
        cw_mm = np.lib.format.open_memmap(cw_path, mode='w+', dtype=np.float32,
                                          shape=(nfr_cw,) + Ecw.shape)
        for k in range(nfr_cw):
            tk = (k / nfr_cw) * (1.0 / f0)
            cw_mm[k] = np.real(Ecw * np.exp(1j * omega * tk)).astype(np.float32)
            cw_frames.append({'x_start': 0.0, 'x_end': L_um, 'y_start': y0_m * 1e6,
                              'y_end': (y0_m + H_total) * 1e6, 'slide_i': 0, 'step': k, 'time': tk})
        cw_mm.flush()
        del cw_mm
    P_in = float(eng.read_monitor(mon_pin)) * dx * dx
    P_out = float(eng.read_monitor(mon_pout)) * dx * dx
    Iavg = (eng.read_monitor(mon_I) / max(1, nacc_I)) if nacc_I else None
    Tr_I = P_out / max(P_in, 1e-30)
    Tr, tr_def, p_si = Tr_I, 'intensity', None
    if Ecw is not None and source_type != 'pulse' and book.has('in') and book.has('out'):
        S_in, S_out = book.flux('in'), book.flux('out')
        p_si = S_in * dx
        if abs(S_in) > 0:
            Tr, tr_def = S_out / S_in, 'poynting'
    emit = _emit_power(book, dx) if source_type != 'pulse' else (None, None)
    eng.close()
    total_time = time.time() - t_total
    print(f'[Result] {label} [full]: total {total_time:.1f}s  Transmission={Tr:.4f} ({tr_def})')
    res = _result_dict(label, P_in, P_out, Tr, frames, total_time, bead_d_um, wg_mat, bead_mat,
                       wg_n, bead_n, wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt,
                       vcsel_offset, source_type, 'full', place_bead, n_aq_l, n_mu_l, n_co_l,
                       frames_memmap=mm_path, t_aq_um=t_aq_um, t_mu_um=t_mucin_um,
                       t_lip_um=t_lipid_um, bead_x_um=bx_um, x_wg_start_um=(input_gap_um or 0.0),
                       x_wg_end_um=L_um, polarization=pol, stack=stack, n_lip=n_lip_l)
    res.update(Iavg=Iavg, Ecw=Ecw, cw_frames=cw_frames, cw_frames_memmap=cw_path,
               cw_rel_change=cw_rel_change, transmission_intensity=Tr_I,
               transmission_def=tr_def, steps_total=steps_total, p_in_si=p_si, p_in_unit='W/m',
               view_primary='transient', cells=Nx * Ny, grid=(Nx, Ny))
    if Ex_cw is not None:
        res['Ex_cw_amp'] = Ex_cw
    _vcsel_result(res, *emit, _vd, _air, x_facet_um, vcsel_waist)
    return res


# ----------------------------------------------------------------------------- stitch
def run_beads_stitched(wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots,
                       wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt, vcsel_offset,
                       source_type, window_w_um, slide_um, place_bead=True, t_aqueous_um=None,
                       t_mucin_um=None, t_lipid_um=0.0, input_gap_um=0.0, length_um=None,
                       bead_x_um=None, bead_y_um=None, wg_n=None, bead_n=None, n_aqueous=None,
                       n_mucin=None, n_cornea=None, n_lipid=None, polarization='s', curved=False,
                       boundary='cpml', cad=None, return_efull=False, stack=None, npml=None,
                       engine_mode=None, beads=None, steps_win_override=None, n_windows=None,
                       window_hook=None, tune_cache=None):
    """2D domain decomposition (reference fdtd2d_core.py:1197-1583): windows of
    ``window_w_um`` every ``slide_um``, each run to CW steady state; the previous window's
    complex field is imposed with a cosine taper over the overlap; inner seams end in a
    CPML extension that is not assembled.  ``steps_win_override`` / ``n_windows`` cap the
    work per window / the number of windows (benchmarks); ``window_hook(w, info)`` is
    called after each window; ``tune_cache`` (a dict) carries the blocking chosen for each
    window shape across calls (the pulse comb's lines share their window shapes)."""
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    bkind = (boundary or 'cpml').lower()
    if bkind not in ('cpml', 'mur1'):
        raise ValueError(f"stitch: unknown boundary '{bkind}' (cpml|mur1)")
    ucpml = bkind == 'cpml'                # the user's choice for the outer edges
    global LAM_LOSS
    pol = 'p' if str(polarization).lower() in ('p', 'tm') else 's'
    kind = 'tm2d' if pol == 'p' else 'te2d'
    t_total = time.time()
    t_aq_um = t_aqueous_um if t_aqueous_um is not None else T_AQ * 1e6
    wg_n = wg_n if wg_n is not None else n_at(wg_mat, lambda_nm)
    bead_n = bead_n if bead_n is not None else n_at(bead_mat, lambda_nm)
    n_aq_l = n_aqueous if n_aqueous is not None else n_at('aqueous', lambda_nm)
    n_mu_l = n_mucin if n_mucin is not None else n_at('mucin', lambda_nm)
    n_co_l = n_cornea if n_cornea is not None else cornea_default(lambda_nm)
    n_lip_l = n_lipid if n_lipid is not None else n_at('lipid', lambda_nm)
    L = (length_um * 1e-6) if length_um is not None else L_DEMO
    L_um = L * 1e6
    T_WG = wg_thickness_um * 1e-6
    LAM = lambda_nm * 1e-9
    LAM_LOSS = LAM
    bx_um = bead_x_um if bead_x_um is not None else L_um / 2.0
    label = f'WG-{wg_mat}_Bead-{bead_mat}_d{bead_d_um:g}_L{lambda_nm:g}' + ('' if place_bead else '_ref')
    dx = dx_nm * 1e-9
    dt = 0.5 * dx / (C0 * np.sqrt(2))
    NP = NPML if npml is None else int(npml)
    _vd = float(input_gap_um or 0.0)
    x_facet_um, _air, _zsrc = _vcsel_geom_2d((NP + 2) if ucpml else 5, dx, _vd)
    input_gap_um = x_facet_um
    y0_m = -TEAR_BUFFER
    H_total = T_WG + AIR_BUFFER + TEAR_BUFFER
    Ny = int(round(H_total / dx))
    Nx_win = int(round(window_w_um * 1e-6 / dx))
    S_cells = int(round(slide_um * 1e-6 / dx))
    O_cells = Nx_win - S_cells
    Nx_full = int(round(L / dx))
    if O_cells < 8:
        O_cells = max(O_cells, 1)
    f0 = C0 / LAM
    omega = 2 * np.pi * f0
    sigma_t = 4 / (2 * np.pi * f0)
    transit = window_w_um * 1e-6 * N_G_FACTOR * wg_n / C0
    steps_win = int(3.0 * transit / dt) + 400
    if steps_win_override:
        steps_win = int(steps_win_override)
    steps_per_T = max(4, int(round((1.0 / f0) / dt)))
    n_acc = steps_per_T * 8
    n_win = (int(np.ceil((Nx_full - Nx_win) / S_cells)) + 1) if Nx_full > Nx_win else 1
    n_win_run = n_win if n_windows is None else min(n_win, int(n_windows))
    print(f'\n===== Bead-Demo [stitch]: {label} =====  {n_win} windows of {Nx_win}x{Ny}, overlap '
          f'{O_cells}, steps/window {steps_win}, pol {pol}, engine {backend.engine_name()}')
    Ch = np.float32(dt / (MU0 * dx))
    src_ix = (NP + 2) if ucpml else 5      # fdtd2d_core.py:1277
    iy0 = int(round((T_WG / 2 + vcsel_offset * 1e-6 - y0_m) / dx))
    n_src = _n_source_2d(src_ix * dx * 1e6, iy0, y0_m, dx, wg_n, T_WG,
                         n_aq_l if stack is None else (float(complex(stack[0]['n']).real) if stack
                                                       else float(n_co_l)), input_gap_um)
    iy_src, vs, vs_ey = _vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, vcsel_offset, vcsel_waist,
                                         lambda_nm, vcsel_tilt, n_src, _zsrc, f0, y0_m=y0_m)
    src = vs_ey if pol == 'p' else vs
    Nyc = (Ny - 1) if pol == 'p' else Ny
    comp = 'Ey' if pol == 'p' else 'Ez'
    wtap = (0.5 * (1.0 + np.cos(np.pi * np.arange(O_cells) / max(O_cells - 1, 1)))).astype(np.float32)
    Efull = np.zeros((Nx_full, Nyc), dtype=np.complex64)
    Ehand = None
    ExI_full = np.zeros((Nx_full, Ny), dtype=np.float32) if pol == 'p' else None
    ExA_full = np.zeros((Nx_full, Ny), dtype=np.complex64) if pol == 'p' else None

    def gx0(w):
        return min(w * S_cells, max(Nx_full - Nx_win, 0))
    EXT = NP + 4
    len_mat = L_um + EXT * dx * 1e6
    iy_lo = int(round((0.0 + TEAR_BUFFER) / dx))
    iy_hi = int(round((T_WG + TEAR_BUFFER) / dx))
    ix_in = int(round(0.1 * Nx_full))
    ix_out = int(round(0.9 * Nx_full))
    ix_in = guard_in_plane(ix_in, max(src_ix, int(round(x_facet_um / (dx * 1e6)))), Nx_full)
    fluxes = FluxAccumulator()
    mode = _engine_mode(engine_mode)
    win_info = []
    tuned = {} if tune_cache is None else tune_cache   # blocking chosen for each window shape
    for w in range(n_win_run):
        g0 = gx0(w)
        x_start_um = g0 * dx * 1e6
        last = w == n_win - 1
        glo = 0 if w == 0 else g0 + O_cells
        ghi = min(g0 + Nx_win, Nx_full)
        fpl = {k: g - g0 for k, g in (('in', ix_in), ('out', ix_out)) if glo <= g < ghi and g - g0 >= 1}
        if w == 0:
            fpl.update(src_up=vs.plane_up, src_dn=vs.plane_dn)
        Nxw = Nx_win + (0 if last else EXT)
        # the window's materials and coefficients (fdtd2d_core.py:1350-1379), computed in
        # compact form on the host and assembled on the device (window_coeffs)
        mp = _material_parts(Nxw, Ny, dx, wg_n, bead_n, T_WG, bx_um, bead_d_um, x_start_um,
                             place_bead, n_aq=n_aq_l, n_mucin=n_mu_l, n_cornea=n_co_l,
                             t_aq_um=t_aq_um, t_mu_um=t_mucin_um, t_lip_um=t_lipid_um,
                             n_lip=n_lip_l, input_gap_um=input_gap_um, length_um=len_mat,
                             curved=curved, bead_y_um=bead_y_um, stack=stack, beads=beads)
        eng = backend.make_engine(kind, (Nxw, Ny), mode)
        packed, ch_tm = window_coeffs(pol, mp, dt, dx)
        eng.build_coeffs(ch_tm if pol == 'p' else Ch, packed)
        del packed
        ms = None if ucpml else MurSet(dx, dt, mur_speeds(_material_rows(mp, 0, Ny)[0], dx, LAM))
        # the CPML carries the seams (inner windows' right end, always) and the outer
        # edges the user chose it for; the outer edges without it get Mur-1
        # (fdtd2d_core.py:1390-1398)
        pml_l = (w == 0) and ucpml
        pml_r = (not last) or ucpml
        bnd = make_boundary('cpml', pol, Nxw, Ny, dx, dt,
                            npml_x=(NP if pml_l else 0, NP if pml_r else 0),
                            npml_y=NP if ucpml else 0, lam_m=LAM)
        bnd.attach(eng)
        if ms is not None:
            attach_mur_2d(eng, pol, ms, x_lo=(w == 0 and not pml_l), x_hi=not pml_r, y=True)
        eng.finalize()
        if w == 0:
            eng.set_source(comp, src.ix, src.amp, src.ph1, src.ph2)
        elif Ehand is not None:
            eng.set_drive(O_cells, wtap, {comp: Ehand}, math='c128')
        wide = eng.wide_acc
        acc_mode = 'c128' if wide else 'c64'
        mon_acc = eng.add_dft(comp, 0, Nxw, acc=acc_mode, phasor='p')
        mon_ex2 = eng.add_sq('tm_ex2') if pol == 'p' else None
        mon_exa = eng.add_dft('ex_node', 0, Nxw, acc=acc_mode, phasor='p') if pol == 'p' else None
        book = PlaneFluxBook(eng, kind)
        for k, il in fpl.items():
            book.add_plane(k, il, carry=fluxes)
        bits = (1 << mon_acc) | sum(1 << book.mon[k] for k in fpl)
        if pol == 'p':
            bits |= (1 << mon_ex2) | (1 << mon_exa)
        ns = np.arange(steps_win)
        t_phys = ns * dt
        tab = new_steps(steps_win)
        fill_phasors(tab, omega, dt, t_phys, cw_envelope(t_phys, sigma_t), src_omega=src.omega)
        win = ns >= steps_win - n_acc
        tab['mask'] = np.where(win, np.uint32(bits), np.uint32(0))
        backend.tune(eng, tab[steps_win // 2], steps_win, cache=tuned,
                     key=(Nxw, Ny, w > 0 and Ehand is not None))
        acc_n = int(win.sum())
        book.note_samples(list(fpl), acc_n)
        tw = time.time()
        chunk = max(1, steps_win // 10)
        for c0 in range(0, steps_win, chunk):
            eng.run(tab[c0:c0 + chunk])
        eng.sync()
        tw = time.time() - tw
        Ew = ((2.0 / max(acc_n, 1)) * eng.read_monitor(mon_acc)).astype(np.complex64)
        if not last:
            off = gx0(w + 1) - g0
            Ehand = Ew[off:off + O_cells, :].copy()
        if w == 0:
            g_lo, l_lo = 0, 0
        else:
            g_lo, l_lo = g0 + O_cells, O_cells
        g_hi = min(g0 + Nx_win, Nx_full)
        l_hi = l_lo + (g_hi - g_lo)
        if g_hi > g_lo:
            Efull[g_lo:g_hi, :] = Ew[l_lo:l_hi, :]
            if ExI_full is not None:
                ExI_full[g_lo:g_hi, :] = eng.read_monitor(mon_ex2)[l_lo:l_hi, :] / max(acc_n, 1)
                ExA_full[g_lo:g_hi, :] = eng.read_monitor(mon_exa)[l_lo:l_hi, :] * (2.0 / max(acc_n, 1))
        fluxes.take(book, list(fpl))
        eng.close()
        info = dict(window=w, cells=Nxw * Ny, steps=steps_win, seconds=tw)
        win_info.append(info)
        if window_hook is not None:
            window_hook(w, info)
        print(f'  [Window {w + 1}/{n_win}] x0={x_start_um:.0f}um  {tw:.2f}s  '
              f'max|E|={np.abs(Ew).max():.3e}', flush=True)
    total_time = time.time() - t_total
    frames = []
    mm_path = None
    nfr = max(1, n_snapshots)
    if save_frames:
        os.makedirs('results', exist_ok=True)
        mm_path = os.path.join('results', f'_stream_beads_{_safe(label)}_stitch.npy')
        ez_mm = np.lib.format.open_memmap(mm_path, mode='w+', dtype=np.float32,
                                          shape=(nfr, Nx_full, Ny))
        for k in range(nfr):
            tk = (k / nfr) * (1.0 / f0)
            fr = np.real(Efull * np.exp(1j * omega * tk)).astype(np.float32)
            ez_mm[k] = _tm_node_field(fr) if pol == 'p' else fr
            frames.append({'x_start': 0.0, 'x_end': L_um, 'y_start': y0_m * 1e6,
                           'y_end': (y0_m + H_total) * 1e6, 'slide_i': 0, 'step': k, 'time': tk})
        ez_mm.flush()
        del ez_mm
    P_in = float(np.sum(np.abs(Efull[ix_in, iy_lo:iy_hi]) ** 2))
    P_out = float(np.sum(np.abs(Efull[ix_out, iy_lo:iy_hi]) ** 2))
    if ExI_full is not None:
        P_in += 2.0 * float(np.sum(ExI_full[ix_in, iy_lo:iy_hi]))
        P_out += 2.0 * float(np.sum(ExI_full[ix_out, iy_lo:iy_hi]))
        for x0 in range(0, Nx_full, 4096):
            x1 = min(x0 + 4096, Nx_full)
            s = 0.5 * np.abs(Efull[x0:x1]) ** 2
            ey2 = np.empty((x1 - x0, Ny), dtype=np.float32)
            ey2[:, 1:-1] = 0.5 * (s[:, :-1] + s[:, 1:])
            ey2[:, 0] = s[:, 0]
            ey2[:, -1] = s[:, -1]
            ExI_full[x0:x1] += ey2
    Tr_I = P_out / max(P_in, 1e-30)
    p_si = None
    if fluxes.has('in') and fluxes.has('out'):
        S_in, S_out = fluxes.flux('in'), fluxes.flux('out')
        p_si = S_in * dx
    else:
        S_in = S_out = 0.0
    Tr = S_out / S_in if abs(S_in) > 0 else Tr_I
    print(f'[Result] {label} [stitch]: total {total_time:.1f}s  Transmission={Tr:.4f}')
    res = _result_dict(label, P_in, P_out, Tr, frames, total_time, bead_d_um, wg_mat, bead_mat,
                       wg_n, bead_n, wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt,
                       vcsel_offset, source_type, 'stitch', place_bead, n_aq_l, n_mu_l, n_co_l,
                       frames_memmap=mm_path, t_aq_um=t_aq_um, t_mu_um=t_mucin_um,
                       t_lip_um=t_lipid_um, bead_x_um=bx_um, x_wg_start_um=(input_gap_um or 0.0),
                       x_wg_end_um=L_um, polarization=pol, stack=stack, n_lip=n_lip_l)
    res.update(transmission_intensity=Tr_I, transmission_def='poynting', view_primary='cw',
               cw_only=True, p_in_si=p_si, p_in_unit='W/m', windows=win_info,
               n_windows=n_win, steps_win=steps_win, Efull=Efull if return_efull else None)
    _vcsel_result(res, *_emit_power(fluxes, dx), _vd, _air, x_facet_um, vcsel_waist)
    if ExI_full is not None:
        res['Iavg'] = ExI_full
        res['Ex_cw_amp'] = ExA_full
    return res


# ----------------------------------------------------------------------------- pulse stitch
_FIXED_INDEX = (('wg_n', None), ('bead_n', None), ('n_aqueous', 'aqueous'),
                ('n_mucin', 'mucin'), ('n_cornea', None), ('n_lipid', 'lipid'))


def run_beads_stitched_pulse(wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots,
                             wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt, vcsel_offset,
                             window_w_um, slide_um, place_bead, pulse_cycles=30.0,
                             pulse_tmax_fs=None, pulse_nfreq=None, pulse_dispersion=True,
                             cad=None, t_eval_s=None, n_aq_tear_offset=None, stack=None, **kw):
    """Transient stitch by Fourier synthesis (reference fdtd2d_core.py:1866-2017).

    The Gaussian source pulse (the full-domain pulse source) is split into the comb of
    lens_sensor_amd.common.fourier_pulse.PulseComb; every comb line is one CW stitch
    (run_beads_stitched) at lambda_k, with the materials at lambda_k unless
    ``pulse_dispersion`` is off (then every index stays at lambda0).  The transient
    frames E(t_j) = (1/pi) Re sum_k i E_k S_k e^{i w_k t_j} dw are summed frequency by
    frequency into a memory-mapped array; the CW frames at the centre line and the
    spectrum T(lambda_k) come with the result, and ``transmission`` is the |S|^2-weighted
    mean.  Keywords run_beads_stitched takes (polarization, boundary, npml, engine_mode,
    beads, ...) pass through."""
    from lens_sensor_amd.common.fourier_pulse import PulseComb, add_frequency
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    t_total = time.time()
    pol = 'p' if str(kw.get('polarization', 's')).lower() in ('p', 'tm') else 's'
    L_um = kw.get('length_um') or L_DEMO * 1e6
    n_core = kw.get('wg_n') or n_at(wg_mat, lambda_nm)
    y0_um, y1_um = -TEAR_BUFFER * 1e6, (wg_thickness_um * 1e-6 + AIR_BUFFER) * 1e6
    # the pulse crosses the structure at a group index about 15 % above n_core
    comb = PulseComb(lambda_nm, pulse_cycles, L_um * 1e-6 * n_core * 1.15 / C0,
                     tmax_fs=pulse_tmax_fs, nfreq=pulse_nfreq)
    if comb.aliased:
        print(f'  [WARN] {pulse_nfreq} frequencies -> repetition period '
              f'{2 * np.pi / comb.dw * 1e15:.0f} fs < t_max {comb.t_max * 1e15:.0f} fs: '
              'temporal aliasing.')
    fixed = {}
    if not pulse_dispersion:
        for key, mat in _FIXED_INDEX:
            if kw.get(key):
                fixed[key] = kw[key]
            elif key == 'wg_n':
                fixed[key] = n_at(wg_mat, lambda_nm)
            elif key == 'bead_n':
                fixed[key] = n_at(bead_mat, lambda_nm)
            elif key == 'n_cornea':
                fixed[key] = cornea_default(lambda_nm)
            else:
                fixed[key] = n_at(mat, lambda_nm)
    base = {k: v for k, v in kw.items() if k not in fixed}
    tear_follows = pulse_dispersion and n_aq_tear_offset is not None
    if t_eval_s is not None:
        t_j = np.asarray(t_eval_s, dtype=np.float64)
    else:
        t_j = np.linspace(0.0, comb.t_max, max(int(n_snapshots), 24))
    label = f'WG-{wg_mat}_Bead-{bead_mat}_d{bead_d_um:g}_L{lambda_nm:g}' + ('' if place_bead else '_ref')
    print(f'\n===== Fourier-Stitch (transient): {label} =====')
    print(f'Pulse: {pulse_cycles:g} periods (s={comb.sp * 1e15:.1f} fs), band '
          f'{comb.lam_nm.min():.1f}..{comb.lam_nm.max():.1f} nm, {len(comb)} frequencies, '
          f'dw/2pi={comb.dw / 2 / np.pi / 1e12:.2f} THz, t_max={comb.t_max * 1e15:.0f} fs, '
          f'{len(t_j)} time frames, dispersion {"on" if pulse_dispersion else "off (n at lambda0)"}')
    if stack is not None:
        print(f'Layers: {describe_stack(stack)}'
              + ('  (material layers follow n(lambda_k))' if pulse_dispersion else ''))
    frames_t = mm_path = None
    spec_T, spec_P = [], []
    r_c = E_c = None
    cell_updates = 0
    amp = np.abs(comb.S).max()
    tuned = {}                           # the lines share their window shapes
    # the frame sums of line k run on a host thread while line k+1 steps on the device
    pool = ThreadPoolExecutor(max_workers=1)
    pending = None
    for k, (wk, lk, Sk) in enumerate(zip(comb.w, comb.lam_nm, comb.S)):
        t1 = time.time()
        args = dict(base)
        if tear_follows:
            args['n_aqueous'] = n_at('aqueous', float(lk)) + float(n_aq_tear_offset)
        stk = stack_at_lambda(stack, float(lk)) if (stack is not None and pulse_dispersion) else stack
        with contextlib.redirect_stdout(io.StringIO()):
            r = run_beads_stitched(wg_mat, bead_mat, bead_d_um, dx_nm, False, 4,
                                   wg_thickness_um, float(lk), vcsel_waist, vcsel_tilt,
                                   vcsel_offset, 'cw', window_w_um, slide_um, place_bead,
                                   return_efull=True, stack=stk, tune_cache=tuned, **fixed,
                                   **args)
        E = np.asarray(r.pop('Efull'))
        if pol == 'p':                       # Ey (Nx, Ny-1) onto the node grid (Nx, Ny)
            En = np.empty((E.shape[0], E.shape[1] + 1), dtype=np.complex64)
            En[:, 1:-1] = 0.5 * (E[:, :-1] + E[:, 1:])
            En[:, 0], En[:, -1] = E[:, 0], E[:, -1]
            E = En
        if frames_t is None:
            os.makedirs('results', exist_ok=True)
            mm_path = os.path.join('results', f'_stream_beads_{_safe(label)}_pulse_{os.getpid()}.npy')
            frames_t = np.lib.format.open_memmap(mm_path, mode='w+', dtype=np.float32,
                                                 shape=(len(t_j),) + E.shape)
            frames_t[:] = 0.0
        if pending is not None:
            pending.result()
        pending = pool.submit(add_frequency, frames_t, E, Sk * comb.dw / np.pi, wk, t_j, True)
        spec_T.append(float(r['transmission']))
        spec_P.append(float(r['P_out']))
        cell_updates += sum(wi['cells'] * wi['steps'] for wi in r.get('windows') or [])
        if k == comb.centre:
            r_c, E_c = r, E
        print(f'  [{k + 1}/{len(comb)}] lambda={lk:.2f} nm  |S|={abs(Sk) / amp:.3f}  '
              f'T={r["transmission"]:.4f}  ({time.time() - t1:.0f}s)', flush=True)
    if pending is not None:
        pending.result()
    pool.shutdown()
    frames_t.flush()
    del frames_t
    frames = [{'x_start': 0.0, 'x_end': L_um, 'y_start': y0_um, 'y_end': y1_um, 'slide_i': 0,
               'step': j, 'time': float(t)} for j, t in enumerate(t_j)]
    # the steady state at the centre line, the analyzer's second view
    wc = comb.w[comb.centre]
    nfr_cw = max(1, n_snapshots)
    cw_path = os.path.join('results', f'_stream_beads_{_safe(label)}_pulse_cw_{os.getpid()}.npy')
    cw_mm = np.lib.format.open_memmap(cw_path, mode='w+', dtype=np.float32,
                                      shape=(nfr_cw,) + E_c.shape)
    cw_frames = []
    for q in range(nfr_cw):
        tq = (q / nfr_cw) * (2 * np.pi / wc)
        cw_mm[q] = np.real(E_c * np.exp(1j * wc * tq)).astype(np.float32)
        cw_frames.append({'x_start': 0.0, 'x_end': L_um, 'y_start': y0_um, 'y_end': y1_um,
                          'slide_i': 0, 'step': q, 'time': tq})
    cw_mm.flush()
    del cw_mm
    res = dict(r_c)
    res.update(scenario=label, method='stitch-pulse', frames=frames, frames_memmap=mm_path,
               cw_frames=cw_frames, cw_frames_memmap=cw_path, view_primary='transient',
               cw_only=False, Iavg=None, source_type='pulse', lambda_nm=lambda_nm,
               total_time=time.time() - t_total,
               spectrum=dict(lam_nm=comb.lam_nm, T=np.array(spec_T), P_out=np.array(spec_P),
                             weight=np.abs(comb.S) / amp),
               pulse=dict(cycles=pulse_cycles, s_fs=comb.sp * 1e15, tmax_fs=comb.t_max * 1e15,
                          nfreq=len(comb), dispersion=bool(pulse_dispersion)),
               cell_updates=cell_updates)
    if stack is not None:                    # the stack at lambda0
        res['stack'] = [dict(layer) for layer in stack]
        res['stack_custom'] = True
        res['layers'] = legacy_layers_tuple(stack)
    res['transmission'] = comb.weighted(spec_T)
    print(f'\n[Result] {label} [Fourier-Stitch]: {len(comb)} frequencies in '
          f'{res["total_time"]:.0f}s  T(weighted)={res["transmission"]:.4f}')
    return res


# ----------------------------------------------------------------------------- time-domain stitch
def run_beads_stitched_td(wg_mat, bead_mat, bead_d_um, dx_nm, save_frames, n_snapshots,
                          wg_thickness_um, lambda_nm, vcsel_waist, vcsel_tilt, vcsel_offset,
                          window_w_um, slide_um, place_bead, pulse_cycles=30.0,
                          pulse_tmax_fs=None, cad=None, t_eval_s=None, t_aqueous_um=None,
                          t_mucin_um=None, t_lipid_um=0.0, input_gap_um=0.0, length_um=None,
                          bead_x_um=None, bead_y_um=None, wg_n=None, bead_n=None,
                          n_aqueous=None, n_mucin=None, n_cornea=None, n_lipid=None,
                          polarization='s', curved=False, boundary='cpml', stack=None,
                          npml=None, engine_mode=None, beads=None, **_ignored):
    """Transient stitch in the time domain with Huygens (TF/SF) planes (reference
    fdtd2d_core.py:1586-1863).

    Every window runs on the same absolute time axis t = n dt, only while signal is
    present (from the first incoming step to the stop test), with CPML on all edges.
    Window w records, at the column where window w+1's TF/SF plane i0 = npml + 2 lies, E(n)
    (TE Ez, TM Ey) at the start of each step and H(n+1/2) (TE Hy, TM Hz) after the H
    update (device plane recorders, read back every TD_ROWS steps); window w+1 replays them
    as the TF/SF boundary (H corrected after the H update, E after the E update).  Frames
    at the steps nearest to t_j, time-integrated core intensity and Poynting flux at 10 % /
    90 % of the length, the pulse energy around the source.  The materials are those at
    lambda0 (no dispersion)."""
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    global LAM_LOSS
    pol = 'p' if str(polarization).lower() in ('p', 'tm') else 's'
    kind = 'tm2d' if pol == 'p' else 'te2d'
    t_total = time.time()
    t_aq_um = t_aqueous_um if t_aqueous_um is not None else T_AQ * 1e6
    wg_n = wg_n if wg_n is not None else n_at(wg_mat, lambda_nm)
    bead_n = bead_n if bead_n is not None else n_at(bead_mat, lambda_nm)
    n_aq_l = n_aqueous if n_aqueous is not None else n_at('aqueous', lambda_nm)
    n_mu_l = n_mucin if n_mucin is not None else n_at('mucin', lambda_nm)
    n_co_l = n_cornea if n_cornea is not None else cornea_default(lambda_nm)
    n_lip_l = n_lipid if n_lipid is not None else n_at('lipid', lambda_nm)
    L = (length_um * 1e-6) if length_um is not None else L_DEMO
    L_um = L * 1e6
    T_WG = wg_thickness_um * 1e-6
    LAM = lambda_nm * 1e-9
    LAM_LOSS = LAM
    bx_um = bead_x_um if bead_x_um is not None else L_um / 2.0
    label = f'WG-{wg_mat}_Bead-{bead_mat}_d{bead_d_um:g}_L{lambda_nm:g}' + ('' if place_bead else '_ref')
    dx = dx_nm * 1e-9
    dt = 0.5 * dx / (C0 * np.sqrt(2))
    NP = NPML if npml is None else int(npml)
    _vd = float(input_gap_um or 0.0)
    x_facet_um, _air, _zsrc = _vcsel_geom_2d(NP + 2, dx, _vd)
    input_gap_um = x_facet_um
    y0_m = -TEAR_BUFFER
    H_total = T_WG + AIR_BUFFER + TEAR_BUFFER
    Ny = int(round(H_total / dx))
    Nx_win = int(round(window_w_um * 1e-6 / dx))
    S_cells = int(round(slide_um * 1e-6 / dx))
    O_cells = Nx_win - S_cells
    Nx_full = int(round(L / dx))
    i0 = NP + 2                              # TF/SF column of the driven windows
    if O_cells < i0 + 4:
        raise ValueError(f'Overlap {O_cells} cells < {i0 + 4}: choose window_w >> slide.')
    f0 = C0 / LAM
    omega = 2 * np.pi * f0
    sp = pulse_cycles / f0
    tc = 3 * sp
    t_max = (pulse_tmax_fs * 1e-15) if pulse_tmax_fs else (tc + 3 * sp + L * wg_n * 1.15 / C0)
    N = int(np.ceil(t_max / dt)) + 1
    n_win = (int(np.ceil((Nx_full - Nx_win) / S_cells)) + 1) if Nx_full > Nx_win else 1

    def gx0(w):
        return min(w * S_cells, max(Nx_full - Nx_win, 0))
    EXT = NP + 4
    len_mat = L_um + EXT * dx * 1e6
    t_j = (np.asarray(t_eval_s, dtype=np.float64) if t_eval_s is not None
           else np.linspace(0.0, t_max, max(int(n_snapshots), 24)))
    n_j = np.clip(np.round(t_j / dt).astype(np.int64), 0, N - 1)
    frame_steps = np.unique(n_j)
    print(f'\n===== Time-domain stitch (Huygens/TF-SF): {label} =====')
    print(f'Pulse {pulse_cycles:g} periods (s={sp * 1e15:.1f} fs), t_max={t_max * 1e15:.0f} fs = '
          f'{N} steps, {n_win} windows of {window_w_um}um (overlap {O_cells * dx * 1e6:.2f}um), '
          f'{len(n_j)} frames, engine {backend.engine_name()}')
    if stack is not None:
        print(f'Layers: {describe_stack(stack)}')
    src_ix = NP + 2
    iy0 = int(round((T_WG / 2 + vcsel_offset * 1e-6 - y0_m) / dx))
    n_src = _n_source_2d(src_ix * dx * 1e6, iy0, y0_m, dx, wg_n, T_WG,
                         n_aq_l if stack is None else (float(complex(stack[0]['n']).real) if stack
                                                       else float(n_co_l)), input_gap_um)
    _iy, vs, vs_ey = _vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, vcsel_offset, vcsel_waist,
                                      lambda_nm, vcsel_tilt, n_src, _zsrc, f0, y0_m=y0_m)
    src = vs_ey if pol == 'p' else vs
    env_t = np.exp(-((np.arange(N) * dt - tc) / sp) ** 2)
    os.makedirs('results', exist_ok=True)
    mm_path = os.path.join('results', f'_stream_beads_{_safe(label)}_td_{os.getpid()}.npy')
    fr_mm = np.lib.format.open_memmap(mm_path, mode='w+', dtype=np.float32,
                                      shape=(len(n_j), Nx_full, Ny))
    fr_mm[:] = 0.0
    Ch = np.float32(dt / (MU0 * dx))
    Nyc = (Ny - 1) if pol == 'p' else Ny
    e_comp, h_comp = ('Ey', 'Hz') if pol == 'p' else ('Ez', 'Hy')
    iy_lo = int(round(TEAR_BUFFER / dx))
    iy_hi = int(round((T_WG + TEAR_BUFFER) / dx))
    ix_in = int(round(0.1 * Nx_full))
    ix_out = int(round(0.9 * Nx_full))
    ix_in = guard_in_plane(ix_in, max(src_ix, int(round(x_facet_um / (dx * 1e6)))), Nx_full)
    mode = _engine_mode(engine_mode)
    rec_in = None
    P = dict(P_in=0.0, P_out=0.0, S_in=0.0, S_out=0.0, S_up=0.0, S_dn=0.0)
    win_info = []
    for w in range(n_win):
        t1 = time.time()
        g0 = gx0(w)
        last = w == n_win - 1
        Nxw = Nx_win + (0 if last else EXT)
        x_start_um = g0 * dx * 1e6
        mp = _material_parts(Nxw, Ny, dx, wg_n, bead_n, T_WG, bx_um, bead_d_um, x_start_um,
                             place_bead, n_aq=n_aq_l, n_mucin=n_mu_l, n_cornea=n_co_l,
                             t_aq_um=t_aq_um, t_mu_um=t_mucin_um, t_lip_um=t_lipid_um,
                             n_lip=n_lip_l, input_gap_um=input_gap_um, length_um=len_mat,
                             curved=curved, bead_y_um=bead_y_um, stack=stack, beads=beads)
        eng = backend.make_engine(kind, (Nxw, Ny), mode)
        packed, ch_tm = window_coeffs(pol, mp, dt, dx)
        eng.build_coeffs(ch_tm if pol == 'p' else Ch, packed)
        del packed, mp
        # CPML on every edge of every window (fdtd2d_core.py:1726)
        make_boundary('cpml', pol, Nxw, Ny, dx, dt, npml_x=(NP, NP), npml_y=NP,
                      lam_m=LAM).attach(eng)
        eng.finalize()
        i_rec = (gx0(w + 1) - g0 + i0) if not last else None
        if w == 0:
            eng.set_source(e_comp, src.ix, src.amp, src.ph1, src.ph2)
            n_start = 0
            n_in_last = int(np.nonzero(env_t > 1e-7)[0][-1])
        else:
            n_start, n_in_last = signal_span(np.abs(rec_in[0]).max(axis=1), N)
            # TF/SF (fdtd2d_core.py:1771, 1784 TE; 1779, 1790 TM): H left of i0 carries the
            # scattered field only, E at i0 the total field
            if pol == 'p':
                tf_h = eng.add_tfsf('Hz', i0 - 1, +1, None, [(0, Ny - 1)], TD_ROWS)
                tf_e = eng.add_tfsf('Ey', i0, +1, 'cay', [(0, Ny - 1)], TD_ROWS)
            else:
                tf_h = eng.add_tfsf('Hy', i0 - 1, -1, None, [(0, Ny)], TD_ROWS)
                tf_e = eng.add_tfsf('Ez', i0, -1, 'ceH', [(1, Ny - 1)], TD_ROWS)
        bits = 0
        rec = None
        if i_rec is not None:
            rec = (np.zeros((N, Nyc), dtype=np.float32), np.zeros((N, Nyc), dtype=np.float32))
            m_re = eng.add_rec(e_comp, i_rec, 0, TD_ROWS)
            m_rh = eng.add_rec(h_comp, i_rec - 1, 1, TD_ROWS)
            bits |= (1 << m_re) | (1 << m_rh)
        a_g, b_g = (0, Nx_win) if w == 0 else (O_cells, Nx_win)
        if last:
            b_g = Nx_full - g0
        sums = {}                                  # name -> monitor id (summed over steps)
        if w == 0:
            sums['S_up'] = eng.add_sx(src.plane_up)
            sums['S_dn'] = eng.add_sx(src.plane_dn)
        for ixg, which in ((ix_in, 'in'), (ix_out, 'out')):
            il = ixg - g0
            if a_g <= il < b_g:
                sums['P_' + which] = eng.add_box(((il, il + 1), (iy_lo, iy_hi)),
                                                 form=1 if pol == 'p' else 0)
                if il >= 1:
                    sums['S_' + which] = eng.add_sx(il)
        for mid in sums.values():
            bits |= 1 << mid
        peak, n_stop, n = 0.0, N, n_start
        while n < N:
            e = min(td_next_event(n, n_in_last, frame_steps), N - 1, n + TD_ROWS - 1)
            k = e - n + 1
            tab = new_steps(k)
            if w == 0:
                env = env_t[n:e + 1]
                fill_phasors(tab, omega, dt, np.arange(n, e + 1) * dt, np.where(env > 1e-9, env, 0.0),
                             src_omega=src.omega)
            tab['mask'] = bits
            tab['row'] = np.arange(k)
            if w > 0:
                eng.load_tfsf(tf_h, rec_in[0][n:e + 1])
                eng.load_tfsf(tf_e, rec_in[1][n:e + 1])
            eng.run(tab)
            if rec is not None:
                rec[0][n:e + 1] = eng.read_monitor(m_re)[:k]
                rec[1][n:e + 1] = eng.read_monitor(m_rh)[:k]
            for jj in np.nonzero(n_j == e)[0]:
                F = eng.read_field(e_comp)
                F = _tm_node_field(F) if pol == 'p' else F
                fr_mm[jj, g0 + a_g:g0 + b_g] = F[a_g:b_g]
            if e > n_in_last and (e - n_in_last) % TD_CHECK == 0:
                cur = eng.field_maxabs(e_comp)
                peak = max(peak, cur)
                if cur < 1e-4 * max(peak, 1e-30):
                    n_stop = e
                    break
            if e % TD_CHECK == 0:
                peak = max(peak, eng.field_maxabs(e_comp))
            n = e + 1
        for name, mid in sums.items():
            P[name] += eng.read_monitor(mid)
        eng.close()
        tw = time.time() - t1
        win_info.append(dict(window=w, cells=Nxw * Ny, steps=n_stop - n_start, seconds=tw))
        print(f'[Window {w + 1}/{n_win}] x0={x_start_um:.1f}um  steps {n_start}..{n_stop} '
              f'({n_stop - n_start} of {N})  {tw:.0f}s', flush=True)
        rec_in = rec
    fr_mm.flush()
    del fr_mm
    total_time = time.time() - t_total
    Tr_I = P['P_out'] / max(P['P_in'], 1e-30)
    S_in, S_out = P['S_in'], P['S_out']
    Tr = S_out / S_in if abs(S_in) > 0 else Tr_I
    tr_def = 'poynting' if abs(S_in) > 0 else 'intensity'
    print(f'\n[Result] {label} [time-domain stitch]: total {total_time:.0f}s  Transmission={Tr:.4f}')
    frames = [{'x_start': 0.0, 'x_end': L_um, 'y_start': y0_m * 1e6, 'y_end': (y0_m + H_total) * 1e6,
               'slide_i': 0, 'step': int(k), 'time': float(k * dt)} for k in n_j]
    res = _result_dict(label, P['P_in'], P['P_out'], Tr, frames, total_time, bead_d_um, wg_mat,
                       bead_mat, wg_n, bead_n, wg_thickness_um, lambda_nm, vcsel_waist,
                       vcsel_tilt, vcsel_offset, 'pulse', 'stitch-td', place_bead, n_aq_l,
                       n_mu_l, n_co_l, frames_memmap=mm_path, t_aq_um=t_aq_um,
                       t_mu_um=t_mucin_um, t_lip_um=t_lipid_um, bead_x_um=bx_um,
                       x_wg_start_um=(input_gap_um or 0.0), x_wg_end_um=L_um, polarization=pol,
                       stack=stack, n_lip=n_lip_l)
    res['transmission_intensity'] = Tr_I
    res['transmission_def'] = tr_def
    _vcsel_result(res, (P['S_dn'] - P['S_up']) * dx, -P['S_up'] * dx, _vd, _air, x_facet_um,
                  vcsel_waist, p_in=S_in * dx, energy=True)
    res.update(view_primary='transient', cw_only=False, windows=win_info, n_windows=n_win,
               cell_updates=sum(wi['cells'] * wi['steps'] for wi in win_info),
               pulse=dict(cycles=pulse_cycles, s_fs=sp * 1e15, tmax_fs=t_max * 1e15,
                          dispersion=False))
    return res


def save_results(result, out_dir='results'):
    """PKL summary and frames NPZ of a 2D result (reference fdtd2d_core.py:2085-2240);
    keys and file names as the reference writes them (lens_sensor_amd.common.results)."""
    from lens_sensor_amd.common import results
    return results.save_results(result, out_dir=out_dir,
                                defaults=dict(t_aq=T_AQ, t_mu=T_MU, n_aq=N_AQ, n_mu=N_MUCIN,
                                              n_lip=N_LIPID))


import inspect as _inspect                                                   # noqa: E402
_RUN_BEADS_ARGS = [k for k in _inspect.signature(_run_beads_body).parameters]
