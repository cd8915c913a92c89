"""3D Yee FDTD solver (planar waveguide, tear-film stack, spherical bead) on the
MI355X step engine.

Entry points keep the reference's signatures (planar_3d/fdtd3d_core.py):
  run_3d           :793-1308   full domain, CW steady state, detectors, flux, <|E|^2>
  run_3d_stitched  :1311-1782  x domain decomposition with phasor handoff

Host-side setup (layer stack, sub-pixel sphere, staggered coefficients, source planes)
is NumPy as in the reference; the whole time loop runs in the step engine (HIP library
through lens_sensor_amd.engine), which keeps every field, psi slab and accumulator in HBM.
Per step the host only contributes table entries (envelope, phasors) computed with the
reference's expressions; snapshots and the CW convergence check are chunk boundaries.

Grid (C order, x slowest, z contiguous):
  Ex (Nx-1,Ny,Nz)  Ey (Nx,Ny-1,Nz)  Ez (Nx,Ny,Nz-1)
  Hx (Nx,Ny-1,Nz-1) Hy (Nx-1,Ny,Nz-1) Hz (Nx-1,Ny-1,Nz)
"""
import contextlib
import io
import os
import sys
import time

import numpy as np

_REPO_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if _REPO_ROOT not in sys.path:
    sys.path.insert(0, _REPO_ROOT)

from lens_sensor_amd.common import backend                                    # noqa: E402
from lens_sensor_amd.common.physics import C0, EPS0, MU0, N_AIR               # noqa: E402
from lens_sensor_amd.common.physics import core_source_mask, SRC_CORE_RAMP_UM  # noqa: E402
from lens_sensor_amd.common.subpixel import sphere_fraction, blend, cell_average_column  # noqa: E402
from lens_sensor_amd.common.flux import PlaneFluxBook, FluxAccumulator, guard_in_plane  # noqa: E402
from lens_sensor_amd.common import vcsel                                      # noqa: E402
from lens_sensor_amd.common.layer_stack import (stack_total_um, legacy_stack,  # noqa: E402
                                                legacy_layers_tuple, stack_at_lambda,
                                                describe as describe_stack)
from lens_sensor_amd.engine.spec import new_steps, fill_phasors, cw_envelope, pulse_envelope  # noqa: E402
from lens_sensor_amd.planar_3d.boundaries3d import (make_boundary_3d, ALL_FACES,  # noqa: E402
                                                    mur_speeds_3d, Sponge3D)

BEAD_SUBPIXEL = True
KAPPA = {}                 # extinction per layer key (wg, lip, aq, mu, co, bead)
SRC_CORE_ONLY = False
LASER_POWER_MW = vcsel.VCSEL_DEFAULTS['power_mw']
LAM_LOSS = 850e-9
N_G_FACTOR = 1.10          # group-index margin on the run time (fdtd3d_core.py:733)
CW_TOL = 0.01              # phasor change between two 8-period windows
CW_MAX_EXTEND = 2.0
NPML = 10                  # reference's CPML thickness (make_boundary_3d default)


def _stack_lossy(stack=None):
    if any(v > 0 for v in (KAPPA or {}).values()):
        return True
    return bool(stack) and any(complex(L['n']).imag > 0 for L in stack)


def _stack_n_below(stack, n_sub):
    return float(complex(stack[0]['n']).real) if stack else float(n_sub)


def _result_stack(stack, t_lip_um, t_aq_um, t_mu_um, n_lip, n_aq, n_mu):
    if stack is not None:
        return [dict(L) for L in stack], True, legacy_layers_tuple(stack)
    kp = dict(KAPPA or {})
    st = legacy_stack(t_lip_um, complex(float(np.real(n_lip)), kp.get('lip', 0.0)),
                      t_aq_um, complex(float(np.real(n_aq)), kp.get('aq', 0.0)),
                      t_mu_um, complex(float(np.real(n_mu)), kp.get('mu', 0.0)))
    return st, False, (t_lip_um * 1e-6, t_aq_um * 1e-6, t_mu_um * 1e-6, n_aq)


# ----------------------------------------------------------------------------- materials
def build_eps3d(Nx, Ny, Nz, dx_um, y0_um, z0_um, wg_n, t_wg_um, t_lip_um, t_aq_um, t_mu_um,
                n_lip, n_aq, n_mu, n_co, bead=None, wg_width_um=None, n_clad_side=N_AIR,
                end_facet_um=0.0, x_wg_start_um=0.0, x0_um=None, x_wg_end_um=None,
                tensor_patch=None, stack=None):
    """eps_r on (Nx,Ny,Nz): layer stack along y (16-sample cell averages), optional
    rectangular channel in z, facets in x, spherical bead with volume-fraction blending.
    Reference: fdtd3d_core.py:119-292.  ``x0_um`` set = global-x (stitch) mode in which
    the core exists only on [x_wg_start, x_wg_end].  ``tensor_patch`` receives the data
    of the anisotropic bead/layer averaging for ``ce_arrays``."""
    global_x = x0_um is not None
    xs_g = (float(x0_um) if global_x else 0.0) + np.arange(Nx) * dx_um
    kp = dict(KAPPA or {})
    lossy = _stack_lossy(stack)
    k_co = kp.get('co', 0.0) if lossy else 0.0
    if lossy:
        wg_n = complex(wg_n, kp.get('wg', 0.0))
        n_lip = complex(n_lip, kp.get('lip', 0.0))
        n_aq = complex(n_aq, kp.get('aq', 0.0))
        n_mu = complex(n_mu, kp.get('mu', 0.0))
        if bead is not None:
            bead = dict(bead, n_c=complex(bead['n'], kp.get('bead', 0.0)))
    edt = np.complex128 if lossy else np.float64
    ys = y0_um + np.arange(Ny) * dx_um
    y_lip_bot = -t_lip_um
    y_aq_bot = y_lip_bot - t_aq_um
    y_mu_bot = y_aq_bot - t_mu_um
    bounds = None
    if stack is not None:
        bounds, yt = [], 0.0
        for L in stack:
            yb = yt - float(L['t_um'])
            nL = complex(L['n'])
            bounds.append((yt, yb, nL if lossy else nL.real))
            yt = yb
        y_mu_bot = yt

    def eps_y(y, core, n_core=None):
        e = np.full(y.shape, N_AIR ** 2, dtype=edt)
        below = y < y_mu_bot
        if hasattr(n_co, 'at_depth'):
            e[below] = ((n_co.at_depth(y_mu_bot - y[below]) + 1j * k_co) ** 2 if lossy
                        else n_co.at_depth(y_mu_bot - y[below]) ** 2)
        else:
            e[below] = (float(n_co) + 1j * k_co) ** 2 if lossy else n_co ** 2
        if bounds is not None:
            for yt_, yb_, n_ in bounds:
                e[(y >= yb_) & (y < yt_)] = n_ ** 2
        else:
            e[(y >= y_mu_bot) & (y < y_aq_bot)] = n_mu ** 2
            e[(y >= y_aq_bot) & (y < y_lip_bot)] = n_aq ** 2
            if t_lip_um > 1e-6:
                e[(y >= y_lip_bot) & (y < 0.0)] = n_lip ** 2
        if core:
            e[(y >= 0.0) & (y < t_wg_um)] = (wg_n if n_core is None else n_core) ** 2
        return e

    col_out = cell_average_column(ys, dx_um, lambda y: eps_y(y, False))
    col_core = cell_average_column(ys, dx_um, lambda y: eps_y(y, True))
    col = col_out if global_x else col_core
    eps = np.broadcast_to(col[None, :, None], (Nx, Ny, Nz)).copy()
    if global_x:
        in_wg_x = xs_g >= (x_wg_start_um or 0.0)
        if x_wg_end_um is not None:
            in_wg_x &= xs_g <= x_wg_end_um
        if in_wg_x.any():
            eps[in_wg_x, :, :] = col_core[None, :, None]
    lz_um = Nz * dx_um
    channel = wg_width_um is not None and 0 < wg_width_um < lz_um
    if channel:
        zs = z0_um + np.arange(Nz) * dx_um
        core_y = (ys >= 0.0) & (ys <= t_wg_um)
        outside_z = np.abs(zs) > (wg_width_um / 2.0)
        yz = core_y[None, :, None] & outside_z[None, None, :]
        if global_x:
            inx = xs_g >= (x_wg_start_um or 0.0)
            if x_wg_end_um is not None:
                inx &= xs_g <= x_wg_end_um
            eps[np.broadcast_to(inx[:, None, None] & yz, eps.shape)] = n_clad_side ** 2
        else:
            eps[np.broadcast_to(yz, eps.shape)] = n_clad_side ** 2
    if not global_x:
        xs_loc = np.arange(Nx) * dx_um
        lx_um = Nx * dx_um
        if end_facet_um and end_facet_um > 0:
            eps[xs_loc >= (lx_um - end_facet_um), :, :] = N_AIR ** 2
        if x_wg_start_um and x_wg_start_um > 0:
            eps[xs_loc < x_wg_start_um, :, :] = N_AIR ** 2
    if tensor_patch is not None:
        # Ey at (i, j+1/2, k) is normal to the layers: exact harmonic mean per (x,z)
        # column of one of four stack types (fdtd3d_core.py:223-246)
        def hcol(core, n_core=None):
            return cell_average_column(ys[:-1], dx_um, lambda y: np.real(eps_y(y, core, n_core)),
                                       harmonic=True, shift=0.5 * dx_um).astype(np.float64)
        T = np.full((Nx, Nz), 0 if global_x else 1, dtype=np.int8)
        if global_x and in_wg_x.any():
            T[in_wg_x, :] = 1
        if channel:
            rows = inx if global_x else np.ones(Nx, dtype=bool)
            T[np.ix_(rows, outside_z)] = 2
        if not global_x:
            if end_facet_um and end_facet_um > 0:
                T[xs_loc >= (lx_um - end_facet_um), :] = 3
            if x_wg_start_um and x_wg_start_um > 0:
                T[xs_loc < x_wg_start_um, :] = 3
        tensor_patch['layer'] = dict(T=T, cols=np.stack([hcol(False), hcol(True),
                                                         hcol(True, n_clad_side),
                                                         np.full(Ny - 1, N_AIR ** 2)]))
    for bd in _bead_list(bead):
        r = bd['d_um'] / 2.0
        xc = bd['x_um']
        yc = bd.get('y_um', -r)
        zc = bd.get('z_um', 0.0)
        x0 = float(x0_um) if global_x else 0.0
        xs = x0 + np.arange(Nx) * dx_um
        zs_ = z0_um + np.arange(Nz) * dx_um
        i0 = max(0, int((xc - r - x0) / dx_um) - 2)
        i1 = min(Nx, int((xc + r - x0) / dx_um) + 3)
        j0 = max(0, int((yc - r - y0_um) / dx_um) - 2)
        j1 = min(Ny, int((yc + r - y0_um) / dx_um) + 3)
        k0 = max(0, int((zc - r - z0_um) / dx_um) - 2)
        k1 = min(Nz, int((zc + r - z0_um) / dx_um) + 3)
        if not (i1 > i0 and j1 > j0 and k1 > k0):
            continue
        sub = eps[i0:i1, j0:j1, k0:k1]
        if BEAD_SUBPIXEL and tensor_patch is not None:
            tensor_patch.setdefault('beads', []).append(dict(
                eps_bg=sub.real.astype(np.float64), i0=i0, j0=j0, k0=k0, xs=xs[i0:i1],
                ys=ys[j0:j1], zs=zs_[k0:k1], xc=xc, yc=yc, zc=zc, r=r, n=bd['n']))
        if BEAD_SUBPIXEL:
            blend(sub, sphere_fraction(xs[i0:i1], ys[j0:j1], zs_[k0:k1], xc, yc, zc, r, dx_um),
                  bd.get('n_c', bd['n']))
        else:
            XX = xs[i0:i1, None, None] - xc
            YY = ys[None, j0:j1, None] - yc
            ZZ = zs_[None, None, k0:k1] - zc
            sub[(XX * XX + YY * YY + ZZ * ZZ) <= r * r] = bd.get('n_c', bd['n']) ** 2
    return eps


def _bead_list(bead):
    """One bead dict (reference) or a list of beads (several per window)."""
    if bead is None:
        return []
    beads = bead if isinstance(bead, (list, tuple)) else [bead]
    return [b for b in beads if b is not None and b.get('d_um', 0) > 0]


def split_eps(eps):
    """eps (real or complex) -> (eps' float32, eps'' float32 or None); fdtd3d_core.py:585-590."""
    eps = np.asarray(eps)
    if np.iscomplexobj(eps):
        return eps.real.astype(np.float32), eps.imag.astype(np.float32)
    return eps, None


def inv_loss(eps, eim, dt, dx, lam_m):
    """inv = dt/(eps0 eps' dx) and D = (1-a)/(1+a), a = w0 dt eps''/(2 eps');
    inv is scaled by 1/(1+a) when lossy.  fdtd3d_core.py:593-603."""
    e = np.asarray(eps)
    inv = np.float32(dt / (EPS0 * dx)) / e
    if eim is None or not np.any(eim > 0):
        return inv, None
    a = (np.float32(np.pi * C0 / lam_m * dt) * np.asarray(eim)) / e
    inv /= (1 + a)
    return inv, ((1 - a) / (1 + a)).astype(np.float32)


def staggered(inv, ax):
    """Node value -> E-component position (+dx/2 along ax): mean of the two nodes
    (harmonic eps along ax).  fdtd3d_core.py:295-302."""
    a = [slice(None)] * 3
    b = [slice(None)] * 3
    a[ax] = slice(None, -1)
    b[ax] = slice(1, None)
    return (np.float32(0.5) * (inv[tuple(a)] + inv[tuple(b)])).astype(np.float32)


def ce_arrays(inv, tp, dt, dx, dx_um, da=None):
    """E update coefficients at the staggered Yee positions (fdtd3d_core.py:305-378):
    exact harmonic layer average for Ey (per column type), and at each bead surface the
    diagonal Farjadpour tensor (1/eps)_ii = n_i^2 <1/eps> + (1 - n_i^2)/<eps>."""
    ces = [staggered(inv, 0), staggered(inv, 1), staggered(inv, 2)]
    tp = tp or {}
    lp = tp.get('layer')
    if lp is not None:
        cy = ces[1]
        hinv = (dt / (EPS0 * dx) / lp['cols']).astype(np.float32)
        T = lp['T'][:, None, :]
        for t in range(hinv.shape[0]):
            np.copyto(cy, hinv[t][None, :, None], where=(T == t))
        if da is not None:
            cy *= staggered(np.float32(0.5) * (1 + da), 1)
    c = dt / (EPS0 * dx)
    for bp in tp.get('beads', []):
        eb = bp['n'] ** 2
        for ax in range(3):
            ce = ces[ax]
            e0 = bp['eps_bg']
            nb = e0.shape[ax]
            e_nx = np.take(e0, np.minimum(np.arange(nb) + 1, nb - 1), axis=ax)
            e_bg_a = 0.5 * (e0 + e_nx)
            e_bg_h = 1.0 / (0.5 * (1.0 / e0 + 1.0 / e_nx))
            if ax == 1 and lp is not None:
                i0_, j0_, k0_ = bp['i0'], bp['j0'], bp['k0']
                Tb = lp['T'][i0_:i0_ + e0.shape[0], k0_:k0_ + e0.shape[2]]
                cols = lp['cols'][:, j0_:j0_ + e0.shape[1]]
                hb = np.transpose(cols[Tb], (0, 2, 1))
                e_bg_h[:, :hb.shape[1], :] = hb
            fac = staggered(np.float32(0.5) * (1 + da), ax) if da is not None else None
            o = [0.0, 0.0, 0.0]
            o[ax] = 0.5 * dx_um
            xs, ys, zs = bp['xs'] + o[0], bp['ys'] + o[1], bp['zs'] + o[2]
            a0, b0, c0 = bp['i0'], bp['j0'], bp['k0']
            na = min(len(xs), ce.shape[0] - a0)
            nbb = min(len(ys), ce.shape[1] - b0)
            nc = min(len(zs), ce.shape[2] - c0)
            if na <= 0 or nbb <= 0 or nc <= 0:
                continue
            xs, ys, zs = xs[:na], ys[:nbb], zs[:nc]
            f = sphere_fraction(xs, ys, zs, bp['xc'], bp['yc'], bp['zc'], bp['r'], dx_um)
            m = f > 0
            if not m.any():
                continue
            X = xs[:, None, None] - bp['xc']
            Y = ys[None, :, None] - bp['yc']
            Z = zs[None, None, :] - bp['zc']
            d = np.maximum(np.sqrt(X * X + Y * Y + Z * Z), 1e-30)
            nc2 = ((X, Y, Z)[ax] / d) ** 2
            ebh = e_bg_h[:na, :nbb, :nc]
            eba = e_bg_a[:na, :nbb, :nc]
            inv_t = nc2 * ((1 - f) / ebh + f / eb) + (1 - nc2) / ((1 - f) * eba + f * eb)
            sl = (slice(a0, a0 + na), slice(b0, b0 + nbb), slice(c0, c0 + nc))
            blk = ce[sl].astype(np.float64)
            val = c * np.broadcast_to(inv_t, f.shape)
            if fac is not None:
                val = val * fac[sl].astype(np.float64)
            blk[m] = val[m]
            ce[sl] = blk.astype(np.float32)
    return tuple(ces)


def stagger_da(da):
    if da is None:
        return None
    return tuple(staggered(da, ax) for ax in range(3))


def _sponge_cells(n_sponge, sponge_um, dx_um, shape=None, warn=True):
    """Absorber thickness in cells (fdtd3d_core.py:381-394); also places the source."""
    n = int(n_sponge) if n_sponge else int(np.ceil(float(sponge_um) / dx_um))
    n = max(8, n)
    if shape is not None:
        cap = max(8, min(s // 6 for s in shape))
        if n > cap and warn:
            print(f'  [WARN] Sponge {n} cells capped to {cap} (grid too small)')
        n = min(n, cap)
    return n


def _abs_cells(bkind, npml, n_sponge, sponge_um, dx_um):
    """Outer absorber thickness in cells for the stack sizing (fdtd3d_core.py:852-854):
    the CPML's cells, none for Mur, the sponge's cells."""
    if bkind == 'cpml':
        return npml
    return 0 if bkind == 'mur1' else _sponge_cells(n_sponge, sponge_um, dx_um)


def _resolve_tear_um(tear_um, t_lip_um, t_aq_um, t_mu_um, n_abs, dx_um, cornea_um=1.5, stack=None):
    if tear_um is not None:
        return float(tear_um)
    absorber = n_abs * dx_um
    layers = stack_total_um(stack) if stack is not None else t_lip_um + t_aq_um + t_mu_um
    tot = layers + cornea_um + absorber
    print(f'  [Auto] tear = {tot:.2f}um (layers {layers:.2f} + cornea {cornea_um:g} + '
          f'absorber {absorber:.2f}um)')
    return tot


def _resolve_pol(polarization, Ny, Nz):
    pol = 'p' if str(polarization).lower() in ('p', 'tm') else 's'
    Nj, Nk = (Ny - 1, Nz) if pol == 'p' else (Ny, Nz - 1)
    return pol, Nj, Nk


def _build_gauss_source(n_sp, t_wg_um, y0_um, z0_um, dx_um, dx, LAM, Nj, Nk, waist_y, waist_z,
                        off_y, off_z, tilt_deg, center_y_um=None, center_z_um=None,
                        n_src_fn=None, pol='s', wg_width_um=None, dt=None, z_prop_um=0.0):
    """Gaussian VCSEL spot as a one-sided Huygens pair on the driven component at
    x = n_sp + 6 (fdtd3d_core.py:494-530)."""
    src_ix = n_sp + 6
    cy = t_wg_um / 2 if center_y_um is None else center_y_um
    cz = 0.0 if center_z_um is None else center_z_um
    yc, zc = cy + off_y, cz + off_z
    j_src = min(max(int(round((yc - y0_um) / dx_um)), 0), Nj - 1)
    k_src = min(max(int(round((zc - z0_um) / dx_um)), 0), Nk - 1)
    yv = y0_um + (np.arange(Nj) + (0.5 if pol == 'p' else 0.0)) * dx_um
    zv = z0_um + (np.arange(Nk) + (0.5 if pol == 's' else 0.0)) * dx_um
    n_here = float(n_src_fn(j_src, k_src)) if n_src_fn is not None else 1.0
    lam_nm = LAM * 1e9
    amp, ph = vcsel.profile_2d(yv, zv, yc, zc, z_prop_um, waist_y, waist_z, lam_nm, n_here, tilt_deg)
    if SRC_CORE_ONLY and center_y_um is None:
        amp = amp * core_source_mask(yv[:, None], t_wg_um, z_um=zv[None, :], wg_width_um=wg_width_um)
    kdx = vcsel.numerical_kx_dx(lam_nm, dx, C0 * dt, n_here, tilt_deg)
    src = vcsel.HuygensSource(src_ix, amp, ph, 2 * np.pi * C0 / LAM, kdx)
    return src_ix, j_src, k_src, src


def _vcsel_geom_3d(n_sp, dx_um, distance_um):
    x_src_um = (n_sp + 6) * dx_um
    air = vcsel.air_path_um(distance_um, dx_um)
    return x_src_um, x_src_um + air, air, vcsel.z_at_source(distance_um, air)


def _n_source_fn_3d(x_src_um, y0_um, dx_um, wg_n, t_wg_um, n_below, x_wg_start_um, pol):
    def fn(j, k):
        if x_wg_start_um and x_src_um < x_wg_start_um:
            return float(N_AIR)
        y = y0_um + (j + (0.5 if pol == 'p' else 0.0)) * dx_um
        if 0.0 <= y < t_wg_um:
            return float(np.real(wg_n))
        return float(N_AIR) if y >= t_wg_um else float(np.real(n_below))
    return fn


def _vcsel_res(p_emit, p_refl, p_in, distance_um, air_um, x_facet_um, waist_um, waist_z_um,
               energy=False):
    """VCSEL figures (reference fdtd3d_core.py:547-561); energy=True: pulse energies (J),
    no source power stored."""
    r = dict(laser_power_mw=LASER_POWER_MW, vcsel_distance_um=float(distance_um or 0.0),
             vcsel_air_um=float(air_um), x_facet_um=float(x_facet_um),
             vcsel_waist_um=float(waist_um), vcsel_waist_z_um=float(waist_z_um),
             p_src_si=None if energy else p_emit)
    if p_emit and p_emit > 0:
        eta = (p_in / p_emit) if p_in is not None else float('nan')
        refl = (p_refl / p_emit) if p_refl is not None else float('nan')
        r.update(coupling_efficiency=eta, facet_reflection=refl)
    return r


def _engine_mode(mode):
    return mode or backend.DEFAULT_MODE


# ----------------------------------------------------------------------------- run_3d
def run_3d(label, wg_n, t_wg_um=5.0, t_lip_um=0.0, t_aq_um=3.5, t_mu_um=0.5,
           n_lip=1.480, n_aq=1.336, n_mu=1.342, n_co=1.376, lam_nm=850.0,
           lx_um=40.0, air_um=3.0, tear_um=None, lz_um=8.0, dx_nm=50.0,
           vcsel_waist_um=3.0, vcsel_waist_z_um=3.0, source_type='cw', n_snapshots=8,
           steps_factor=2.0, bead=None, n_sponge=None, sponge_um=1.2, sponge_alpha=0.3,
           det_in_um=None, det_out_um=None, vol_dtype='float32', calibrate_steps=0,
           wg_width_um=None, n_clad_side=N_AIR, save_vector=False, check_resources=False,
           pec_faces=(), end_facet_um=0.0, vcsel_tilt_deg=0.0, vcsel_offset_y_um=0.0,
           vcsel_offset_z_um=0.0, input_gap_um=0.0, polarization='s', allow_large=False,
           boundary='cpml', cad=None, pulse_cycles=30.0, stack=None,
           npml=None, engine_mode=None, steps_override=None, cw_volume=True):
    """Full-domain 3D run (reference fdtd3d_core.py:793-1308).  Extra keywords of this
    build: ``npml`` (CPML cells, reference fixed 10), ``engine_mode`` ('ref32' for the
    reference's float32 semantics, 'fp64acc' for FP64 phase/accumulators),
    ``steps_override`` (run exactly this many steps, no CW extension), ``cw_volume``
    (False: CW phasor only on the z-centre monitor plane, no volume accumulator)."""
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    if save_vector:
        raise NotImplementedError('vector snapshots are not part of the device path')
    t_total = time.time()
    dx = dx_nm * 1e-9
    dx_um = dx_nm / 1000.0
    dt = 0.5 * dx / (C0 * np.sqrt(3))
    global LAM_LOSS
    LAM = lam_nm * 1e-9
    LAM_LOSS = LAM
    f0 = C0 / LAM
    npml = NPML if npml is None else int(npml)
    bkind = str(boundary or 'cpml').lower()
    if bkind not in ('cpml', 'sponge', 'mur1'):
        raise ValueError(f"unknown boundary '{bkind}' (sponge|cpml|mur1)")
    tear_um = _resolve_tear_um(tear_um, t_lip_um, t_aq_um, t_mu_um, _abs_cells(bkind, npml, n_sponge,
                                                                             sponge_um, dx_um),
                               dx_um, stack=stack)
    ly_um = t_wg_um + air_um + tear_um
    y0_um = -tear_um
    z0_um = -lz_um / 2.0
    Nx, Ny, Nz = int(round(lx_um / dx_um)), int(round(ly_um / dx_um)), int(round(lz_um / dx_um))
    n_sp = _sponge_cells(n_sponge, sponge_um, dx_um, (Nx, Ny, Nz), warn=bkind == 'sponge')
    cells = Nx * Ny * Nz
    print(f'\n===== 3D-FDTD: {label} =====  {Nx} x {Ny} x {Nz} ({cells / 1e6:.1f}M cells), '
          f'dx={dx_nm}nm, dt={dt * 1e15:.4f} fs, engine {backend.engine_name()}')
    pec_faces = tuple(pec_faces or ())
    absorbing = tuple(f for f in ALL_FACES if f not in pec_faces)
    bnd = make_boundary_3d(boundary, dx, dt, absorbing, shape=(Nx, Ny, Nz), npml=npml, lam_m=LAM,
                           pec_faces=pec_faces)
    _vd = float(input_gap_um or 0.0)
    x_src_um, x_wg_start_um, _air, _zsrc = _vcsel_geom_3d(n_sp, dx_um, _vd)
    tp = {}
    eps = build_eps3d(Nx, Ny, Nz, dx_um, y0_um, z0_um, wg_n, t_wg_um, t_lip_um, t_aq_um, t_mu_um,
                      n_lip, n_aq, n_mu, n_co, bead, wg_width_um=wg_width_um,
                      n_clad_side=n_clad_side, end_facet_um=end_facet_um,
                      x_wg_start_um=x_wg_start_um, tensor_patch=tp, stack=stack)
    eps, eim = split_eps(eps)
    if bkind == 'mur1':
        # Mur with the local speeds (x faces n_eff, y/z faces n of the cells),
        # fdtd3d_core.py:940-947
        bnd.set_speeds(mur_speeds_3d(eps, dx, LAM))
    inv, da = inv_loss(eps, eim, dt, dx, LAM)
    del eps, eim
    cex, cey, cez = ce_arrays(inv, tp, dt, dx, dx_um, da=da)
    da = stagger_da(da)
    del inv
    Ch = np.float32(dt / (MU0 * dx))
    pol, Nj, Nk = _resolve_pol(polarization, Ny, Nz)
    src_comp = 'Ey' if pol == 'p' else 'Ez'
    src_ix, j_src, k_src, vs = _build_gauss_source(
        n_sp, t_wg_um, y0_um, z0_um, dx_um, dx, LAM, Nj, Nk, vcsel_waist_um, vcsel_waist_z_um,
        vcsel_offset_y_um, vcsel_offset_z_um, vcsel_tilt_deg,
        n_src_fn=_n_source_fn_3d(x_src_um, y0_um, dx_um, wg_n, t_wg_um,
                                 n_aq if stack is None else _stack_n_below(stack, n_co),
                                 x_wg_start_um, pol),
        pol=pol, wg_width_um=wg_width_um, dt=dt, z_prop_um=_zsrc)

    eng = backend.make_engine('yee3d', (Nx, Ny, Nz), _engine_mode(engine_mode))
    eng.set_coeffs(Ch, cex=cex, cey=cey, cez=cez,
                   dax=None if da is None else da[0], day=None if da is None else da[1],
                   daz=None if da is None else da[2])
    del cex, cey, cez
    if bkind == 'cpml':
        bnd.attach(eng)
    elif bkind == 'mur1':
        bnd.attach(eng, (Nx, Ny, Nz), pec_faces=pec_faces)
    else:                                    # graded sponge (fdtd3d_core.py:999, 1133-1135)
        Sponge3D(n_sp, sponge_alpha, absorbing).attach(eng, (Nx, Ny, Nz), dx, dt,
                                                      pec_faces=pec_faces)
    eng.finalize()
    eng.set_source(src_comp, src_ix, vs.amp, vs.ph1, vs.ph2)

    if det_in_um is None:
        det_in_um = 0.2 * lx_um
    if det_out_um is None:
        det_out_um = 0.8 * lx_um
    j_wg0 = int(round((0.0 - y0_um) / dx_um))
    j_wg1 = int(round((t_wg_um - y0_um) / dx_um))
    hw = max(2, int(round(1.0 / dx_um)))

    def xrange_at(xu):
        c = int(round(xu / dx_um))
        return max(0, c - hw), min(Nx, c + hw)
    iin0, iin1 = xrange_at(det_in_um)
    iout0, iout1 = xrange_at(det_out_um)
    ci = (iin0 + iin1) // 2
    sh = guard_in_plane(ci, max(src_ix, int(round(x_wg_start_um / dx_um))), Nx) - ci
    iin0, iin1 = iin0 + sh, min(Nx, iin1 + sh)
    det_in_um = 0.5 * (iin0 + iin1) * dx_um

    omega = 2 * np.pi * f0
    sigma_t = 4 / (2 * np.pi * f0)
    steps_total = int((24.0 / (2 * np.pi * f0) + steps_factor * lx_um * 1e-6 * N_G_FACTOR * wg_n / C0) / dt)
    if steps_override:
        steps_total = int(steps_override)
    steps_max = int(CW_MAX_EXTEND * steps_total)
    snap_every = max(1, steps_total // max(1, n_snapshots))
    n_avg0 = steps_total // 2
    steps_per_T = max(4, int(round((1.0 / f0) / dt)))
    n_acc_cw = min(steps_total, 8 * steps_per_T)
    kc = Nk // 2
    conv = (source_type != 'pulse') and not calibrate_steps and not steps_override
    m_I = max(1, steps_per_T // 8)
    cw_rel_change = float('nan')

    acc_full = 'c128' if eng.wide_acc else 'c64p'
    zs = (n_sp, -n_sp)
    mon_pin = eng.add_box(((iin0, iin1), (j_wg0, j_wg1), zs))
    mon_pout = eng.add_box(((iout0, iout1), (j_wg0, j_wg1), zs))
    mon_I = eng.add_sq('3d_p' if pol == 'p' else '3d_s', Nx)
    mon_cw = (eng.add_dft(src_comp, 0, Nx, acc=acc_full) if cw_volume
              else eng.add_dft(src_comp, 0, Nx, kplane=kc, acc=acc_full))
    mon_pa = eng.add_dft(src_comp, 0, Nx, kplane=kc, acc=acc_full) if conv else None
    mon_pb = eng.add_dft(src_comp, 0, Nx, kplane=kc, acc=acc_full) if conv else None
    book = PlaneFluxBook(eng, 'yee3d')
    for key, il in (('in', max(1, (iin0 + iin1) // 2)), ('out', max(1, (iout0 + iout1) // 2)),
                    ('src_up', vs.plane_up), ('src_dn', vs.plane_dn)):
        book.add_plane(key, il)
    flux_bits = sum(1 << book.mon[k] for k in book.keys())

    n_snaps = 0 if calibrate_steps else len([n for n in range(steps_total) if (n + 1) % snap_every == 0])
    vol_dt = np.float16 if vol_dtype == 'float16' else np.float32
    vol_buf = np.zeros((n_snaps, Nx, Nj, Nk), dtype=vol_dt) if n_snaps else None
    snap_steps, snap_times = [], []
    ksnap = 0
    nacc_cw = n_pl_a = n_pl_b = n_I = 0
    t_start = time.time()
    if calibrate_steps:
        warmup = 5
        n_run = min(steps_total, warmup + calibrate_steps)
    else:
        n_run = steps_total

    row = new_steps(1)
    fill_phasors(row, omega, dt, np.array([n_avg0 * dt]), np.array([1.0]), src_omega=vs.omega)
    backend.tune(eng, row[0], n_run)

    def masks(n0, n1):
        nonlocal nacc_cw, n_pl_a, n_pl_b, n_I
        ns = np.arange(n0, n1)
        mk = np.full(len(ns), (1 << mon_pin) | (1 << mon_pout), dtype=np.uint32)
        sel_I = (ns >= n_avg0) & (((ns - n_avg0) % m_I) == 0)
        mk[sel_I] |= np.uint32(1 << mon_I)
        in_cw = ns >= steps_total - n_acc_cw
        mk[in_cw] |= np.uint32((1 << mon_cw) | flux_bits | ((1 << mon_pb) if conv else 0))
        cnt_cw = int(in_cw.sum())
        if conv:
            in_a = (~in_cw) & (ns >= steps_total - 2 * n_acc_cw)
            mk[in_a] |= np.uint32(1 << mon_pa)
            n_pl_a += int(in_a.sum())
            n_pl_b += cnt_cw
        nacc_cw += cnt_cw
        n_I += int(sel_I.sum())
        book.note_samples(book.keys(), cnt_cw)
        return mk

    def run_until(n0, n1):
        tab = new_steps(n1 - n0)
        t_phys = np.arange(n0, n1) * dt
        env = (pulse_envelope(t_phys, f0, pulse_cycles) if source_type == 'pulse'
               else cw_envelope(t_phys, sigma_t))
        fill_phasors(tab, omega, dt, t_phys, env, src_omega=vs.omega)
        tab['mask'] = masks(n0, n1)
        eng.run(tab)

    t_cal0 = None
    n = 0
    while n < n_run:
        if calibrate_steps and n == 0:
            run_until(0, min(warmup, n_run))
            eng.sync()
            t_cal0 = time.time()
            n = min(warmup, n_run)
            continue
        # next event: snapshot boundary or end of the current run
        stop = n_run
        if vol_buf is not None and ksnap < n_snaps:
            nxt = ((n // snap_every) + 1) * snap_every      # step index n' with (n'+1) % every == 0 -> n'+1
            stop = min(stop, nxt)
        run_until(n, stop)
        n = stop
        if vol_buf is not None and ksnap < n_snaps and n % snap_every == 0:
            vol_buf[ksnap] = eng.read_field(src_comp).astype(vol_dt)
            snap_steps.append(n)
            snap_times.append(time.time() - t_start)
            ksnap += 1
        if conv and n == steps_total and n_pl_a > 0 and n_pl_b > 0:
            if eng.wide_acc:                     # FP64 windows: reduced on the device
                num, den = eng.monitor_change(mon_pb, mon_pa, n_pl_b, n_pl_a)
                cw_rel_change = float(np.sqrt(num) / max(float(np.sqrt(den)), 1e-30))
            else:                                # the reference's arithmetic on the host
                a = eng.read_monitor(mon_pb) / n_pl_b
                b = eng.read_monitor(mon_pa) / n_pl_a
                cw_rel_change = float(np.linalg.norm(a - b) / max(float(np.linalg.norm(a)), 1e-30))
            if cw_rel_change > CW_TOL and steps_total + n_acc_cw <= steps_max:
                print(f'  [CW] not yet in steady state (change {100 * cw_rel_change:.2f}%) '
                      f'-> +{n_acc_cw} steps')
                mon_pa, mon_pb = mon_pb, mon_pa
                n_pl_a, n_pl_b = n_pl_b, 0
                eng.zero_monitor(mon_pb)
                eng.zero_monitor(mon_cw)
                nacc_cw = 0
                book.reset()
                steps_total += n_acc_cw
                n_run += n_acc_cw
            else:
                print(f'  [CW] phasor change over the last 2 DFT windows (z centre): '
                      f'{100 * cw_rel_change:.3f}%')

    if calibrate_steps:
        eng.sync()
        n_meas = n_run - warmup
        dt_cal = time.time() - t_cal0
        rate = cells * n_meas / max(dt_cal, 1e-9)
        print(f'[CALIBRATION] {n_meas} steps in {dt_cal:.3f}s -> {rate / 1e6:.0f} Mcell/s')
        eng.close()
        return dict(scenario=label, dim='3D', calibrated=True, per_step_s=dt_cal / max(1, n_meas),
                    cell_update_rate=rate, steps_total=steps_total,
                    est_full_s=dt_cal / max(1, n_meas) * steps_total, cells=cells, dx_nm=dx_nm)

    Iavg_np = eng.read_monitor(mon_I) / max(1, n_I)
    P_in = float(eng.read_monitor(mon_pin)) * dx ** 3
    P_out = float(eng.read_monitor(mon_pout)) * dx ** 3
    Tr_I = P_out / max(P_in, 1e-30)
    Tr, tr_def, p_si = Tr_I, 'intensity', None
    if nacc_cw > 0 and source_type != 'pulse' and book.has('in') and book.has('out'):
        S = [book.flux('in'), book.flux('out')]
        p_si = S[0] * dx * dx
        if abs(S[0]) > 0:
            Tr, tr_def = S[1] / S[0], 'poynting'
    cw_vol = None
    Ecw = None
    if nacc_cw > 0:
        Ecw = (2.0 / nacc_cw) * eng.read_monitor(mon_cw)
    if Ecw is not None and cw_volume and not calibrate_steps:
        # steady-state volume set of the driven component: Re(E_cw e^{i w t_k}) at the
        # snapshot count of phases over one period (reference fdtd3d_core.py:1268-1280)
        nfr_cw = max(1, ksnap if ksnap else n_snapshots)
        cw_vol = np.empty((nfr_cw,) + Ecw.shape, dtype=vol_dt)
        for k in range(nfr_cw):
            tk = (k / nfr_cw) * (1.0 / f0)
            cw_vol[k] = np.real(Ecw * np.exp(1j * omega * tk)).astype(vol_dt)
    cw_ok = source_type != 'pulse' and book.has('src_up') and book.has('src_dn')
    p_emit = ((book.flux('src_dn') - book.flux('src_up')) * dx * dx) if cw_ok else None
    p_refl = (-book.flux('src_up') * dx * dx) if cw_ok else None
    eng.close()
    total_time = time.time() - t_total
    print(f'[Result] {label} [3D]: total {total_time:.1f}s  P_in={P_in:.3e}  P_out={P_out:.3e}  '
          f'Transmission={Tr:.4f} ({tr_def})')
    rs = _result_stack(stack, t_lip_um, t_aq_um, t_mu_um, n_lip, n_aq, n_mu)
    vres = _vcsel_res(p_emit, p_refl, p_si, _vd, _air, x_wg_start_um, vcsel_waist_um,
                      vcsel_waist_z_um)
    return dict(**vres, scenario=label, dim='3D', P_in=P_in, P_out=P_out, p_in_si=p_si,
                p_in_unit='W', transmission=Tr, vols_array=(vol_buf[:ksnap] if vol_buf is not None
                                                            else None),
                vols_ex=None, vols_ey=None, vector=False, snap_steps=snap_steps,
                snap_times=snap_times, Iavg=Iavg_np, Ecw=Ecw, cw_vol=cw_vol,
                view_primary='transient', f0=f0, transmission_intensity=Tr_I,
                transmission_def=tr_def, cw_rel_change=cw_rel_change, total_time=total_time,
                steps_total=steps_total, dt=dt, dx_nm=dx_nm, lx_um=lx_um, ly_um=ly_um,
                lz_um=lz_um, y0_um=y0_um, z0_um=z0_um, t_wg_um=t_wg_um, wg_n=wg_n,
                layers=rs[2], stack=rs[0], stack_custom=rs[1], n_lip=n_lip, n_aq=n_aq,
                n_mu=n_mu, n_co=n_co, lam_nm=lam_nm, source_type=source_type,
                bead=(bead if bead else {}), det_in_um=det_in_um, det_out_um=det_out_um,
                x_wg_start_um=x_wg_start_um, end_facet_um=(end_facet_um or 0.0),
                x_src_um=x_src_um, polarization=pol, cells=cells,
                # the snapshot volumes stay in host memory (no temporary stream files);
                # no CAD geometry on the device path
                _stream_paths=[], cad=None)


# ----------------------------------------------------------------------------- stitch
def run_3d_stitched(label, wg_n, window_w_um=20.0, slide_um=12.0, t_wg_um=5.0, t_lip_um=0.0,
                    t_aq_um=3.5, t_mu_um=0.5, n_lip=1.480, n_aq=1.336, n_mu=1.342, n_co=1.376,
                    lam_nm=850.0, lx_um=40.0, air_um=3.0, tear_um=None, lz_um=8.0, dx_nm=50.0,
                    vcsel_waist_um=3.0, vcsel_waist_z_um=3.0, source_type='cw', n_snapshots=8,
                    bead=None, n_sponge=None, sponge_um=1.2, sponge_alpha=0.3, det_in_um=None,
                    det_out_um=None, vol_dtype='float32', wg_width_um=None, n_clad_side=N_AIR,
                    vcsel_tilt_deg=0.0, vcsel_offset_y_um=0.0, vcsel_offset_z_um=0.0,
                    input_gap_um=0.0, polarization='s', boundary='cpml', cad=None,
                    return_efull=False, stack=None, npml=None, engine_mode=None,
                    tune_cache=None):
    """3D x-domain decomposition (reference fdtd3d_core.py:1311-1782): windows run to CW
    steady state one after another; the previous window's complex E (all three
    components) is imposed with a cosine taper in the overlap; inner seams end in a
    non-assembled CPML extension; the assembled Efull/Iavg skip overlap and extension.
    ``tune_cache`` (a dict) carries the blocking chosen for each window shape across calls
    (the pulse comb's lines share their windows)."""
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    t_total = time.time()
    dx = dx_nm * 1e-9
    dx_um = dx_nm / 1000.0
    dt = 0.5 * dx / (C0 * np.sqrt(3))
    global LAM_LOSS
    LAM = lam_nm * 1e-9
    LAM_LOSS = LAM
    f0 = C0 / LAM
    omega = 2 * np.pi * f0
    sigma_t = 4 / (2 * np.pi * f0)
    npml = NPML if npml is None else int(npml)
    bkind = str(boundary or 'cpml').lower()
    if bkind not in ('cpml', 'sponge', 'mur1'):
        raise ValueError(f"unknown boundary '{bkind}' (sponge|cpml|mur1)")
    tear_um = _resolve_tear_um(tear_um, t_lip_um, t_aq_um, t_mu_um,
                               _abs_cells(bkind, npml, n_sponge, sponge_um, dx_um), dx_um,
                               stack=stack)
    ly_um = t_wg_um + air_um + tear_um
    y0_um = -tear_um
    z0_um = -lz_um / 2.0
    Ny, Nz = int(round(ly_um / dx_um)), int(round(lz_um / dx_um))
    Nx_full = int(round(lx_um / dx_um))
    Nx_win = min(int(round(window_w_um / dx_um)), Nx_full)
    S_cells = max(1, min(int(round(slide_um / dx_um)), Nx_win))
    O_cells = Nx_win - S_cells
    n_sp = _sponge_cells(n_sponge, sponge_um, dx_um, (Nx_win, Ny, Nz), warn=bkind == 'sponge')
    _vd = float(input_gap_um or 0.0)
    x_src_um, x_wg_start_um, _air, _zsrc = _vcsel_geom_3d(n_sp, dx_um, _vd)
    EXT = max(n_sp, npml) + 4
    x_wg_end_um = lx_um + EXT * dx_um
    n_win = (int(np.ceil((Nx_full - Nx_win) / S_cells)) + 1) if Nx_full > Nx_win else 1

    def gx0(w):
        return min(w * S_cells, max(Nx_full - Nx_win, 0))
    transit = window_w_um * 1e-6 * N_G_FACTOR * wg_n / C0
    steps_win = int(3.0 * transit / dt) + 400
    steps_per_T = max(4, int(round((1.0 / f0) / dt)))
    n_acc = steps_per_T * 8
    print(f'\n===== 3D-FDTD [stitch]: {label} =====  {n_win} windows of {Nx_win}x{Ny}x{Nz}, '
          f'overlap {O_cells}, steps/window {steps_win}, engine {backend.engine_name()}')
    Ch = np.float32(dt / (MU0 * dx))
    pol, Nj, Nk = _resolve_pol(polarization, Ny, Nz)
    src_comp = 'Ey' if pol == 'p' else 'Ez'
    src_ix, j_src, k_src, vs = _build_gauss_source(
        n_sp, t_wg_um, y0_um, z0_um, dx_um, dx, LAM, Nj, Nk, vcsel_waist_um, vcsel_waist_z_um,
        vcsel_offset_y_um, vcsel_offset_z_um, vcsel_tilt_deg,
        n_src_fn=_n_source_fn_3d(x_src_um, y0_um, dx_um, wg_n, t_wg_um,
                                 n_aq if stack is None else _stack_n_below(stack, n_co),
                                 x_wg_start_um, pol),
        pol=pol, wg_width_um=wg_width_um, dt=dt, z_prop_um=_zsrc)
    wtap = (0.5 * (1.0 + np.cos(np.pi * np.arange(O_cells) / max(O_cells - 1, 1)))).astype(np.float32)
    if det_in_um is None:
        det_in_um = 0.2 * lx_um
    if det_out_um is None:
        det_out_um = 0.8 * lx_um
    j_wg0 = int(round((0.0 - y0_um) / dx_um))
    j_wg1 = int(round((t_wg_um - y0_um) / dx_um))
    vol_dt = np.float16 if vol_dtype == 'float16' else np.float32
    Efull = np.zeros((Nx_full, Nj, Nk), dtype=np.complex64)
    Iavg = np.zeros((Nx_full, Nj, Nk), dtype=np.float32)
    m_I = max(1, int(round((1.0 / f0) / dt)) // 8)
    Ehand = None
    gi = guard_in_plane(min(max(int(round(det_in_um / dx_um)), 1), Nx_full - 2),
                        max(src_ix, int(round(x_wg_start_um / dx_um))), Nx_full)
    det_in_um = gi * dx_um
    go = min(max(int(round(det_out_um / dx_um)), 1), Nx_full - 2)
    fluxes = FluxAccumulator()
    mode = _engine_mode(engine_mode)
    tuned = {} if tune_cache is None else tune_cache
    for w in range(n_win):
        g0 = gx0(w)
        x0_um = g0 * dx_um
        last = w == n_win - 1
        Nxw = Nx_win + (0 if last else EXT)
        glo = 0 if w == 0 else g0 + O_cells
        ghi = Nx_full if last else min(g0 + Nx_win, Nx_full)
        fpl = {k: g - g0 for k, g in (('in', gi), ('out', go)) if glo <= g < ghi and g - g0 >= 1}
        if w == 0:
            fpl.update(src_up=vs.plane_up, src_dn=vs.plane_dn)
        off = (gx0(w + 1) - g0) if not last else 0
        tp = {}
        eps = build_eps3d(Nxw, Ny, Nz, dx_um, y0_um, z0_um, wg_n, t_wg_um, t_lip_um, t_aq_um,
                          t_mu_um, n_lip, n_aq, n_mu, n_co, bead, wg_width_um=wg_width_um,
                          n_clad_side=n_clad_side, x0_um=x0_um, x_wg_start_um=x_wg_start_um,
                          x_wg_end_um=x_wg_end_um, tensor_patch=tp, stack=stack)
        eps, eim = split_eps(eps)
        msp = mur_speeds_3d(eps, dx, LAM) if bkind == 'mur1' else None
        inv, da = inv_loss(eps, eim, dt, dx, LAM)
        del eps, eim
        cex, cey, cez = ce_arrays(inv, tp, dt, dx, dx_um, da=da)
        da = stagger_da(da)
        del inv
        # outer faces (the user's absorber): xmin only in window 0 (the others are driven),
        # xmax only in the last window (inner seams end in their own CPML)
        aface = tuple(f for f in ALL_FACES if not ((w > 0 and f == 'xmin') or (not last and f == 'xmax')))
        if bkind == 'cpml':                  # one CPML for the outer faces and the seam
            pml = make_boundary_3d('cpml', dx, dt, aface + (() if last else ('xmax',)),
                                   shape=(Nxw, Ny, Nz), npml=npml, lam_m=LAM)
        else:                                # seam CPML on xmax only; sponge / Mur outside
            pml = make_boundary_3d('cpml', dx, dt, () if last else ('xmax',),
                                   shape=(Nxw, Ny, Nz), npml=npml, lam_m=LAM)
        eng = backend.make_engine('yee3d', (Nxw, Ny, Nz), mode)
        eng.set_coeffs(Ch, cex=cex, cey=cey, cez=cez,
                       dax=None if da is None else da[0], day=None if da is None else da[1],
                       daz=None if da is None else da[2])
        del cex, cey, cez
        pml.attach(eng)
        if bkind == 'mur1':                  # fdtd3d_core.py:1611-1612, 1633-1634
            make_boundary_3d('mur1', dx, dt, aface, speeds=msp).attach_faces(eng, (Nxw, Ny, Nz),
                                                                             pec_after=False)
        elif bkind == 'sponge':              # fdtd3d_core.py:1635-1637
            Sponge3D(n_sp, sponge_alpha, aface).attach_faces(eng)
        eng.finalize()
        if w == 0:
            eng.set_source(src_comp, src_ix, vs.amp, vs.ph1, vs.ph2)
        elif Ehand is not None:
            eng.set_drive(O_cells, wtap, dict(zip(('Ex', 'Ey', 'Ez'), Ehand)), math='c64')
        wide = eng.wide_acc
        mon_acc = eng.add_dft(src_comp, 0, Nx_win, acc='c128' if wide else 'split32', phasor='q')
        mon_hand = ([eng.add_dft(c, off, off + O_cells, acc='c128' if wide else 'c64', phasor='p')
                     for c in ('Ex', 'Ey', 'Ez')] if not last else [])
        mon_I = eng.add_sq('3d_p' if pol == 'p' else '3d_s', Nx_win)
        book = PlaneFluxBook(eng, 'yee3d')
        for k, il in fpl.items():
            book.add_plane(k, il, carry=fluxes)
        acc_bits = (1 << mon_acc) | sum(1 << m for m in mon_hand) | sum(1 << book.mon[k] for k in fpl)
        tab = new_steps(steps_win)
        ns = np.arange(steps_win)
        t_phys = ns * dt
        fill_phasors(tab, omega, dt, t_phys, cw_envelope(t_phys, sigma_t), src_omega=vs.omega)
        win = ns >= steps_win - n_acc
        mk = np.zeros(steps_win, dtype=np.uint32)
        mk[win] |= np.uint32(acc_bits)
        selI = win & (((ns - (steps_win - n_acc)) % m_I) == 0)
        mk[selI] |= np.uint32(1 << mon_I)
        tab['mask'] = mk
        acc_n = int(win.sum())
        n_I = int(selI.sum())
        book.note_samples(list(fpl), acc_n)
        # blocking tuned once per window shape and set-up (source window, driven windows)
        backend.tune(eng, tab[steps_win // 2], steps_win, cache=tuned,
                     key=(Nxw, Ny, Nz, w > 0 and Ehand is not None))
        chunk = max(1, steps_win // 10)
        for c0 in range(0, steps_win, chunk):
            eng.run(tab[c0:c0 + chunk])
        print(f'  [Window {w + 1}/{n_win}] {steps_win} steps done', flush=True)
        sc = np.float32(2.0 / max(acc_n, 1))
        A = eng.read_monitor(mon_acc)
        Ew = np.empty((Nx_win, Nj, Nk), dtype=np.complex64)
        Ew.real = (A.real.astype(np.float32) * sc)
        Ew.imag = (A.imag.astype(np.float32) * sc)
        if mon_hand:
            Ehand = [(eng.read_monitor(m).astype(np.complex64)) * np.complex64(2.0 / max(acc_n, 1))
                     for m in mon_hand]
        acc_I = eng.read_monitor(mon_I)
        fluxes.take(book, list(fpl))
        eng.close()
        if w == 0:
            g_lo, l_lo = 0, 0
        else:
            l_lo = O_cells
            g_lo = g0 + l_lo
        g_hi = Nx_full if last else min(g0 + Nx_win, Nx_full)
        l_hi = l_lo + (g_hi - g_lo)
        if g_hi > g_lo:
            Efull[g_lo:g_hi] = Ew[l_lo:l_hi]
            Iavg[g_lo:g_hi] = acc_I[l_lo:l_hi].astype(np.float32) * np.float32(1.0 / max(n_I, 1))
    total_time = time.time() - t_total
    nfr = max(1, n_snapshots)
    snap_times = [(k / nfr) * (1.0 / f0) for k in range(nfr)]
    snap_steps = [k + 1 for k in range(nfr)]
    cosk = [float(np.cos(omega * tk)) for tk in snap_times]
    sink = [float(np.sin(omega * tk)) for tk in snap_times]
    vol_buf = np.empty((nfr, Nx_full, Nj, Nk), dtype=vol_dt)
    for k in range(nfr):
        vol_buf[k] = (Efull.real * cosk[k] - Efull.imag * sink[k]).astype(vol_dt)
    iin = min(max(int(round(det_in_um / dx_um)), 0), Nx_full - 1)
    iout = min(max(int(round(det_out_um / dx_um)), 0), Nx_full - 1)
    P_in = 2.0 * float(np.sum(np.asarray(Iavg[iin, j_wg0:j_wg1, n_sp:-n_sp]))) * dx ** 3
    P_out = 2.0 * float(np.sum(np.asarray(Iavg[iout, j_wg0:j_wg1, n_sp:-n_sp]))) * dx ** 3
    Tr_I = P_out / max(P_in, 1e-30)
    Tr, tr_def, p_si = Tr_I, 'intensity', None
    if fluxes.has('in') and fluxes.has('out'):
        S = [fluxes.flux('in'), fluxes.flux('out')]
        p_si = S[0] * dx * dx
        if abs(S[0]) > 0:
            Tr, tr_def = S[1] / S[0], 'poynting'
    print(f'[Result] {label} [3D stitch]: total {total_time:.1f}s  Transmission={Tr:.4f} ({tr_def})')
    rs = _result_stack(stack, t_lip_um, t_aq_um, t_mu_um, n_lip, n_aq, n_mu)
    cw_ok = source_type != 'pulse' and fluxes.has('src_up') and fluxes.has('src_dn')
    vres = _vcsel_res(((fluxes.flux('src_dn') - fluxes.flux('src_up')) * dx * dx) if cw_ok else None,
                      (-fluxes.flux('src_up') * dx * dx) if cw_ok else None, p_si, _vd, _air,
                      x_wg_start_um, vcsel_waist_um, vcsel_waist_z_um)
    return dict(**vres, scenario=label, dim='3D', P_in=P_in, P_out=P_out, p_in_si=p_si,
                p_in_unit='W', transmission=Tr, vols_array=vol_buf, vols_ex=None, vols_ey=None,
                vector=False, snap_steps=snap_steps, snap_times=snap_times, Iavg=Iavg,
                total_time=total_time, steps_total=steps_win * n_win, dt=dt, dx_nm=dx_nm,
                lx_um=lx_um, ly_um=ly_um, lz_um=lz_um, y0_um=y0_um, z0_um=z0_um,
                t_wg_um=t_wg_um, wg_n=wg_n, layers=rs[2], stack=rs[0], stack_custom=rs[1],
                n_lip=n_lip, n_aq=n_aq, n_mu=n_mu, n_co=n_co, lam_nm=lam_nm,
                source_type=source_type, bead=(bead if bead else {}), det_in_um=det_in_um,
                det_out_um=det_out_um, x_wg_start_um=x_wg_start_um, end_facet_um=0.0,
                x_src_um=x_src_um, polarization=pol, view_primary='cw', cw_only=True,
                Efull=(Efull if return_efull else None), transmission_intensity=Tr_I,
                transmission_def=tr_def, n_windows=n_win, cells_per_window=Nx_win * Ny * Nz,
                _stream_paths=[], cad=None)


def run_3d_stitched_td(label, wg_n, window_w_um=20.0, slide_um=12.0, t_wg_um=5.0, t_lip_um=0.0,
                       t_aq_um=3.5, t_mu_um=0.5, n_lip=1.480, n_aq=1.336, n_mu=1.342,
                       n_co=1.376, lam_nm=850.0, lx_um=40.0, air_um=3.0, tear_um=None,
                       lz_um=8.0, dx_nm=50.0, vcsel_waist_um=3.0, vcsel_waist_z_um=3.0,
                       source_type='pulse', n_snapshots=8, bead=None, n_sponge=None,
                       sponge_um=1.2, sponge_alpha=0.3, det_in_um=None, det_out_um=None,
                       vol_dtype='float32', wg_width_um=None, n_clad_side=N_AIR,
                       vcsel_tilt_deg=0.0, vcsel_offset_y_um=0.0, vcsel_offset_z_um=0.0,
                       input_gap_um=0.0, polarization='s', boundary='cpml', cad=None,
                       pulse_cycles=30.0, pulse_tmax_fs=None, t_eval_s=None, stack=None,
                       npml=None, engine_mode=None):
    """Transient 3D stitch in the time domain with Huygens (TF/SF) planes (reference
    fdtd3d_core.py:1785-2084), the 3D form of run_beads_stitched_td.

    Window w records Ey, Ez(n) at the start of each step and Hy, Hz(n+1/2) after the H
    update at the column of window w+1's TF/SF plane i0 = n_sp + 2 (device plane
    recorders into memory-mapped series); window w+1 replays them: Hy[i0-1] -= Ch Ez_inc,
    Hz[i0-1] += Ch Ey_inc after the H update, Ey[i0] += ce_y Hz_inc, Ez[i0] -= ce_z Hy_inc
    after the E update.  CPML on all faces (boundary 'cpml', when the TF/SF plane lies
    behind it), else the graded sponge on all faces.  Output: transient volume frames of
    the driven component, <|E|^2> over the whole time axis, time-integrated core
    intensity and Poynting flux at the detector planes, pulse energy around the source.
    No dispersion."""
    from lens_sensor_amd.common.td_stitch import TD_CHECK, TD_ROWS, td_next_event, signal_span
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    t_total = time.time()
    dx = dx_nm * 1e-9
    dx_um = dx_nm / 1000.0
    dt = 0.5 * dx / (C0 * np.sqrt(3))
    global LAM_LOSS
    LAM = lam_nm * 1e-9
    LAM_LOSS = LAM
    f0 = C0 / LAM
    omega = 2 * np.pi * f0
    npml = NPML if npml is None else int(npml)
    bkind = str(boundary or 'cpml').lower()
    tear_um = _resolve_tear_um(tear_um, t_lip_um, t_aq_um, t_mu_um,
                               _abs_cells(bkind, npml, n_sponge, sponge_um, dx_um), dx_um,
                               stack=stack)
    ly_um = t_wg_um + air_um + tear_um
    y0_um = -tear_um
    z0_um = -lz_um / 2.0
    Ny, Nz = int(round(ly_um / dx_um)), int(round(lz_um / dx_um))
    Nx_full = int(round(lx_um / dx_um))
    Nx_win = min(int(round(window_w_um / dx_um)), Nx_full)
    S_cells = max(1, min(int(round(slide_um / dx_um)), Nx_win))
    O_cells = Nx_win - S_cells
    n_sp = _sponge_cells(n_sponge, sponge_um, dx_um, (Nx_win, Ny, Nz), warn=bkind == 'sponge')
    i0 = n_sp + 2                                # TF/SF plane of the driven windows
    if O_cells < i0 + 4:
        raise ValueError(f'Overlap {O_cells} cells < {i0 + 4}: choose window_w >> slide.')
    _vd = float(input_gap_um or 0.0)
    x_src_um, x_wg_start_um, _air, _zsrc = _vcsel_geom_3d(n_sp, dx_um, _vd)
    EXT = n_sp + 4
    x_wg_end_um = lx_um + EXT * dx_um
    n_win = (int(np.ceil((Nx_full - Nx_win) / S_cells)) + 1) if Nx_full > Nx_win else 1

    def gx0(w):
        return min(w * S_cells, max(Nx_full - Nx_win, 0))
    sp = pulse_cycles / f0
    tc = 3 * sp
    t_max = (pulse_tmax_fs * 1e-15) if pulse_tmax_fs else (tc + 3 * sp + lx_um * 1e-6 * wg_n * 1.15 / C0)
    N = int(np.ceil(t_max / dt)) + 1
    t_j = (np.asarray(t_eval_s, dtype=np.float64) if t_eval_s is not None
           else np.linspace(0.0, t_max, max(int(n_snapshots), 8)))
    n_j = np.clip(np.round(t_j / dt).astype(np.int64), 0, N - 1)
    frame_steps = np.unique(n_j)
    pol, Nj, Nk = _resolve_pol(polarization, Ny, Nz)
    src_comp = 'Ey' if pol == 'p' else 'Ez'
    print(f'\n===== 3D time-domain stitch (Huygens/TF-SF): {label} =====')
    print(f'Domain {lx_um:.1f}x{ly_um:.1f}x{lz_um:.1f}um, {n_win} windows of {window_w_um}um '
          f'(overlap {O_cells * dx_um:.1f}um), grid/window {Nx_win}x{Ny}x{Nz}, Sponge {n_sp}, '
          f'{N} steps, {len(n_j)} frames, engine {backend.engine_name()}')
    src_ix, j_src, k_src, vs = _build_gauss_source(
        n_sp, t_wg_um, y0_um, z0_um, dx_um, dx, LAM, Nj, Nk, vcsel_waist_um, vcsel_waist_z_um,
        vcsel_offset_y_um, vcsel_offset_z_um, vcsel_tilt_deg,
        n_src_fn=_n_source_fn_3d(x_src_um, y0_um, dx_um, wg_n, t_wg_um,
                                 n_aq if stack is None else _stack_n_below(stack, n_co),
                                 x_wg_start_um, pol),
        pol=pol, wg_width_um=wg_width_um, dt=dt, z_prop_um=_zsrc)
    env_t = np.exp(-((np.arange(N) * dt - tc) / sp) ** 2)
    if det_in_um is None:
        det_in_um = 0.2 * lx_um
    if det_out_um is None:
        det_out_um = 0.8 * lx_um
    iin = int(round(det_in_um / dx_um))
    iout = int(round(det_out_um / dx_um))
    j_wg0 = int(round((0.0 - y0_um) / dx_um))
    j_wg1 = int(round((t_wg_um - y0_um) / dx_um))
    iin = guard_in_plane(iin, max(src_ix, int(round(x_wg_start_um / dx_um))), Nx_full)
    vol_dt = np.float16 if vol_dtype == 'float16' else np.float32
    os.makedirs('results', exist_ok=True)
    safe = ''.join(c if (c.isalnum() or c in '-_.') else '_' for c in str(label))
    paths = []
    p = os.path.join('results', f'_stream3d_{safe}_td_ez.npy')
    paths.append(p)
    vol_buf = np.lib.format.open_memmap(p, mode='w+', dtype=vol_dt, shape=(len(n_j), Nx_full, Nj, Nk))
    vol_buf[:] = 0
    p = os.path.join('results', f'_stream3d_{safe}_td_Iavg.npy')
    paths.append(p)
    Iavg = np.lib.format.open_memmap(p, mode='w+', dtype=np.float32, shape=(Nx_full, Nj, Nk))
    Iavg[:] = 0
    m_I = max(1, int(round((1.0 / f0) / dt)) // 8)
    n_I_glob = max(1, (N + m_I - 1) // m_I)        # mean over the whole time axis
    Ch = np.float32(dt / (MU0 * dx))
    plane = {'Ey': (Ny - 1, Nz), 'Ez': (Ny, Nz - 1), 'Hy': (Ny, Nz - 1), 'Hz': (Ny - 1, Nz)}
    mode = _engine_mode(engine_mode)
    rec_in = None
    P = dict(P_in=0.0, P_out=0.0, S_in=0.0, S_out=0.0, S_up=0.0, S_dn=0.0)
    for w in range(n_win):
        t1 = time.time()
        g0 = gx0(w)
        last = w == n_win - 1
        Nxw = Nx_win + (0 if last else EXT)
        x0_um = g0 * dx_um
        tp = {}
        eps = build_eps3d(Nxw, Ny, Nz, dx_um, y0_um, z0_um, wg_n, t_wg_um, t_lip_um, t_aq_um,
                          t_mu_um, n_lip, n_aq, n_mu, n_co, bead, wg_width_um=wg_width_um,
                          n_clad_side=n_clad_side, x0_um=x0_um, x_wg_start_um=x_wg_start_um,
                          x_wg_end_um=x_wg_end_um, tensor_patch=tp, stack=stack)
        eps, eim = split_eps(eps)
        inv, da = inv_loss(eps, eim, dt, dx, LAM)
        del eps, eim
        cex, cey, cez = ce_arrays(inv, tp, dt, dx, dx_um, da=da)
        da = stagger_da(da)
        del inv
        eng = backend.make_engine('yee3d', (Nxw, Ny, Nz), mode)
        eng.set_coeffs(Ch, cex=cex, cey=cey, cez=cez,
                       dax=None if da is None else da[0], day=None if da is None else da[1],
                       daz=None if da is None else da[2])
        del cex, cey, cez
        # CPML on all faces (the TF/SF plane behind it), else the sponge on all faces
        # (fdtd3d_core.py:1958-1965)
        use_pml = bkind == 'cpml' and i0 >= 11
        if bkind == 'cpml' and not use_pml:
            print('  [WARN] TF/SF plane too close to the boundary for CPML -> sponge')
        if use_pml:
            make_boundary_3d('cpml', dx, dt, ALL_FACES, shape=(Nxw, Ny, Nz), npml=npml,
                             lam_m=LAM).attach(eng)
        else:
            make_boundary_3d('cpml', dx, dt, (), shape=(Nxw, Ny, Nz), npml=npml,
                             lam_m=LAM).attach(eng)
            Sponge3D(n_sp, sponge_alpha, ALL_FACES).attach_faces(eng)
        eng.finalize()
        i_rec = (gx0(w + 1) - g0 + i0) if not last else None
        if w == 0:
            eng.set_source(src_comp, src_ix, vs.amp, vs.ph1, vs.ph2)
            n_start, n_in_last = 0, int(np.nonzero(env_t > 1e-7)[0][-1])
        else:
            # per step max|Ez_inc| + max|Ey_inc|, summed in float32 (fdtd3d_core.py:1951-1952)
            amp = np.empty(N, dtype=np.float64)
            for a in range(0, N, 1024):
                b = min(a + 1024, N)
                amp[a:b] = (np.abs(rec_in['Ez'][a:b]).max(axis=(1, 2))
                            + np.abs(rec_in['Ey'][a:b]).max(axis=(1, 2)))
            n_start, n_in_last = signal_span(amp, N)
            tfsf = {'Ez': eng.add_tfsf('Hy', i0 - 1, -1, None, [(0, Ny), (0, Nz - 1)], TD_ROWS),
                    'Ey': eng.add_tfsf('Hz', i0 - 1, +1, None, [(0, Ny - 1), (0, Nz)], TD_ROWS),
                    'Hz': eng.add_tfsf('Ey', i0, +1, 'cey', [(0, Ny - 1), (1, Nz - 1)], TD_ROWS),
                    'Hy': eng.add_tfsf('Ez', i0, -1, 'cez', [(1, Ny - 1), (0, Nz - 1)], TD_ROWS)}
        rec, mrec = None, {}
        if i_rec is not None:
            rec = {}
            for c, shp in plane.items():
                pth = os.path.join('results', f'_stream3d_{safe}_td_rec{w}_{c}.npy')
                rec[c] = np.lib.format.open_memmap(pth, mode='w+', dtype=np.float32, shape=(N,) + shp)
                rec[c][:] = 0
                paths.append(pth)
                mrec[c] = eng.add_rec(c, i_rec if c[0] == 'E' else i_rec - 1,
                                      0 if c[0] == 'E' else 1, TD_ROWS)
        a_g = 0 if w == 0 else O_cells
        b_g = (Nx_full - g0) if last else Nx_win
        mon_I = eng.add_sq('3d_p' if pol == 'p' else '3d_s', Nx_win)
        sums = {}
        if w == 0:
            sums['S_up'] = eng.add_sx(vs.plane_up)
            sums['S_dn'] = eng.add_sx(vs.plane_dn)
        for ixg, which in ((iin, 'in'), (iout, 'out')):
            il = ixg - g0
            if a_g <= il < b_g:
                sums['P_' + which] = eng.add_box(((il, il + 1), (j_wg0, j_wg1), (n_sp, -n_sp)))
                if il >= 1:
                    sums['S_' + which] = eng.add_sx(il)
        bits = sum(1 << m for m in list(sums.values()) + list(mrec.values()))
        peak, n_stop, n = 0.0, N, n_start
        while n < N:
            e = min(td_next_event(n, n_in_last, frame_steps), N - 1, n + TD_ROWS - 1)
            k = e - n + 1
            ns = np.arange(n, e + 1)
            tab = new_steps(k)
            if w == 0:
                env = env_t[n:e + 1]
                fill_phasors(tab, omega, dt, ns * dt, np.where(env > 1e-9, env, 0.0),
                             src_omega=vs.omega)
            mk = np.full(k, bits, dtype=np.uint32)
            mk[ns % m_I == 0] |= np.uint32(1 << mon_I)
            tab['mask'] = mk
            tab['row'] = np.arange(k)
            if w > 0:
                for c, tid in tfsf.items():
                    eng.load_tfsf(tid, rec_in[c][n:e + 1])
            eng.run(tab)
            for c, mid in mrec.items():
                rec[c][n:e + 1] = eng.read_monitor(mid)[:k]
            hit = np.nonzero(n_j == e)[0]
            if len(hit):
                blk = eng.read_field(src_comp)[a_g:b_g].astype(vol_dt)
                for jj in hit:
                    vol_buf[jj, g0 + a_g:g0 + b_g] = blk
            if e % TD_CHECK == 0 or (e > n_in_last and (e - n_in_last) % TD_CHECK == 0):
                cur = eng.field_maxabs(src_comp)
                peak = max(peak, cur)
                if e > n_in_last and cur < 1e-4 * max(peak, 1e-30):
                    n_stop = e
                    break
            n = e + 1
        for name, mid in sums.items():
            P[name] += eng.read_monitor(mid)
        acc_I = eng.read_monitor(mon_I)
        Iavg[g0 + a_g:g0 + b_g] = acc_I[a_g:b_g] * np.float32(1.0 / n_I_glob)
        eng.close()
        print(f'[Window {w + 1}/{n_win}] x0={x0_um:.1f}um  steps {n_start}..{n_stop} '
              f'({n_stop - n_start} of {N})  {time.time() - t1:.0f}s', flush=True)
        if rec_in is not None:                   # the previous window's series are used up
            for c in list(rec_in):
                pth = rec_in[c].filename
                del rec_in[c]
                try:
                    os.remove(pth)
                    paths.remove(pth)
                except (OSError, ValueError):
                    pass
        rec_in = rec
    vol_buf.flush()
    Iavg.flush()
    P_in, P_out = P['P_in'] * dx ** 3, P['P_out'] * dx ** 3
    Tr_I = P_out / max(P_in, 1e-30)
    S_in, S_out = P['S_in'], P['S_out']
    Tr = S_out / S_in if abs(S_in) > 0 else Tr_I
    tr_def = 'poynting' if abs(S_in) > 0 else 'intensity'
    total_time = time.time() - t_total
    print(f'\n[Result] {label} [3D time-domain stitch]: total {total_time:.0f}s  '
          f'Transmission={Tr:.4f} ({tr_def})  [intensity: {Tr_I:.4f}]  ({n_win} windows)')
    rs = _result_stack(stack, t_lip_um, t_aq_um, t_mu_um, n_lip, n_aq, n_mu)
    ea = dx * dx * dt                              # area x time -> energy [J]
    vres = _vcsel_res((P['S_dn'] - P['S_up']) * ea, -P['S_up'] * ea, S_in * ea, _vd, _air,
                      x_wg_start_um, vcsel_waist_um, vcsel_waist_z_um, energy=True)
    return dict(**vres, scenario=label, dim='3D', P_in=P_in, P_out=P_out, transmission=Tr,
                vols_array=vol_buf, vols_ex=None, vols_ey=None, vector=False,
                _stream_paths=paths, snap_steps=[int(k) for k in n_j],
                snap_times=[float(k * dt) for k in n_j], Iavg=Iavg, total_time=total_time,
                steps_total=N, dt=dt, dx_nm=dx_nm, lx_um=lx_um, ly_um=ly_um, lz_um=lz_um,
                y0_um=y0_um, z0_um=z0_um, t_wg_um=t_wg_um, wg_n=wg_n, layers=rs[2],
                stack=rs[0], stack_custom=rs[1], n_lip=n_lip, n_aq=n_aq, n_mu=n_mu, n_co=n_co,
                lam_nm=lam_nm, source_type='pulse', bead=(bead if bead else {}),
                det_in_um=iin * dx_um, det_out_um=iout * dx_um, x_wg_start_um=x_wg_start_um,
                end_facet_um=0.0, x_src_um=x_src_um, polarization=pol,
                view_primary='transient', cw_only=False, cad=None, cad_region=None,
                transmission_intensity=Tr_I, transmission_def=tr_def)


def _index_at(spec, lam_nm):
    """One entry of a dispersion map at lam_nm: a material name, 'cornea_layered' (the
    default cornea) or (material, offset)."""
    from lens_sensor_amd.common.physics import n_at, cornea_default
    if spec == 'cornea_layered':
        return cornea_default(lam_nm)
    if isinstance(spec, (tuple, list)):
        return n_at(spec[0], lam_nm) + float(spec[1])
    return n_at(spec, lam_nm)


def run_3d_stitched_pulse(label, wg_n, window_w_um=20.0, slide_um=12.0, pulse_cycles=30.0,
                          pulse_tmax_fs=None, pulse_nfreq=None, pulse_dispersion=True,
                          disp_map=None, t_eval_s=None, cad=None, stack=None, **kw):
    """Transient 3D stitch by Fourier synthesis (reference fdtd3d_core.py:2087-2215).

    One CW run_3d_stitched per line of the pulse comb
    (lens_sensor_amd.common.fourier_pulse.PulseComb) on the driven component; the frames
    E(t_j) = (1/pi) Re sum_k i E_k S_k e^{i w_k t_j} dw are summed frequency by frequency
    (complex64, x blocks of 64 planes) into a memory-mapped volume array.  ``disp_map``
    {'wg_n' | 'bead' | run_3d_stitched keyword: material, 'cornea_layered' or
    (material, offset)} sets those indices from n(lambda_k) when ``pulse_dispersion`` is
    on; material layers of ``stack`` follow n(lambda_k) as well.  The result also holds
    the CW volume at the centre line (8 phases), the time-averaged |E|^2 of the frames
    and the spectrum T(lambda_k); ``transmission`` is the |S|^2-weighted mean."""
    from lens_sensor_amd.common.fourier_pulse import PulseComb, add_frequency
    if cad is not None:
        raise NotImplementedError('CAD geometry import is outside the device path')
    t_total = time.time()
    lam_nm = kw.pop('lam_nm', 850.0)
    lx = kw.get('lx_um', 40.0)
    comb = PulseComb(lam_nm, pulse_cycles, lx * 1e-6 * wg_n * 1.15 / C0,
                     tmax_fs=pulse_tmax_fs, nfreq=pulse_nfreq)
    if comb.aliased:
        print(f'  [WARN] {pulse_nfreq} frequencies -> repetition period '
              f'{2 * np.pi / comb.dw * 1e15:.0f} fs < t_max {comb.t_max * 1e15:.0f} fs: '
              'temporal aliasing.')
    n_snap = int(kw.pop('n_snapshots', 8))
    t_j = (np.asarray(t_eval_s, dtype=np.float64) if t_eval_s is not None
           else np.linspace(0.0, comb.t_max, max(n_snap, 8)))
    print(f'\n===== 3D-Fourier-Stitch (transient): {label} =====')
    print(f'Pulse {pulse_cycles:g} periods (s={comb.sp * 1e15:.1f} fs), band '
          f'{comb.lam_nm.min():.1f}..{comb.lam_nm.max():.1f} nm, {len(comb)} frequencies, '
          f't_max={comb.t_max * 1e15:.0f} fs, {len(t_j)} frames, '
          f'dispersion {"on" if pulse_dispersion else "off"}')
    if stack is not None:
        print(f'Layers: {describe_stack(stack)}'
              + ('  (material layers follow n(lambda_k))' if pulse_dispersion else ''))
    disp = dict(disp_map or {}) if pulse_dispersion else {}
    tuned = {}                                   # the lines share their window shapes
    fr = fr_path = None
    spec_T = []
    r_c = E_c = None
    amp = np.abs(comb.S).max()
    os.makedirs('results', exist_ok=True)
    safe = ''.join(c if (c.isalnum() or c in '-_.') else '_' for c in str(label))
    for k, (wk, lk, Sk) in enumerate(zip(comb.w, comb.lam_nm, comb.S)):
        t1 = time.time()
        args, wg_k = dict(kw), wg_n
        for par, mat in disp.items():
            if par == 'wg_n':
                wg_k = _index_at(mat, lk)
            elif par == 'bead':
                if args.get('bead'):
                    args['bead'] = dict(args['bead'], n=_index_at(mat, lk))
            else:
                args[par] = _index_at(mat, lk)
        stk = stack_at_lambda(stack, float(lk)) if (stack is not None and pulse_dispersion) else stack
        with contextlib.redirect_stdout(io.StringIO()):
            r = run_3d_stitched(f'{label}_f{k}', wg_k, window_w_um=window_w_um,
                                slide_um=slide_um, lam_nm=float(lk), n_snapshots=1,
                                return_efull=True, stack=stk, tune_cache=tuned, **args)
        E = r.pop('Efull')
        if fr is None:
            fr_path = os.path.join('results', f'_stream3d_{safe}_fourier_ez.npy')
            fr = np.lib.format.open_memmap(fr_path, mode='w+', dtype=np.float32,
                                           shape=(len(t_j),) + E.shape)
            fr[:] = 0.0
        add_frequency(fr, E, Sk * comb.dw / np.pi, wk, t_j, wide=False)
        spec_T.append(float(r['transmission']))
        r['vols_array'] = r['Iavg'] = None
        if k == comb.centre:
            r_c, E_c = r, E
        del E
        print(f'  [{k + 1}/{len(comb)}] lambda={lk:.2f} nm  |S|={abs(Sk) / amp:.3f}  '
              f'T={r["transmission"]:.4f}  ({time.time() - t1:.0f}s)', flush=True)
    fr.flush()
    # the steady state at the centre line (8 phases) and the frames' mean |E|^2
    nfr_cw = 8
    cw_path = os.path.join('results', f'_stream3d_{safe}_fourier_cwvol.npy')
    cw_vol = np.lib.format.open_memmap(cw_path, mode='w+', dtype=np.float32,
                                       shape=(nfr_cw,) + E_c.shape)
    for q in range(nfr_cw):
        cw_vol[q] = np.real(E_c * np.complex64(np.exp(2j * np.pi * q / nfr_cw))).astype(np.float32)
    cw_vol.flush()
    ia_path = os.path.join('results', f'_stream3d_{safe}_fourier_Iavg.npy')
    Iavg = np.lib.format.open_memmap(ia_path, mode='w+', dtype=np.float32, shape=E_c.shape)
    for x0 in range(0, E_c.shape[0], 64):
        Iavg[x0:x0 + 64] = np.mean(np.asarray(fr[:, x0:x0 + 64], dtype=np.float32) ** 2, axis=0)
    Iavg.flush()
    res = dict(r_c)
    dt_c = r_c['dt']
    res.update(scenario=label, vols_array=fr, vols_ex=None, vols_ey=None, vector=False,
               _stream_paths=[fr_path, cw_path, ia_path],
               snap_steps=[int(round(t / dt_c)) for t in t_j], snap_times=[float(t) for t in t_j],
               Iavg=Iavg, cw_vol=cw_vol, view_primary='transient', cw_only=False,
               source_type='pulse', lam_nm=lam_nm, total_time=time.time() - t_total,
               transmission=comb.weighted(spec_T),
               spectrum=dict(lam_nm=comb.lam_nm, T=np.array(spec_T), weight=np.abs(comb.S) / amp),
               pulse=dict(cycles=pulse_cycles, tmax_fs=comb.t_max * 1e15, nfreq=len(comb),
                          dispersion=bool(pulse_dispersion)))
    if stack is not None:                     # the stack at lambda0
        res.update(stack=[dict(layer) for layer in stack], stack_custom=True,
                   layers=legacy_layers_tuple(stack))
    print(f'\n[Result] {label} [3D-Fourier-Stitch]: {len(comb)} frequencies in '
          f'{res["total_time"]:.0f}s  T(weighted)={res["transmission"]:.4f}')
    return res


def save_results_3d(result, out_dir='results', prefix='fdtd3d', material_names=None):
    """PKL summary, z mid-plane frames NPZ and volume NPZ of a 3D result (reference
    fdtd3d_core.py:2218-2423); keys and file names as the reference writes them
    (lens_sensor_amd.common.results)."""
    from lens_sensor_amd.common import results
    return results.save_results_3d(result, out_dir=out_dir, prefix=prefix,
                                   material_names=material_names)
